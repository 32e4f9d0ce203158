"""reference_records = 1: the reference's own norm_type=1 output (src/norm.cu:560-600) -- one
record per (angle, block, template, 256-element chunk of the correlation index) whose largest z
reaches the threshold, the pixel chosen by get_peak_pos's tree (src/kernels.cu:766-793) -- against
the FP64 oracle's restatement of that second pass.

Tolerance rule: a record may be missing on one side only when the oracle's chunk maximum lies
within tau = 1e-4 max(1, |z|) of the threshold, and its pixel may differ only when the oracle's
chunk maximum and the chunk's next value are closer than tau; scores agree within tau.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("l,chunk", [(320, 0), (320, 4), (256, 0)])
def test_reference_records_match_the_oracle(gs, orc, l, chunk):
    """l = 320 runs the tiled second pass (chunk parts merged on the host; chunk 4: two launches
    per angle, the second partial), l = 256 the generic one."""
    import synth

    cfg = gs.SearchConfig(angpix=1.0, phistep=20.0, kk=3.0, energy=300.0, cs=2.7, highres=8.0,
                          lowres=100.0, diameter=128.0, threshold=5.0, window_size=l, overlap=0,
                          norm_type=1, invert=0, phase_flip=0, nms_radius=64.0, reference_records=1)
    ds = synth.make_dataset(2300 + l, 1280, 960, n_templates=6, tsize=128, n_particles=6, snr=0.5,
                            defocus=1.8, dfdiff=0.003, dfang=15.0)
    o = orc.search_norm_records(cfg, ds.image, ds.templates, ds.ctf)
    with gs.Engine(cfg) as eng:
        if chunk:
            eng.set_chunk(chunk)
        eng.load_templates(ds.templates)
        g = eng.search(ds.image, ds.ctf)
    assert len(o) > 20
    chunks = l * l // 256
    key = lambda r: (int(r["angle_index"]), int(r["block"]), int(r["template_id"]), int(r["pixel"]) // 256)
    gk = [key(r) for r in g]
    assert gk == sorted(gk), "records in the reference's order: angle, block, template, chunk"
    ok = {(int(r["angle_index"]), int(r["block"]), int(r["template_id"]), int(r["chunk"])): r for r in o}
    gd = dict(zip(gk, g))
    for k in set(ok) ^ set(gd):
        r = ok.get(k)
        assert r is not None and abs(r["score"] - cfg.threshold) < 1e-4 * max(1.0, abs(r["score"])), k
    both = set(ok) & set(gd)
    assert len(both) >= 0.99 * len(ok)
    near_ties = 0
    for k in both:
        a, b = ok[k], gd[k]
        tau = 1e-4 * max(1.0, abs(float(a["score"])))
        assert abs(float(b["score"]) - float(a["score"])) < tau, (k, b["score"], a["score"])
        if int(b["pixel"]) != int(a["pixel"]):
            assert a["margin"] < tau, (k, int(b["pixel"]), int(a["pixel"]), a["margin"])
            near_ties += 1
        assert int(b["pixel"]) // 256 == k[3] and k[3] < chunks
        assert b["x"] == a["x"] or int(b["pixel"]) != int(a["pixel"])
    print(f"l={l} chunk={chunk}: {len(o)} reference records, {len(both)} matched, {near_ties} near-tie pixels")
