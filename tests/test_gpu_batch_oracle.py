"""The batched forms (C2: gs_search_batch; C4: gs_search_batch_refine) against the FP64 oracle
at the batch configuration's real frame sizes: 3710 x 3838 (whole-frame transforms of
7 * 53 x 2 * 19 * 101 points, 12 x 12 blocks with clamped last ones) and 4096 x 4096 (13 x 13),
with 256-px and 128-px template stacks of uneven counts, 4 templates per correlation launch so
that each set crosses launches, and graph replay on.  Template and angle counts are reduced
(phistep 30, 12 angles) so that the oracle finishes in seconds; mu and sigma run over the same
hypotheses on both sides.  Refined picks are compared with the oracle refiner from the same
starts (0.05 px, 0.1 deg, 5 A; z within 1e-4).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = np.array([0.05, 0.05, 0.1, 5.0])


def _cfg(gs):
    return gs.SearchConfig(angpix=1.0, phistep=30.0, kk=3.0, energy=300.0, cs=2.7, highres=8.0,
                           lowres=100.0, diameter=128.0, threshold=5.0, window_size=320, overlap=0,
                           norm_type=1, invert=0, phase_flip=0, nms_radius=64.0)


@pytest.fixture(scope="module")
def items():
    import synth

    a = synth.make_dataset(2401, 3710, 3838, n_templates=6, tsize=256, n_particles=40, snr=0.4,
                           defocus=2.6, dfdiff=0.01, dfang=60.0)
    b = synth.make_dataset(2402, 4096, 4096, n_templates=5, tsize=128, n_particles=40, snr=0.4,
                           defocus=0.9)
    return [a, b]


def _key(r):
    return [(int(p["x"]), int(p["y"]), int(p["template_id"]), int(p["angle_index"])) for p in r]


@pytest.mark.timeout(600)
def test_batch_and_batch_refine_match_the_oracle(gs, orc, items):
    cfg = _cfg(gs)
    batch = [{"image": d.image, "ctf": d.ctf, "templates": d.templates} for d in items]
    eng = gs.Engine(cfg)
    eng.set_chunk(4)
    eng.set_graph(True)
    got = eng.search_batch(batch)
    recs, refs = eng.search_batch_refine(batch, 4, iterations=5)
    eng.close()
    for i, d in enumerate(items):
        opicks, st = orc.search_norm(cfg, d.image, d.templates, d.ctf, want_stats=True)
        assert len(opicks) > 0
        assert _key(got[i]) == _key(opicks), i
        np.testing.assert_allclose(got[i]["score"], opicks["score"], rtol=1e-4)
        assert _key(recs[i]) == _key(opicks)
        # refinement of the first 4 picks from zero offsets, against the oracle refiner with the
        # oracle's own mu and sigma
        r = refs[i]
        assert len(r) == min(4, len(opicks))
        pk = np.zeros(len(r), dtype=orc.PEAK_DTYPE)
        for n in ("block", "pixel", "template_id"):
            pk[n] = opicks[n][:len(r)] if n != "pixel" else opicks["pixel"][:len(r)]
        angles = orc.angles(cfg.phistep)
        pk["theta"] = angles[opicks["angle_index"][:len(r)]]
        l = cfg.window_size
        q = pk["block"].astype(np.int64) * l * l + pk["pixel"]
        R = orc.Refiner(cfg, d.image, d.templates, d.ctf)
        R.float_geometry(True)
        o, zo, _, z0 = R.run(pk, st["mu"][q], st["sigma"][q], iters=5, step=0.1, scale=(2.0, 2.0, 3.0, 500.0))
        R.close()
        np.testing.assert_allclose(r["score0"], z0, rtol=1e-4)
        np.testing.assert_allclose(r["score"], zo, rtol=1e-4)
        for n, tol in zip(("dx", "dy", "theta", "ddf"), TOL):
            assert np.max(np.abs(r["peak"][n] - o[n])) <= tol, (i, n)
