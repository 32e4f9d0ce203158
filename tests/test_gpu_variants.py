"""Every result-invariant option of the norm_type=1 correlation kernels against the FP64 oracle:
column-pass variants 0-13, row-pass variants 0-9, templates per launch (including a partial last
launch) and the two-stream pipeline.  Each must give the oracle's per-pixel z of the best
record within 1e-4 and the oracle's pick list."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(gs, orc):
    import synth

    cfg = gs.SearchConfig(angpix=1.0, phistep=10.0, kk=3.0, energy=300.0, cs=2.7, highres=8.0,
                          lowres=100.0, diameter=128.0, threshold=5.0, window_size=320, overlap=0,
                          norm_type=1, invert=0, phase_flip=0, nms_radius=64.0)
    ds = synth.make_dataset(2101, 1280, 960, n_templates=6, tsize=128, n_particles=6, snr=0.5,
                            defocus=1.7, dfdiff=0.004, dfang=20.0)
    picks, st = orc.search_norm(cfg, ds.image, ds.templates, ds.ctf, want_stats=True)
    z = (st["best"] - st["mu"]) / st["sigma"]
    return cfg, ds, picks, z


def _key(a):
    return [(int(p["x"]), int(p["y"]), int(p["template_id"]), int(p["angle_index"])) for p in a]


@pytest.mark.parametrize("col,row,chunk,overlap", [
    (0, 0, 0, 0), (1, 0, 0, 0), (2, 0, 0, 0), (3, 0, 0, 0), (4, 0, 0, 0), (5, 0, 0, 0), (6, 0, 0, 0),
    (7, 6, 0, 0),
    (8, 7, 0, 0), (8, 7, 4, 0),  # row-packed inter-pass layout (also with a partial last launch)
    (8, 8, 0, 0), (8, 8, 4, 0),  # the same with the lazy argmax (running maximum across launches)
    (7, 9, 0, 0), (7, 9, 4, 0),  # unpacked layout, packing in the row pass, lazy argmax
    (9, 7, 0, 0), (9, 7, 4, 0),  # the same with 8 templates per workgroup (one partial chunk)
    (10, 7, 0, 0), (10, 7, 4, 0),  # row-packed, lane twiddles loaded into registers
    (11, 8, 0, 0), (11, 8, 4, 0),  # 9 with the block columns held and the next template prefetched
    (12, 8, 0, 0), (13, 8, 4, 0),  # 8 with the column group the fastest index of the order
    (0, 1, 0, 0), (0, 2, 0, 0), (0, 3, 0, 0), (0, 4, 0, 0), (0, 5, 0, 0), (0, 6, 0, 0),
    (3, 6, 4, 0),   # eight templates per iteration with a partial last iteration (4 + 2 templates)
    (3, 1, 4, 0),   # two launches per angle, the second partial (4 + 2 templates)
    (4, 4, 4, 0),   # the same with the bank-conflict-free layouts of both passes
    (0, 1, 2, 1),   # three launches per angle through the double-buffered two-stream pipeline
    (5, 5, 2, 1),
    (8, 8, 2, 1),   # row-packed layout and running maximum through the two-stream pipeline
])
def test_variant_matches_oracle(gs, case, col, row, chunk, overlap):
    cfg, ds, picks, z_o = case
    print(f"variant col={col} row={row} chunk={chunk} overlap={overlap}", flush=True)
    eng = gs.Engine(cfg)
    eng.set_variant(0, col)
    eng.set_variant(1, row)
    eng.set_variant(2, overlap)
    if chunk:
        eng.set_chunk(chunk)
    eng.load_templates(ds.templates)
    g = eng.search(ds.image, ds.ctf)
    best = eng.stage(6, np.float32).astype(np.float64)
    z_g = (best - eng.stage(4, np.float64)) / eng.stage(5, np.float64)
    assert np.max(np.abs(z_g - z_o) / np.maximum(1.0, np.abs(z_o))) < 1e-4
    assert len(picks) > 0
    assert _key(g) == _key(picks)
    np.testing.assert_allclose(g["score"], picks["score"], rtol=1e-4)
    eng.close()


def test_packed_layout_needs_its_row_pass(gs, case):
    """Column variants 8 to 13 write the row-packed layout that only row variants 7 and 8 read; a
    mismatched pair is rejected before any kernel runs."""
    cfg, ds, _, _ = case
    eng = gs.Engine(cfg)
    eng.load_templates(ds.templates)
    for col, row in ((8, 6), (9, 6), (11, 6), (7, 7), (7, 8), (8, 9)):
        eng.set_variant(0, col)
        eng.set_variant(1, row)
        with pytest.raises(gs.GisspaError, match="go together"):
            eng.search(ds.image, ds.ctf)
    eng.close()
