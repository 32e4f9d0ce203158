"""Window sizes other than 320 and 720 (the generic correlation path: runtime-planned transforms
instead of the tiled kernels), and the refinement at all three kinds of window, against the FP64
oracle, through the C ABI.

The reference plans its transforms for any window_size (src/norm.cu:93-116); so does this
build.  Same bar as the l = 320 parity tests: per-pixel z of the best record within 1e-4,
identical picks with scores within 1e-4 (norm_type=1); identical records (norm_type=0);
refinement within the stated tolerances.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _key(a):
    return [(int(p["x"]), int(p["y"]), int(p["template_id"]), int(p["angle_index"])) for p in a]


@pytest.mark.parametrize("l", [256, 300])
def test_norm_window_size(gs, orc, l):
    import synth

    print(f"norm_type=1 window {l}", flush=True)
    cfg = gs.SearchConfig(phistep=10.0, window_size=l, norm_type=1, nms_radius=64.0)
    ds = synth.make_dataset(40 + l, 640, 576, n_templates=3, tsize=128, n_particles=3, snr=0.5,
                            defocus=1.5, dfdiff=0.02, dfang=30.0)
    opicks, st = orc.search_norm(cfg, ds.image, ds.templates, ds.ctf, want_stats=True)
    eng = gs.Engine(cfg)
    eng.load_templates(ds.templates)
    g = eng.search(ds.image, ds.ctf)
    mu = eng.stage(4, np.float64)
    sd = eng.stage(5, np.float64)
    best = eng.stage(6, np.float32).astype(np.float64)
    assert mu.size == st["mu"].size and mu.size % (l * l) == 0
    z_o = (st["best"] - st["mu"]) / st["sigma"]
    z_g = (best - mu) / sd
    assert np.max(np.abs(z_g - z_o) / np.maximum(1.0, np.abs(z_o))) < 1e-4
    arg = eng.stage(7, np.int32)
    same = arg == st["arg"]
    assert same.mean() > 0.999, same.mean()
    assert np.all(same[z_o >= cfg.threshold])
    assert len(opicks) > 0
    assert _key(g) == _key(opicks)
    np.testing.assert_allclose(g["score"], opicks["score"], rtol=1e-4)
    # the autotuner has nothing to choose for the generic path and keeps the settings
    col, row, chunk = eng.autotune()
    assert (col, row) == (12, 8)
    eng.close()


def test_nonorm_window_size(gs, orc):
    import synth

    print("norm_type=0 window 360", flush=True)
    cfg = gs.SearchConfig(phistep=30.0, window_size=360, norm_type=0, phase_flip=1, threshold=3.0,
                          nms_radius=0.0)
    ds = synth.make_dataset(7, 720, 720, n_templates=2, tsize=128, n_particles=2, snr=0.5, defocus=1.5)
    opicks, orecs = orc.search_nonorm(cfg, ds.image, ds.templates, ds.ctf)
    eng = gs.Engine(cfg)
    eng.load_templates(ds.templates)
    g = eng.search(ds.image, ds.ctf)
    assert len(orecs) > 0
    ko = sorted((int(p["angle_index"]), int(p["template_id"])) for p in orecs)
    kg = sorted((int(p["angle_index"]), int(p["template_id"])) for p in g)
    assert kg == ko
    oi = {(int(p["angle_index"]), int(p["template_id"])): p for p in orecs}
    for p in g:
        q = oi[(int(p["angle_index"]), int(p["template_id"]))]
        assert int(p["pixel"]) == int(q["pixel"])
        assert abs(p["score"] - q["score"]) <= 1e-4 * abs(q["score"])
        assert abs(p["local_x"] - q["local_x"]) < 1e-3 and abs(p["local_y"] - q["local_y"]) < 1e-3
    eng.close()


@pytest.mark.parametrize("l,npk", [(256, 4), (320, 5), (720, 5)])
def test_refinement_window_size(gs, orc, l, npk):
    """Refinement at a window without LDS plane transforms (256: runtime FFT, k_ref_rotate,
    k_ref_dot) and at both fused sizes (320, 720: k_ref_rotate_rows, k_ref_cols_dot) with an odd
    pick count, so that the last B plane holds one pick."""
    import synth

    print(f"refinement window {l}, {npk} picks", flush=True)
    cfg = gs.SearchConfig(angpix=1.0, phistep=10.0, kk=3.0, energy=300.0, cs=2.7, highres=8.0,
                          lowres=100.0, diameter=128.0, threshold=5.0, window_size=l, overlap=0,
                          norm_type=1, invert=0, phase_flip=0, nms_radius=64.0)
    ds = synth.make_dataset(2102, 768, 768, n_templates=4, tsize=128, n_particles=5, snr=0.5,
                            defocus=1.7, dfdiff=0.004, dfang=20.0)
    eng = gs.Engine(cfg)
    eng.load_templates(ds.templates)
    recs = eng.search(ds.image, ds.ctf)
    assert len(recs) >= 4
    mu, sigma = eng.stage(4, np.float64), eng.stage(5, np.float64)
    pk = eng.peaks_from_records(recs)
    z, _ = eng.score_and_grad(pk)
    np.testing.assert_allclose(z, recs["score"], rtol=2e-5)
    k = min(npk, len(recs))
    if npk % 2 and k % 2 == 0:
        k -= 1  # keep the pick count odd
    assert k >= 3
    rng = np.random.default_rng(9)
    pk = eng.peaks_from_records(recs[:k])
    pk["dx"] = rng.uniform(-0.5, 0.5, len(pk))
    pk["dy"] = rng.uniform(-0.5, 0.5, len(pk))
    pk["theta"] += rng.uniform(-2.5, 2.5, len(pk))
    pk["ddf"] = rng.uniform(-300, 300, len(pk))
    q = pk["block"].astype(np.int64) * l * l + pk["pixel"]
    R = orc.Refiner(cfg, ds.image, ds.templates, ds.ctf)
    R.float_geometry(True)
    o = np.zeros(len(pk), dtype=orc.PEAK_DTYPE)
    for n in pk.dtype.names:
        o[n] = pk[n]
    z, g = eng.score_and_grad(pk)
    zo, go, _ = R.eval(o, mu[q], sigma[q])
    np.testing.assert_allclose(z, zo, rtol=1e-4)
    for c, atol in enumerate((1e-3, 1e-3, 1e-3, 1e-6)):
        np.testing.assert_allclose(g[:, c], go[:, c], rtol=2e-3, atol=atol)
    out = eng.refine(pk)
    ref, zr, _, _ = R.run(o, mu[q], sigma[q], iters=20, step=0.1, scale=(2.0, 2.0, 3.0, 500.0))
    for c, n in enumerate(("dx", "dy", "theta", "ddf")):
        assert np.max(np.abs(out["peak"][n] - ref[n])) <= (0.05, 0.05, 0.1, 5.0)[c], n
    np.testing.assert_allclose(out["score"], zr, rtol=1e-4)
    R.close()
    eng.close()
