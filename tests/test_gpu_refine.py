"""GPU adjoint refinement (gs_score_and_grad, gs_refine, gs_select) against the FP64 oracle.

Tolerances: z within 1e-4 relative; gradient within 2e-3 of the oracle's per component (plus
a small absolute floor per unit); refined parameters within 0.05 px, 0.1 deg and 5 Angstrom of
the oracle's refinement from the same start with the same ascent.  mu and sigma are the
GPU search's (frozen constants of the score); the oracle search's agree to 1e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = np.array([0.05, 0.05, 0.1, 5.0])
GRAD_ATOL = np.array([1e-3, 1e-3, 1e-3, 1e-6])


@pytest.fixture(scope="module")
def case(gs, orc):
    import synth

    cfg = gs.SearchConfig(angpix=1.0, phistep=10.0, kk=3.0, energy=300.0, cs=2.7, highres=8.0,
                          lowres=100.0, diameter=128.0, threshold=5.0, window_size=320, overlap=0,
                          norm_type=1, invert=0, phase_flip=0, nms_radius=64.0)
    ds = synth.make_dataset(2101, 1280, 960, n_templates=6, tsize=128, n_particles=6, snr=0.5,
                            defocus=1.7, dfdiff=0.004, dfang=20.0)
    eng = gs.Engine(cfg)
    eng.load_templates(ds.templates)
    recs = eng.search(ds.image, ds.ctf)
    mu = eng.stage(4, np.float64)
    sigma = eng.stage(5, np.float64)
    R = orc.Refiner(cfg, ds.image, ds.templates, ds.ctf)
    R.float_geometry(True)  # the HIP path samples the rotation in float, as the search does
    yield dict(cfg=cfg, ds=ds, eng=eng, recs=recs, mu=mu, sigma=sigma, R=R)
    R.close()
    eng.close()


def _frozen(case, pk):
    q = pk["block"].astype(np.int64) * 320 * 320 + pk["pixel"]
    return case["mu"][q], case["sigma"][q]


def _perturbed(case, k, seed):
    rng = np.random.default_rng(seed)
    pk = case["eng"].peaks_from_records(case["recs"][:k])
    pk["dx"] = rng.uniform(-0.5, 0.5, len(pk))
    pk["dy"] = rng.uniform(-0.5, 0.5, len(pk))
    pk["theta"] += rng.uniform(-2.5, 2.5, len(pk))
    pk["ddf"] = rng.uniform(-300, 300, len(pk))
    return pk


def _as_oracle(orc, pk):
    o = np.zeros(len(pk), dtype=orc.PEAK_DTYPE)
    for n in pk.dtype.names:
        o[n] = pk[n]
    return o


def test_zero_offsets_reproduce_the_search_score(case):
    eng, recs = case["eng"], case["recs"]
    assert len(recs) >= 6
    pk = eng.peaks_from_records(recs)
    z, g = eng.score_and_grad(pk)
    np.testing.assert_allclose(z, recs["score"], rtol=2e-5)
    assert np.all(np.isfinite(g))


def test_score_and_gradient_match_the_oracle(case, orc):
    eng = case["eng"]
    pk = _perturbed(case, 8, seed=3)
    z, g = eng.score_and_grad(pk)
    mu, sg = _frozen(case, pk)
    zo, go, _ = case["R"].eval(_as_oracle(orc, pk), mu, sg)
    np.testing.assert_allclose(z, zo, rtol=1e-4)
    for c in range(4):
        np.testing.assert_allclose(g[:, c], go[:, c], rtol=2e-3, atol=GRAD_ATOL[c])


def test_refinement_matches_the_oracle(case, orc):
    eng = case["eng"]
    pk = _perturbed(case, 8, seed=4)
    out = eng.refine(pk)  # defaults: 20 steps of 0.1, scale (2 px, 2 px, 3 deg, 500 A)
    mu, sg = _frozen(case, pk)
    ref, zo, go, z0o = case["R"].run(_as_oracle(orc, pk), mu, sg, iters=20, step=0.1,
                                     scale=(2.0, 2.0, 3.0, 500.0))
    got = out["peak"]
    for c, n in enumerate(("dx", "dy", "theta", "ddf")):
        assert np.max(np.abs(got[n] - ref[n])) <= TOL[c], n
    np.testing.assert_allclose(out["score"], zo, rtol=1e-4)
    np.testing.assert_allclose(out["score0"], z0o, rtol=1e-4)
    assert np.all(out["score"] >= out["score0"] - 1e-6)
    assert eng.refine_timing() > 0


def test_refinement_through_torch_autograd(case, gs):
    """A torch.autograd.Function whose forward is gs_score_and_grad and whose backward returns
    its gradient; ascent driven by torch gives gs_refine's result."""
    torch = pytest.importorskip("torch")
    eng = case["eng"]
    pk0 = _perturbed(case, 4, seed=5)

    class ZScore(torch.autograd.Function):
        @staticmethod
        def forward(ctx, params):
            pk = pk0.copy()
            p = params.detach().cpu().numpy()
            pk["dx"], pk["dy"], pk["theta"], pk["ddf"] = p[:, 0], p[:, 1], p[:, 2], p[:, 3]
            z, g = eng.score_and_grad(pk)
            ctx.save_for_backward(torch.from_numpy(g))
            return torch.from_numpy(z)

        @staticmethod
        def backward(ctx, gz):
            (g,) = ctx.saved_tensors
            return gz[:, None] * g

    params = torch.tensor(np.stack([pk0["dx"], pk0["dy"], pk0["theta"], pk0["ddf"]], 1),
                          dtype=torch.float64, requires_grad=True)
    s2 = torch.tensor([2.0, 2.0, 3.0, 500.0], dtype=torch.float64) ** 2
    for _ in range(20):
        z = ZScore.apply(params)
        params.grad = None
        z.sum().backward()
        with torch.no_grad():
            params += 0.1 * s2 * params.grad
    out = eng.refine(pk0)
    p = params.detach().numpy()
    for c, n in enumerate(("dx", "dy", "theta", "ddf")):
        np.testing.assert_allclose(out["peak"][n], p[:, c], rtol=1e-9, atol=1e-9 * max(1.0, TOL[c]))


def test_select_reproduces_and_extends_the_records(case):
    eng, recs = case["eng"], case["recs"]
    cfg = case["cfg"]
    same = eng.select(cfg.threshold, cfg.nms_radius)
    assert len(same) == len(recs)
    for n in recs.dtype.names:
        assert np.array_equal(same[n], recs[n]), n
    more = eng.select(cfg.threshold - 1.0, 0.0)
    assert len(more) > len(recs)
    key = {(int(r["block"]), int(r["pixel"])) for r in more}
    assert all((int(r["block"]), int(r["pixel"])) in key for r in recs)
    # the run's own records are unchanged
    again = eng.records()
    assert np.array_equal(again["score"], recs["score"])


def test_refinement_beyond_one_device_batch(case, orc):
    """More picks than two device batches of the refinement (2048 each, one per stream lane):
    the first two batches run side by side on two streams, the third reuses the first lane's
    buffers with a smaller count.  Picks of the second and third batches must come out exactly
    as a separate call on those picks alone, and a sample of every batch must agree with the
    oracle."""
    eng = case["eng"]
    base = _perturbed(case, 6, seed=11)
    k = 4200
    pk = np.concatenate([base] * (k // len(base) + 1))[:k].copy()
    rng = np.random.default_rng(12)
    pk["dx"] += rng.uniform(-0.2, 0.2, k)
    pk["theta"] += rng.uniform(-0.5, 0.5, k)
    out = eng.refine(pk, iterations=3)
    for lo, hi in ((4096, 4200), (2048, 2100)):
        part = eng.refine(pk[lo:hi].copy(), iterations=3)
        assert np.array_equal(out["score"][lo:hi], part["score"])
        assert np.array_equal(out["peak"][lo:hi], part["peak"])
    head = eng.refine(pk[:8].copy(), iterations=3)
    np.testing.assert_allclose(out["score"][:8], head["score"], rtol=1e-6)
    idx = np.array([0, 2047, 2048, 4095, 4096, 4199])
    mu, sg = _frozen(case, pk[idx])
    o, zo, _, _ = case["R"].run(_as_oracle(orc, pk[idx]), mu, sg, iters=3, step=0.1,
                                scale=(2.0, 2.0, 3.0, 500.0))
    np.testing.assert_allclose(out["score"][idx], zo, rtol=1e-4)
    for n, tol in zip(("dx", "dy", "theta", "ddf"), TOL):
        assert np.max(np.abs(out["peak"][n][idx] - o[n])) <= tol
