"""Pins of the FP64 refinement oracle (orc_refine_*).

The refinement has no reference counterpart, so its score and gradient are pinned by:
* the search itself: at zero offsets and theta on the angle grid, z is the oracle search's score;
* the search's own stage functions: at integer shifts, any theta and any defocus offset, c is
  the correlation plane built by template preprocessing at the offset defocus, RotateTemplate,
  compute_corner_CCG and the inverse transform (orc_refine_plane);
* central differences for the gradient (steps below the kinks of the bilinear interpolation and
  of |CTF|: 1e-3 px, 1e-3 deg, 0.05 Angstrom);
* the planted truth: picks of noise-level particles ascend to the particles' true in-plane
  angle (continuous in the generator) and integer position.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def case(orc):
    import gisspa
    import synth

    cfg = gisspa.SearchConfig(angpix=1.0, phistep=10.0, kk=3.0, energy=300.0, cs=2.7, highres=8.0,
                              lowres=100.0, diameter=128.0, threshold=5.0, window_size=320,
                              overlap=0, norm_type=1, invert=0, phase_flip=0, nms_radius=64.0)
    ds = synth.make_dataset(2101, 1280, 960, n_templates=6, tsize=128, n_particles=6, snr=0.5,
                            defocus=1.7, dfdiff=0.004, dfang=20.0)
    picks, st = orc.search_norm(cfg, ds.image, ds.templates, ds.ctf, want_stats=True)
    R = orc.Refiner(cfg, ds.image, ds.templates, ds.ctf)
    ang = orc.angles(cfg.phistep)
    pk = np.zeros(len(picks), dtype=orc.PEAK_DTYPE)
    pk["block"] = picks["block"]
    pk["pixel"] = picks["pixel"]
    pk["template_id"] = picks["template_id"]
    pk["theta"] = ang[picks["angle_index"]]
    q = picks["block"] * 320 * 320 + picks["pixel"]
    yield dict(cfg=cfg, ds=ds, picks=picks, st=st, R=R, pk=pk, mu=st["mu"][q], sigma=st["sigma"][q])
    R.close()


def _perturbed(case, k=6, seed=0):
    rng = np.random.default_rng(seed)
    pk = case["pk"][:k].copy()
    pk["dx"] = rng.uniform(-0.5, 0.5, k)
    pk["dy"] = rng.uniform(-0.5, 0.5, k)
    pk["theta"] += rng.uniform(-2.5, 2.5, k)
    pk["ddf"] = rng.uniform(-300, 300, k)
    return pk


def test_zero_offsets_reproduce_the_search_score(case):
    R = case["R"]
    assert len(case["pk"]) >= 6
    R.float_geometry(True)
    z, _, c = R.eval(case["pk"], case["mu"], case["sigma"])
    R.float_geometry(False)
    np.testing.assert_allclose(z, case["picks"]["score"], rtol=1e-6)
    q = case["picks"]["block"] * 320 * 320 + case["picks"]["pixel"]
    np.testing.assert_allclose(c, case["st"]["best"][q], rtol=1e-12)


@pytest.mark.parametrize("shift,dtheta,ddf", [
    ((1, 0), 0.0, 0.0), ((0, -2), 0.0, 0.0), ((-1, 1), 3.7, 0.0), ((0, 0), -6.2, 150.0),
    ((2, -1), 1.3, -250.0),
])
def test_integer_shift_matches_the_search_plane(case, shift, dtheta, ddf):
    """c at an integer shift equals the search path's correlation plane at that pixel, for
    any theta and defocus offset (rotation geometry in float on both sides)."""
    R = case["R"]
    R.float_geometry(True)
    pk = case["pk"][:3].copy()
    pk["theta"] += dtheta
    pk["ddf"] = ddf
    pk["dx"], pk["dy"] = shift
    _, _, c = R.eval(pk, case["mu"][:3], case["sigma"][:3])
    R.float_geometry(False)
    for i in range(3):
        plane = R.plane(int(pk["block"][i]), int(pk["template_id"][i]), float(pk["theta"][i]), ddf)
        px = (int(pk["pixel"][i]) % 320 + shift[0]) % 320
        py = (int(pk["pixel"][i]) // 320 + shift[1]) % 320
        assert abs(c[i] - plane[py, px]) <= 1e-10 * np.abs(plane).max()


@pytest.mark.parametrize("name,col,h,rtol", [
    ("dx", 0, 1e-3, 1e-5), ("dy", 1, 1e-3, 1e-5), ("theta", 2, 1e-3, 2e-3), ("ddf", 3, 0.05, 1e-4),
])
def test_gradient_matches_central_differences(case, name, col, h, rtol):
    R = case["R"]
    pk = _perturbed(case)
    mu, sg = case["mu"][:6], case["sigma"][:6]
    _, g, _ = R.eval(pk, mu, sg)
    a, b = pk.copy(), pk.copy()
    a[name] += h
    b[name] -= h
    fd = (R.eval(a, mu, sg)[0] - R.eval(b, mu, sg)[0]) / (2 * h)
    np.testing.assert_allclose(g[:, col], fd, rtol=rtol, atol=1e-3 * np.abs(fd).max())


def test_ascent_recovers_the_planted_pose(case):
    """Planted particles have integer positions, the micrograph defocus and a continuous
    in-plane angle: refinement from perturbed starting points moves to within 0.3 px and 0.5
    deg of them, and z does not decrease."""
    R, ds = case["R"], case["ds"]
    picks = case["picks"]
    pk = _perturbed(case, seed=1)
    k = len(pk)
    ref, z, g, z0 = R.run(pk, case["mu"][:k], case["sigma"][:k], iters=20, step=0.1,
                          scale=(2.0, 2.0, 3.0, 500.0))
    assert np.all(z >= z0 - 1e-9)
    checked = 0
    for i in range(k):
        d = np.hypot(ds.positions[:, 0] - picks["x"][i], ds.positions[:, 1] - picks["y"][i])
        j = int(d.argmin())
        if d[j] > 0 or ds.particle_templates[j] != picks["template_id"][i]:
            continue
        # template rotation theta of the search equals the generator's in-plane angle psi
        dth = (ref["theta"][i] - ds.particle_psi[j] + 180) % 360 - 180
        assert abs(dth) < 0.5, (i, ref["theta"][i], ds.particle_psi[j])
        assert abs(ref["dx"][i]) < 0.3 and abs(ref["dy"][i]) < 0.3
        checked += 1
    assert checked >= 4
