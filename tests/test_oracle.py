"""Pinning the FP64 oracle (CPU only).

The reference ships no tests, golden vectors or fixtures for this path, and cannot run here
(CUDA + cuFFT), so each oracle function is pinned by what the mathematics fixes: brute-force
DFTs, numpy's independent FFT, closed forms (textbook CTF, band edges, geometry), invariances
(affine invariance of the standardisation, scale invariance of whitening, spectral
normalisation) and brute-force spatial correlation / greedy NMS on tiny inputs.
"""
import math

import numpy as np
import pytest

import oracle as orc_mod  # noqa: F401  (path from conftest)


def dft_bruteforce(x, sign):
    n = len(x)
    k = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return W @ x


SMOOTH = sorted({2 ** a * 3 ** b * 5 ** c for a in range(7) for b in range(4) for c in range(3)
                 if 2 ** a * 3 ** b * 5 ** c <= 64})


@pytest.mark.parametrize("n", SMOOTH + [7, 11, 13, 31, 53, 101])
def test_fft_matches_bruteforce_dft(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        got = orc.fft1d(x, sign)
        ref = dft_bruteforce(x, sign)
        assert np.max(np.abs(got - ref)) < 1e-11 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n", [320, 360, 720, 3710, 3838, 4092, 5760])
def test_fft_matches_numpy_and_parseval(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = orc.fft1d(x, -1)
    assert np.max(np.abs(got - np.fft.fft(x))) < 1e-12 * n
    assert abs(np.sum(np.abs(got) ** 2) / n - np.sum(np.abs(x) ** 2)) < 1e-9 * np.sum(np.abs(x) ** 2)
    # impulse <-> constant, and the shift theorem
    d = np.zeros(n, complex)
    d[3] = 1
    assert np.allclose(orc.fft1d(d, -1), np.exp(-2j * np.pi * 3 * np.arange(n) / n), atol=1e-12)


def test_fft2_matches_numpy(orc):
    rng = np.random.default_rng(2)
    x = rng.standard_normal((40, 36)) + 1j * rng.standard_normal((40, 36))
    assert np.allclose(orc.fft2d(x, -1), np.fft.fft2(x), atol=1e-10)
    assert np.allclose(orc.fft2d(x, 1), np.fft.ifft2(x) * x.size, atol=1e-10)


def test_radial_bin_exact(orc):
    # floor(sqrt(n) + 1/2) == (isqrt(4n) + 1) // 2 for integers n
    pts = [(dx, dy) for dx in range(0, 120) for dy in range(0, 120, 7)]
    rng = np.random.default_rng(0)
    pts += [tuple(int(v) for v in rng.integers(0, 3000, 2)) for _ in range(2000)]
    for dx, dy in pts:
        n = dx * dx + dy * dy
        assert orc.radial_bin(dx, dy) == (math.isqrt(4 * n) + 1) // 2 - 1


def textbook_ctf(s, df_um, cs_mm, kv, A):
    """CTF = sqrt(1-A^2) sin chi + A cos chi with chi = pi lambda df s^2 - pi/2 Cs lambda^3 s^4
    (df, Cs in Angstrom, relativistic electron wavelength)."""
    lam = 12.2639 / math.sqrt(kv * 1000 + 0.97845 * kv * kv)
    chi = math.pi * lam * df_um * 1e4 * s * s - 0.5 * math.pi * cs_mm * 1e7 * lam ** 3 * s ** 4
    return math.sqrt(1 - A * A) * math.sin(chi) + A * math.cos(chi)


def test_ctf_matches_textbook_isotropic(orc):
    nx = ny = 320
    apix = 1.0
    for x, y in [(0, 160), (5, 170), (40, 100), (100, 260), (150, 10), (160, 160)]:
        # the reference evaluates CTF_AST at (x, (y + ny/2) % ny): dy = y - ny/2
        X = min(x, nx - x)
        s = math.hypot(X / (nx * apix), (y - ny // 2) / (ny * apix))
        ref = textbook_ctf(s, 2.0, 2.7, 300.0, 0.07)
        got = orc.ctf(x, y, nx, ny, apix, 2.0, 0.0, 0.0, mode=1)
        assert abs(got - ref) < 1e-6, (x, y, got, ref)
        assert abs(orc.ctf(x, y, nx, ny, apix, 2.0, 0.0, 0.0, mode=2) - abs(ref)) < 1e-6
        assert orc.ctf(x, y, nx, ny, apix, 2.0, 0.0, 0.0, mode=0) == (1.0 if ref > 0 else -1.0)


def test_ctf_astigmatism_direction(orc):
    # astigmatic defocus: df(phi) = def - dfdiff * cos(2 (dfang - phi)) (reference convention:
    # dfu = -def + dfdiff, dfv = -def - dfdiff, df_ast = (dfu + dfv)/2 + dfdiff cos(2(dfang - ag)))
    nx = ny = 512
    for dfang in (0.0, 30.0, 75.0):
        for x, y in [(40, 256 + 30), (10, 256 + 90), (60, 256 - 20)]:
            ag = math.atan2(y - ny // 2, x)
            df = 2.0 - 0.1 * math.cos(2 * (dfang * 3.1415926535 / 180 - ag))
            s = math.hypot(x / nx, (y - ny // 2) / ny)
            ref = textbook_ctf(s, df, 2.7, 300.0, 0.07)
            got = orc.ctf(x, y, nx, ny, 1.0, 2.0, 0.1, dfang, mode=1)
            assert abs(got - ref) < 1e-6


def test_block_offsets_closed_form(orc):
    assert list(orc.block_offsets(5760, 320, 0)) == list(range(0, 5760, 320))
    o = orc.block_offsets(4092, 320, 0)
    assert list(o) == list(range(0, 3521, 320)) + [3772]
    assert len(orc.block_offsets(5760, 320, 0)) * len(o) == 234
    assert list(orc.block_offsets(1000, 320, 40)) == [0, 280, 560, 680]
    assert list(orc.block_offsets(320, 320, 0)) == [0]


def test_tile_geometry(orc):
    ulx, uly = orc.tile_geometry(1440, 1440)
    assert list(zip(ulx, uly)) == [(0, 0), (0, 720), (720, 0), (720, 720)]
    ulx, uly = orc.tile_geometry(1800, 720)
    # cols = 3: centre of the middle column = x_step*(2j+1)/2 = 600*3/2 = 900
    assert list(ulx) == [0, 540, 1080] and list(uly) == [0, 0, 0]


def test_angle_loop_float_accumulation(orc):
    for step in (5.0, 10.0, 0.7, 7.3):
        ref = []
        e = np.float32(0)
        while e < np.float32(360):
            ref.append(float(e))
            e = np.float32(e + np.float32(step))
        assert list(orc.angles(step)) == ref


def test_clip_float_sequential_sums(orc):
    rng = np.random.default_rng(4)
    W, H = 48, 40
    img = rng.standard_normal((H, W)).astype(np.float32)
    img[3, 4] = 9.0
    img[10, 11] = -8.5
    f32 = np.float32
    s = f32(0)
    s2 = f32(0)
    for v in img.ravel():
        s = f32(s + f32(f32(v / f32(W)) / f32(H)))
        s2 = f32(s2 + f32(f32(f32(v * v) / f32(W)) / f32(H)))
    var = f32(np.sqrt(f32(s2 - f32(s * s))))
    up, low = int(f32(s + f32(6) * var)), int(f32(s - f32(6) * var))
    ref = img.copy()
    ref[(img > up) | (img < low)] = s
    assert np.array_equal(orc.clip_image(img, 0), ref)
    assert np.array_equal(orc.clip_image(img, 1), -ref)


def brute_corr(I, T, l):
    f = np.fft.ifft2(I)
    g = np.fft.ifft2(T)
    out = np.zeros((l, l), complex)
    for py in range(l):
        for px in range(l):
            sh = np.roll(np.roll(f, -(py + l // 2), axis=0), -(px + l // 2), axis=1)
            out[py, px] = l * l * np.sum(sh * np.conj(g))
    return out * (-1 if (l // 2) % 2 else 1)


@pytest.mark.parametrize("l", [16, 18])
def test_correlation_shift_and_sign(orc, l):
    rng = np.random.default_rng(l)
    I = rng.standard_normal((l, l)) + 1j * rng.standard_normal((l, l))
    T = rng.standard_normal((l, l)) + 1j * rng.standard_normal((l, l))
    assert np.allclose(orc.correlate(I, T), brute_corr(I, T, l), atol=1e-9)


def greedy_nms(c, R):
    c = sorted(c, key=lambda p: (-p[2], p[1], p[0]))
    kept = []
    for p in c:
        if all((p[0] - q[0]) ** 2 + (p[1] - q[1]) ** 2 >= R * R for q in kept):
            kept.append(p)
    return kept


def test_nms_bruteforce(orc):
    rng = np.random.default_rng(9)
    n = 300
    a = np.zeros(n, dtype=orc.PICK_DTYPE)
    a["x"] = rng.integers(0, 800, n)
    a["y"] = rng.integers(0, 800, n)
    a["score"] = rng.random(n).astype(np.float32) * 10
    got = orc.nms(a, 64.0)
    ref = greedy_nms([(float(p["x"]), float(p["y"]), float(p["score"])) for p in a], 64.0)
    assert [(p["x"], p["y"]) for p in got] == [(p[0], p[1]) for p in ref]


def _cfg(**kw):
    import gisspa

    return gisspa.SearchConfig(**kw)


def _spectral_norm(F, l):
    ky, kx = np.meshgrid(np.arange(l), np.arange(l), indexing="ij")
    fx, fy = np.minimum(kx, l - kx), np.minimum(ky, l - ky)
    r = (np.floor(np.sqrt(fx ** 2 + fy ** 2) + 0.5) - 1).astype(int)
    m = (r >= 0) & (r < l // 2) & (kx <= l // 2)
    return math.sqrt(np.sum(np.abs(F[m]) ** 2) / m.sum() ** 2)


@pytest.mark.parametrize("norm_type", [1, 0])
def test_template_prep_invariants(orc, norm_type):
    import synth

    l = 320 if norm_type else 720
    cfg = _cfg(window_size=l, norm_type=norm_type)
    rng = np.random.default_rng(1)
    # templates filling the whole box, so an affine map of the template is an affine map of the
    # padded plane: the edge-mean / box-sigma standardisation must remove it exactly
    t = rng.standard_normal((2, l, l)).astype(np.float32) + 0.3
    a = orc.prep_templates(cfg, t, (1.5, 0.01, 20.0))
    b = orc.prep_templates(cfg, (2.5 * t.astype(np.float64) - 4.0).astype(np.float32), (1.5, 0.01, 20.0))
    assert np.allclose(a, b, atol=1e-5 * np.max(np.abs(a)))
    for k in range(2):
        F = np.fft.fft2(a[k]) / (l * l) if norm_type else a[k]
        assert abs(_spectral_norm(F, l) - 1.0) < 1e-9
        # template band-pass: nothing at or beyond hr + 8 (hr = l*apix/highres)
        ky, kx = np.meshgrid(np.arange(l), np.arange(l), indexing="ij")
        r = np.floor(np.hypot(np.minimum(kx, l - kx), np.minimum(ky, l - ky)) + 0.5) - 1
        assert np.max(np.abs(F[r >= l / 8.0 + 8])) < 1e-12 * np.max(np.abs(F))
        assert abs(F[0, 0]) < 1e-12 * np.max(np.abs(F))


def test_image_prep_invariants(orc):
    cfg = _cfg(window_size=320, norm_type=1)
    rng = np.random.default_rng(3)
    img = rng.standard_normal((400, 480)).astype(np.float32)
    a = orc.prep_image(cfg, img, (1.5, 0.0, 0.0))
    b = orc.prep_image(cfg, img * np.float32(4.0), (1.5, 0.0, 0.0))  # exact in float32
    assert np.allclose(a, b, atol=1e-9 * np.max(np.abs(a)))  # whitening removes global scale
    assert abs(a.mean()) < 1e-12 * np.max(np.abs(a))          # DC zeroed (set_0Hz_to_0_at_RI)
    F = np.fft.fft2(a)
    ky, kx = np.meshgrid(np.arange(400), np.arange(480), indexing="ij")
    r = np.floor(np.hypot(np.minimum(kx, 480 - kx), np.minimum(ky, 400 - ky)) + 0.5) - 1
    hr, lr = 480 / 8.0, 480 / 100.0
    assert np.max(np.abs(F[(r >= hr + 8) | (r < lr - 8)])) < 1e-9 * np.max(np.abs(F))


def test_rotation_identity_and_quarter_turn(orc):
    l = 320
    cfg = _cfg(window_size=l, norm_type=1, diameter=128.0)
    rng = np.random.default_rng(5)
    P = rng.standard_normal((1, l, l))
    y, x = np.mgrid[0:l, 0:l]
    r = np.hypot(x - l // 2, y - l // 2)
    rin, rout = 1.5 * 128 / 2, 1.5 * 128 / 2 + 8
    mask = np.where(r > rout, 0.0, np.where(r >= rin, 0.5 * np.cos(3.1415926535 * (r - rin) / 8) + 0.5, 1.0))

    def expected(sp):
        v = sp * mask
        v = (v - v.mean()) / v.std()
        return np.fft.fft2(v) / (l * l)

    got0 = orc.rotate_templates(cfg, P.astype(complex), 0.0)[0]
    assert np.allclose(got0, expected(P[0]), atol=1e-12)
    got90 = orc.rotate_templates(cfg, P.astype(complex), 90.0)[0]
    rot = np.zeros((l, l))
    rot[:, 1:] = P[0][l - np.arange(1, l), :].T  # out[j][i] = in[l - i][j] (i >= 1)
    assert np.allclose(got90, expected(rot), atol=1e-8)


def test_norm_search_finds_planted_particles(orc):
    import synth

    cfg = _cfg(phistep=10.0, window_size=320, norm_type=1, nms_radius=64.0)
    ds = synth.make_dataset(11, 640, 640, n_templates=3, tsize=128, n_particles=3, snr=0.6, defocus=1.5)
    picks, _ = orc.search_norm(cfg, ds.image, ds.templates, ds.ctf)
    # z is standardised over only 3 x 36 hypotheses per pixel (max z < sqrt(107)), so a particle
    # near a block edge can stay under threshold; every pick must still be a planted particle
    # found with its own template, and most particles must be found.
    found = 0
    for p in picks:
        d = np.hypot(ds.positions[:, 0] - p["x"], ds.positions[:, 1] - p["y"])
        j = int(np.argmin(d))
        assert d[j] <= 3 and ds.particle_templates[j] == p["template_id"]
        found += 1
    assert found >= 2
