"""Pins for the oracle's whitening, Wiener/CTF weighting, image filter and norm_type=0 score.

Each check comes from what the mathematics fixes, computed independently of the oracle's code:
per-bin mean powers over the half plane (brute-force integer bins), the textbook CTF and the
reference's signal/noise curves in closed form, the large-kk limit of the image filter, the
standard deviation of the plane without its peak pixel, and the position of a planted blob
after a real rotation of the plane.  Each check fails under a one-line change of the oracle
function it pins (binning the whole plane, a dropped kk, a dropped square root, an off-by-one
bin, `>=` for `>` on the exp(-100 s^2) cut, rc = l^2 instead of l^2 - 1, a flipped sign of the
back-rotation).
"""
import math

import numpy as np
import pytest

# signal and noise curves of Parameters (include/EMReader/DataReader2.h:76-77), float values
A_, B_, B2_ = np.float32(-10.81), np.float32(1.0), np.float32(0.32)
BF_, BF2_, BF3_ = np.float32(-18.17), np.float32(-15.22), np.float32(1.72)
AMP = float(np.float32(0.07))


def _cfg(**kw):
    import gisspa

    return gisspa.SearchConfig(**kw)


def bins(nx, ny):
    """Radial bin floor(hypot(fx, fy) + 1/2) - 1 of every pixel, by exact integer arithmetic."""
    y, x = np.mgrid[0:ny, 0:nx]
    fx, fy = np.minimum(x, nx - x), np.minimum(y, ny - y)
    n = fx.astype(np.int64) ** 2 + fy.astype(np.int64) ** 2
    k = np.array([math.isqrt(int(v)) for v in n.ravel()]).reshape(n.shape)
    return (k + (n > k * k + k)) - 1, x, y, fx, fy


def signed_freq(nx, ny):
    y, x = np.mgrid[0:ny, 0:nx]
    return np.where(x > nx // 2, x - nx, x), np.where(y >= ny // 2, y - ny, y)


def textbook_ctf(s, phi, defocus_um, dfdiff_um, dfang_deg, cs_mm=2.7, kv=300.0, A=AMP):
    """sqrt(1 - A^2) sin chi + A cos chi, chi = pi lambda df s^2 - pi/2 Cs lambda^3 s^4 with the
    astigmatic defocus df(phi) = defocus - dfdiff cos(2 (dfang - phi)) (Angstrom)."""
    lam = 12.2639 / math.sqrt(kv * 1000 + 0.97845 * kv * kv)
    df = (defocus_um - dfdiff_um * np.cos(2 * (dfang_deg * 3.1415926535 / 180 - phi))) * 1e4
    chi = np.pi * lam * df * s * s - 0.5 * np.pi * cs_mm * 1e7 * lam ** 3 * s ** 4
    return math.sqrt(1 - A * A) * np.sin(chi) + A * np.cos(chi)


def signal(s):
    return np.exp(float(BF_) * s * s + float(BF2_) * s + float(BF3_))


def noise(s):
    return np.exp(float(A_) * s * s + float(B_) * s + float(B2_))


def raised_cosine_lowpass(r, hr):
    """Template band of apply_weighting_function: 1 below hr, cosine edge over 8 bins, else 0
    (and 0 for the DC bin r = -1)."""
    out = np.where((r >= 0) & (r < hr), 1.0, 0.0)
    edge = (r >= hr) & (r < hr + 8)
    return np.where(edge, 0.5 * np.cos(3.1415926535 * (r - hr) / 16) + 0.5, out)


def random_spectrum(rng, nx, ny, g, r, x, right=3.0):
    """|F| = g(r) on the half plane x <= nx/2, right * g(r) on x > nx/2, random phases."""
    mag = np.where((r >= 0) & (r < len(g)), g[np.clip(r, 0, len(g) - 1)], 1.0)
    mag = np.where(x > nx // 2, right * mag, mag)
    return mag * np.exp(2j * np.pi * rng.random((ny, nx)))


# ------------------------------------------------------------------------- whitening
@pytest.mark.parametrize("l", [32, 320])
def test_whitening_unit_mean_power_per_half_plane_bin(orc, l):
    rng = np.random.default_rng(l)
    r, x, y, fx, fy = bins(l, l)
    g = np.exp(rng.uniform(-3, 3, l))  # strongly varying radial amplitude
    F = random_spectrum(rng, l, l, g, r, x)
    out = orc.whiten_template(F)
    ratio = out / F
    # a positive real factor per pixel (the phase is kept)
    assert np.max(np.abs(ratio.imag)) < 1e-12 * np.max(np.abs(ratio))
    assert np.min(ratio.real) > 0
    half = x <= l // 2
    for k in range(l // 2):
        sel = r == k
        if not sel.any():
            continue
        # every pixel of the bin, on both halves, is divided by the same factor ...
        f = ratio.real[sel]
        assert np.ptp(f) < 1e-12 * f.max()
        # ... which makes the mean power over the half plane x <= l/2 of the bin exactly 1
        assert abs(np.mean(np.abs(out[sel & half]) ** 2) - 1.0) < 1e-12
    # the DC bin (r = -1) and the corners (r >= l/2) are left alone
    keep = (r < 0) | (r >= l // 2)
    assert np.allclose(out[keep], F[keep], rtol=0, atol=0)


def test_whitening_of_white_noise_is_flat(orc):
    l = 64
    rng = np.random.default_rng(7)
    F = np.fft.fft2(rng.standard_normal((l, l))) / l ** 2
    out = orc.whiten_template(F)
    r, x, *_ = bins(l, l)
    m = [np.mean(np.abs(out[(r == k) & (x <= l // 2)]) ** 2) for k in range(l // 2)]
    assert np.allclose(m, 1.0, atol=1e-12)


# ------------------------------------------------------------------------- Wiener weight
@pytest.mark.parametrize("apix,l,kk,dfdiff", [(1.0, 320, 0.0, 0.0), (1.0, 320, 3.0, 0.0),
                                              (1.3, 256, 3.0, 0.05), (1.0, 320, 30.0, 0.005)])
def test_template_weight_closed_form(orc, apix, l, kk, dfdiff):
    """On a flat spectrum the weight is band(r) |CTF(s)| sqrt(S / ((kk + 1) N + kk S CTF^2)),
    with the textbook CTF at the pixel's signed frequency, s = |k| / (l apix) from the continuous
    radius, S and N the reference's signal and noise curves, band the raised-cosine low-pass
    on the integer bin."""
    defocus, dfang = 2.0, 30.0
    cfg = _cfg(angpix=apix, kk=kk, highres=8.0, lowres=100.0, window_size=l)
    out = orc.weight_template(cfg, np.ones((l, l), complex), (defocus, dfdiff, dfang))
    assert np.max(np.abs(out.imag)) == 0
    r, x, y, fx, fy = bins(l, l)
    kx, ky = signed_freq(l, l)
    s = np.hypot(fx, fy) * float(np.float32(1.0) / (np.float32(apix) * np.float32(l)))
    C = textbook_ctf(np.hypot(kx, ky) / (l * apix), np.arctan2(ky, kx), defocus, dfdiff, dfang)
    hr = float(np.float32(l) * np.float32(apix) / np.float32(8.0))
    band = raised_cosine_lowpass(r, hr)
    S, N = signal(s), noise(s)
    ref = band * np.abs(C) * np.sqrt(S / ((kk + 1) * N + kk * S * C * C))
    inb = (r >= 0) & (r < l // 2)
    ref = np.where(inb, ref, band)
    assert np.max(np.abs(out.real - ref)) < 2e-6 * np.max(ref)
    # Wiener identity f^2 ((kk + 1) N / S + kk CTF^2) = band^2 CTF^2 on the band
    sel = inb & (band > 0)
    lhs = out.real[sel] ** 2 * ((kk + 1) * N[sel] / S[sel] + kk * C[sel] ** 2)
    rhs = band[sel] ** 2 * C[sel] ** 2
    assert np.max(np.abs(lhs - rhs)) < 1e-5 * np.max(rhs)
    # the weight vanishes towards a CTF zero and at DC
    assert out.real[0, 0] == 0
    near0 = sel & (np.abs(C) < 0.02)
    if near0.any():
        assert np.max(out.real[near0]) < 0.05 * np.max(out.real)


# ------------------------------------------------------------------------- image filter
@pytest.mark.parametrize("kk", [3.0, 1e6])
def test_image_filter_closed_form(orc, kk):
    """whiten_filter_weight_Img after do_phase_flip on a non-square frame: sign(CTF) times
    band(r) E(r) sqrt((S CTF^2 + N) / ((kk + 1) N + kk S CTF^2)) / sqrt(P_r), where P_r is the mean
    power of bin r over the half plane x <= nx/2, E = exp(-100 s^2) only for r > l apix / 6
    (l = max(nx, ny)), and s = |(fx / nx, fy / ny)| / apix.  As kk grows the CTF-dependent
    factor tends to the flat 1 / sqrt(kk), which isolates band, E and the whitening."""
    nx, ny, apix = 96, 80, 1.0
    defocus, dfdiff, dfang = 1.5, 0.02, 40.0
    cfg = _cfg(angpix=apix, kk=kk, highres=6.0, lowres=40.0, window_size=96)
    rng = np.random.default_rng(12)
    r, x, y, fx, fy = bins(nx, ny)
    l = max(nx, ny)
    g = np.exp(rng.uniform(-2, 2, l))
    F = random_spectrum(rng, nx, ny, g, r, x)
    out = orc.filter_image_spectrum(cfg, F, (defocus, dfdiff, dfang))
    kx, ky = signed_freq(nx, ny)
    s = np.sqrt((fx / nx) ** 2 + (fy / ny) ** 2) / apix
    # CTF_AST takes the astigmatism angle of the integer frequency indices (k:297), which on a
    # non-square frame is not the angle of the physical frequency
    C = textbook_ctf(np.sqrt((kx / nx) ** 2 + (ky / ny) ** 2) / apix, np.arctan2(ky, kx),
                     defocus, dfdiff, dfang)
    sign = np.where(C > 0, 1.0, -1.0)
    hr, lr = l * apix / 6.0, l * apix / 40.0
    band = np.where((r < hr) & (r >= lr), 1.0, 0.0)
    band = np.where((r >= hr) & (r < hr + 8), 0.5 * np.cos(3.1415926535 * (r - hr) / 16) + 0.5, band)
    band = np.where((r >= lr - 8) & (r < lr) & (r >= 0), 0.5 * np.cos(3.1415926535 * (lr - r) / 16) + 0.5, band)
    assert l * apix / 6 == 16.0 and (r == 16).any()  # the exp cut is strict: bin 16 keeps E = 1
    E = np.where(r > l * apix / 6, np.exp(-100 * s * s), 1.0)
    Pr = np.where((r >= 0) & (r < l // 2), g[np.clip(r, 0, l - 1)] ** 2, 1.0)
    S, N = signal(s), noise(s)
    inb = (r >= 0) & (r < l // 2)
    if kk > 1e5:
        # flat limit: out sqrt(kk) = sign band E / sqrt(P_r) (1 + O(1/kk))
        ref = np.where(inb, sign * band * E / np.sqrt(Pr), 0.0)
        got = out / F * math.sqrt(kk)
        assert np.max(np.abs(got.imag)) < 1e-9
        assert np.max(np.abs(got.real - ref)) < 1e-5 * np.max(np.abs(ref))
    else:
        W = np.sqrt((S * C * C + N) / ((kk + 1) * N + kk * S * C * C))
        ref = np.where(inb, sign * band * E * W / np.sqrt(Pr), sign * band)
        got = out / F
        assert np.max(np.abs(got.real - ref)) < 2e-6 * np.max(np.abs(ref))
    assert out[0, 0] == 0  # the DC bin is outside the band


def test_prep_image_is_filtered_spectrum(orc):
    """prep_image = real part of the inverse transform of the filtered spectrum of the frame
    (scaled by 1/N), with the DC term zeroed (norm_type=1)."""
    nx, ny = 96, 80
    cfg = _cfg(angpix=1.0, kk=3.0, highres=6.0, lowres=40.0, window_size=96)
    rng = np.random.default_rng(13)
    img = rng.standard_normal((ny, nx)).astype(np.float32)
    ctf = (1.5, 0.02, 40.0)
    F = np.fft.fft2(img.astype(np.float64)) / (nx * ny)
    G = orc.filter_image_spectrum(cfg, F, ctf)
    G[0, 0] = 0
    ref = np.real(np.fft.ifft2(G)) * (nx * ny)
    assert np.allclose(orc.prep_image(cfg, img, ctf), ref, atol=1e-12 * np.max(np.abs(ref)))


# ------------------------------------------------------------------------- norm_type=0 score
def test_nonorm_plane_score_bruteforce(orc):
    l, d_m = 64, 40.0
    rng = np.random.default_rng(21)
    Cp = rng.standard_normal((l, l))
    Cp[30, 41] = 9.0           # planted interior peak
    Cp[2, 5] = 50.0            # larger value inside the border (x < d_m/4): not a peak
    Cp[60, 20] = 40.0          # y > l - d_m/4: border too
    st = orc.nonorm_plane(Cp.astype(complex), d_m)
    yy, xx = np.mgrid[0:l, 0:l]
    interior = ~((xx < d_m / 4) | (xx > l - d_m / 4) | (yy < d_m / 4) | (yy > l - d_m / 4))
    assert st[0] == 9.0 and int(st[1]) == 30 * l + 41
    assert math.isclose(st[2], Cp.sum(), rel_tol=1e-12)
    assert math.isclose(st[3], (Cp ** 2).sum(), rel_tol=1e-12)
    # score: peak over the population standard deviation of the l^2 - 1 other values
    rest = np.delete(Cp.ravel(), 30 * l + 41)
    assert math.isclose(orc.nonorm_score(st, l), 9.0 / np.std(rest), rel_tol=1e-12)
    # first maximum in raster order among equal interior values
    Cp[35, 20] = 9.0
    st2 = orc.nonorm_plane(Cp.astype(complex), d_m)
    assert int(st2[1]) == 30 * l + 41
    assert interior[30, 41] and interior[35, 20]
    # a plane with no positive interior value has peak 0 and score 0
    st3 = orc.nonorm_plane(-np.abs(Cp).astype(complex), d_m)
    assert st3[0] == 0.0 and orc.nonorm_score(st3, l) == 0.0


@pytest.mark.parametrize("theta", [90.0, 37.5, 200.0, 315.0])
def test_nonorm_center_undoes_block_rotation(orc, theta):
    """A blob at (X, Y) of a block, rotated by rotate_subIMG, correlates best at its rotated
    position; the back-rotated centre of that position is (X, Y) plus the block offset."""
    l, ov = 64, 8
    yy, xx = np.mgrid[0:l, 0:l]
    for X, Y in [(20.0, 41.0), (45.0, 15.0), (33.0, 30.0)]:
        blob = np.exp(-((xx - X) ** 2 + (yy - Y) ** 2) / (2 * 1.5 ** 2))
        rot = orc.rotate_plane(blob.astype(complex), theta).real
        pos = int(np.argmax(rot))
        for bi, bj in [(0, 0), (2, 1)]:
            cx, cy = orc.nonorm_center(pos, bi, bj, l, ov, theta)
            assert abs(cx - (bi * (l - ov) + X)) <= 1.0 and abs(cy - (bj * (l - ov) + Y)) <= 1.0, \
                (theta, X, Y, cx, cy)
    # a quarter turn is a permutation: (px, py) -> (py, l - px)
    if theta == 90.0:
        for pos in (0, 5 * l + 7, 63 * l + 1):
            cx, cy = orc.nonorm_center(pos, 0, 0, l, 0, 90.0)
            assert abs(cx - pos // l) < 1e-4 and abs(cy - (l - pos % l)) < 1e-4
