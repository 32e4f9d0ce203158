"""Seeded synthetic cryo-EM inputs for tests and benchmarks.

A density of isotropic 3-D Gaussian blobs projects in closed form to a sum of 2-D Gaussians of
the same width, so templates and particles are rendered exactly without interpolation.
Particles are projections at template orientations with a random in-plane angle, filtered by
the signed CTF of the micrograph (the reference's CTF_AST formula, src/kernels.cu:290-317) and
added to N(0,1) noise at the requested SNR = var(signal inside the particle discs) / var(noise).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

PI_REF = 3.1415926535


def blob_density(seed: int, n_blobs: int = 120, sigma: float = 3.0, radius: float = 64.0):
    """Blob centres (Angstrom) uniform in a ball of `radius`, and their common width."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n_blobs, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius * rng.random(n_blobs) ** (1.0 / 3.0)
    return d * r[:, None], sigma


def euler_matrix(alt: float, az: float, phi: float) -> np.ndarray:
    """ZXZ rotation Rz(phi) Rx(alt) Rz(az), angles in degrees."""
    a, b, c = np.deg2rad([az, alt, phi])

    def rz(t):
        return np.array([[math.cos(t), -math.sin(t), 0], [math.sin(t), math.cos(t), 0], [0, 0, 1]])

    def rx(t):
        return np.array([[1, 0, 0], [0, math.cos(t), -math.sin(t)], [0, math.sin(t), math.cos(t)]])

    return rz(c) @ rx(b) @ rz(a)


def random_orientations(n: int, rng) -> np.ndarray:
    """(alt, az, psi) uniform on SO(3)."""
    alt = np.rad2deg(np.arccos(1 - 2 * rng.random(n)))
    az = 360.0 * rng.random(n)
    psi = 360.0 * rng.random(n)
    return np.stack([alt, az, psi], axis=1)


def grid_orientations(n: int) -> np.ndarray:
    """(alt, az, 0) on a Fibonacci sphere: a near-uniform grid of projection directions."""
    k = np.arange(n) + 0.5
    alt = np.rad2deg(np.arccos(1 - 2 * k / n))
    az = np.mod(np.rad2deg(np.pi * (1 + 5 ** 0.5) * k), 360.0)
    return np.stack([alt, az, np.zeros(n)], axis=1)


def project(centres: np.ndarray, sigma: float, euler, size: int, apix: float) -> np.ndarray:
    """Projection of the blob density at orientation `euler` onto a size x size grid."""
    R = euler_matrix(*euler)
    c = centres @ R.T
    coords = (np.arange(size) - size // 2) * apix
    gx = np.exp(-((coords[None, :] - c[:, 0:1]) ** 2) / (2 * sigma ** 2))
    gy = np.exp(-((coords[None, :] - c[:, 1:2]) ** 2) / (2 * sigma ** 2))
    return (gy.T @ gx).astype(np.float64)


def ctf_grid(nx: int, ny: int, apix: float, defocus: float, dfdiff: float, dfang: float,
             energy: float = 300.0, cs: float = 2.7, ampconst: float = 0.07) -> np.ndarray:
    """Signed CTF on the unshifted FFT grid (ny, nx), in the reference's convention."""
    lam = 12.2639 / math.sqrt(energy * 1000.0 + 0.97845 * energy * energy)
    fx = np.fft.fftfreq(nx)[None, :] / apix
    fy = np.fft.fftfreq(ny)[:, None] / apix
    s2 = fx ** 2 + fy ** 2
    ag = np.arctan2(fy, np.abs(fx) + 0 * fy)
    dfu, dfv = -defocus + dfdiff, -defocus - dfdiff
    df = 0.5 * (dfu + dfv + 2 * dfdiff * np.cos(2 * (dfang * PI_REF / 180 - ag)))
    gamma = -2 * PI_REF * (cs * 2.5e6 * lam ** 3 * s2 ** 2 + df * 5000.0 * lam * s2)
    return math.sqrt(1 - ampconst ** 2) * np.sin(gamma) + ampconst * np.cos(gamma)


@dataclass
class Dataset:
    image: np.ndarray          # (H, W) float32
    templates: np.ndarray      # (n, w, w) float32
    eulers: np.ndarray         # (n, 3) alt, az, psi of the templates
    positions: np.ndarray      # (k, 2) particle centres x, y
    particle_templates: np.ndarray  # (k,) template index of each particle
    particle_psi: np.ndarray   # (k,) in-plane angle of each particle (degrees)
    ctf: tuple                 # (defocus um, dfdiff um, dfang deg)


def place_positions(rng, W, H, k, margin, min_dist):
    pts = []
    tries = 0
    while len(pts) < k and tries < 200000:
        tries += 1
        p = np.array([rng.uniform(margin, W - margin), rng.uniform(margin, H - margin)])
        if all(np.hypot(*(p - q)) >= min_dist for q in pts):
            pts.append(p)
    return np.round(np.array(pts)).astype(np.int64)


def make_dataset(seed: int, W: int, H: int, n_templates: int, tsize: int, n_particles: int,
                 snr: float, defocus: float = 1.5, dfdiff: float = 0.0, dfang: float = 0.0,
                 apix: float = 1.0, grid: bool = False, diameter: float = 128.0,
                 noise: bool = True) -> Dataset:
    rng = np.random.default_rng(seed)
    centres, sigma = blob_density(seed + 7919)
    eul = grid_orientations(n_templates) if grid else random_orientations(n_templates, rng)
    eul[:, 2] = 0.0
    templates = np.stack([project(centres, sigma, e, tsize, apix) for e in eul]).astype(np.float32)
    pos = place_positions(rng, W, H, n_particles, margin=diameter + tsize // 2, min_dist=diameter)
    ptmpl = rng.integers(0, n_templates, size=len(pos))
    ppsi = 360.0 * rng.random(len(pos))
    sig = np.zeros((H, W))
    half = tsize // 2
    for (x, y), t, psi in zip(pos, ptmpl, ppsi):
        e = (eul[t, 0], eul[t, 1], psi)
        patch = project(centres, sigma, e, tsize, apix)
        sig[y - half:y - half + tsize, x - half:x - half + tsize] += patch
    F = np.fft.fft2(sig) * ctf_grid(W, H, apix, defocus, dfdiff, dfang)
    sig = np.real(np.fft.ifft2(F))
    if len(pos):
        yy, xx = np.mgrid[0:H, 0:W]
        disc = np.zeros((H, W), dtype=bool)
        r = diameter / 2
        for x, y in pos:
            y0, y1 = max(0, int(y - r)), min(H, int(y + r) + 1)
            x0, x1 = max(0, int(x - r)), min(W, int(x + r) + 1)
            disc[y0:y1, x0:x1] |= (xx[y0:y1, x0:x1] - x) ** 2 + (yy[y0:y1, x0:x1] - y) ** 2 <= r * r
        v = sig[disc].var()
        if v > 0:
            sig *= math.sqrt(snr / v)
    img = sig + (rng.normal(size=(H, W)) if noise else 0.0)
    return Dataset(img.astype(np.float32), templates, eul, pos, ptmpl, ppsi, (defocus, dfdiff, dfang))


def write_mrc(path: str, data: np.ndarray, apix: float = 1.0):
    """MRC2014 mode 2 (float32) writer: one image (ny, nx) or a stack (nz, ny, nx)."""
    a = np.ascontiguousarray(data, dtype=np.float32)
    if a.ndim == 2:
        a = a[None]
    nz, ny, nx = a.shape
    hdr = np.zeros(256, dtype=np.int32)
    hdr[0:3] = (nx, ny, nz)
    hdr[3] = 2
    hdr[7:10] = (nx, ny, nz)
    hv = hdr.view(np.float32)
    hv[10:13] = (nx * apix, ny * apix, nz * apix)
    hv[13:16] = 90.0
    hdr[16:19] = (1, 2, 3)
    hv[19], hv[20], hv[21] = a.min(), a.max(), a.mean()
    hdr[52] = int.from_bytes(b"MAP ", "little")
    hdr[53] = 0x00004144
    hdr[55] = 0
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(a.tobytes())


def read_mrc(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(1024), dtype=np.int32)
        nx, ny, nz, mode = hdr[0], hdr[1], hdr[2], hdr[3]
        f.seek(1024 + int(hdr[23]))
        dt = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16}[int(mode)]
        a = np.frombuffer(f.read(), dtype=dt, count=int(nx) * int(ny) * int(nz))
    return a.reshape(nz, ny, nx).astype(np.float32)
