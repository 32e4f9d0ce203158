"""CPU checks of the data formats either side of the search and of the synthetic generator.

* MRC reading in the executable (modes 0, 1, 2, 6, either byte order; emdata.cpp:412-677 is the
  reference's reader, whose mode 0 maps unsigned bytes to v / 100 - 1.28): dimensions and
  checksums of files written here with numpy; a micrograph is the section of its LST index
  (image.cpp:9, emdata.cpp:516-524).
* EMAN HDF5 stacks (emdata.cpp:192-220: EMAN2 /MDF/images/i/image, EMAN1 root datasets "i" under
  a "num_dataset" attribute), written here through the HDF5 C library (skipped without it).
* Template generation: the closed-form projector of isotropic Gaussian blobs against numerical
  integration along the beam, rotation invariance of a centred blob, and the ZXZ Euler
  convention; the harness CTF against the oracle's CTF_AST.
"""
import os
import subprocess

import numpy as np
import pytest

import hdf5_fixture

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gisspa_v2.0_amd", "lib", "gisspa_main")


def _write_mrc(path, a, mode, big_endian=False):
    nz, ny, nx = a.shape
    dt = {0: "i1", 1: "i2", 2: "f4", 6: "u2"}[mode]
    bo = ">" if big_endian else "<"
    hdr = np.zeros(256, dtype=bo + "i4")
    hdr[0:3] = (nx, ny, nz)
    hdr[3] = mode
    hdr[23] = 64  # extended header bytes, skipped by the reader
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(b"\x00" * 64)
        f.write(np.ascontiguousarray(a, dtype=bo + dt if dt != "i1" else dt).tobytes())


@pytest.mark.parametrize("mode,big", [(0, False), (1, False), (2, False), (6, False), (2, True), (1, True)])
def test_mrc_reader_modes(tmp_path, mode, big):
    if not os.path.exists(EXE):
        pytest.skip("executable not built")
    rng = np.random.default_rng(mode + 10 * big)
    shape = (3, 17, 29)
    if mode == 0:
        a = rng.integers(-128, 128, size=shape)
    elif mode == 1:
        a = rng.integers(-30000, 30000, size=shape)
    elif mode == 6:
        a = rng.integers(0, 65535, size=shape)
    else:
        a = rng.standard_normal(shape).astype(np.float32)
    p = tmp_path / "x.mrc"
    _write_mrc(p, a, mode, big)
    out = subprocess.run([EXE, "--mrc-info", str(p)], capture_output=True, text=True, check=True).stdout.split()
    assert out[0] == "mrc" and [int(v) for v in out[1:4]] == [29, 17, 3]
    if mode == 0:  # the reference's conversion of (unsigned) bytes
        v = ((a.astype(np.int64) & 0xFF) / 100.0 - 1.28).astype(np.float32).astype(np.float64)
    else:
        v = a.astype(np.float32).astype(np.float64)
    assert float(out[4]) == pytest.approx(v.sum(), rel=1e-12, abs=1e-9)
    assert float(out[5]) == pytest.approx((v * v).sum(), rel=1e-12)
    # one section as a micrograph is read: index in range selects it, any other index section 0
    for idx, sec in ((0, 0), (2, 2), (3, 0), (-1, 0)):
        out = subprocess.run([EXE, "--image-info", str(p), str(idx)], capture_output=True, text=True,
                             check=True).stdout.split()
        assert out[0] == "image" and [int(x) for x in out[1:3]] == [29, 17]
        assert float(out[3]) == pytest.approx(v[sec].sum(), rel=1e-12, abs=1e-9)
        assert float(out[4]) == pytest.approx((v[sec] * v[sec]).sum(), rel=1e-12)


@pytest.mark.parametrize("eman1", [False, True])
def test_hdf5_stacks(tmp_path, eman1):
    if not os.path.exists(EXE):
        pytest.skip("executable not built")
    stack = np.random.default_rng(5 + eman1).standard_normal((3, 17, 29)).astype(np.float32)
    p = tmp_path / ("t1.hdf" if eman1 else "t2.hdf")
    hdf5_fixture.write(p, stack, eman1)
    # EMAN2 dims are [ny][nx]; EMAN1 as the reference reads them: nx = dims[0], ny = dims[1]
    nx, ny = (17, 29) if eman1 else (29, 17)
    v = stack.astype(np.float64)
    for i in range(3):
        out = subprocess.run([EXE, "--image-info", str(p), str(i)], capture_output=True, text=True,
                             check=True).stdout.split()
        assert out[0] == "image" and [int(x) for x in out[1:3]] == [nx, ny]
        assert float(out[3]) == pytest.approx(v[i].sum(), rel=1e-12, abs=1e-9)
        assert float(out[4]) == pytest.approx((v[i] * v[i]).sum(), rel=1e-12)
    out = subprocess.run([EXE, "--stack-info", str(p), "3"], capture_output=True, text=True,
                         check=True).stdout.split()
    assert out[0] == "stack" and [int(x) for x in out[1:4]] == [nx, ny, 3]
    assert float(out[4]) == pytest.approx(v.sum(), rel=1e-12, abs=1e-9)
    # a stack shorter than asked for is an error, not a silent short read
    r = subprocess.run([EXE, "--stack-info", str(p), "4"], capture_output=True, text=True)
    assert r.returncode != 0 and "no image 3" in r.stderr


def test_synth_mrc_round_trip(tmp_path):
    import synth

    a = np.random.default_rng(1).standard_normal((2, 5, 7)).astype(np.float32)
    synth.write_mrc(str(tmp_path / "s.mrc"), a)
    assert np.array_equal(synth.read_mrc(str(tmp_path / "s.mrc")), a)


def test_projection_matches_numerical_integration():
    """project() of off-centre blobs equals the line integral of the rotated 3-D density along z
    (up to the constant sqrt(2 pi) sigma it omits)."""
    import synth

    rng = np.random.default_rng(3)
    centres = rng.uniform(-20, 20, size=(4, 3))
    sigma, size, apix = 3.0, 48, 1.5
    euler = (37.0, 112.0, 201.0)
    got = synth.project(centres, sigma, euler, size, apix)
    R = synth.euler_matrix(*euler)
    c = centres @ R.T
    coords = (np.arange(size) - size // 2) * apix
    z = np.linspace(-60, 60, 2401)
    X, Y = np.meshgrid(coords, coords)
    ref = np.zeros((size, size))
    for cx, cy, cz in c:
        rho = np.exp(-((X[..., None] - cx) ** 2 + (Y[..., None] - cy) ** 2 + (z - cz) ** 2) / (2 * sigma ** 2))
        ref += np.trapezoid(rho, z, axis=-1)
    np.testing.assert_allclose(got * np.sqrt(2 * np.pi) * sigma, ref, rtol=1e-6, atol=1e-9)


def test_projection_rotation_invariance_and_euler_convention():
    import synth

    centred = np.zeros((1, 3))
    a = synth.project(centred, 2.5, (0, 0, 0), 33, 1.0)
    b = synth.project(centred, 2.5, (71.0, 23.0, 150.0), 33, 1.0)
    np.testing.assert_allclose(a, b, atol=1e-12)
    # ZXZ: alt = 90 about x maps (x, y, z) -> (x, -z, y); the projection keeps (x, -z)
    blob = np.array([[6.0, 0.0, 9.0]])
    p = synth.project(blob, 1.0, (90.0, 0.0, 0.0), 41, 1.0)
    iy, ix = np.unravel_index(np.argmax(p), p.shape)
    assert (ix - 20, iy - 20) == (6, -9)
    # in-plane phi = 90 about z maps (x, y) -> (-y, x)
    p = synth.project(np.array([[7.0, 0.0, 0.0]]), 1.0, (0.0, 0.0, 90.0), 41, 1.0)
    iy, ix = np.unravel_index(np.argmax(p), p.shape)
    assert (ix - 20, iy - 20) == (0, 7)


def test_grid_orientations_uniform_directions():
    import synth

    e = synth.grid_orientations(1000)
    v = np.stack([np.sin(np.deg2rad(e[:, 0])) * np.cos(np.deg2rad(e[:, 1])),
                  np.sin(np.deg2rad(e[:, 0])) * np.sin(np.deg2rad(e[:, 1])),
                  np.cos(np.deg2rad(e[:, 0]))], 1)
    assert np.all(np.abs(v.mean(0)) < 2e-3)  # no preferred direction
    np.testing.assert_allclose((v ** 2).mean(0), 1 / 3, atol=2e-3)


def test_harness_ctf_matches_oracle(orc):
    """synth.ctf_grid on the unshifted grid equals CTF_AST at the reference's index convention."""
    import synth

    apix = 1.25
    # CTF_AST takes the astigmatism angle atan2 in index units (k:294-300), which is the
    # physical angle only on square grids; the harness uses frequency units, so the astigmatic
    # case is compared on a square grid.
    for nx, ny, defocus, dfdiff, dfang in ((96, 80, 1.5, 0.0, 0.0), (96, 96, 2.2, 0.01, 35.0)):
        g = synth.ctf_grid(nx, ny, apix, defocus, dfdiff, dfang, cs=2.5)
        for y in range(0, ny, 7):
            for x in range(0, nx, 5):
                # CTF_AST reads x1 = x and y1 = (y + ny/2) mod ny (k:290-317)
                v = orc.ctf(x, (y + ny // 2) % ny, nx, ny, apix, defocus, dfdiff, dfang, cs=2.5, mode=1)
                if dfdiff == 0.0 or x <= nx // 2:
                    # (apix and cs exact in float: the oracle holds config values in float)
                    assert g[y, x] == pytest.approx(v, abs=1e-9), (x, y)
