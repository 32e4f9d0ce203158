"""gs_fft2_c2c_noalloc: the 2-D transform of gs_fft2_c2c on caller-owned device memory with its
work buffer in the library's static device array, checked against gs_fft2_c2c (and numpy) on
consecutive calls of decreasing size, plus its size limit.

The test's own device buffers and stream come from the HIP runtime that libgisspa.so links
(through ctypes): torch bundles a second HIP runtime, which cannot share the device with it in
one process."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fn(gs):
    fn = gs.load().gs_fft2_c2c_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    return fn


class Hip:
    """The few HIP runtime calls the test needs, from the runtime the library already loaded."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamSynchronize",
                     "hipStreamDestroy", "hipSetDevice"):
            getattr(self.lib, name).restype = C.c_int

    def ok(self, st):
        assert st == 0, st

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.lib.hipMalloc(C.byref(p), C.c_size_t(nbytes)))
        return p

    def h2d(self, dst, a):
        self.ok(self.lib.hipMemcpy(dst, C.c_void_p(a.ctypes.data), C.c_size_t(a.nbytes), 1))

    def d2h(self, a, src):
        self.ok(self.lib.hipMemcpy(C.c_void_p(a.ctypes.data), src, C.c_size_t(a.nbytes), 2))


def test_noalloc_transform_matches_allocating_one(gs):
    fn = _fn(gs)
    hip = Hip()
    hip.ok(hip.lib.hipSetDevice(0))
    stream = C.c_void_p()
    hip.ok(hip.lib.hipStreamCreate(C.byref(stream)))
    rng = np.random.default_rng(17)
    # the work buffer keeps the previous call's contents: each call is smaller than the last
    for batch, ny, nx, sign in [(2, 320, 320, -1), (3, 120, 144, +1), (2, 36, 50, -1)]:
        x = (rng.standard_normal((batch, ny, nx)) + 1j * rng.standard_normal((batch, ny, nx))).astype(np.complex64)
        ref = gs.fft2_c2c(x, sign)
        d = hip.malloc(x.nbytes)
        try:
            hip.h2d(d, x)
            st = fn(0, d, ny, nx, batch, sign, stream)
            assert st == 0, gs.load().gs_last_error()
            hip.ok(hip.lib.hipStreamSynchronize(stream))
            got = np.empty_like(x)
            hip.d2h(got, d)
        finally:
            hip.lib.hipFree(d)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-6, (batch, ny, nx)
        xd = x.astype(np.complex128)
        npy = np.fft.fft2(xd) if sign < 0 else np.fft.ifft2(xd) * ny * nx
        assert np.linalg.norm(got - npy) / np.linalg.norm(npy) < 1e-6
    hip.lib.hipStreamDestroy(stream)


def test_noalloc_rejects_sizes_beyond_its_buffer(gs):
    fn = _fn(gs)
    hip = Hip()
    d = hip.malloc(256)
    try:
        st = fn(0, d, 4096, 4096, 1, -1, None)
    finally:
        hip.lib.hipFree(d)
    assert st == 1  # GS_ERR_INVALID, before any work is queued
    assert b"exceed the work buffer" in gs.load().gs_last_error()
