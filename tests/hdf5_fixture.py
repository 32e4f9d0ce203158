"""Writes EMAN HDF5 image stacks through the HDF5 C library (ctypes), for format tests; skips
the calling test when no libhdf5 can be loaded."""
import ctypes

import numpy as np
import pytest


def _lib():
    for name in ("libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"):
        try:
            h = ctypes.CDLL(name)
            break
        except OSError:
            h = None
    if h is None:
        pytest.skip("no HDF5 library to write the fixture")
    hid = ctypes.c_int64
    h.H5open()
    h.H5Fcreate.restype = hid
    h.H5Fcreate.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid, hid]
    h.H5Gcreate2.restype = hid
    h.H5Gcreate2.argtypes = [hid, ctypes.c_char_p, hid, hid, hid]
    h.H5Screate_simple.restype = hid
    h.H5Screate_simple.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_void_p]
    h.H5Screate.restype = hid
    h.H5Screate.argtypes = [ctypes.c_int]
    h.H5Dcreate2.restype = hid
    h.H5Dcreate2.argtypes = [hid, ctypes.c_char_p, hid, hid, hid, hid, hid]
    h.H5Dwrite.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
    h.H5Acreate2.restype = hid
    h.H5Acreate2.argtypes = [hid, ctypes.c_char_p, hid, hid, hid, hid]
    h.H5Awrite.argtypes = [hid, hid, ctypes.c_void_p]
    for f in ("H5Dclose", "H5Sclose", "H5Gclose", "H5Fclose", "H5Aclose"):
        getattr(h, f).argtypes = [hid]
    return h, hid


def write(path, stack, eman1):
    """stack [n][a][b] float32 as an EMAN2 file (image i = /MDF/images/i/image, dims [a][b]) or an
    EMAN1 file (root attribute num_dataset = n, image i = root dataset "i")."""
    h, hid = _lib()
    ftype = hid.in_dll(h, "H5T_IEEE_F32LE_g").value
    mfloat = hid.in_dll(h, "H5T_NATIVE_FLOAT_g").value
    f = h.H5Fcreate(str(path).encode(), 2, 0, 0)  # truncate
    assert f >= 0
    groups = []
    if eman1:
        sp = h.H5Screate(0)  # scalar
        at = h.H5Acreate2(f, b"num_dataset", hid.in_dll(h, "H5T_STD_I32LE_g").value, sp, 0, 0)
        n = ctypes.c_int(len(stack))
        h.H5Awrite(at, hid.in_dll(h, "H5T_NATIVE_INT_g").value, ctypes.byref(n))
        h.H5Aclose(at)
        h.H5Sclose(sp)
    else:
        groups = [h.H5Gcreate2(f, b"/MDF", 0, 0, 0), h.H5Gcreate2(f, b"/MDF/images", 0, 0, 0)]
    for i, a in enumerate(stack):
        a = np.ascontiguousarray(a, dtype=np.float32)
        dims = (ctypes.c_ulonglong * 2)(*a.shape)
        sp = h.H5Screate_simple(2, dims, None)
        if eman1:
            d = h.H5Dcreate2(f, str(i).encode(), ftype, sp, 0, 0, 0)
        else:
            g = h.H5Gcreate2(f, f"/MDF/images/{i}".encode(), 0, 0, 0)
            d = h.H5Dcreate2(g, b"image", ftype, sp, 0, 0, 0)
        assert d >= 0
        assert h.H5Dwrite(d, mfloat, 0, 0, 0, a.ctypes.data) >= 0
        h.H5Dclose(d)
        h.H5Sclose(sp)
        if not eman1:
            h.H5Gclose(g)
    for g in reversed(groups):
        h.H5Gclose(g)
    h.H5Fclose(f)
