"""ctypes binding of the FP64 oracle (oracle/lib/libgisspa_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the smoke check (to check the GPU path's
results) and by the CPU-baseline leg of bench.py (timed as the CPU reference point).  The product
path (libgisspa.so) never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GISSPA_ORACLE_LIB") or os.path.join(HERE, "lib", "libgisspa_oracle.so")


class Config(C.Structure):
    _fields_ = [
        ("angpix", C.c_float), ("phistep", C.c_float), ("kk", C.c_float), ("energy", C.c_float),
        ("cs", C.c_float), ("highres", C.c_float), ("lowres", C.c_float), ("diameter", C.c_float),
        ("threshold", C.c_float), ("window_size", C.c_int32), ("overlap", C.c_int32),
        ("norm_type", C.c_int32), ("invert", C.c_int32), ("phase_flip", C.c_int32),
        ("nms_radius", C.c_float),
    ]


PICK_DTYPE = np.dtype([
    ("x", "<f4"), ("y", "<f4"), ("template_id", "<i4"), ("psi", "<f4"), ("score", "<f4"),
    ("angle_index", "<i4"), ("block", "<i4"), ("pixel", "<i4"), ("local_x", "<f4"),
    ("local_y", "<f4"),
])

_P = C.c_void_p
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_fft1d": (C.c_int, [_P, C.c_int, C.c_int]),
        "orc_fft2d": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
        "orc_ctf": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_double] * 7 + [C.c_int]),
        "orc_radial_bin": (C.c_int, [C.c_int, C.c_int]),
        "orc_block_offsets": (C.c_int, [C.c_int, C.c_int, C.c_int, _P, C.c_int]),
        "orc_tile_geometry": (C.c_int, [C.c_int, C.c_int, _P, _P, C.c_int]),
        "orc_angles": (C.c_int, [C.c_float, _P, C.c_int]),
        "orc_prep_templates": (C.c_int, [C.POINTER(Config), _P, C.c_int, C.c_int, C.c_int, _P, _P]),
        "orc_clip_image": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
        "orc_prep_image": (C.c_int, [C.POINTER(Config), _P, C.c_int, C.c_int, _P, _P]),
        "orc_rotate_templates": (C.c_int, [C.POINTER(Config), _P, C.c_int, C.c_float, _P]),
        "orc_search_norm": (C.c_int, [C.POINTER(Config), _P, C.c_int, C.c_int, _P, C.c_int, C.c_int,
                                      C.c_int, _P, _P, C.c_int, _P, _P, _P, _P, C.c_int, C.c_int]),
        "orc_search_nonorm": (C.c_int, [C.POINTER(Config), _P, C.c_int, C.c_int, _P, C.c_int, C.c_int,
                                        C.c_int, _P, _P, C.c_int, _P, _P, C.c_int]),
        "orc_search_norm_blocks": (C.c_int, [C.POINTER(Config), _P, C.c_int, C.c_int, _P, C.c_int, C.c_int,
                                             C.c_int, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P,
                                             _P, _P]),
        "orc_search_norm_records": (C.c_int, [C.POINTER(Config), _P, C.c_int, C.c_int, _P, C.c_int, C.c_int,
                                              C.c_int, _P, _P, C.c_int, C.c_int, C.c_int, _P, C.c_int]),
        "orc_whiten_template": (C.c_int, [_P, C.c_int]),
        "orc_weight_template": (C.c_int, [C.POINTER(Config), _P, _P, C.c_int]),
        "orc_filter_image_spectrum": (C.c_int, [C.POINTER(Config), _P, _P, C.c_int, C.c_int]),
        "orc_rotate_plane": (C.c_int, [_P, _P, C.c_int, C.c_float]),
        "orc_nonorm_plane": (C.c_int, [_P, C.c_int, C.c_float, _P]),
        "orc_nonorm_score": (C.c_double, [_P, C.c_int]),
        "orc_nonorm_center": (None, [C.c_longlong, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, _P, _P]),
        "orc_nms": (C.c_int, [_P, C.c_int, C.c_float]),
        "orc_correlate": (C.c_int, [_P, _P, C.c_int, _P]),
        "orc_num_threads": (C.c_int, []),
        "orc_set_num_threads": (None, [C.c_int]),
        "orc_refine_create": (_P, [C.POINTER(Config), _P, C.c_int, C.c_int, _P, C.c_int, C.c_int,
                                   C.c_int, _P]),
        "orc_refine_free": (None, [_P]),
        "orc_refine_geometry": (None, [_P, C.c_int]),
        "orc_refine_eval": (C.c_int, [_P, _P, C.c_int, _P, _P, _P, _P, _P]),
        "orc_refine_run": (C.c_int, [_P, _P, C.c_int, _P, _P, C.c_int, C.c_double, _P, _P, _P, _P]),
        "orc_refine_plane": (C.c_int, [_P, C.c_int, C.c_int, C.c_double, C.c_double, _P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def config(cfg) -> Config:
    """Oracle config from any object with the gs_params field names."""
    return Config(*[getattr(cfg, f[0]) for f in Config._fields_])


def _ok(st):
    if st < 0:
        raise RuntimeError("oracle call failed")
    return st


def fft1d(x: np.ndarray, sign: int) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    _ok(load().orc_fft1d(a.ctypes.data, a.shape[-1], sign))
    return a


def fft2d(x: np.ndarray, sign: int) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    _ok(load().orc_fft2d(a.ctypes.data, a.shape[0], a.shape[1], sign))
    return a


def ctf(x, y, nx, ny, apix, defocus, dfdiff, dfang, energy=300.0, cs=2.7, ampconst=0.07, mode=1):
    return load().orc_ctf(x, y, nx, ny, apix, defocus, dfdiff, dfang, energy, cs, ampconst, mode)


def radial_bin(dx, dy):
    return load().orc_radial_bin(dx, dy)


def block_offsets(length, l, overlap):
    out = np.zeros(4096, dtype=np.int32)
    n = load().orc_block_offsets(length, l, overlap, out.ctypes.data, out.size)
    return out[:n].copy()


def tile_geometry(W, H):
    ulx = np.zeros(4096, dtype=np.int32)
    uly = np.zeros(4096, dtype=np.int32)
    n = load().orc_tile_geometry(W, H, ulx.ctypes.data, uly.ctypes.data, 4096)
    return ulx[:n].copy(), uly[:n].copy()


def angles(phistep):
    out = np.zeros(100000, dtype=np.float32)
    n = load().orc_angles(phistep, out.ctypes.data, out.size)
    return out[:n].copy()


def prep_templates(cfg, tmpl, ctf3):
    t = np.ascontiguousarray(tmpl, dtype=np.float32)
    l = cfg.window_size
    out = np.zeros((t.shape[0], l, l), dtype=np.complex128)
    c = config(cfg)
    d = np.ascontiguousarray(ctf3, dtype=np.float64)
    _ok(load().orc_prep_templates(C.byref(c), t.ctypes.data, t.shape[0], t.shape[2], t.shape[1],
                                  d.ctypes.data, out.ctypes.data))
    return out


def clip_image(img, invert=0):
    a = np.ascontiguousarray(img, dtype=np.float32).copy()
    load().orc_clip_image(a.ctypes.data, a.shape[1], a.shape[0], invert)
    return a


def prep_image(cfg, img, ctf3):
    a = np.ascontiguousarray(img, dtype=np.float32)
    out = np.zeros(a.shape, dtype=np.float64)
    c = config(cfg)
    d = np.ascontiguousarray(ctf3, dtype=np.float64)
    _ok(load().orc_prep_image(C.byref(c), a.ctypes.data, a.shape[1], a.shape[0], d.ctypes.data,
                              out.ctypes.data))
    return out


def rotate_templates(cfg, prepped, theta):
    p = np.ascontiguousarray(prepped, dtype=np.complex128)
    out = np.zeros_like(p)
    c = config(cfg)
    _ok(load().orc_rotate_templates(C.byref(c), p.ctypes.data, p.shape[0], theta, out.ctypes.data))
    return out


def search_norm(cfg, image, tmpl, ctf3, want_stats=False, n_angles=0, n_templates=0, max_picks=1 << 20):
    img = np.ascontiguousarray(image, dtype=np.float32)
    t = np.ascontiguousarray(tmpl, dtype=np.float32)
    c = config(cfg)
    d = np.ascontiguousarray(ctf3, dtype=np.float64)
    picks = np.zeros(max_picks, dtype=PICK_DTYPE)
    l = cfg.window_size
    nb = len(block_offsets(img.shape[1], l, cfg.overlap)) * len(block_offsets(img.shape[0], l, cfg.overlap))
    stats = None
    ptrs = [None] * 4
    if want_stats:
        stats = dict(mu=np.zeros(nb * l * l), sigma=np.zeros(nb * l * l), best=np.zeros(nb * l * l),
                     arg=np.zeros(nb * l * l, dtype=np.int32))
        ptrs = [stats[k].ctypes.data for k in ("mu", "sigma", "best", "arg")]
    n = _ok(load().orc_search_norm(C.byref(c), img.ctypes.data, img.shape[1], img.shape[0], t.ctypes.data,
                                   t.shape[0], t.shape[2], t.shape[1], d.ctypes.data, picks.ctypes.data,
                                   max_picks, *ptrs, n_angles, n_templates))
    return picks[:min(n, max_picks)].copy(), stats


REFREC_DTYPE = np.dtype([
    ("x", "<f4"), ("y", "<f4"), ("template_id", "<i4"), ("angle_index", "<i4"), ("block", "<i4"),
    ("chunk", "<i4"), ("pixel", "<i4"), ("score", "<f4"), ("margin", "<f8"),
])


def search_norm_blocks(cfg, image, tmpl, ctf3, blocks, n_angles=0, n_templates=0):
    """norm_type=1 statistics of the listed blocks: dict of mu, sigma, best, second (float64)
    and arg, arg2 (int32), each [len(blocks)][l*l]."""
    img = np.ascontiguousarray(image, dtype=np.float32)
    t = np.ascontiguousarray(tmpl, dtype=np.float32)
    c = config(cfg)
    d = np.ascontiguousarray(ctf3, dtype=np.float64)
    bl = np.ascontiguousarray(blocks, dtype=np.int32)
    L2 = cfg.window_size ** 2
    out = {k: np.zeros((bl.size, L2)) for k in ("mu", "sigma", "best", "second")}
    out.update({k: np.zeros((bl.size, L2), dtype=np.int32) for k in ("arg", "arg2")})
    _ok(load().orc_search_norm_blocks(C.byref(c), img.ctypes.data, img.shape[1], img.shape[0], t.ctypes.data,
                                      t.shape[0], t.shape[2], t.shape[1], d.ctypes.data, bl.ctypes.data, bl.size,
                                      n_angles, n_templates,
                                      *[out[k].ctypes.data for k in ("mu", "sigma", "best", "second", "arg",
                                                                     "arg2")]))
    return out


def search_norm_records(cfg, image, tmpl, ctf3, blocks=None, n_angles=0, n_templates=0, max_out=1 << 21):
    """The reference's own norm_type=1 records (REFREC_DTYPE), in its order."""
    img = np.ascontiguousarray(image, dtype=np.float32)
    t = np.ascontiguousarray(tmpl, dtype=np.float32)
    c = config(cfg)
    d = np.ascontiguousarray(ctf3, dtype=np.float64)
    bl = np.ascontiguousarray([] if blocks is None else blocks, dtype=np.int32)
    recs = np.zeros(max_out, dtype=REFREC_DTYPE)
    n = _ok(load().orc_search_norm_records(C.byref(c), img.ctypes.data, img.shape[1], img.shape[0],
                                           t.ctypes.data, t.shape[0], t.shape[2], t.shape[1], d.ctypes.data,
                                           bl.ctypes.data if bl.size else None, bl.size, n_angles, n_templates,
                                           recs.ctypes.data, max_out))
    if n > max_out:
        raise RuntimeError(f"{n} records exceed max_out={max_out}")
    return recs[:n].copy()


def whiten_template(F):
    a = np.ascontiguousarray(F, dtype=np.complex128).copy()
    _ok(load().orc_whiten_template(a.ctypes.data, a.shape[0]))
    return a


def weight_template(cfg, F, ctf3):
    a = np.ascontiguousarray(F, dtype=np.complex128).copy()
    c = config(cfg)
    d = np.ascontiguousarray(ctf3, dtype=np.float64)
    _ok(load().orc_weight_template(C.byref(c), d.ctypes.data, a.ctypes.data, a.shape[0]))
    return a


def filter_image_spectrum(cfg, F, ctf3):
    a = np.ascontiguousarray(F, dtype=np.complex128).copy()
    c = config(cfg)
    d = np.ascontiguousarray(ctf3, dtype=np.float64)
    _ok(load().orc_filter_image_spectrum(C.byref(c), d.ctypes.data, a.ctypes.data, a.shape[1], a.shape[0]))
    return a


def rotate_plane(P, theta):
    a = np.ascontiguousarray(P, dtype=np.complex128)
    out = np.zeros_like(a)
    _ok(load().orc_rotate_plane(a.ctypes.data, out.ctypes.data, a.shape[0], theta))
    return out


def nonorm_plane(Cplane, d_m):
    a = np.ascontiguousarray(Cplane, dtype=np.complex128)
    st = np.zeros(4)
    _ok(load().orc_nonorm_plane(a.ctypes.data, a.shape[0], d_m, st.ctypes.data))
    return st


def nonorm_score(stats4, l):
    st = np.ascontiguousarray(stats4, dtype=np.float64)
    return load().orc_nonorm_score(st.ctypes.data, l)


def nonorm_center(pos, bi, bj, l, overlap, theta):
    cx, cy = C.c_float(), C.c_float()
    load().orc_nonorm_center(int(pos), bi, bj, l, overlap, theta, C.byref(cx), C.byref(cy))
    return cx.value, cy.value


def search_nonorm(cfg, image, tmpl, ctf3, max_out=1 << 18):
    img = np.ascontiguousarray(image, dtype=np.float32)
    t = np.ascontiguousarray(tmpl, dtype=np.float32)
    c = config(cfg)
    d = np.ascontiguousarray(ctf3, dtype=np.float64)
    recs = np.zeros(max_out, dtype=PICK_DTYPE)
    picks = np.zeros(max_out, dtype=PICK_DTYPE)
    nrec = C.c_int()
    n = _ok(load().orc_search_nonorm(C.byref(c), img.ctypes.data, img.shape[1], img.shape[0], t.ctypes.data,
                                     t.shape[0], t.shape[2], t.shape[1], d.ctypes.data, recs.ctypes.data,
                                     max_out, C.byref(nrec), picks.ctypes.data, max_out))
    return picks[:n].copy(), recs[:nrec.value].copy()


def nms(cands: np.ndarray, radius: float) -> np.ndarray:
    a = np.ascontiguousarray(cands, dtype=PICK_DTYPE).copy()
    n = load().orc_nms(a.ctypes.data, a.size, radius)
    return a[:n].copy()


def correlate(I: np.ndarray, T: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(I, dtype=np.complex128)
    b = np.ascontiguousarray(T, dtype=np.complex128)
    out = np.zeros_like(a)
    _ok(load().orc_correlate(a.ctypes.data, b.ctypes.data, a.shape[0], out.ctypes.data))
    return out


def num_threads():
    return load().orc_num_threads()


def set_num_threads(n):
    load().orc_set_num_threads(n)


PEAK_DTYPE = np.dtype({
    "names": ["block", "pixel", "template_id", "dx", "dy", "theta", "ddf"],
    "formats": ["<i4", "<i4", "<i4", "<f8", "<f8", "<f8", "<f8"],
    "offsets": [0, 4, 8, 16, 24, 32, 40],
    "itemsize": 48,
})


class Refiner:
    """FP64 adjoint refinement context of one micrograph (orc_refine_*)."""

    def __init__(self, cfg, image, tmpl, ctf3):
        img = np.ascontiguousarray(image, dtype=np.float32)
        t = np.ascontiguousarray(tmpl, dtype=np.float32)
        self._c = config(cfg)
        d = np.ascontiguousarray(ctf3, dtype=np.float64)
        self.l = cfg.window_size
        self._h = load().orc_refine_create(C.byref(self._c), img.ctypes.data, img.shape[1], img.shape[0],
                                           t.ctypes.data, t.shape[0], t.shape[2], t.shape[1], d.ctypes.data)
        if not self._h:
            raise RuntimeError("orc_refine_create failed")

    def close(self):
        if self._h:
            load().orc_refine_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def float_geometry(self, on: bool):
        load().orc_refine_geometry(self._h, 1 if on else 0)

    def eval(self, peaks, mu, sigma):
        """Returns z (k,), grad (k, 4) per (px, px, deg, Angstrom), c (k,)."""
        p = np.ascontiguousarray(peaks, dtype=PEAK_DTYPE)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        sg = np.ascontiguousarray(sigma, dtype=np.float64)
        z = np.zeros(len(p))
        g = np.zeros((len(p), 4))
        c = np.zeros(len(p))
        _ok(load().orc_refine_eval(self._h, p.ctypes.data, len(p), mu.ctypes.data, sg.ctypes.data,
                                   z.ctypes.data, g.ctypes.data, c.ctypes.data))
        return z, g, c

    def run(self, peaks, mu, sigma, iters=20, step=0.1, scale=(1.0, 1.0, 1.0, 1.0)):
        """Returns refined peaks, final z, final grad, starting z."""
        p = np.ascontiguousarray(peaks, dtype=PEAK_DTYPE).copy()
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        sg = np.ascontiguousarray(sigma, dtype=np.float64)
        sc = np.ascontiguousarray(scale, dtype=np.float64)
        z = np.zeros(len(p))
        z0 = np.zeros(len(p))
        g = np.zeros((len(p), 4))
        _ok(load().orc_refine_run(self._h, p.ctypes.data, len(p), mu.ctypes.data, sg.ctypes.data, iters,
                                  step, sc.ctypes.data, z.ctypes.data, g.ctypes.data, z0.ctypes.data))
        return p, z, g, z0

    def plane(self, block, t, theta, ddf=0.0):
        out = np.zeros(self.l * self.l)
        _ok(load().orc_refine_plane(self._h, block, t, theta, ddf, out.ctypes.data))
        return out.reshape(self.l, self.l)
