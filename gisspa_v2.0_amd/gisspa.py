"""ctypes binding of libgisspa.so (the C ABI declared in include/gisspa.h).

The search itself runs in the HIP library; this module only marshals numpy arrays across the
C ABI.  It raises if the library has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, fields

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libgisspa.so")

GS_STAGE_TEMPLATES, GS_STAGE_IMAGE, GS_STAGE_BLOCKS, GS_STAGE_ROTATED = 0, 1, 2, 3
GS_STAGE_MU, GS_STAGE_SIGMA, GS_STAGE_BEST, GS_STAGE_ARG, GS_STAGE_PLANES = 4, 5, 6, 7, 8


class Params(C.Structure):
    """gs_params: the reference config keys (include/EMReader/DataReader2.h:17-44)."""

    _fields_ = [
        ("angpix", C.c_float), ("phistep", C.c_float), ("kk", C.c_float), ("energy", C.c_float),
        ("cs", C.c_float), ("highres", C.c_float), ("lowres", C.c_float), ("diameter", C.c_float),
        ("threshold", C.c_float), ("window_size", C.c_int32), ("overlap", C.c_int32),
        ("norm_type", C.c_int32), ("invert", C.c_int32), ("phase_flip", C.c_int32),
        ("nms_radius", C.c_float), ("reference_records", C.c_int32),
    ]


class Ctf(C.Structure):
    _fields_ = [("defocus", C.c_double), ("dfdiff", C.c_double), ("dfang", C.c_double)]


class Record(C.Structure):
    _fields_ = [
        ("x", C.c_float), ("y", C.c_float), ("template_id", C.c_int32), ("psi", C.c_float),
        ("score", C.c_float), ("angle_index", C.c_int32), ("block", C.c_int32),
        ("pixel", C.c_int32), ("local_x", C.c_float), ("local_y", C.c_float),
    ]


class BatchItem(C.Structure):
    _fields_ = [
        ("image", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32), ("ctf", Ctf),
        ("templates", C.c_void_p), ("n_templates", C.c_int32), ("template_width", C.c_int32),
        ("template_height", C.c_int32), ("alt_az", C.c_void_p),
    ]


class Peak(C.Structure):
    _fields_ = [
        ("block", C.c_int32), ("pixel", C.c_int32), ("template_id", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("theta", C.c_double), ("ddf", C.c_double),
    ]


class RefineOpts(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("step", C.c_double), ("scale", C.c_double * 4)]


class Refined(C.Structure):
    _fields_ = [("peak", Peak), ("score", C.c_double), ("score0", C.c_double),
                ("grad", C.c_double * 4)]


# gs_peak as a numpy record (same layout as the C struct)
PEAK_DTYPE = np.dtype({
    "names": ["block", "pixel", "template_id", "dx", "dy", "theta", "ddf"],
    "formats": ["<i4", "<i4", "<i4", "<f8", "<f8", "<f8", "<f8"],
    "offsets": [0, 4, 8, 16, 24, 32, 40],
    "itemsize": 48,
})
REFINED_DTYPE = np.dtype({
    "names": ["peak", "score", "score0", "grad"],
    "formats": [PEAK_DTYPE, "<f8", "<f8", ("<f8", (4,))],
    "offsets": [0, 48, 56, 64],
    "itemsize": 96,
})


RECORD_DTYPE = np.dtype([
    ("x", "<f4"), ("y", "<f4"), ("template_id", "<i4"), ("psi", "<f4"), ("score", "<f4"),
    ("angle_index", "<i4"), ("block", "<i4"), ("pixel", "<i4"), ("local_x", "<f4"),
    ("local_y", "<f4"),
])

# Entry points declared in include/gisspa.h, with ctypes signatures.
_P = C.c_void_p
SIGNATURES = {
    "gs_create": (C.c_int, [C.POINTER(Params), C.c_int, C.POINTER(_P)]),
    "gs_load_templates": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, _P]),
    "gs_set_image": (C.c_int, [_P, _P, C.c_int, C.c_int, C.POINTER(Ctf)]),
    "gs_run": (C.c_int, [_P]),
    "gs_get_records": (C.c_int, [_P, C.POINTER(_P), C.POINTER(C.c_size_t)]),
    "gs_search": (C.c_int, [_P, _P, C.c_int, C.c_int, C.POINTER(Ctf), C.POINTER(_P), C.POINTER(C.c_size_t)]),
    "gs_format_record": (C.c_int, [_P, C.POINTER(Record), C.c_int, C.c_char_p, C.POINTER(Ctf), C.c_int, C.c_char_p, C.c_size_t]),
    "gs_free": (None, [_P]),
    "gs_destroy": (None, [_P]),
    "gs_last_error": (C.c_char_p, []),
    "gs_stage_bytes": (C.c_size_t, [_P, C.c_int]),
    "gs_fetch_stage": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "gs_set_limits": (C.c_int, [_P, C.c_int, C.c_int]),
    "gs_fft_c2c": (C.c_int, [C.c_int, _P, C.c_int, C.c_int, C.c_int]),
    "gs_fft2_c2c": (C.c_int, [C.c_int, _P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "gs_fft2_c2c_noalloc": (C.c_int, [C.c_int, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "gs_last_timing": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "gs_set_profiling": (C.c_int, [_P, C.c_int]),
    "gs_kernel_timing": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "gs_set_chunk": (C.c_int, [_P, C.c_int]),
    "gs_set_variant": (C.c_int, [_P, C.c_int, C.c_int]),
    "gs_copy_bandwidth": (C.c_int, [C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "gs_search_batch": (C.c_int, [_P, C.POINTER(BatchItem), C.c_int, C.POINTER(_P), C.POINTER(C.c_size_t)]),
    "gs_set_graph": (C.c_int, [_P, C.c_int]),
    "gs_graph_stats": (C.c_int, [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "gs_autotune": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "gs_select": (C.c_int, [_P, C.c_float, C.c_float, C.POINTER(_P), C.POINTER(C.c_size_t)]),
    "gs_refine_defaults": (RefineOpts, []),
    "gs_peak_from_record": (C.c_int, [_P, C.POINTER(Record), C.POINTER(Peak)]),
    "gs_score_and_grad": (C.c_int, [_P, _P, C.c_size_t, _P, _P]),
    "gs_refine": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(RefineOpts), _P]),
    "gs_refine_timing": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "gs_search_batch_refine": (C.c_int, [_P, C.POINTER(BatchItem), C.c_int, C.c_int, C.POINTER(RefineOpts),
                                         C.POINTER(_P), C.POINTER(C.c_size_t), C.POINTER(_P),
                                         C.POINTER(C.c_size_t)]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load libgisspa.so; raises OSError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} not built: run make -C gisspa_v2.0_amd")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class GisspaError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"gisspa status {status}: {msg}")
        self.status = status


def _check(st: int):
    if st != 0:
        raise GisspaError(st, load().gs_last_error().decode(errors="replace"))


@dataclass
class SearchConfig:
    """Search parameters with the reference's key names and defaults
    (include/EMReader/DataReader2.h:17-44)."""

    angpix: float = 1.0
    phistep: float = 10.0
    kk: float = 3.0
    energy: float = 300.0
    cs: float = 2.7
    highres: float = 8.0
    lowres: float = 100.0
    diameter: float = 128.0
    threshold: float = 7.0
    window_size: int = 320
    overlap: int = 0
    norm_type: int = 1
    invert: int = 0
    phase_flip: int = 0
    nms_radius: float = 64.0
    reference_records: int = 0

    def to_c(self) -> Params:
        return Params(**{f.name: getattr(self, f.name) for f in fields(self)})


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class Engine:
    """One search engine bound to a device (gs_create ... gs_destroy)."""

    def __init__(self, cfg: SearchConfig, device: int = 0):
        lib = load()
        self.cfg = cfg
        self._h = C.c_void_p()
        self._params = cfg.to_c()
        _check(lib.gs_create(C.byref(self._params), device, C.byref(self._h)))

    def close(self):
        if self._h:
            load().gs_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def load_templates(self, tmpl: np.ndarray, alt_az: np.ndarray | None = None):
        t = _f32(tmpl)
        assert t.ndim == 3
        self._tmpl = t
        aa = None if alt_az is None else _f32(alt_az).reshape(-1, 2)
        self._altaz = aa
        _check(load().gs_load_templates(self._h, t.ctypes.data, t.shape[0], t.shape[2], t.shape[1],
                                        None if aa is None else aa.ctypes.data))

    def set_image(self, image: np.ndarray, ctf=(1.5, 0.0, 0.0)):
        img = _f32(image)
        self._ctf = Ctf(*ctf)
        _check(load().gs_set_image(self._h, img.ctypes.data, img.shape[1], img.shape[0], C.byref(self._ctf)))

    def run(self):
        _check(load().gs_run(self._h))

    @staticmethod
    def _take_records(p, n) -> np.ndarray:
        try:
            buf = (C.c_char * (n.value * RECORD_DTYPE.itemsize)).from_address(p.value) if n.value else b""
            return np.frombuffer(bytes(buf), dtype=RECORD_DTYPE).copy()
        finally:
            load().gs_free(p)

    def records(self) -> np.ndarray:
        p = C.c_void_p()
        n = C.c_size_t()
        _check(load().gs_get_records(self._h, C.byref(p), C.byref(n)))
        return self._take_records(p, n)

    def select(self, threshold: float, nms_radius: float) -> np.ndarray:
        """Records of the last run at another threshold / NMS radius (gs_select)."""
        p = C.c_void_p()
        n = C.c_size_t()
        _check(load().gs_select(self._h, threshold, nms_radius, C.byref(p), C.byref(n)))
        return self._take_records(p, n)

    def peaks_from_records(self, recs) -> np.ndarray:
        out = np.zeros(len(recs), dtype=PEAK_DTYPE)
        pk = Peak()
        for i, rec in enumerate(recs):
            r = Record(*[rec[name] for name in RECORD_DTYPE.names])
            _check(load().gs_peak_from_record(self._h, C.byref(r), C.byref(pk)))
            out[i] = (pk.block, pk.pixel, pk.template_id, pk.dx, pk.dy, pk.theta, pk.ddf)
        return out

    def score_and_grad(self, peaks):
        """z (k,) and dz/d(dx, dy, theta, ddf) (k, 4) of picks (gs_score_and_grad)."""
        p = np.ascontiguousarray(peaks, dtype=PEAK_DTYPE)
        z = np.zeros(len(p))
        g = np.zeros((len(p), 4))
        _check(load().gs_score_and_grad(self._h, p.ctypes.data, len(p), z.ctypes.data, g.ctypes.data))
        return z, g

    def refine(self, peaks, iterations=None, step=None, scale=None) -> np.ndarray:
        """Gradient ascent of picks (gs_refine); returns a REFINED_DTYPE array."""
        p = np.ascontiguousarray(peaks, dtype=PEAK_DTYPE)
        o = load().gs_refine_defaults()
        if iterations is not None:
            o.iterations = iterations
        if step is not None:
            o.step = step
        if scale is not None:
            for i in range(4):
                o.scale[i] = scale[i]
        out = np.zeros(len(p), dtype=REFINED_DTYPE)
        _check(load().gs_refine(self._h, p.ctypes.data, len(p), C.byref(o), out.ctypes.data))
        return out

    def refine_timing(self) -> float:
        ms = C.c_double()
        _check(load().gs_refine_timing(self._h, C.byref(ms)))
        return ms.value

    def search(self, image, ctf=(1.5, 0.0, 0.0)) -> np.ndarray:
        self.set_image(image, ctf)
        self.run()
        return self.records()

    def set_limits(self, n_angles: int = 0, n_templates: int = 0):
        _check(load().gs_set_limits(self._h, n_angles, n_templates))

    def set_profiling(self, on: bool):
        _check(load().gs_set_profiling(self._h, 1 if on else 0))

    def set_chunk(self, tc: int):
        _check(load().gs_set_chunk(self._h, tc))

    def set_variant(self, kernel: int, variant: int):
        _check(load().gs_set_variant(self._h, kernel, variant))

    def set_graph(self, on: bool):
        _check(load().gs_set_graph(self._h, 1 if on else 0))

    def graph_stats(self):
        a, b = C.c_longlong(), C.c_longlong()
        _check(load().gs_graph_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self):
        """Returns (column-pass variant, row-pass variant, templates per launch) chosen."""
        c, r, k = C.c_int(), C.c_int(), C.c_int()
        _check(load().gs_autotune(self._h, C.byref(c), C.byref(r), C.byref(k)))
        return c.value, r.value, k.value

    def search_batch_refine(self, items, top_k: int, iterations=None, step=None, scale=None):
        """gs_search_batch_refine: returns (record arrays per item, REFINED_DTYPE arrays per item)."""
        arr, keep = self._batch_array(items)
        o = load().gs_refine_defaults()
        if iterations is not None:
            o.iterations = iterations
        if step is not None:
            o.step = step
        if scale is not None:
            for i in range(4):
                o.scale[i] = scale[i]
        counts = (C.c_size_t * len(items))()
        rcounts = (C.c_size_t * len(items))()
        p, q = C.c_void_p(), C.c_void_p()
        _check(load().gs_search_batch_refine(self._h, arr, len(items), top_k, C.byref(o), C.byref(p), counts,
                                             C.byref(q), rcounts))
        try:
            total = sum(counts)
            buf = (C.c_char * (total * RECORD_DTYPE.itemsize)).from_address(p.value) if total else b""
            allrec = np.frombuffer(bytes(buf), dtype=RECORD_DTYPE).copy()
            rtotal = sum(rcounts)
            rbuf = (C.c_char * (rtotal * REFINED_DTYPE.itemsize)).from_address(q.value) if rtotal else b""
            allref = np.frombuffer(bytes(rbuf), dtype=REFINED_DTYPE).copy()
        finally:
            load().gs_free(p)
            load().gs_free(q)
        self._ctf = Ctf(*items[-1].get("ctf", (1.5, 0.0, 0.0)))
        recs, refs, o1, o2 = [], [], 0, 0
        for c, rc in zip(counts, rcounts):
            recs.append(allrec[o1:o1 + c])
            refs.append(allref[o2:o2 + rc])
            o1 += c
            o2 += rc
        return recs, refs

    @staticmethod
    def _batch_array(items):
        keep = []
        arr = (BatchItem * len(items))()
        for i, it in enumerate(items):
            img = _f32(it["image"])
            keep.append(img)
            b = arr[i]
            b.image = img.ctypes.data
            b.height, b.width = img.shape
            b.ctf = Ctf(*it.get("ctf", (1.5, 0.0, 0.0)))
            t = it.get("templates")
            if t is not None:
                t = _f32(t)
                keep.append(t)
                b.templates = t.ctypes.data
                b.n_templates, b.template_height, b.template_width = t.shape
                aa = it.get("alt_az")
                if aa is not None:
                    aa = _f32(aa).reshape(-1, 2)
                    keep.append(aa)
                    b.alt_az = aa.ctypes.data
        return arr, keep

    def search_batch(self, items) -> list:
        """items: sequence of dicts with 'image', 'ctf' and optionally 'templates', 'alt_az'.
        Returns one record array per item (gs_search_batch)."""
        arr, keep = self._batch_array(items)
        counts = (C.c_size_t * len(items))()
        p = C.c_void_p()
        _check(load().gs_search_batch(self._h, arr, len(items), C.byref(p), counts))
        try:
            total = sum(counts)
            buf = (C.c_char * (total * RECORD_DTYPE.itemsize)).from_address(p.value) if total else b""
            allrec = np.frombuffer(bytes(buf), dtype=RECORD_DTYPE).copy()
        finally:
            load().gs_free(p)
        if items:
            self._ctf = Ctf(*items[-1].get("ctf", (1.5, 0.0, 0.0)))
        out, o = [], 0
        for c in counts:
            out.append(allrec[o:o + c])
            o += c
        return out

    def kernel_timing(self, which: int):
        avg = C.c_double()
        cnt = C.c_longlong()
        _check(load().gs_kernel_timing(self._h, which, C.byref(avg), C.byref(cnt)))
        return avg.value, cnt.value

    def timing(self):
        a, b = C.c_double(), C.c_double()
        _check(load().gs_last_timing(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def stage(self, which: int, dtype) -> np.ndarray:
        nbytes = load().gs_stage_bytes(self._h, which)
        if nbytes == 0:
            raise GisspaError(5, f"stage {which} unavailable")
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        _check(load().gs_fetch_stage(self._h, which, out.ctypes.data, nbytes))
        return out

    def format_record(self, rec, image_index: int, path: str, tile_local: bool = False) -> str:
        r = Record(*[rec[name] for name in RECORD_DTYPE.names])
        buf = C.create_string_buffer(1024)
        load().gs_format_record(self._h, C.byref(r), image_index, path.encode(), C.byref(self._ctf),
                                1 if tile_local else 0, buf, 1024)
        return buf.value.decode()


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 20, device: int = 0) -> float:
    """Measured streaming-copy rate in GB/s (read + write)."""
    g = C.c_double()
    _check(load().gs_copy_bandwidth(device, nbytes, reps, C.byref(g)))
    return g.value


def fft_c2c(x: np.ndarray, sign: int, device: int = 0) -> np.ndarray:
    """Batched 1-D complex transform along the last axis on the GPU (unnormalised)."""
    a = np.array(x, dtype=np.complex64, order="C", copy=True)
    n = a.shape[-1]
    _check(load().gs_fft_c2c(device, a.ctypes.data, n, a.size // n, sign))
    return a


def fft2_c2c(x: np.ndarray, sign: int, device: int = 0) -> np.ndarray:
    """Batched 2-D complex transform over the last two axes on the GPU (unnormalised)."""
    a = np.array(x, dtype=np.complex64, order="C", copy=True)
    ny, nx = a.shape[-2:]
    _check(load().gs_fft2_c2c(device, a.ctypes.data, ny, nx, a.size // (nx * ny), sign))
    return a


if __name__ == "__main__":
    import sys

    if "--copy-probe" in sys.argv:
        # one 1 GiB streaming copy (1 GiB read + 1 GiB write): the calibration run for the
        # FETCH_SIZE / WRITE_SIZE traffic counters
        print(f"copy {copy_bandwidth(1 << 30, 1):.1f} GB/s")
