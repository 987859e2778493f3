"""ctypes view of oracle/liboracle.so — the CPU checker used by the tests.

TEST INFRASTRUCTURE ONLY (also used by __graft_entry__.smoke and bench.py's
cpu_baseline leg).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")

_D, _F, _U32, _U64, _I, _P = C.c_double, C.c_float, C.c_uint32, C.c_uint64, C.c_int, C.c_void_p
_SIG = {
    "or_mandelbrot_d": (_F, [_U32, _D, _D]),
    "or_mandelbrot_f": (_F, [_U32, _F, _F]),
    "or_julia_d": (_F, [_U32, _D, _D, _D, _D]),
    "or_julia_f": (_F, [_U32, _F, _F, _D, _D]),
    "or_escape_d": (_U32, [_U32, _D, _D, _D, _D, _D, C.POINTER(_D), C.POINTER(_I)]),
    "or_escape_f": (_U32, [_U32, _F, _F, _F, _F, _F, C.POINTER(_F), C.POINTER(_I)]),
    "or_smooth": (_D, [_U32, _D, _D]),
    "or_splitmix64": (_U64, [_U64]),
    "or_jitter": (None, [_U64, _U64, C.POINTER(_D), C.POINTER(_D)]),
    "or_tile_point_d": (None, [_P, _U32, _U32, _U32, C.POINTER(_D), C.POINTER(_D)]),
    "or_tile_point_f": (None, [_P, _U32, _U32, _U32, C.POINTER(_F), C.POINTER(_F)]),
    "or_escape_tile_rows": (_U64, [_P, _I, _U32, _U32, _P, _P, _I]),
    "or_render_iterations": (C.c_uint64, []),
    "or_render_main": (None, [_I, _I, _D, _D, _P, C.c_int64, _U32, _U32, _D, _D, _D, _D, _U32, _F,
                              _U32]),
    "or_render_advanced": (None, [_I, _I, _D, _D, _P, C.c_int64, _U32, _U32, _D, _D, _D, _D, _U32,
                                  _F, _U32, _D, _D, _D, _D, _P, C.c_int64, _U32, _U32]),
    "or_foveation": (None, [_U32, _U32, _U32, _U32, _F, C.POINTER(_F), C.POINTER(_I)]),
    "or_compose": (None, [_P, C.c_int64, _P, _U32, _U32, _U32, _P, _U32, _F, _I]),
    "or_compose_smooth": (None, [_P, _U32, _P, _U32, _U32, _U32, _P, _U32]),
    "or_render_undersampled": (None, [_P, C.c_int64, _U32, _U32, _F, _F, _F, _F, _U32, _U32]),
    "or_default_palette": (None, [_P]),
    "or_argb_to_rgba": (_U32, [_U32]),
    "or_dd_add": (None, [_D, _D, _D, _D, C.POINTER(_D), C.POINTER(_D)]),
    "or_reference_orbit_f": (_U32, [_D, _D, _D, _D, _U32, _D, _P]),
    "or_reference_orbit_d": (_U32, [_D, _D, _D, _D, _U32, _D, _P]),
    "or_perturb_pixel_f": (_U32, [_P, _U32, _F, _F, _U32, _F, _U32, _F, C.POINTER(_F), C.POINTER(_I)]),
    "or_perturb_pixel_d": (_U32, [_P, _U32, _D, _D, _U32, _D, _U32, _D, C.POINTER(_D), C.POINTER(_I)]),
    "or_deep_offset": (None, [_U32, _U32, _U32, _U32, _D, _D, C.POINTER(_D), C.POINTER(_D)]),
    "or_deep_direct_d": (_U32, [_D, _D, _D, _D, _U32, _D, C.POINTER(_D), C.POINTER(_I)]),
    "or_deep_direct_dd": (_U32, [_D, _D, _D, _D, _D, _D, _U32, _D, C.POINTER(_D), C.POINTER(_I)]),
    "or_deep_pixels": (_U64, [_I, _P, _P, _U64, _U32, _U32, _D, _D, _D, _D, _D, _D, _U32, _D, _P,
                              _U32, _D, _D, _U32, _F, _P, _P, _P, _I]),
}

DEEP_PERTURB_F32, DEEP_DIRECT_F64, DEEP_DIRECT_DD, DEEP_PERTURB_F64 = 0, 1, 2, 3

PIXEL_INFO_DTYPE = np.dtype([("value", "<f4"), ("weight", "<f4"), ("is_reused", "u1"),
                             ("pad", "u1", (3,)), ("weight_of_new_samples", "<f4")])

_lib = None


def bind(path: str) -> C.CDLL:
    """Load an oracle build at `path` and declare the signatures of the symbols it
    exports (a build of chaos_oracle.cpp alone has no deep-zoom symbols)."""
    l = C.CDLL(path)
    for n, (r, a) in _SIG.items():
        try:
            f = getattr(l, n)
        except AttributeError:
            continue
        f.restype = r
        f.argtypes = a
    return l


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise ImportError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        _lib = bind(ORACLE_LIB)
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def escape_tile(tile, precision: int, rows=None, threads: int = 0):
    """Counts/smooth planes of a tile (tile.out_offset must be 0, pitch = width).
    Returns (counts uint32[h,w], smooth float64[h,w], iterations executed)."""
    h, w = tile.height, tile.width
    r0, r1 = (0, h) if rows is None else rows
    counts = np.zeros((h, tile.out_pitch), dtype=np.uint32)
    smooth = np.zeros((h, tile.out_pitch), dtype=np.float64)
    it = lib().or_escape_tile_rows(C.byref(tile), precision, r0, r1, _ptr(counts), _ptr(smooth),
                                   threads)
    return counts[:, :w], smooth[:, :w], int(it)


def render_iterations() -> int:
    """Iterations executed by the oracle's computeFractal since the last call (resets)."""
    return int(lib().or_render_iterations())


def render_main(precision, kind, julia_c, w, h, seg, max_iter, max_ss, flags):
    out = np.zeros((h, w), dtype=PIXEL_INFO_DTYPE)
    lib().or_render_main(precision, kind, julia_c[0], julia_c[1], _ptr(out), w * 16, w, h, *seg,
                         max_iter, max_ss, flags)
    return out


def render_advanced(precision, kind, julia_c, w, h, seg, max_iter, max_ss, flags, old_seg,
                    prev: np.ndarray, focus):
    out = np.zeros((h, w), dtype=PIXEL_INFO_DTYPE)
    prev = np.ascontiguousarray(prev)
    lib().or_render_advanced(precision, kind, julia_c[0], julia_c[1], _ptr(out), w * 16, w, h, *seg,
                             max_iter, max_ss, flags, *old_seg, _ptr(prev), w * 16, focus[0], focus[1])
    return out


def compose(pix: np.ndarray, palette: np.ndarray, max_ss: float, visualise: bool):
    h, w = pix.shape
    pix = np.ascontiguousarray(pix)
    pal = np.ascontiguousarray(palette, dtype=np.uint32)
    out = np.zeros((h, w), dtype=np.uint32)
    lib().or_compose(_ptr(pix), w * 16, _ptr(out), w, w, h, _ptr(pal), pal.size, max_ss,
                     1 if visualise else 0)
    return out


def compose_smooth(smooth: np.ndarray, palette: np.ndarray):
    h, w = smooth.shape
    smooth = np.ascontiguousarray(smooth, dtype=np.float64)
    pal = np.ascontiguousarray(palette, dtype=np.uint32)
    out = np.zeros((h, w), dtype=np.uint32)
    lib().or_compose_smooth(_ptr(smooth), w, _ptr(out), w, w, h, _ptr(pal), pal.size)
    return out


def render_undersampled(w, h, seg_f, max_iter, level):
    out = np.zeros((h, w), dtype=PIXEL_INFO_DTYPE)
    lib().or_render_undersampled(_ptr(out), w * 16, w, h, *seg_f, max_iter, level)
    return out


def default_palette() -> np.ndarray:
    out = np.zeros(1536, dtype=np.uint32)
    lib().or_default_palette(_ptr(out))
    return out


# ---------------------------------------------------------------------------
# deep zoom
# ---------------------------------------------------------------------------
def dd_add(a_hi, a_lo, b_hi, b_lo=0.0):
    hi, lo = C.c_double(), C.c_double()
    lib().or_dd_add(a_hi, a_lo, b_hi, b_lo, C.byref(hi), C.byref(lo))
    return hi.value, lo.value


def reference_orbit(cx, cy, max_iter, bailout_sq, precision=64, lo=(0.0, 0.0)):
    """Double-double reference orbit rounded to the delta type: (entries (M+1, 4), M)."""
    dt = np.float64 if precision == 64 else np.float32
    buf = np.zeros((max_iter + 1) * 4, dtype=dt)
    f = lib().or_reference_orbit_d if precision == 64 else lib().or_reference_orbit_f
    M = f(cx, lo[0], cy, lo[1], max_iter, bailout_sq, buf.ctypes.data_as(C.c_void_p))
    return buf[: (M + 1) * 4].reshape(-1, 4).copy(), M


def deep_pixels(method, px, py, desc, orbit=None, M=0, ref=(0.0, 0.0), threads=None):
    """Pixels (px, py) of the deep frame `desc` (a cu DeepDesc-like object) by `method`:
    returns (counts u32, mag f64, state i32, total iterations)."""
    px = np.ascontiguousarray(px, dtype=np.uint32)
    py = np.ascontiguousarray(py, dtype=np.uint32)
    n = px.size
    counts = np.zeros(n, np.uint32)
    mag = np.zeros(n, np.float64)
    state = np.zeros(n, np.int32)
    if orbit is None:
        orbit = np.zeros(4, np.float64)
    orbit = np.ascontiguousarray(orbit)
    threads = threads or (os.cpu_count() or 1)
    total = lib().or_deep_pixels(
        method, px.ctypes.data_as(C.c_void_p), py.ctypes.data_as(C.c_void_p), n, desc.width,
        desc.height, desc.centre_x, desc.centre_x_lo, desc.centre_y, desc.centre_y_lo, desc.pixel_w,
        desc.pixel_h, desc.max_iter, desc.bailout_sq, orbit.ctypes.data_as(C.c_void_p), M, ref[0],
        ref[1], desc.flags, desc.glitch_eps, counts.ctypes.data_as(C.c_void_p),
        mag.ctypes.data_as(C.c_void_p), state.ctypes.data_as(C.c_void_p), threads)
    return counts, mag, state, int(total)
