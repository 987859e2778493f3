"""ctypes binding of include/chaosultra.h (the C ABI of the render path).

The library is loaded from this package's lib/ directory; a missing library is an
error (there is no fallback).  Status codes become the Python counterparts of the
reference's exception classes.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libchaosultra_hip.so")
KERNELS_DIR = os.path.join(PKG, "kernels")

# ---- status codes ---------------------------------------------------------
CU_OK = 0
CU_ERR_ILLEGAL_ARGUMENT = 1
CU_ERR_INITIALIZATION = 2
CU_ERR_RENDERER = 3
CU_ERR_ILLEGAL_STATE = 4
CU_ERR_DEVICE = 5


class ChaosError(RuntimeError):
    status = CU_ERR_DEVICE


class IllegalArgumentError(ChaosError, ValueError):
    """java.lang.IllegalArgumentException"""
    status = CU_ERR_ILLEGAL_ARGUMENT


class CudaInitializationError(ChaosError):
    """cudarenderer.CudaInitializationException"""
    status = CU_ERR_INITIALIZATION


class FractalRendererError(ChaosError):
    """rendering.FractalRendererException"""
    status = CU_ERR_RENDERER


class IllegalStateError(ChaosError):
    """java.lang.IllegalStateException"""
    status = CU_ERR_ILLEGAL_STATE


class DeviceError(ChaosError):
    """jcuda.CudaException"""
    status = CU_ERR_DEVICE


_ERRORS = {c.status: c for c in (IllegalArgumentError, CudaInitializationError,
                                 FractalRendererError, IllegalStateError, DeviceError)}

# ---- enums ----------------------------------------------------------------
STATE_NOT_INITIALIZED, STATE_READY_TO_RENDER = 0, 1
PRECISION_SINGLE, PRECISION_DOUBLE, PRECISION_TOO_BIG = 0, 1, 2
FLAG_ADAPTIVE_SS = 1 << 0
FLAG_VISUALISE_SAMPLE_COUNT = 1 << 1
FLAG_FOVEATION = 1 << 2
FLAG_SAMPLE_REUSE = 1 << 3
FLAG_ZOOMING = 1 << 4
FLAG_ZOOMING_IN = 1 << 5
FRACTAL_MANDELBROT, FRACTAL_JULIA = 0, 1
SAMPLE_CENTRE, SAMPLE_JITTER, SAMPLE_CORNER = 0, 1, 2
ITER_SLOTS = 1024  # CU_ITER_SLOTS


class RenderingModel(C.Structure):
    """cu_rendering_model — RenderingModel as a struct."""
    _fields_ = [
        ("left_bottom_x", C.c_double), ("left_bottom_y", C.c_double),
        ("right_top_x", C.c_double), ("right_top_y", C.c_double),
        ("max_iterations", C.c_int32),
        ("max_super_sampling", C.c_float),
        ("use_adaptive_super_sampling", C.c_int32),
        ("visualise_sample_count", C.c_int32),
        ("use_foveated_rendering", C.c_int32),
        ("mouse_focus_x", C.c_int32), ("mouse_focus_y", C.c_int32),
        ("zooming", C.c_int32), ("zooming_in", C.c_int32),
        ("use_sample_reuse", C.c_int32),
        ("sample_reuse_cache_dirty", C.c_int32),
        ("floating_point_precision", C.c_int32),
        ("error_count", C.c_int32),
        ("last_error", C.c_char * 256),
        ("fractal_custom_params", C.c_char * 256),
    ]

    def set_segment(self, lbx, lby, rtx, rty):
        self.left_bottom_x, self.left_bottom_y = lbx, lby
        self.right_top_x, self.right_top_y = rtx, rty

    def segment(self):
        return (self.left_bottom_x, self.left_bottom_y, self.right_top_x, self.right_top_y)


class TileDesc(C.Structure):
    """cu_tile_desc (128 bytes)."""
    _fields_ = [
        ("left", C.c_double), ("top", C.c_double), ("pixel_w", C.c_double), ("pixel_h", C.c_double),
        ("left_f", C.c_float), ("top_f", C.c_float), ("pixel_w_f", C.c_float), ("pixel_h_f", C.c_float),
        ("julia_cx", C.c_double), ("julia_cy", C.c_double),
        ("bailout_sq", C.c_double),
        ("seed", C.c_uint64),
        ("out_offset", C.c_uint64),
        ("work_offset", C.c_uint64),
        ("width", C.c_uint32), ("height", C.c_uint32),
        ("out_pitch", C.c_uint32),
        ("max_iter", C.c_uint32),
        ("kind", C.c_uint32),
        ("spp", C.c_uint32),
        ("pattern", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


assert C.sizeof(TileDesc) == 128


class Tune(C.Structure):
    _fields_ = [("block_steps", C.c_uint32), ("waves_per_cu", C.c_uint32),
                ("chunk", C.c_uint32), ("refill_lanes", C.c_uint32)]


class FrameTune(C.Structure):
    """cu_frame_tune"""
    _fields_ = [("block_steps", C.c_uint32), ("handoff_lanes", C.c_uint32),
                ("tail_passes", C.c_uint32), ("wave_shape", C.c_uint32),
                ("drain_refill", C.c_uint32)]


class DeepDesc(C.Structure):
    """cu_deep_desc (80 bytes)"""
    _fields_ = [("centre_x", C.c_double), ("centre_y", C.c_double),
                ("centre_x_lo", C.c_double), ("centre_y_lo", C.c_double),
                ("pixel_w", C.c_double), ("pixel_h", C.c_double),
                ("bailout_sq", C.c_double),
                ("width", C.c_uint32), ("height", C.c_uint32),
                ("max_iter", C.c_uint32), ("flags", C.c_uint32),
                ("max_references", C.c_uint32), ("glitch_eps", C.c_float)]


assert C.sizeof(DeepDesc) == 80


class DeepReport(C.Structure):
    """cu_deep_report"""
    _fields_ = [("iterations", C.c_ulonglong), ("reference_length", C.c_uint32),
                ("references", C.c_uint32), ("glitched", C.c_uint32), ("reserved", C.c_uint32),
                ("orbit_ms", C.c_float), ("render_ms", C.c_float), ("total_ms", C.c_float),
                ("reserved2", C.c_float)]


class ServeConfig(C.Structure):
    """cu_serve_config"""
    _fields_ = [("hbm_budget", C.c_uint64), ("deep_width", C.c_double), ("slots", C.c_uint32),
                ("deep_method", C.c_uint32), ("max_super_sampling", C.c_float),
                ("shallow_policy", C.c_uint32)]


class FrameRequest(C.Structure):
    """cu_frame_request"""
    _fields_ = [("centre_x", C.c_double), ("centre_y", C.c_double), ("width", C.c_double),
                ("tag", C.c_uint64), ("max_iter", C.c_uint32), ("reserved", C.c_uint32)]


class FrameDone(C.Structure):
    """cu_frame_done"""
    _fields_ = [("tag", C.c_uint64), ("session", C.c_uint32), ("path", C.c_uint32),
                ("submit_s", C.c_double), ("start_s", C.c_double), ("done_s", C.c_double),
                ("iterations", C.c_ulonglong), ("checksum", C.c_ulonglong),
                ("status", C.c_int32), ("reserved", C.c_uint32)]


class ServeReport(C.Structure):
    """cu_serve_report"""
    _fields_ = [("iterations", C.c_ulonglong), ("frames", C.c_uint64),
                ("bytes_in_use", C.c_uint64), ("peak_bytes", C.c_uint64),
                ("sessions_admitted", C.c_uint32), ("sessions_waiting", C.c_uint32)]


assert C.sizeof(ServeConfig) == 32 and C.sizeof(FrameRequest) == 40
assert C.sizeof(FrameDone) == 64 and C.sizeof(ServeReport) == 40

DEEP_DIRECT_F64 = 1
DEEP_PERTURB_F32 = 32
DEEP_PERTURB_F64 = 64
SERVE_SHALLOW_FP64, SERVE_SHALLOW_REFERENCE_SWITCH = 0, 1
DEEP_STATIC_MAX_ITER = 65535
DEEP_REBASE = 1
DEEP_REUSE_ORBIT = 2

# name -> (restype, argtypes)
_P = C.c_void_p
_I = C.c_int
_U32 = C.c_uint32
_U64 = C.c_uint64
_SZ = C.c_size_t
_D = C.c_double
SIGNATURES = {
    "cu_last_error": (C.c_char_p, []),
    "cu_init": (_I, [_I]),
    "cu_device_count": (_I, [C.POINTER(_I)]),
    "cu_module_load": (_I, [C.c_char_p, C.POINTER(_P)]),
    "cu_module_unload": (_I, [_P]),
    "cu_module_get_function": (_I, [_P, C.c_char_p, C.POINTER(_P)]),
    "cu_module_get_global": (_I, [_P, C.c_char_p, C.POINTER(_P), C.POINTER(_SZ)]),
    "cu_module_write_global": (_I, [_P, C.c_char_p, _P, _SZ]),
    "cu_mem_alloc": (_I, [C.POINTER(_P), _SZ]),
    "cu_mem_alloc_pitch": (_I, [C.POINTER(_P), C.POINTER(_SZ), _SZ, _SZ, C.c_uint]),
    "cu_mem_free": (_I, [_P]),
    "cu_host_alloc": (_I, [C.POINTER(_P), _SZ]),
    "cu_host_free": (_I, [_P]),
    "cu_memcpy_htod": (_I, [_P, _P, _SZ]),
    "cu_memcpy_dtoh": (_I, [_P, _P, _SZ]),
    "cu_memcpy_dtod": (_I, [_P, _P, _SZ]),
    "cu_memcpy_2d_dtoh": (_I, [_P, _SZ, _P, _SZ, _SZ, _SZ]),
    "cu_memset_u32": (_I, [_P, _U32, _SZ]),
    "cu_launch_kernel": (_I, [_P, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint,
                              C.c_uint, C.POINTER(_P)]),
    "cu_ctx_synchronize": (_I, []),
    "cu_stream_create": (_I, [C.POINTER(_P)]),
    "cu_stream_destroy": (_I, [_P]),
    "cu_stream_synchronize": (_I, [_P]),
    "cu_stream_query": (_I, [_P, C.POINTER(_I)]),
    "cu_provider_create": (_I, [C.c_char_p, C.POINTER(C.c_int32)]),
    "cu_provider_destroy": (_I, [C.c_int32]),
    "cu_provider_available_fractals": (_I, [C.c_int32, C.c_char_p, _SZ]),
    "cu_provider_get_default_renderer": (_I, [C.c_int32, C.POINTER(C.c_int32)]),
    "cu_provider_get_renderer": (_I, [C.c_int32, C.c_char_p, _I, C.POINTER(C.c_int32)]),
    "cu_renderer_initialize_rendering": (_I, [C.c_int32, _I, _I, C.POINTER(_U32), _I]),
    "cu_renderer_free_rendering_resources": (_I, [C.c_int32]),
    "cu_renderer_get_state": (_I, [C.c_int32, C.POINTER(_I)]),
    "cu_renderer_get_width": (_I, [C.c_int32, C.POINTER(_I)]),
    "cu_renderer_get_height": (_I, [C.c_int32, C.POINTER(_I)]),
    "cu_renderer_launch_debug_kernel": (_I, [C.c_int32]),
    "cu_renderer_render_fast": (_I, [C.c_int32, C.POINTER(RenderingModel)]),
    "cu_renderer_render_quality": (_I, [C.c_int32, C.POINTER(RenderingModel)]),
    "cu_renderer_close": (_I, [C.c_int32]),
    "cu_renderer_set_fractal_custom_params": (_I, [C.c_int32, C.c_char_p]),
    "cu_renderer_get_fractal_name": (_I, [C.c_int32, C.c_char_p, _SZ]),
    "cu_renderer_supply_default_values": (_I, [C.c_int32, C.POINTER(RenderingModel), _I, _I]),
    "cu_renderer_iterations": (_I, [_I, C.POINTER(C.c_ulonglong)]),
    "cu_renderer_read_output": (_I, [C.c_int32, C.POINTER(_U32), _SZ]),
    "cu_renderer_read_pixel_info": (_I, [C.c_int32, _I, _P, _SZ]),
    "cu_set_debug_mode": (_I, [_I]),
    "cu_plane_segment_from_center": (_I, [_D, _D, _D, _I, _I, C.POINTER(_D)]),
    "cu_precision_for_segment": (_I, [_D, _D, _D, _D, _I, _I, C.POINTER(_I)]),
    "cu_launch_geometry": (_I, [_I, _I, C.POINTER(C.c_uint), C.POINTER(C.c_uint)]),
    "cu_default_palette": (_I, [C.POINTER(_U32), _SZ]),
    "cu_palette_from_argb": (_I, [C.POINTER(_U32), C.POINTER(_U32), _SZ]),
    "cu_palette_load_or_default": (_I, [C.c_char_p, C.POINTER(_U32), _SZ, C.POINTER(_SZ),
                                        C.POINTER(C.c_int)]),
    "cu_image_save_png": (_I, [C.c_char_p, C.POINTER(_U32), C.c_int, C.c_int]),
    "cu_tile_init": (_I, [C.POINTER(TileDesc), _D, _D, _D, _D, _U32, _U32, _U32, _D, _U32, _D, _D,
                          _U32, _U32, _U64]),
    "cu_tile_fused_safe": (_I, [C.POINTER(TileDesc), _I, C.POINTER(_I)]),
    "cu_engine_create": (_I, [C.c_char_p, C.POINTER(_P)]),
    "cu_engine_destroy": (_I, [_P]),
    "cu_escape_tile": (_I, [_P, _I, _I, C.POINTER(TileDesc), _P, _P, _P, _P]),
    "cu_escape_tiles_persistent": (_I, [_P, _I, _I, _P, _U32, _U64, _P, _P, _P, _P,
                                        C.POINTER(Tune), _P]),
    "cu_escape_tiles_persistent_capture": (_I, [_P, _I, _I, _P, _U32, _U64, _P, _P, _P, _P,
                                                C.POINTER(Tune), C.POINTER(_P)]),
    "cu_graph_launch": (_I, [_P, _P]),
    "cu_graph_last_ms": (_I, [_P, C.POINTER(C.c_float)]),
    "cu_graph_destroy": (_I, [_P]),
    "cu_compose_smooth": (_I, [_P, _P, _U32, _P, _U32, _U32, _U32, _P, _U32, _P]),
    "cu_engine_last_kernel_ms": (_I, [_P, C.POINTER(C.c_float)]),
    "cu_iters_total": (_I, [_P, C.POINTER(C.c_ulonglong)]),
    "cu_frame_create": (_I, [_P, C.POINTER(TileDesc), _U32, _I, _I, C.POINTER(_P)]),
    "cu_frame_destroy": (_I, [_P]),
    "cu_frame_render": (_I, [_P, C.POINTER(FrameTune), _P, _U32, _P, _P]),
    "cu_frame_outputs": (_I, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_U64)]),
    "cu_frame_iterations": (_I, [_P, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]),
    "cu_frame_last_ms": (_I, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "cu_frame_capture": (_I, [_P, C.POINTER(FrameTune), _P, _U32, _P]),
    "cu_frame_replay": (_I, [_P, _P]),
    "cu_engine_configure": (_I, [_P, C.POINTER(Tune)]),
    "cu_deep_create": (_I, [_P, C.POINTER(DeepDesc), _I, C.POINTER(_P)]),
    "cu_deep_destroy": (_I, [_P]),
    "cu_deep_retarget": (_I, [_P, C.POINTER(DeepDesc)]),
    "cu_deep_render": (_I, [_P, _P, _U32, _P, _P]),
    "cu_deep_outputs": (_I, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "cu_deep_stats": (_I, [_P, C.POINTER(DeepReport)]),
    "cu_deep_orbit": (_I, [_P, _P, _U32, C.POINTER(_U32)]),
    "cu_deep_render_static": (_I, [_I, C.POINTER(DeepDesc), _P, _P, _P, _U32, _P, _P, _P]),
    "cu_serve_create": (_I, [C.c_char_p, C.POINTER(ServeConfig), C.POINTER(_P)]),
    "cu_serve_destroy": (_I, [_P]),
    "cu_serve_open": (_I, [_P, _U32, _U32, _P, _U32, _U32, C.POINTER(_U32)]),
    "cu_serve_submit": (_I, [_P, _U32, C.POINTER(FrameRequest)]),
    "cu_serve_close": (_I, [_P, _U32]),
    "cu_serve_poll": (_I, [_P, C.POINTER(FrameDone), _U32, C.POINTER(_U32), _I]),
    "cu_serve_read_rgba": (_I, [_P, _U32, _P, _SZ]),
    "cu_serve_last_failure": (_I, [_P, C.c_char_p, _SZ]),
    "cu_serve_stats": (_I, [_P, C.POINTER(ServeReport)]),
    "cu_serve_session_times": (_I, [_P, _U32, C.POINTER(_D), C.POINTER(_D)]),
}

_lib = None


def load() -> C.CDLL:
    """Load lib/libchaosultra_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                              f"(or python chaos-ultra_amd/build.py) first")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    return load().cu_last_error().decode(errors="replace")


def check(rc: int) -> None:
    if rc != CU_OK:
        raise _ERRORS.get(rc, DeviceError)(last_error())


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
