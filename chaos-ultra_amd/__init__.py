"""chaos-ultra for MI355X — Python view of the C ABI (include/chaosultra.h).

Import with ``importlib.import_module("chaos-ultra_amd")`` (the directory name has a
hyphen).  Everything here forwards to lib/libchaosultra_hip.so; device work runs in
the gfx950 code objects under kernels/.  The classes mirror the reference's plugin
surface: FractalRendererProvider / FractalRenderer (rendering/FractalRenderer.java,
rendering/FractalRendererProvider.java) and the RenderingModel it renders from.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import (  # noqa: F401  (re-exported)
    ServeConfig, FrameRequest, FrameDone, ServeReport,
    KERNELS_DIR, LIB_PATH, ChaosError, CudaInitializationError, DeviceError, FractalRendererError,
    IllegalArgumentError, IllegalStateError, RenderingModel, TileDesc, Tune, FrameTune,
    DeepDesc, DeepReport, DEEP_DIRECT_F64, DEEP_PERTURB_F32, DEEP_PERTURB_F64, DEEP_REBASE,
    DEEP_REUSE_ORBIT,
    FLAG_ADAPTIVE_SS, FLAG_FOVEATION, FLAG_SAMPLE_REUSE, FLAG_VISUALISE_SAMPLE_COUNT,
    FLAG_ZOOMING, FLAG_ZOOMING_IN, FRACTAL_JULIA, FRACTAL_MANDELBROT, PRECISION_DOUBLE,
    PRECISION_SINGLE, PRECISION_TOO_BIG, SAMPLE_CENTRE, SAMPLE_CORNER, SAMPLE_JITTER,
    STATE_NOT_INITIALIZED, STATE_READY_TO_RENDER,
)

MAX_SUPER_SAMPLING = 64  # FractalRenderer.MAX_SUPER_SAMPLING

PIXEL_INFO_DTYPE = np.dtype([("value", "<f4"), ("weight", "<f4"), ("is_reused", "u1"),
                             ("pad", "u1", (3,)), ("weight_of_new_samples", "<f4")])
assert PIXEL_INFO_DTYPE.itemsize == 16


def init(device: int = 0) -> None:
    abi.call("cu_init", device)


def device_count() -> int:
    n = C.c_int(0)
    abi.call("cu_device_count", C.byref(n))
    return n.value


def synchronize() -> None:
    abi.call("cu_ctx_synchronize")


# ---------------------------------------------------------------------------
# device memory
# ---------------------------------------------------------------------------
class DeviceBuffer:
    """A hipMalloc'd buffer freed on close()/garbage collection."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        abi.call("cu_mem_alloc", C.byref(self.ptr), C.c_size_t(self.nbytes))

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        b.upload(arr)
        return b

    @property
    def address(self) -> int:
        return self.ptr.value or 0

    def upload(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("array larger than buffer")
        abi.call("cu_memcpy_htod", self.ptr, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes))

    def download(self, dtype, count: int | None = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize if count is None else int(count)
        out = np.empty(n, dtype=dtype)
        abi.call("cu_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes))
        return out

    def fill_u32(self, value: int = 0) -> None:
        abi.call("cu_memset_u32", self.ptr, C.c_uint32(value), C.c_size_t(self.nbytes // 4))

    def close(self) -> None:
        if self.ptr and self.ptr.value:
            abi.call("cu_mem_free", self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostBuffer:
    """Page-locked host memory (cu_host_alloc) freed on close()/garbage collection;
    array() views it as a numpy array, download_from() copies a device buffer in."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        abi.call("cu_host_alloc", C.byref(self.ptr), C.c_size_t(self.nbytes))

    def array(self, dtype) -> np.ndarray:
        dtype = np.dtype(dtype)
        raw = (C.c_ubyte * self.nbytes).from_address(self.ptr.value)
        return np.frombuffer(raw, dtype=dtype, count=self.nbytes // dtype.itemsize)

    def download_from(self, dev: DeviceBuffer, nbytes: int | None = None) -> None:
        n = min(self.nbytes, dev.nbytes) if nbytes is None else int(nbytes)
        if n > self.nbytes or n > dev.nbytes:
            raise ValueError("copy larger than a buffer")
        abi.call("cu_memcpy_dtoh", self.ptr, dev.ptr, C.c_size_t(n))

    def close(self) -> None:
        if self.ptr and self.ptr.value:
            abi.call("cu_host_free", self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class IterationCounter(DeviceBuffer):
    """CU_ITER_SLOTS device u64 slots receiving the iterations executed by a launch."""

    def __init__(self):
        super().__init__(abi.ITER_SLOTS * 8)
        self.clear()

    def clear(self) -> None:
        self.fill_u32(0)

    def total(self) -> int:
        t = C.c_ulonglong(0)
        abi.call("cu_iters_total", self.ptr, C.byref(t))
        return int(t.value)


class Stream:
    def __init__(self):
        self.handle = C.c_void_p()
        abi.call("cu_stream_create", C.byref(self.handle))

    def synchronize(self) -> None:
        abi.call("cu_stream_synchronize", self.handle)

    def done(self) -> bool:
        d = C.c_int(0)
        abi.call("cu_stream_query", self.handle, C.byref(d))
        return bool(d.value)

    def close(self) -> None:
        if self.handle and self.handle.value:
            abi.call("cu_stream_destroy", self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# renderer layer
# ---------------------------------------------------------------------------
def plane_segment_from_center(cx: float, cy: float, zoom: float, width: int, height: int):
    """Model.setPlaneSegmentFromCenter: zoom is the segment height."""
    out = (C.c_double * 4)()
    abi.call("cu_plane_segment_from_center", C.c_double(cx), C.c_double(cy), C.c_double(zoom),
             width, height, out)
    return tuple(out)


def precision_for_segment(lbx, lby, rtx, rty, width, height) -> int:
    p = C.c_int(0)
    abi.call("cu_precision_for_segment", C.c_double(lbx), C.c_double(lby), C.c_double(rtx),
             C.c_double(rty), width, height, C.byref(p))
    return p.value


def launch_geometry(width: int, height: int):
    gx, gy = C.c_uint(0), C.c_uint(0)
    abi.call("cu_launch_geometry", width, height, C.byref(gx), C.byref(gy))
    return gx.value, gy.value


def default_palette() -> np.ndarray:
    out = np.zeros(1536, dtype=np.uint32)
    abi.call("cu_default_palette", out.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_size_t(1536))
    return out


def palette_from_argb(argb) -> np.ndarray:
    argb = np.ascontiguousarray(argb, dtype=np.uint32)
    out = np.zeros_like(argb)
    abi.call("cu_palette_from_argb", argb.ctypes.data_as(C.POINTER(C.c_uint32)),
             out.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_size_t(argb.size))
    return out


def load_palette(path: str | None, max_entries: int = 1 << 16):
    """ImageHelpers.loadColorPaletteOrDefault: (RGBA palette, loaded-from-file flag)."""
    out = np.zeros(max(max_entries, 1536), dtype=np.uint32)
    n, loaded = C.c_size_t(), C.c_int()
    abi.call("cu_palette_load_or_default", (path or "").encode(),
             out.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_size_t(out.size), C.byref(n),
             C.byref(loaded))
    return out[:n.value].copy(), bool(loaded.value)


def save_png(path: str, rgba: np.ndarray) -> None:
    """ImageHelpers.saveImageToFile(..., "png") of an (h, w) array of RGBA words."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint32)
    h, w = rgba.shape
    abi.call("cu_image_save_png", path.encode(), rgba.ctypes.data_as(C.POINTER(C.c_uint32)), w, h)


class FractalRenderer:
    """FractalRenderer (rendering/FractalRenderer.java) backed by a native renderer handle."""

    def __init__(self, handle: int):
        self.handle = handle

    def initialize_rendering(self, width: int, height: int, palette: np.ndarray) -> None:
        self._palette = np.ascontiguousarray(palette, dtype=np.uint32)
        abi.call("cu_renderer_initialize_rendering", self.handle, width, height,
                 self._palette.ctypes.data_as(C.POINTER(C.c_uint32)), int(self._palette.size))

    def free_rendering_resources(self) -> None:
        abi.call("cu_renderer_free_rendering_resources", self.handle)

    def get_state(self) -> int:
        s = C.c_int(0)
        abi.call("cu_renderer_get_state", self.handle, C.byref(s))
        return s.value

    def get_width(self) -> int:
        v = C.c_int(0)
        abi.call("cu_renderer_get_width", self.handle, C.byref(v))
        return v.value

    def get_height(self) -> int:
        v = C.c_int(0)
        abi.call("cu_renderer_get_height", self.handle, C.byref(v))
        return v.value

    def launch_debug_kernel(self) -> None:
        abi.call("cu_renderer_launch_debug_kernel", self.handle)

    def render_fast(self, model: RenderingModel) -> None:
        abi.call("cu_renderer_render_fast", self.handle, C.byref(model))

    def render_quality(self, model: RenderingModel) -> None:
        abi.call("cu_renderer_render_quality", self.handle, C.byref(model))

    def close(self) -> None:
        abi.call("cu_renderer_close", self.handle)

    def set_fractal_custom_params(self, text: str) -> None:
        abi.call("cu_renderer_set_fractal_custom_params", self.handle, text.encode())

    def get_fractal_name(self) -> str:
        buf = C.create_string_buffer(256)
        abi.call("cu_renderer_get_fractal_name", self.handle, buf, C.c_size_t(256))
        return buf.value.decode()

    def supply_default_values(self, model: RenderingModel, canvas_width: int, canvas_height: int) -> None:
        abi.call("cu_renderer_supply_default_values", self.handle, C.byref(model), canvas_width,
                 canvas_height)

    def iterations(self) -> int:
        """Iterations executed by this renderer's module since it was loaded."""
        n = C.c_ulonglong()
        abi.call("cu_renderer_iterations", self.handle, C.byref(n))
        return n.value

    def read_output(self) -> np.ndarray:
        w, h = self.get_width(), self.get_height()
        out = np.zeros(w * h, dtype=np.uint32)
        abi.call("cu_renderer_read_output", self.handle, out.ctypes.data_as(C.POINTER(C.c_uint32)),
                 C.c_size_t(out.size))
        return out.reshape(h, w)

    def read_pixel_info(self, primary: bool = True) -> np.ndarray:
        w, h = self.get_width(), self.get_height()
        out = np.zeros(w * h, dtype=PIXEL_INFO_DTYPE)
        abi.call("cu_renderer_read_pixel_info", self.handle, 1 if primary else 0,
                 out.ctypes.data_as(C.c_void_p), C.c_size_t(out.nbytes))
        return out.reshape(h, w)


class FractalRendererProvider:
    """CudaFractalRendererProvider counterpart: modules are <kernels_dir>/<name>.hsaco."""

    def __init__(self, kernels_dir: str = KERNELS_DIR):
        h = C.c_int32(0)
        abi.call("cu_provider_create", kernels_dir.encode(), C.byref(h))
        self.handle = h.value

    def get_available_fractals(self) -> set:
        buf = C.create_string_buffer(1024)
        abi.call("cu_provider_available_fractals", self.handle, buf, C.c_size_t(1024))
        return set(buf.value.decode().split(",")) if buf.value else set()

    def get_default_renderer(self) -> FractalRenderer:
        h = C.c_int32(0)
        abi.call("cu_provider_get_default_renderer", self.handle, C.byref(h))
        return FractalRenderer(h.value)

    def get_renderer(self, name: str, force_reload: bool = False) -> FractalRenderer:
        h = C.c_int32(0)
        abi.call("cu_provider_get_renderer", self.handle, name.encode(), 1 if force_reload else 0,
                 C.byref(h))
        return FractalRenderer(h.value)

    def close(self) -> None:
        if self.handle:
            abi.call("cu_provider_destroy", self.handle)
            self.handle = 0


def set_debug_mode(enabled: bool) -> None:
    abi.call("cu_set_debug_mode", 1 if enabled else 0)


# ---------------------------------------------------------------------------
# escape-family layer
# ---------------------------------------------------------------------------
def make_tile(lbx, lby, rtx, rty, width, height, max_iter, bailout_sq=4.0,
              kind=FRACTAL_MANDELBROT, julia_c=(0.0, 0.0), spp=1, pattern=SAMPLE_CENTRE,
              seed=0) -> TileDesc:
    t = TileDesc()
    abi.call("cu_tile_init", C.byref(t), C.c_double(lbx), C.c_double(lby), C.c_double(rtx),
             C.c_double(rty), width, height, max_iter, C.c_double(bailout_sq), kind,
             C.c_double(julia_c[0]), C.c_double(julia_c[1]), spp, pattern, C.c_uint64(seed))
    return t


def tile_fused_safe(tile: TileDesc, precision: int) -> bool:
    s = C.c_int(0)
    abi.call("cu_tile_fused_safe", C.byref(tile), precision, C.byref(s))
    return bool(s.value)


class Engine:
    """The escape-family module (kernels/escape_family.hsaco)."""

    def __init__(self, kernels_dir: str = KERNELS_DIR):
        self.handle = C.c_void_p()
        abi.call("cu_engine_create", kernels_dir.encode(), C.byref(self.handle))

    def escape_tile(self, precision, tile: TileDesc, counts: DeviceBuffer, smooth: DeviceBuffer | None,
                    iters: DeviceBuffer | None, variant: int = -1, stream: Stream | None = None):
        abi.call("cu_escape_tile", self.handle, precision, variant, C.byref(tile), counts.ptr,
                 smooth.ptr if smooth else None, iters.ptr if iters else None,
                 stream.handle if stream else None)

    def escape_tiles_persistent(self, precision, variant, tiles: DeviceBuffer, n_tiles: int,
                                total: int, counts: DeviceBuffer, smooth: DeviceBuffer | None,
                                iters: DeviceBuffer | None, queue: DeviceBuffer,
                                tune: Tune | None = None, stream: Stream | None = None):
        abi.call("cu_escape_tiles_persistent", self.handle, precision, variant, tiles.ptr, n_tiles,
                 C.c_uint64(total), counts.ptr, smooth.ptr if smooth else None,
                 iters.ptr if iters else None, queue.ptr, C.byref(tune) if tune else None,
                 stream.handle if stream else None)

    def compose_smooth(self, smooth: DeviceBuffer, smooth_pitch, rgba: DeviceBuffer, rgba_pitch,
                       width, height, palette: DeviceBuffer, palette_length,
                       stream: Stream | None = None):
        abi.call("cu_compose_smooth", self.handle, smooth.ptr, smooth_pitch, rgba.ptr, rgba_pitch,
                 width, height, palette.ptr, palette_length, stream.handle if stream else None)

    def configure(self, tune: Tune) -> None:
        abi.call("cu_engine_configure", self.handle, C.byref(tune))

    def last_kernel_ms(self) -> float:
        ms = C.c_float(0)
        abi.call("cu_engine_last_kernel_ms", self.handle, C.byref(ms))
        return ms.value

    def close(self) -> None:
        if self.handle and self.handle.value:
            abi.call("cu_engine_destroy", self.handle)
            self.handle = C.c_void_p()


class Frame:
    """The straggler-compaction frame pipeline over a set of tiles (cu_frame)."""

    def __init__(self, engine: Engine, tiles, precision: int = 64, variant: int = -1):
        tiles = list(tiles)
        arr = (TileDesc * len(tiles))(*tiles)
        self.handle = C.c_void_p()
        abi.call("cu_frame_create", engine.handle, arr, len(tiles), precision, variant,
                 C.byref(self.handle))
        counts, smooth, total = C.c_void_p(), C.c_void_p(), C.c_uint64()
        abi.call("cu_frame_outputs", self.handle, C.byref(counts), C.byref(smooth), C.byref(total))
        self.d_counts, self.d_smooth, self.total_pixels = counts.value, smooth.value, total.value

    def render(self, tune: FrameTune | None = None, palette: DeviceBuffer | None = None,
               palette_length: int = 0, rgba: DeviceBuffer | None = None,
               stream: Stream | None = None) -> None:
        abi.call("cu_frame_render", self.handle, C.byref(tune) if tune else None,
                 palette.ptr if palette else None, palette_length, rgba.ptr if rgba else None,
                 stream.handle if stream else None)

    def capture(self, tune: FrameTune | None = None, palette: DeviceBuffer | None = None,
                palette_length: int = 0, rgba: DeviceBuffer | None = None) -> None:
        """Capture one render into a HIP graph (replayed by replay())."""
        abi.call("cu_frame_capture", self.handle, C.byref(tune) if tune else None,
                 palette.ptr if palette else None, palette_length, rgba.ptr if rgba else None)

    def replay(self, stream: Stream | None = None) -> None:
        abi.call("cu_frame_replay", self.handle, stream.handle if stream else None)

    def counts(self) -> np.ndarray:
        out = np.empty(self.total_pixels, dtype=np.uint32)
        abi.call("cu_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.d_counts),
                 C.c_size_t(out.nbytes))
        return out

    def smooth(self) -> np.ndarray:
        out = np.empty(self.total_pixels, dtype=np.float64)
        abi.call("cu_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.d_smooth),
                 C.c_size_t(out.nbytes))
        return out

    def iterations(self) -> int:
        """Iterations executed by the last render, all passes."""
        t = C.c_ulonglong(0)
        abi.call("cu_frame_iterations", self.handle, C.byref(t), None)
        return int(t.value)

    def iterations_pass1(self) -> int:
        """Iterations executed by pass 1 of the last render."""
        t, p = C.c_ulonglong(0), C.c_ulonglong(0)
        abi.call("cu_frame_iterations", self.handle, C.byref(t), C.byref(p))
        return int(p.value)

    def last_ms(self):
        a, b = C.c_float(0), C.c_float(0)
        abi.call("cu_frame_last_ms", self.handle, C.byref(a), C.byref(b))
        return a.value, b.value

    def close(self) -> None:
        if self.handle and self.handle.value:
            abi.call("cu_frame_destroy", self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# deep zoom (cu_deep_*)
# ---------------------------------------------------------------------------
def make_deep(centre_x, centre_y, width_plane, width, height, max_iter, bailout_sq=256.0,
              rebase=True, max_references=8, glitch_eps=1e-6, centre_lo=(0.0, 0.0)) -> DeepDesc:
    """A deep frame of `width` x `height` square pixels spanning `width_plane` of the
    real axis, centred on (centre_x + centre_lo[0], centre_y + centre_lo[1])."""
    pw = width_plane / width
    return DeepDesc(centre_x, centre_y, centre_lo[0], centre_lo[1], pw, pw, bailout_sq, width,
                    height, max_iter, DEEP_REBASE if rebase else 0, max_references, glitch_eps)


class DeepFrame:
    """One deep-zoom frame (cu_deep): direct FP64, or FP32/FP64 perturbation around a
    double-double reference orbit computed on the GPU."""

    def __init__(self, engine: Engine, desc: DeepDesc, method: int = DEEP_PERTURB_F64):
        self.desc, self.method = desc, method
        self.handle = C.c_void_p()
        abi.call("cu_deep_create", engine.handle, C.byref(desc), method, C.byref(self.handle))
        counts, smooth, glitch = C.c_void_p(), C.c_void_p(), C.c_void_p()
        abi.call("cu_deep_outputs", self.handle, C.byref(counts), C.byref(smooth), C.byref(glitch))
        self._planes = (counts.value, smooth.value, glitch.value)
        self.pixels = desc.width * desc.height

    def render(self, palette: DeviceBuffer | None = None, palette_length: int = 0,
               rgba: DeviceBuffer | None = None, stream: Stream | None = None) -> None:
        abi.call("cu_deep_render", self.handle, palette.ptr if palette else None, palette_length,
                 rgba.ptr if rgba else None, stream.handle if stream else None)

    def retarget(self, desc: DeepDesc) -> None:
        """Another view of the same size (max_iter no larger than at creation)."""
        abi.call("cu_deep_retarget", self.handle, C.byref(desc))
        self.desc = desc

    def _download(self, which: int, dtype) -> np.ndarray:
        out = np.empty(self.pixels, dtype=dtype)
        abi.call("cu_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), C.c_void_p(self._planes[which]),
                 C.c_size_t(out.nbytes))
        return out.reshape(self.desc.height, self.desc.width)

    def counts(self) -> np.ndarray:
        return self._download(0, np.uint32)

    def smooth(self) -> np.ndarray:
        return self._download(1, np.float64)

    def glitched(self) -> np.ndarray:
        return self._download(2, np.uint8)

    def report(self) -> DeepReport:
        r = DeepReport()
        abi.call("cu_deep_stats", self.handle, C.byref(r))
        return r

    def orbit(self) -> np.ndarray:
        """The primary reference orbit of the last render, shape (M + 1, 4)."""
        dt = np.float32 if self.method == DEEP_PERTURB_F32 else np.float64
        buf = np.empty((self.desc.max_iter + 1) * 4, dtype=dt)
        M = C.c_uint32(0)
        abi.call("cu_deep_orbit", self.handle, buf.ctypes.data_as(C.c_void_p), self.desc.max_iter + 1,
                 C.byref(M))
        return buf[: (M.value + 1) * 4].reshape(-1, 4)

    def close(self) -> None:
        if self.handle and self.handle.value:
            abi.call("cu_deep_destroy", self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
