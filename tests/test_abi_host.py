"""The C ABI without a device: exports, host-side model logic, host-forcible errors."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def header_symbols():
    with open(os.path.join(ROOT, "include", "chaosultra.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cu_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(cu):
    lib = cu.abi.load()
    syms = header_symbols()
    assert len(syms) > 50
    missing = [s for s in syms if not hasattr(lib, s)]
    assert missing == []


def test_binding_covers_header(cu):
    assert sorted(cu.abi.SIGNATURES) == header_symbols()


def test_code_objects_built(cu):
    for name in ("mandelbrot", "julia", "escape_family"):
        p = os.path.join(cu.KERNELS_DIR, f"{name}.hsaco")
        assert os.path.getsize(p) > 1000
        with open(p, "rb") as f:
            head = f.read(64)
        # hipcc --genco writes a clang offload bundle holding the gfx950 code object
        assert head.startswith(b"__CLANG_OFFLOAD_BUNDLE__") or head.startswith(b"\x7fELF")
        with open(p, "rb") as f:
            assert b"gfx950" in f.read()


def test_module_load_missing_file_is_illegal_argument(cu):
    m = C.c_void_p()
    rc = cu.abi.load().cu_module_load(b"/nonexistent/mandelbrot.hsaco", C.byref(m))
    assert rc == cu.abi.CU_ERR_ILLEGAL_ARGUMENT
    assert "Invalid module file name" in cu.abi.last_error()


def test_unknown_fractal_is_illegal_argument(cu):
    p = cu.FractalRendererProvider()
    try:
        assert p.get_available_fractals() == {"mandelbrot", "julia"}
        with pytest.raises(cu.IllegalArgumentError, match="Unknown fractal: newton"):
            p.get_renderer("newton")
    finally:
        p.close()


def test_stale_handles_are_illegal_state(cu):
    with pytest.raises(cu.IllegalStateError):
        cu.FractalRenderer(987654).get_state()
    with pytest.raises(cu.IllegalStateError):
        cu.abi.call("cu_provider_destroy", 987654)


def test_plane_segment_from_center(cu):
    # Model.setPlaneSegmentFromCenter: zoom is the height, width = zoom * W / H
    assert cu.plane_segment_from_center(-0.5, 0.0, 3.0, 1024, 1024) == (-2.0, -1.5, 1.0, 1.5)
    lbx, lby, rtx, rty = cu.plane_segment_from_center(0.0, 0.0, 2.0, 1920, 1080)
    assert (lby, rty) == (-1.0, 1.0)
    assert rtx - lbx == pytest.approx(2.0 * 1920 / 1080, rel=1e-15)
    with pytest.raises(cu.IllegalArgumentError):
        cu.plane_segment_from_center(0, 0, 1, 0, 10)


def test_precision_switch(cu):
    P = cu
    # 1024 px over width 3: far above float ulp -> single
    assert cu.precision_for_segment(-2, -1.5, 1, 1.5, 1024, 1024) == P.PRECISION_SINGLE
    # C2: pixel 2.44e-8 < ulp_f32(0.7436) = 5.96e-8 -> double
    seg = cu.plane_segment_from_center(-0.743643887037151, 0.131825904205330, 1e-4, 4096, 4096)
    assert cu.precision_for_segment(*seg, 4096, 4096) == P.PRECISION_DOUBLE
    # pixel below the double ulp of the left edge -> tooBig
    seg = cu.plane_segment_from_center(-0.743643887037151, 0.131825904205330, 1e-14, 1024, 1024)
    assert cu.precision_for_segment(*seg, 1024, 1024) == P.PRECISION_TOO_BIG
    # the test is on the ulp of the *left/bottom* coordinate: a segment near 0 stays single
    assert cu.precision_for_segment(0.0, 0.0, 1e-6, 1e-6, 10, 10) == P.PRECISION_SINGLE


def test_launch_geometry_limits(cu):
    assert cu.launch_geometry(1024, 1000) == (32, 32)
    assert cu.launch_geometry(0, 10) == (0, 0)
    assert cu.launch_geometry(65535 * 32, 1) == (65535, 1)
    with pytest.raises(cu.FractalRendererError, match="width must be smaller than"):
        cu.launch_geometry(65535 * 32 + 1, 1)
    with pytest.raises(cu.FractalRendererError, match="height must be smaller than"):
        cu.launch_geometry(1, 65535 * 32 + 1)
    with pytest.raises(cu.FractalRendererError, match="width"):  # no int overflow on the way
        cu.launch_geometry(2 ** 31 - 1, 32)


def test_default_palette_matches_png_and_oracle(cu, oracle):
    with open(os.path.join(GOLDEN, "palettes_argb.json")) as f:
        pal = json.load(f)
    png = cu.palette_from_argb(np.array(pal["defaultPalette.png"], dtype=np.uint32))
    assert np.array_equal(cu.default_palette(), png)
    assert np.array_equal(cu.default_palette(), oracle.default_palette())


def test_palette_conversion_matches_oracle(cu, oracle):
    rng = np.random.default_rng(3)
    argb = rng.integers(0, 2 ** 32, 500, dtype=np.uint64).astype(np.uint32)
    want = np.array([oracle.lib().or_argb_to_rgba(int(v)) for v in argb], dtype=np.uint32)
    assert np.array_equal(cu.palette_from_argb(argb), want)


def test_tile_init_pixel_sizes(cu):
    t = cu.make_tile(-2.0, -1.5, 1.0, 1.5, 1024, 1024, 1000)
    assert (t.left, t.top, t.pixel_w, t.pixel_h) == (-2.0, 1.5, 3 / 1024, 3 / 1024)
    assert (t.left_f, t.top_f, t.pixel_w_f) == (-2.0, 1.5, np.float32(3 / 1024))
    seg = cu.plane_segment_from_center(-0.743643887037151, 0.131825904205330, 1e-4, 4096, 4096)
    t = cu.make_tile(*seg, 4096, 4096, 5000, 256.0, spp=4, pattern=cu.SAMPLE_JITTER, seed=1)
    assert t.pixel_w == (seg[2] - seg[0]) / 4096
    f32 = np.float32
    assert np.float32(t.pixel_w_f) == (f32(seg[2]) - f32(seg[0])) / f32(4096)


@pytest.mark.parametrize("bad", [dict(width=0), dict(spp=0), dict(max_iter=0), dict(kind=7),
                                 dict(pattern=9), dict(bailout_sq=float("nan")),
                                 dict(lbx=float("inf"))])
def test_tile_init_rejects(cu, bad):
    args = dict(lbx=-2.0, lby=-1.5, rtx=1.0, rty=1.5, width=8, height=8, max_iter=10,
                bailout_sq=4.0, kind=0, spp=1, pattern=0)
    args.update(bad)
    with pytest.raises(cu.IllegalArgumentError):
        cu.make_tile(args["lbx"], args["lby"], args["rtx"], args["rty"], args["width"],
                     args["height"], args["max_iter"], args["bailout_sq"], kind=args["kind"],
                     spp=args["spp"], pattern=args["pattern"])


def test_fused_safety_rule(cu):
    t = cu.make_tile(-2.0, -1.5, 1.0, 1.5, 64, 64, 10)
    assert cu.tile_fused_safe(t, 64) and cu.tile_fused_safe(t, 32)
    t = cu.make_tile(-1e-320, -1.5, 1.0, 1.5, 64, 64, 10)       # subnormal origin
    assert not cu.tile_fused_safe(t, 64)
    t = cu.make_tile(0.0, 0.0, 1.0, 1.0, 64, 64, 10)            # FP32 needs a nonzero origin
    assert cu.tile_fused_safe(t, 64) and not cu.tile_fused_safe(t, 32)


def test_init_without_device_is_initialization_error(cu):
    if cu.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(cu.CudaInitializationError):
        cu.init(0)


def test_replay_host_built(cu):
    """The native call-stack replay host is built beside the library and checks
    its arguments before touching a device."""
    import subprocess
    exe = os.path.join(os.path.dirname(cu.abi.LIB_PATH), "chaosultra_replay")
    assert os.access(exe, os.X_OK)
    p = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=30)
    assert p.returncode == 2 and "usage" in p.stdout


def test_host_alloc_arguments(cu):
    """cu_host_alloc rejects a NULL out-pointer and a zero size before any device
    call; cu_host_free(NULL) is a no-op."""
    p = C.c_void_p()
    with pytest.raises(cu.IllegalArgumentError):
        cu.abi.call("cu_host_alloc", None, C.c_size_t(64))
    with pytest.raises(cu.IllegalArgumentError):
        cu.abi.call("cu_host_alloc", C.byref(p), C.c_size_t(0))
    assert not p.value
    cu.abi.call("cu_host_free", None)
