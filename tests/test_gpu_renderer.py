"""The FractalRenderer / FractalRendererProvider surface on the GPU against the oracle.

renderQuality and renderFast run the module kernels (fractalRenderMain*,
fractalRenderAdvanced*, compose): the renderer layer launches their 32 x 4-thread
form (*Strip), a host binding the driver layer the reference-named kernels with
the reference's 32 x 32 geometry.  Their pixel_info_t buffers and composed RGBA8
images must equal the oracle's lock-step restatement bit for bit, whichever
geometry computed them.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture()
def provider(gpu):
    p = gpu.FractalRendererProvider()
    yield p
    p.close()


def model_for(cu, seg, max_iter, max_ss, adaptive=True):
    m = cu.RenderingModel()
    m.set_segment(*seg)
    m.max_iterations = max_iter
    m.max_super_sampling = max_ss
    m.use_adaptive_super_sampling = 1 if adaptive else 0
    return m


def fields_equal(got, want):
    for f in ("value", "weight", "is_reused", "weight_of_new_samples"):
        a, b = got[f], want[f]
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            y, x = bad[0]
            raise AssertionError(f"{f} differs at {len(bad)} pixels, first ({x},{y}): "
                                 f"{a[y, x]} vs {b[y, x]}")


@pytest.mark.parametrize("w,h", [(256, 192), (200, 120), (33, 7)])
@pytest.mark.parametrize("max_ss,adaptive", [(1.0, True), (5.0, True), (5.0, False), (7.4, True),
                                            (12.0, True), (24.0, True)])
def test_render_quality_mandelbrot(gpu, provider, oracle, w, h, max_ss, adaptive):
    cu = gpu
    r = provider.get_default_renderer()
    pal = cu.default_palette()
    r.initialize_rendering(w, h, pal)
    try:
        seg = cu.plane_segment_from_center(-0.7453, 0.1127, 0.01, w, h)
        m = model_for(cu, seg, 800, max_ss, adaptive)
        r.render_quality(m)
        assert m.error_count == 0, m.last_error
        prec = 32 if m.floating_point_precision == cu.PRECISION_SINGLE else 64
        got = r.read_pixel_info(primary=True)
        flags = cu.FLAG_ADAPTIVE_SS if adaptive else 0
        want = oracle.render_main(prec, cu.FRACTAL_MANDELBROT, (0, 0), w, h, seg, 800, max_ss, flags)
        fields_equal(got, want)
        img = r.read_output()
        assert np.array_equal(img, oracle.compose(want, pal, max_ss, False))
    finally:
        r.free_rendering_resources()


@pytest.mark.parametrize("w,h,max_ss,adaptive", [(3840, 2160, 4.0, False),
                                                 (1920, 1080, 16.0, True),
                                                 (1920, 1080, 1.0, False)])
def test_render_quality_c2_view_full_size(gpu, provider, oracle, w, h, max_ss, adaptive):
    """The module kernels at full size on the 4096^2 frame's view (centre
    (-0.743643887037151, 0.131825904205330), width 1e-4, maxIter 5000, FP64 by the
    precision switch): every pixel's fields, the composed image and the executed
    iterations equal the oracle's."""
    cu = gpu
    r = provider.get_default_renderer()
    pal = cu.default_palette()
    r.initialize_rendering(w, h, pal)
    try:
        seg = cu.plane_segment_from_center(-0.743643887037151, 0.131825904205330, 1e-4 * h / w,
                                           w, h)
        m = model_for(cu, seg, 5000, max_ss, adaptive)
        before = r.iterations()
        r.render_quality(m)
        assert m.error_count == 0, m.last_error
        assert m.floating_point_precision == cu.PRECISION_DOUBLE
        got = r.read_pixel_info(primary=True)
        flags = cu.FLAG_ADAPTIVE_SS if adaptive else 0
        oracle.render_iterations()
        want = oracle.render_main(64, cu.FRACTAL_MANDELBROT, (0, 0), w, h, seg, 5000, max_ss, flags)
        assert r.iterations() - before == oracle.render_iterations()
        fields_equal(got, want)
        assert np.array_equal(r.read_output(), oracle.compose(want, pal, max_ss, False))
    finally:
        r.free_rendering_resources()


@pytest.mark.parametrize("zoom,prec", [(3.0, 32), (1e-6, 64), (1e-15, 64)])
def test_render_quality_precision_switch(gpu, provider, oracle, zoom, prec):
    cu = gpu
    r = provider.get_default_renderer()
    w, h = 96, 64
    r.initialize_rendering(w, h, cu.default_palette())
    try:
        seg = cu.plane_segment_from_center(-0.743643887037151, 0.131825904205330, zoom, w, h)
        m = model_for(cu, seg, 600, 3.0)
        r.render_quality(m)
        want_p = {3.0: cu.PRECISION_SINGLE, 1e-6: cu.PRECISION_DOUBLE, 1e-15: cu.PRECISION_TOO_BIG}[zoom]
        assert m.floating_point_precision == want_p
        want = oracle.render_main(prec, cu.FRACTAL_MANDELBROT, (0, 0), w, h, seg, 600, 3.0,
                                  cu.FLAG_ADAPTIVE_SS)
        fields_equal(r.read_pixel_info(), want)
    finally:
        r.free_rendering_resources()


def test_render_quality_julia_custom_params(gpu, provider, oracle):
    cu = gpu
    r = provider.get_renderer("julia")
    assert r.get_fractal_name() == "julia"
    w, h = 160, 96
    r.initialize_rendering(w, h, cu.default_palette())
    try:
        m = cu.RenderingModel()
        r.supply_default_values(m, w, h)
        assert m.max_iterations == 900 and m.fractal_custom_params == b"-0.4;0.6"
        r.set_fractal_custom_params("-0.8;0.156")
        seg = cu.plane_segment_from_center(0.0, 0.0, 2.5, w, h)
        m = model_for(cu, seg, 500, 4.0)
        r.render_quality(m)
        prec = 32 if m.floating_point_precision == cu.PRECISION_SINGLE else 64
        want = oracle.render_main(prec, cu.FRACTAL_JULIA, (-0.8, 0.156), w, h, seg, 500, 4.0,
                                  cu.FLAG_ADAPTIVE_SS)
        fields_equal(r.read_pixel_info(), want)
        with pytest.raises(cu.IllegalArgumentError):
            r.set_fractal_custom_params("abc;1")
    finally:
        r.free_rendering_resources()


def test_render_quality_julia_full_hd(gpu, provider, oracle):
    """The Julia module at 1920x1080 on a view narrow enough for the FP64 kernel
    (c = -0.7269 + 0.1889i, width 2e-5 about (0.4, 0.05), where counts run from
    about 200 to 2000; maxIter 4000, 8 adaptive samples): every pixel, the composed
    image and the executed iterations equal the oracle's."""
    cu = gpu
    r = provider.get_renderer("julia")
    w, h = 1920, 1080
    pal = cu.default_palette()
    r.initialize_rendering(w, h, pal)
    try:
        r.set_fractal_custom_params("-0.7269;0.1889")
        seg = cu.plane_segment_from_center(0.4, 0.05, 2e-5 * h / w, w, h)
        m = model_for(cu, seg, 4000, 8.0, True)
        before = r.iterations()
        r.render_quality(m)
        assert m.error_count == 0, m.last_error
        assert m.floating_point_precision == cu.PRECISION_DOUBLE
        oracle.render_iterations()
        want = oracle.render_main(64, cu.FRACTAL_JULIA, (-0.7269, 0.1889), w, h, seg, 4000, 8.0,
                                  cu.FLAG_ADAPTIVE_SS)
        assert r.iterations() - before == oracle.render_iterations()
        fields_equal(r.read_pixel_info(primary=True), want)
        assert np.array_equal(r.read_output(), oracle.compose(want, pal, 8.0, False))
    finally:
        r.free_rendering_resources()


def zoom_segment(seg, focus, w, h, coeff):
    """Zoom the segment about the pixel `focus` (RenderingController.zoomAt's geometry)."""
    lbx, lby, rtx, rty = seg
    fx = lbx + (rtx - lbx) * focus[0] / w
    fy = rty - (rty - lby) * focus[1] / h
    return (fx - (fx - lbx) * coeff, fy - (fy - lby) * coeff, fx + (rtx - fx) * coeff,
            fy + (rty - fy) * coeff)


@pytest.mark.parametrize("foveation", [True, False])
@pytest.mark.parametrize("max_ss", [1.0, 4.0, 0.6])
def test_render_fast_reuse_and_foveation(gpu, provider, oracle, foveation, max_ss):
    cu = gpu
    r = provider.get_default_renderer()
    w, h = 192, 128
    pal = cu.default_palette()
    r.initialize_rendering(w, h, pal)
    try:
        seg = cu.plane_segment_from_center(-0.7453, 0.1127, 0.02, w, h)
        m = model_for(cu, seg, 700, max(max_ss, 1.0))
        m.use_sample_reuse = 1
        m.use_foveated_rendering = 1 if foveation else 0
        r.render_quality(m)
        prev = r.read_pixel_info(primary=True)
        focus = (120, 40)
        for step in range(3):
            new = zoom_segment(m.segment(), focus, w, h, 0.977)
            old = m.segment()
            m.set_segment(*new)
            m.max_super_sampling = max_ss
            m.zooming, m.zooming_in = 1, 1
            m.mouse_focus_x, m.mouse_focus_y = focus
            r.render_fast(m)
            assert m.error_count == 0, m.last_error
            prec = 32 if m.floating_point_precision == cu.PRECISION_SINGLE else 64
            flags = (cu.FLAG_ADAPTIVE_SS | cu.FLAG_SAMPLE_REUSE | cu.FLAG_ZOOMING |
                     cu.FLAG_ZOOMING_IN | (cu.FLAG_FOVEATION if foveation else 0))
            want = oracle.render_advanced(prec, cu.FRACTAL_MANDELBROT, (0, 0), w, h, new, 700,
                                          max_ss, flags, old, prev, focus)
            got = r.read_pixel_info(primary=True)  # buffers were switched: output is primary
            fields_equal(got, want)
            assert np.array_equal(r.read_output(), oracle.compose(want, pal, max_ss, False))
            prev = got
    finally:
        r.free_rendering_resources()


def test_visualise_sample_count(gpu, provider, oracle):
    cu = gpu
    r = provider.get_default_renderer()
    w, h = 64, 64
    pal = cu.default_palette()
    r.initialize_rendering(w, h, pal)
    try:
        seg = cu.plane_segment_from_center(-0.7453, 0.1127, 0.05, w, h)
        m = model_for(cu, seg, 300, 9.0)
        m.visualise_sample_count = 1
        r.render_quality(m)
        want = oracle.render_main(32 if m.floating_point_precision == 0 else 64,
                                  cu.FRACTAL_MANDELBROT, (0, 0), w, h, seg, 300, 9.0,
                                  cu.FLAG_ADAPTIVE_SS)
        assert np.array_equal(r.read_output(), oracle.compose(want, pal, 9.0, True))
        m.visualise_sample_count = 0  # written to the module global again on change
        r.render_quality(m)
        assert np.array_equal(r.read_output(), oracle.compose(want, pal, 9.0, False))
    finally:
        r.free_rendering_resources()


def test_lifecycle(gpu, provider):
    cu = gpu
    r = provider.get_default_renderer()
    assert r.get_state() == cu.STATE_NOT_INITIALIZED
    m = model_for(cu, (-2, -1, 1, 1), 100, 1.0)
    with pytest.raises(cu.IllegalStateError, match="initialized first"):
        r.render_quality(m)
    r.initialize_rendering(32, 32, cu.default_palette())
    assert r.get_state() == cu.STATE_READY_TO_RENDER
    assert (r.get_width(), r.get_height()) == (32, 32)
    with pytest.raises(cu.IllegalStateError, match="Already initialized"):
        r.initialize_rendering(32, 32, cu.default_palette())
    m.max_iterations = 0
    with pytest.raises(cu.IllegalArgumentError, match="maxIterations"):
        r.render_quality(m)
    m.max_iterations = 10
    m.left_bottom_x = float("nan")
    with pytest.raises(cu.IllegalArgumentError, match="finite"):
        r.render_quality(m)
    r.launch_debug_kernel()
    r.free_rendering_resources()
    with pytest.raises(cu.IllegalStateError, match="Already free"):
        r.free_rendering_resources()
    # same name without force returns the active renderer; force reloads the module
    assert provider.get_default_renderer().handle == r.handle
    r2 = provider.get_renderer("mandelbrot", force_reload=True)
    assert r2.handle != r.handle
    r2.initialize_rendering(64, 48, cu.default_palette())
    m = model_for(cu, (-2, -1, 1, 1), 100, 2.0)
    r2.render_quality(m)
    assert r2.read_output().shape == (48, 64)
    # switching fractals closes the active renderer and its module
    r3 = provider.get_renderer("julia")
    assert r2.get_state() == cu.STATE_NOT_INITIALIZED
    r3.close()


def test_driver_layer_under_sampled_kernel(gpu, oracle):
    """fractalRenderUnderSampled through the bare driver calls (module, function,
    pitched allocation, launch with JCuda-style parameter pointers)."""
    cu = gpu
    L = cu.abi.load()
    mod = C.c_void_p()
    cu.abi.check(L.cu_module_load(f"{cu.KERNELS_DIR}/mandelbrot.hsaco".encode(), C.byref(mod)))
    try:
        fn = C.c_void_p()
        cu.abi.check(L.cu_module_get_function(mod, b"fractalRenderUnderSampled", C.byref(fn)))
        with pytest.raises(cu.IllegalArgumentError, match="not found"):
            cu.abi.call("cu_module_get_function", mod, b"fractalRenderMainHalf", C.byref(fn))
        w, h, level = 100, 70, 4
        dptr, pitch = C.c_void_p(), C.c_size_t()
        cu.abi.check(L.cu_mem_alloc_pitch(C.byref(dptr), C.byref(pitch), w * 16, h, 16))
        zero = np.zeros(pitch.value * h // 4, dtype=np.uint32)
        cu.abi.check(L.cu_memcpy_htod(dptr, zero.ctypes.data_as(C.c_void_p), zero.nbytes))
        args = [C.c_void_p(dptr.value), C.c_long(pitch.value), C.c_uint(w), C.c_uint(h),
                C.c_float(-2.0), C.c_float(-1.0), C.c_float(1.0), C.c_float(1.0), C.c_uint(300),
                C.c_uint(level)]
        params = (C.c_void_p * len(args))(*[C.cast(C.pointer(a), C.c_void_p) for a in args])
        gx, gy = (w + 31) // 32, (h + 31) // 32
        cu.abi.check(L.cu_launch_kernel(fn, gx, gy, 1, 32, 32, 1, 0, params))
        cu.abi.check(L.cu_ctx_synchronize())
        got = np.zeros((h, w), dtype=cu.PIXEL_INFO_DTYPE)
        cu.abi.check(L.cu_memcpy_2d_dtoh(got.ctypes.data_as(C.c_void_p), w * 16, dptr, pitch.value,
                                         w * 16, h))
        cu.abi.check(L.cu_mem_free(dptr))
        want = oracle.render_undersampled(w, h, (-2.0, -1.0, 1.0, 1.0), 300, level)
        fields_equal(got, want)
        # module globals: size-checked writes
        cu.abi.call("cu_module_write_global", mod, b"VISUALIZE_SAMPLE_COUNT",
                    C.byref(C.c_uint8(1)), C.c_size_t(1))
        with pytest.raises(cu.IllegalArgumentError, match="allocated to size 1"):
            cu.abi.call("cu_module_write_global", mod, b"VISUALIZE_SAMPLE_COUNT",
                        C.byref(C.c_double(1.0)), C.c_size_t(8))
    finally:
        cu.abi.check(L.cu_module_unload(mod))


class _PointU(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint)]


class _RectD(C.Structure):
    _fields_ = [("lbx", C.c_double), ("lby", C.c_double), ("rtx", C.c_double), ("rty", C.c_double)]


class _RectF(C.Structure):
    _fields_ = [("lbx", C.c_float), ("lby", C.c_float), ("rtx", C.c_float), ("rty", C.c_float)]


@pytest.mark.parametrize("prec,max_ss,flags_adaptive", [(64, 4.0, False), (64, 16.0, True),
                                                         (32, 5.0, True)])
def test_main_kernel_geometries_agree(gpu, oracle, prec, max_ss, flags_adaptive):
    """fractalRenderMain* through the bare driver calls with the reference's
    32 x 32-thread blocks over ceil(W/32) x ceil(H/32), and its strip form with
    32 x 4-thread blocks over ceil(W/32) x ceil(H/4): the same 8x4 vote groups, so
    the same pixel_info_t, equal to the oracle's (a height that is not a multiple
    of 32 leaves the last reference block partly outside the image)."""
    cu = gpu
    L = cu.abi.load()
    w, h = 480, 270
    if prec == 64:
        seg = cu.plane_segment_from_center(-0.743643887037151, 0.131825904205330, 1e-4 * h / w, w, h)
        max_iter, rect, name = 3000, _RectD(*seg), "fractalRenderMainDouble"
    else:
        seg = cu.plane_segment_from_center(-0.5, 0.0, 2.0, w, h)
        max_iter, rect, name = 1600, _RectF(*seg), "fractalRenderMainFloat"
    flags = cu.FLAG_ADAPTIVE_SS if flags_adaptive else 0
    mod = C.c_void_p()
    cu.abi.check(L.cu_module_load(f"{cu.KERNELS_DIR}/mandelbrot.hsaco".encode(), C.byref(mod)))
    dptr, pitch = C.c_void_p(), C.c_size_t()
    cu.abi.check(L.cu_mem_alloc_pitch(C.byref(dptr), C.byref(pitch), w * 16, h, 16))
    try:
        got = {}
        for kernel, rows in ((name, 32), (name + "Strip", 4)):
            fn = C.c_void_p()
            cu.abi.check(L.cu_module_get_function(mod, kernel.encode(), C.byref(fn)))
            zero = np.zeros(pitch.value * h // 4, dtype=np.uint32)
            cu.abi.check(L.cu_memcpy_htod(dptr, zero.ctypes.data_as(C.c_void_p), zero.nbytes))
            args = [C.c_void_p(dptr.value), C.c_long(pitch.value), _PointU(w, h), rect,
                    C.c_uint(max_iter), C.c_float(max_ss), C.c_uint(flags)]
            params = (C.c_void_p * len(args))(*[C.cast(C.pointer(a), C.c_void_p) for a in args])
            cu.abi.check(L.cu_launch_kernel(fn, (w + 31) // 32, (h + rows - 1) // rows, 1, 32, rows,
                                            1, 0, params))
            cu.abi.check(L.cu_ctx_synchronize())
            out = np.zeros((h, w), dtype=cu.PIXEL_INFO_DTYPE)
            cu.abi.check(L.cu_memcpy_2d_dtoh(out.ctypes.data_as(C.c_void_p), w * 16, dptr,
                                             pitch.value, w * 16, h))
            got[rows] = out
        want = oracle.render_main(prec, cu.FRACTAL_MANDELBROT, (0, 0), w, h, seg, max_iter, max_ss,
                                  flags)
        fields_equal(got[32], want)
        fields_equal(got[4], want)
    finally:
        cu.abi.check(L.cu_mem_free(dptr))
        cu.abi.check(L.cu_module_unload(mod))


@pytest.mark.parametrize("adaptive", [True, False])
def test_renderer_iteration_count_matches_oracle(gpu, provider, oracle, adaptive):
    """cu_renderer_iterations: the module's executed-iteration total grows by
    exactly the oracle's count for renderQuality and for a renderFast frame
    (reuse + foveation), the figure bench.py turns into Giga-iterations/s."""
    cu = gpu
    r = provider.get_default_renderer()
    w, h = 200, 136
    pal = cu.default_palette()
    r.initialize_rendering(w, h, pal)
    try:
        seg = cu.plane_segment_from_center(-0.7453, 0.1127, 0.02, w, h)
        m = model_for(cu, seg, 900, 6.0, adaptive)
        m.use_sample_reuse = 1
        m.use_foveated_rendering = 1
        flags = cu.FLAG_ADAPTIVE_SS if adaptive else 0
        before = r.iterations()
        r.render_quality(m)
        prec = 32 if m.floating_point_precision == cu.PRECISION_SINGLE else 64
        oracle.render_iterations()
        prev = oracle.render_main(prec, cu.FRACTAL_MANDELBROT, (0, 0), w, h, seg, 900, 6.0, flags)
        want = oracle.render_iterations()
        assert want > 0 and r.iterations() - before == want
        new = zoom_segment(seg, (50, 90), w, h, 0.977)
        m.set_segment(*new)
        m.zooming, m.zooming_in = 1, 1
        m.mouse_focus_x, m.mouse_focus_y = 50, 90
        before = r.iterations()
        r.render_fast(m)
        prec = 32 if m.floating_point_precision == cu.PRECISION_SINGLE else 64
        oracle.render_advanced(prec, cu.FRACTAL_MANDELBROT, (0, 0), w, h, new, 900, 6.0,
                               flags | cu.FLAG_SAMPLE_REUSE | cu.FLAG_ZOOMING | cu.FLAG_ZOOMING_IN |
                               cu.FLAG_FOVEATION, seg, prev, (50, 90))
        assert r.iterations() - before == oracle.render_iterations()
    finally:
        r.free_rendering_resources()
