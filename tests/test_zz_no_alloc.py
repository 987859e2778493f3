"""cu_deep_render_static, the deep render that allocates nothing, against
cu_deep_render on the same frames.

Two calls in a row through the C ABI, the second with a smaller frame and a
smaller iteration limit into the same output planes: counts, smooth values,
RGBA and iteration totals must equal those of a cu_deep frame created for the
same description (bit-exact: the same kernels' code on the same orbit).
Arguments beyond the static scratch are refused before any device work.
"""
import ctypes as C

import numpy as np
import pytest

C4_CENTRE = (-0.743643887037151, 0.131825904205330)


def _call_static(cu, method, desc, counts, smooth, pal, rgba, iters):
    return cu.abi.load().cu_deep_render_static(
        method, C.byref(desc), counts.ptr, smooth.ptr, pal.ptr, pal.nbytes // 4, rgba.ptr,
        iters.ptr, None)


def test_static_refuses_beyond_scratch(cu):
    """Checked before any device call (no GPU needed)."""
    L = cu.abi.load()
    d = cu.make_deep(*C4_CENTRE, 1e-10, 64, 32, cu.abi.DEEP_STATIC_MAX_ITER + 1, 4.0)
    rc = L.cu_deep_render_static(cu.DEEP_PERTURB_F32, C.byref(d), C.c_void_p(8), None, None, 0,
                                 None, None, None)
    assert rc == cu.abi.CU_ERR_ILLEGAL_ARGUMENT
    assert "CU_DEEP_STATIC_MAX_ITER" in cu.abi.last_error()
    d = cu.make_deep(*C4_CENTRE, 1e-10, 64, 32, 1000, 4.0, rebase=False)
    rc = L.cu_deep_render_static(cu.DEEP_PERTURB_F64, C.byref(d), C.c_void_p(8), None, None, 0,
                                 None, None, None)
    assert rc == cu.abi.CU_ERR_ILLEGAL_ARGUMENT and "CU_DEEP_REBASE" in cu.abi.last_error()
    d = cu.make_deep(*C4_CENTRE, 1e-10, 64, 32, 1000, 4.0)
    rc = L.cu_deep_render_static(cu.DEEP_DIRECT_F64, C.byref(d), C.c_void_p(8), None, None, 0,
                                 None, None, None)
    assert rc == cu.abi.CU_ERR_ILLEGAL_ARGUMENT and "method" in cu.abi.last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["f32", "f64"])
def test_static_equals_deep_render(gpu, engine, method):
    cu = gpu
    m = cu.DEEP_PERTURB_F32 if method == "f32" else cu.DEEP_PERTURB_F64
    pal = ((np.arange(256, dtype=np.uint32) * 0x00010305) | 0xFF000000).astype(np.uint32)
    d_pal = cu.DeviceBuffer.from_array(pal)
    W0, H0 = 320, 200
    counts = cu.DeviceBuffer(W0 * H0 * 4)
    smooth = cu.DeviceBuffer(W0 * H0 * 8)
    rgba = cu.DeviceBuffer(W0 * H0 * 4)
    iters = cu.IterationCounter()
    frames = [cu.make_deep(*C4_CENTRE, 1e-10, W0, H0, 6000, 4.0),
              cu.make_deep(*C4_CENTRE, 1e-6, 96, 64, 3000, 256.0)]
    try:
        for k, desc in enumerate(frames):
            print(f"frame {k}: {desc.width}x{desc.height} max_iter {desc.max_iter} method {method}",
                  flush=True)
            n = desc.width * desc.height
            iters.clear()
            rc = _call_static(cu, m, desc, counts, smooth, d_pal, rgba, iters)
            assert rc == cu.abi.CU_OK, cu.abi.last_error()
            cu.synchronize()
            got_c = counts.download(np.uint32)[:n]
            got_s = smooth.download(np.float64)[:n]
            got_rgba = rgba.download(np.uint32)[:n]
            got_it = iters.total()

            f = cu.DeepFrame(engine, desc, m)
            want_rgba_buf = cu.DeviceBuffer(n * 4)
            try:
                f.render(d_pal, 256, want_rgba_buf)
                cu.synchronize()
                want_c, want_s = f.counts().ravel(), f.smooth().ravel()
                want_rgba = want_rgba_buf.download(np.uint32)
                want_it = f.report().iterations
            finally:
                f.close()
                want_rgba_buf.close()
            assert np.array_equal(got_c, want_c), f"frame {k}: counts differ"
            assert np.array_equal(got_s, want_s), f"frame {k}: smooth values differ"
            assert np.array_equal(got_rgba, want_rgba), f"frame {k}: RGBA differs"
            assert got_it == want_it
            assert (got_c > 0).any()
    finally:
        for b in (counts, smooth, rgba, iters, d_pal):
            b.close()
