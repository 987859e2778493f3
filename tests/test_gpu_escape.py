"""Escape-family kernels on the GPU against the oracle.

Integer outputs (escape counts, iterations executed) must be bit-exact; the
smooth value must agree within 1e-12 relative (the device's log/log2 differ
from glibc's by a few ulp at most).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMOOTH_RTOL = 1e-12
C2_CENTER = (-0.743643887037151, 0.131825904205330)


def run_tile(cu, engine, tile, precision, variant):
    w, h = tile.width, tile.height
    counts = cu.DeviceBuffer(w * h * 4)
    smooth = cu.DeviceBuffer(w * h * 8)
    iters = cu.IterationCounter()
    engine.escape_tile(precision, tile, counts, smooth, iters, variant=variant)
    cu.synchronize()
    return (counts.download(np.uint32).reshape(h, w), smooth.download(np.float64).reshape(h, w),
            iters.total())


def assert_smooth_close(got, want):
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1.0)
    assert rel.max() <= SMOOTH_RTOL, f"smooth off by {rel.max():.3e}"


@pytest.mark.parametrize("precision,variant", [(64, 0), (64, 1), (32, 0), (32, 1)])
def test_c1_frame_bit_exact_vs_oracle(gpu, engine, oracle, precision, variant):
    """The C1 frame (1024 x 1024, maxIter 1000) against the oracle, whose C1 frame
    equals a numpy brute force of the reference loop (tests/test_oracle_pins.py)."""
    cu = gpu
    tile = cu.make_tile(-2.0, -1.5, 1.0, 1.5, 1024, 1024, 1000, 4.0)
    counts, smooth, it = run_tile(cu, engine, tile, precision, variant)
    want_c, want_s, want_it = oracle.escape_tile(tile, precision)
    assert np.array_equal(counts, want_c)
    assert np.array_equal(counts, counts[::-1])  # dyadic pixel centres: mirror-symmetric
    assert it == want_it
    assert_smooth_close(smooth, want_s)


@pytest.mark.parametrize("precision", [64, 32])
def test_julia_frame_bit_exact_vs_oracle(gpu, engine, oracle, precision):
    cu = gpu
    tile = cu.make_tile(-1.5, -1.5, 1.5, 1.5, 512, 512, 900, 4.0, kind=cu.FRACTAL_JULIA,
                        julia_c=(-0.7269, 0.1889))
    counts, _, _ = run_tile(cu, engine, tile, precision, -1)
    want_c, _, _ = oracle.escape_tile(tile, precision)
    assert np.array_equal(counts, want_c)


@pytest.mark.parametrize("precision,variant", [(64, 0), (64, 1), (32, 0), (32, 1)])
def test_c2_jitter_smooth_vs_oracle(gpu, engine, oracle, precision, variant):
    """C2's sampling at reduced size: 4 seeded-jitter samples, bailout 256, maxIter 5000."""
    cu = gpu
    seg = cu.plane_segment_from_center(*C2_CENTER, 1e-4 * 256 / 4096 * 8, 256, 256)
    tile = cu.make_tile(*seg, 256, 256, 5000, 256.0, spp=4, pattern=cu.SAMPLE_JITTER, seed=42)
    counts, smooth, it = run_tile(cu, engine, tile, precision, variant)
    want_c, want_s, want_it = oracle.escape_tile(tile, precision)
    assert np.array_equal(counts, want_c)
    assert it == want_it
    assert_smooth_close(smooth, want_s)


@pytest.mark.parametrize("spp", [1, 2, 3, 8, 16, 32, 64])
def test_sample_counts_and_lane_layouts(gpu, engine, oracle, spp):
    """spp a power of two runs one sample per lane (wave = 64/spp pixels, samples
    summed across lanes in order); other spp run the samples of a pixel in one lane."""
    cu = gpu
    seg = cu.plane_segment_from_center(-0.7453, 0.1127, 0.004, 67, 45)
    tile = cu.make_tile(*seg, 67, 45, 700, 256.0, spp=spp, pattern=cu.SAMPLE_JITTER, seed=spp)
    for precision in (64, 32):
        counts, smooth, it = run_tile(cu, engine, tile, precision, -1)
        want_c, want_s, want_it = oracle.escape_tile(tile, precision)
        assert np.array_equal(counts, want_c)
        assert it == want_it
        assert_smooth_close(smooth, want_s)


def test_edge_budgets_and_radii(gpu, engine, oracle):
    """max_iter below the block length, bailout < 4 (no unchecked blocks), a Julia c
    outside the bailout disc, corner sampling, odd sizes."""
    cu = gpu
    cases = [
        cu.make_tile(-2.0, -1.5, 1.0, 1.5, 37, 29, 1, 4.0),
        cu.make_tile(-2.0, -1.5, 1.0, 1.5, 37, 29, 7, 4.0),
        cu.make_tile(-2.0, -1.5, 1.0, 1.5, 41, 33, 300, 2.0),
        cu.make_tile(-1.0, -1.0, 1.0, 1.0, 45, 19, 200, 4.0, kind=cu.FRACTAL_JULIA, julia_c=(2.5, 0.5)),
        cu.make_tile(-2.0, -1.5, 1.0, 1.5, 64, 64, 500, 1e6, pattern=cu.SAMPLE_CORNER, spp=3),
        cu.make_tile(-0.75, -0.01, -0.74, 0.01, 17, 300, 20000, 4.0),
    ]
    for t in cases:
        for precision in (64, 32):
            counts, smooth, it = run_tile(cu, engine, t, precision, 0)
            want_c, want_s, want_it = oracle.escape_tile(t, precision)
            assert np.array_equal(counts, want_c)
            assert it == want_it
            assert_smooth_close(smooth, want_s)


def mixed_tiles(cu, rng, n, max_side=96):
    tiles = []
    offset = 0
    for k in range(n):
        w = int(rng.integers(1, max_side))
        h = int(rng.integers(1, max_side))
        julia = rng.random() < 0.3
        cx, cy = (-0.75, 0.1) if not julia else (0.0, 0.0)
        width = float(3 * (1e-4 / 3) ** rng.random())
        seg = cu.plane_segment_from_center(cx + rng.uniform(-0.05, 0.05), cy, width, w, h)
        jc = (0.8 * np.sqrt(rng.random()) * np.cos(2 * np.pi * rng.random()),
              0.8 * np.sqrt(rng.random()) * np.sin(2 * np.pi * rng.random()))
        t = cu.make_tile(*seg, w, h, int(256 * 16 ** rng.random()),
                         [4.0, 256.0][k % 2], kind=cu.FRACTAL_JULIA if julia else cu.FRACTAL_MANDELBROT,
                         julia_c=jc, spp=1 + k % 3, pattern=[cu.SAMPLE_CENTRE, cu.SAMPLE_JITTER][k % 2],
                         seed=1000 + k)
        t.out_offset = offset
        t.work_offset = offset
        offset += w * h
        tiles.append(t)
    return tiles, offset


def run_persistent(cu, engine, tiles, total, precision, variant, tune):
    arr = (cu.TileDesc * len(tiles))(*tiles)
    d_tiles = cu.DeviceBuffer(len(tiles) * 128)
    import ctypes as C
    cu.abi.call("cu_memcpy_htod", d_tiles.ptr, C.cast(arr, C.c_void_p), C.c_size_t(len(tiles) * 128))
    counts = cu.DeviceBuffer(total * 4)
    smooth = cu.DeviceBuffer(total * 8)
    iters = cu.IterationCounter()
    queue = cu.DeviceBuffer(8)
    engine.escape_tiles_persistent(precision, variant, d_tiles, len(tiles), total, counts, smooth,
                                   iters, queue, tune)
    cu.synchronize()
    return counts.download(np.uint32), smooth.download(np.float64), iters.total()


def oracle_tiles(oracle, tiles, total, precision):
    counts = np.zeros(total, dtype=np.uint32)
    smooth = np.zeros(total, dtype=np.float64)
    it = 0
    for t in tiles:
        off = t.out_offset
        t0 = type(t).from_buffer_copy(t)
        t0.out_offset = 0
        c, s, i = oracle.escape_tile(t0, precision)
        n = t.width * t.height
        counts[off:off + n] = c.ravel()
        smooth[off:off + n] = s.ravel()
        it += i
    return counts, smooth, it


@pytest.mark.parametrize("tune", [(16, 32, 256, 16), (8, 8, 1, 1), (32, 16, 7, 64), (16, 4, 33, 40)])
@pytest.mark.parametrize("precision", [64, 32])
def test_persistent_matches_oracle(gpu, engine, oracle, tune, precision):
    cu = gpu
    rng = np.random.default_rng(tune[0] * 1000 + tune[2] + precision)
    tiles, total = mixed_tiles(cu, rng, 24)
    tn = cu.Tune(*tune)
    counts, smooth, it = run_persistent(cu, engine, tiles, total, precision, 0, tn)
    want_c, want_s, want_it = oracle_tiles(oracle, tiles, total, precision)
    assert np.array_equal(counts, want_c)
    assert it == want_it
    assert_smooth_close(smooth, want_s)


def test_persistent_fused_equals_exact_on_safe_tiles(gpu, engine, oracle):
    cu = gpu
    rng = np.random.default_rng(77)
    tiles, total = mixed_tiles(cu, rng, 40)
    assert all(cu.tile_fused_safe(t, 64) for t in tiles)
    a = run_persistent(cu, engine, tiles, total, 64, 1, None)
    want_c, want_s, want_it = oracle_tiles(oracle, tiles, total, 64)
    assert np.array_equal(a[0], want_c)
    assert a[2] == want_it


def test_persistent_one_tile_equals_grid_kernel(gpu, engine):
    cu = gpu
    seg = cu.plane_segment_from_center(*C2_CENTER, 1e-4, 512, 512)
    t = cu.make_tile(*seg, 512, 512, 5000, 256.0, spp=4, pattern=cu.SAMPLE_JITTER, seed=9)
    g_counts, g_smooth, g_it = run_tile(cu, engine, t, 64, 1)
    p_counts, p_smooth, p_it = run_persistent(cu, engine, [t], 512 * 512, 64, 1, None)
    assert np.array_equal(p_counts.reshape(512, 512), g_counts)
    assert np.array_equal(p_smooth.reshape(512, 512), g_smooth)
    assert p_it == g_it


def test_compose_smooth_vs_oracle(gpu, engine, oracle):
    cu = gpu
    rng = np.random.default_rng(4)
    w, h = 333, 77
    smooth = rng.uniform(0, 5000, (h, w))
    smooth[0, :10] = 0
    pal = np.arange(256, dtype=np.uint32) * 0x010203 + 0xFF000000
    d_s = cu.DeviceBuffer.from_array(smooth)
    d_p = cu.DeviceBuffer.from_array(pal)
    d_o = cu.DeviceBuffer(w * h * 4)
    engine.compose_smooth(d_s, w, d_o, w, w, h, d_p, 256)
    cu.synchronize()
    got = d_o.download(np.uint32).reshape(h, w)
    assert np.array_equal(got, oracle.compose_smooth(smooth, pal))
    # a palette longer than the LDS stage is read from global memory
    pal2 = (np.arange(5000, dtype=np.uint32) * 2654435761).astype(np.uint32)
    d_p2 = cu.DeviceBuffer.from_array(pal2)
    engine.compose_smooth(d_s, w, d_o, w, w, h, d_p2, 5000)
    cu.synchronize()
    assert np.array_equal(d_o.download(np.uint32).reshape(h, w), oracle.compose_smooth(smooth, pal2))


def test_compose_smooth_known_slots(gpu, engine, oracle):
    """Palette slots of values the oracle's known answers fix: half-integers round
    away from zero, values at and past L wrap, values below zero take slot L - 1."""
    import oracle_pins as P
    cu = gpu
    vals = np.array([v for v, _ in P.COMPOSE_KNOWN] + [-1e9, -0.5, 0.49999], np.float64)
    want_slots = [k for _, k in P.COMPOSE_KNOWN] + [255, 255, 255]
    smooth = np.tile(vals, (3, 1))
    pal = (0x10000 + np.arange(256)).astype(np.uint32)
    d_s = cu.DeviceBuffer.from_array(smooth)
    d_p = cu.DeviceBuffer.from_array(pal)
    d_o = cu.DeviceBuffer(smooth.size * 4)
    engine.compose_smooth(d_s, vals.size, d_o, vals.size, vals.size, 3, d_p, 256)
    cu.synchronize()
    got = d_o.download(np.uint32).reshape(3, -1)
    assert list(got[1] - 0x10000) == want_slots
    assert np.array_equal(got, oracle.compose_smooth(smooth, pal))
