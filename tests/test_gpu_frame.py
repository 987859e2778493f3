"""The straggler-compaction frame pipeline on the GPU against the oracle.

Whatever the hand-off threshold, block length, number of tail passes or the
persistent drain of the queue, every sample must come out exactly as the
oracle's sequential loop computes it: escape counts and iteration totals
bit-exact, smooth values within 1e-12 relative.
"""
import numpy as np
import pytest

from test_gpu_escape import C2_CENTER, assert_smooth_close, mixed_tiles, oracle_tiles

pytestmark = pytest.mark.gpu


def render(cu, engine, tiles, precision, variant, tune, palette=None):
    f = cu.Frame(engine, tiles, precision, variant)
    try:
        rgba = None
        if palette is not None:
            pal = cu.DeviceBuffer.from_array(palette)
            rgba = cu.DeviceBuffer(f.total_pixels * 4)
            f.render(tune, pal, palette.size, rgba)
        else:
            f.render(tune)
        cu.synchronize()
        out = (f.counts(), f.smooth(), f.iterations())
        if rgba is not None:
            out = out + (rgba.download(np.uint32),)
        return out
    finally:
        f.close()


# (block_steps, handoff_lanes, tail_passes, drain_refill): drain_refill > 0 runs
# what the tail passes (all handing off) leave in one persistent launch
@pytest.mark.parametrize("tune", [(16, 16, 3, 0), (8, 16, 3, 0), (16, 1, 1, 0), (16, 32, 8, 0),
                                  (8, 4, 2, 0), (16, 24, 0, 16), (8, 32, 0, 64), (16, 8, 0, 1),
                                  (16, 24, 2, 16), (8, 8, 7, 32), (16, 24, 0, 40), (16, 16, 0, 48)])
@pytest.mark.parametrize("precision", [64, 32])
def test_c2_frame_pipeline_vs_oracle(gpu, engine, oracle, tune, precision):
    cu = gpu
    seg = cu.plane_segment_from_center(*C2_CENTER, 1e-4 / 8, 512, 384)
    tile = cu.make_tile(*seg, 512, 384, 5000, 256.0, spp=4, pattern=cu.SAMPLE_JITTER, seed=3)
    b, h, p, d = tune
    counts, smooth, it = render(cu, engine, [tile], precision, -1, cu.FrameTune(b, h, p, 0, d))
    want_c, want_s, want_it = oracle.escape_tile(tile, precision)
    assert np.array_equal(counts.reshape(384, 512), want_c)
    assert it == want_it
    assert_smooth_close(smooth.reshape(384, 512), want_s)


@pytest.mark.parametrize("precision", [64, 32])
def test_c2_full_frame_bands_vs_oracle(gpu, engine, oracle, precision):
    """The 4096x4096 4-spp frame itself (SplitMix64 seed 20261202, the benchmark's
    tile and tune) through the pipeline with compose; three row bands (top, centre,
    bottom: 384 rows each) run through the oracle on the host's threads and must
    match the GPU frame's counts, smooth values and palette colours."""
    cu = gpu
    seg = cu.plane_segment_from_center(*C2_CENTER, 1e-4, 4096, 4096)
    tile = cu.make_tile(*seg, 4096, 4096, 5000, 256.0, spp=4, pattern=cu.SAMPLE_JITTER,
                        seed=20261202)
    pal = ((np.arange(256, dtype=np.uint32) * 0x00010305) | 0xFF000000).astype(np.uint32)
    counts, smooth, it, rgba = render(cu, engine, [tile], precision, -1,
                                      cu.FrameTune(16, 24, 0, 0, 40), palette=pal)
    counts, smooth, rgba = (a.reshape(4096, 4096) for a in (counts, smooth, rgba))
    for r0 in (0, 1856, 3712):
        want_c, want_s, _ = oracle.escape_tile(tile, precision, rows=(r0, r0 + 384))
        assert np.array_equal(counts[r0:r0 + 384], want_c[r0:r0 + 384]), r0
        assert_smooth_close(smooth[r0:r0 + 384], want_s[r0:r0 + 384])
        want_rgba = oracle.compose_smooth(want_s[r0:r0 + 384], pal)
        assert np.array_equal(rgba[r0:r0 + 384], want_rgba), r0


def test_pass1_iterations_are_its_own(gpu, engine, oracle):
    """cu_frame_iterations' pass-1 figure counts the iterations pass 1 itself ran,
    including the steps of the samples it hands on, and the tail passes count only
    their own steps: the two add up to the oracle's total, and the later pass 1
    hands its stragglers on (1 lane left instead of 24), the more of the work it
    has done."""
    cu = gpu
    seg = cu.plane_segment_from_center(*C2_CENTER, 1e-4 / 8, 512, 384)
    tile = cu.make_tile(*seg, 512, 384, 5000, 256.0, spp=4, pattern=cu.SAMPLE_JITTER, seed=3)
    _, _, want_it = oracle.escape_tile(tile, 64)
    f = cu.Frame(engine, [tile], 64, -1)
    try:
        first = {}
        for handoff in (1, 24):
            f.render(cu.FrameTune(16, handoff, 3, 0, 0))
            cu.synchronize()
            assert f.iterations() == want_it
            first[handoff] = f.iterations_pass1()
    finally:
        f.close()
    assert 0.5 * want_it < first[24] < first[1] < want_it


def test_rgba_read_back_into_page_locked_memory(gpu, engine):
    """A frame's RGBA image copied into cu_host_alloc memory equals the same image
    copied into ordinary host memory; sizes past either buffer are refused."""
    cu = gpu
    seg = cu.plane_segment_from_center(*C2_CENTER, 1e-4 / 8, 256, 192)
    tile = cu.make_tile(*seg, 256, 192, 2000, 256.0, spp=4, pattern=cu.SAMPLE_JITTER, seed=3)
    palette = ((np.arange(256, dtype=np.uint32) * 0x00010305) | 0xFF000000).astype(np.uint32)
    *_, rgba = render(cu, engine, [tile], 64, -1, cu.FrameTune(16, 24, 3, 0, 0), palette)
    dev = cu.DeviceBuffer.from_array(rgba)
    host = cu.HostBuffer(rgba.nbytes)
    try:
        host.download_from(dev)
        assert np.array_equal(host.array(np.uint32), rgba)
        with pytest.raises(ValueError):
            host.download_from(dev, rgba.nbytes + 4)
    finally:
        host.close()
        dev.close()


def test_c1_frame_pipeline_matches_oracle(gpu, engine, oracle):
    cu = gpu
    tile = cu.make_tile(-2.0, -1.5, 1.0, 1.5, 1024, 1024, 1000, 4.0)
    counts, _, _ = render(cu, engine, [tile], 64, 0, cu.FrameTune(16, 16, 3, 0))
    want_c, _, _ = oracle.escape_tile(tile, 64)
    assert np.array_equal(counts.reshape(want_c.shape), want_c)


@pytest.mark.parametrize("precision", [64, 32])
def test_mixed_tiles_pipeline_vs_oracle(gpu, engine, oracle, precision):
    cu = gpu
    rng = np.random.default_rng(11 + precision)
    tiles, total = mixed_tiles(cu, rng, 40, max_side=120)
    for t in tiles:  # the pipeline needs spp a power of two
        t.spp = [1, 2, 4][t.spp % 3]
    want_c, want_s, want_it = oracle_tiles(oracle, tiles, total, precision)
    for tune in (cu.FrameTune(16, 8, 4, 0), cu.FrameTune(8, 16, 0, 2, 24)):
        counts, smooth, it = render(cu, engine, tiles, precision, 0, tune)
        assert np.array_equal(counts, want_c)
        assert it == want_it
        assert_smooth_close(smooth, want_s)


def test_fused_compose(gpu, engine, oracle):
    cu = gpu
    seg = cu.plane_segment_from_center(*C2_CENTER, 1e-3, 300, 200)
    tile = cu.make_tile(*seg, 300, 200, 2000, 256.0, spp=4, pattern=cu.SAMPLE_JITTER, seed=5)
    pal = ((np.arange(256, dtype=np.uint32) * 0x00010305) | 0xFF000000).astype(np.uint32)
    counts, smooth, it, rgba = render(cu, engine, [tile], 64, -1, None, palette=pal)
    want_c, want_s, _ = oracle.escape_tile(tile, 64)
    assert np.array_equal(counts.reshape(200, 300), want_c)
    assert np.array_equal(rgba.reshape(200, 300), oracle.compose_smooth(want_s, pal))


@pytest.mark.parametrize("precision", [64, 32])
def test_wide_julia_negative_smooth_compose(gpu, engine, oracle, precision):
    """A Julia view 40 wide: orbits starting at |z0| > 7 leave at step 0 with a
    negative smooth value, which takes the palette slot of 0 on both sides."""
    cu = gpu
    tile = cu.make_tile(-20.0, -20.0, 20.0, 20.0, 160, 160, 300, 4.0, kind=cu.FRACTAL_JULIA,
                        julia_c=(-0.4, 0.6))
    pal = ((np.arange(256, dtype=np.uint32) * 0x00010305) | 0xFF000000).astype(np.uint32)
    counts, smooth, it, rgba = render(cu, engine, [tile], precision, -1, None, palette=pal)
    want_c, want_s, _ = oracle.escape_tile(tile, precision)
    assert (want_s < -0.5).sum() > 1000
    assert np.array_equal(counts.reshape(160, 160), want_c)
    assert np.array_equal(rgba.reshape(160, 160), oracle.compose_smooth(want_s, pal))


def test_frame_rejects(gpu, engine):
    cu = gpu
    t = cu.make_tile(-2, -1, 1, 1, 16, 16, 10, spp=3)
    with pytest.raises(cu.IllegalArgumentError, match="power of two"):
        cu.Frame(engine, [t])
    t = cu.make_tile(-2, -1, 1, 1, 16, 16, 10)
    f = cu.Frame(engine, [t])
    try:
        with pytest.raises(cu.IllegalArgumentError, match="handoff_lanes"):
            f.render(cu.FrameTune(16, 40, 2, 0))
        with pytest.raises(cu.IllegalArgumentError, match="tail_passes"):
            f.render(cu.FrameTune(16, 16, 0, 0, 0))
        with pytest.raises(cu.IllegalArgumentError, match="drain_refill"):
            f.render(cu.FrameTune(16, 16, 0, 0, 65))
        with pytest.raises(cu.IllegalArgumentError, match="before a drain"):
            f.render(cu.FrameTune(16, 16, 8, 0, 16))
        with pytest.raises(cu.IllegalStateError):
            f.iterations()
    finally:
        f.close()


@pytest.mark.parametrize("spp,w,h", [(4, 101, 37), (8, 64, 40), (16, 48, 24), (64, 16, 12), (1, 77, 33)])
def test_pipeline_spp_and_ragged_sizes(gpu, engine, oracle, spp, w, h):
    """Every power-of-two spp, and frames whose pixel count is not a multiple of
    4 (the combine's general path) or whose width is not a multiple of the wave
    block, against the oracle; hand-off after the first lanes finish so that many
    pixels are completed by the tail passes and the combine."""
    cu = gpu
    seg = cu.plane_segment_from_center(*C2_CENTER, 1e-4, w, h)
    tile = cu.make_tile(*seg, w, h, 3000, 256.0, spp=spp, pattern=cu.SAMPLE_JITTER, seed=spp)
    want_c, want_s, want_it = oracle.escape_tile(tile, 64)
    pal = ((np.arange(256, dtype=np.uint32) * 0x00010305) | 0xFF000000).astype(np.uint32)
    for tune in (cu.FrameTune(16, 32, 2, 0), cu.FrameTune(8, 16, 0, 0, 8)):
        counts, smooth, it, rgba = render(cu, engine, [tile], 64, -1, tune, palette=pal)
        assert np.array_equal(counts.reshape(h, w), want_c)
        assert it == want_it
        assert_smooth_close(smooth.reshape(h, w), want_s)
        assert np.array_equal(rgba.reshape(h, w), oracle.compose_smooth(want_s, pal))
