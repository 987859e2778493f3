"""The uneven-tile batch (256 seeded tiles, 64..1024 px, maxIter 256..65536, 30 %
Julia) on the GPU.

* a subset of the batch through the frame pipeline, every wave shape, against
  the oracle: counts bit-exact, smooth within 1e-12 relative;
* the whole batch in one pipeline render equals each tile rendered alone by the
  grid kernel, and the persistent lane-refill kernel;
* a captured render replayed several times, with other work in between, gives
  the eager result every time;
* the autotuner's pick renders the same pixels.
"""
import importlib

import numpy as np
import pytest

from test_gpu_escape import assert_smooth_close, oracle_tiles

pytestmark = pytest.mark.gpu

wl = importlib.import_module("chaos-ultra_amd.workloads")
at = importlib.import_module("chaos-ultra_amd.autotune")

_BATCH = {}


def batch(cu):
    if "tiles" not in _BATCH:
        _BATCH["tiles"], _BATCH["info"] = wl.uneven_tiles(cu, 256)
    return _BATCH["tiles"], _BATCH["info"]


def light_subset(cu, n=12):
    """The n tiles of the batch with the least side^2 * maxIter."""
    tiles, info = batch(cu)
    order = np.argsort([i["side"] ** 2 * i["max_iter"] for i in info])[:n]
    return [tiles[k] for k in order]


def with_offsets(tiles):
    offset = 0
    for t in tiles:
        t.out_offset = t.work_offset = offset
        t.out_pitch = t.width
        offset += t.width * t.height
    return offset


def render_frame(cu, engine, tiles, tune, precision=64):
    f = cu.Frame(engine, tiles, precision, -1)
    try:
        f.render(tune)
        cu.synchronize()
        return f.counts(), f.smooth(), f.iterations()
    finally:
        f.close()


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
def test_subset_pipeline_vs_oracle(gpu, engine, oracle, shape):
    cu = gpu
    tiles = light_subset(cu)
    counts, smooth, it = render_frame(cu, engine, tiles, cu.FrameTune(16, 16, 5, shape))
    total = with_offsets(tiles)
    want_c, want_s, want_it = oracle_tiles(oracle, tiles, total, 64)
    assert np.array_equal(counts, want_c)
    assert it == want_it
    assert_smooth_close(smooth, want_s)


def test_batch_pipeline_equals_grid_and_persistent(gpu, engine):
    cu = gpu
    tiles, _ = batch(cu)
    counts, smooth, it = render_frame(cu, engine, tiles, cu.FrameTune(16, 16, 5, 0))
    total = with_offsets(tiles)
    off = 0
    it_grid = 0
    for t in tiles:
        n = t.width * t.height
        c = cu.DeviceBuffer(n * 4)
        s = cu.DeviceBuffer(n * 8)
        ic = cu.IterationCounter()
        t1 = type(t).from_buffer_copy(t)
        t1.out_offset = 0
        engine.escape_tile(64, t1, c, s, ic)
        assert np.array_equal(c.download(np.uint32), counts[off:off + n])
        assert np.array_equal(s.download(np.float64), smooth[off:off + n])
        it_grid += ic.total()
        off += n
    assert it_grid == it
    pb = at.PersistentBatch(cu, tiles)
    pb.render(engine, cu.Tune(16, 32, 256, 16))
    cu.synchronize()
    assert pb.total == total
    assert np.array_equal(pb.counts.download(np.uint32), counts)
    assert pb.iters.total() == it


def test_captured_render_replays(gpu, engine):
    cu = gpu
    tiles = light_subset(cu, 24)
    pal = cu.DeviceBuffer.from_array(((np.arange(64, dtype=np.uint32) * 0x030201) | 0xFF000000)
                                     .astype(np.uint32))
    f = cu.Frame(engine, tiles, 64, -1)
    other = cu.Frame(engine, light_subset(cu, 6), 64, -1)
    try:
        rgba = cu.DeviceBuffer(f.total_pixels * 4)
        tn = cu.FrameTune(8, 24, 3, 1)
        f.render(tn, pal, 64, rgba)
        cu.synchronize()
        want = (f.counts(), f.smooth(), rgba.download(np.uint32), f.iterations())
        f.capture(tn, pal, 64, rgba)
        for k in range(4):
            rgba.fill_u32(0)
            other.render(cu.FrameTune(16, 8, 2, k % 4))  # eager work between replays
            cu.synchronize()
            f.replay()
            cu.synchronize()
            got = (f.counts(), f.smooth(), rgba.download(np.uint32), f.iterations())
            assert np.array_equal(got[0], want[0])
            assert np.array_equal(got[1], want[1])
            assert np.array_equal(got[2], want[2])
            assert got[3] == want[3]
            assert f.last_ms()[0] == 0.0 and f.last_ms()[1] > 0.0
    finally:
        f.close()
        other.close()


def test_autotuner_pick_renders_the_same_pixels(gpu, engine):
    cu = gpu
    tiles = light_subset(cu, 32)
    res = at.tune_batch(cu, engine, tiles, repeats=1, warmup=0)
    assert res["engine"] in ("pipeline", "persistent") and res["ms"] > 0
    assert len(res["pipeline"]["table"]) >= 10
    p = res["pipeline"]["best"]
    got, _, _ = render_frame(cu, engine, tiles, at.frame_tune(cu, p))
    want, _, _ = render_frame(cu, engine, tiles, cu.FrameTune(16, 16, 5, 0))
    assert np.array_equal(got, want)


def test_autotuner_fp32_searches_both_engines(gpu, engine):
    """The FP32 search times the persistent kernel too, and whichever engine wins
    renders the batch's pixels exactly as the FP32 pipeline does."""
    cu = gpu
    tiles = light_subset(cu, 16)
    res = at.tune_batch(cu, engine, tiles, precision=32, repeats=1, warmup=0)
    assert "persistent" in res and len(res["persistent"]["table"]) >= 6
    want, want_s, want_it = render_frame(cu, engine, tiles, cu.FrameTune(16, 16, 5, 0), 32)
    q = res["persistent"]["best"]
    batch = at.PersistentBatch(cu, tiles, 32)
    batch.render(engine, cu.Tune(q["block_steps"], q["waves_per_cu"], q["chunk"],
                                 q["refill_lanes"]))
    cu.synchronize()
    assert np.array_equal(batch.counts.download(np.uint32), want)
    assert batch.iters.total() == want_it


def test_persistent_capture_replays(gpu, engine):
    """The autotuned engine's persistent launch as one captured graph: replays with
    eager work between them give the eager launch's pixels and iteration count."""
    cu = gpu
    at = importlib.import_module("chaos-ultra_amd.autotune")
    tiles = light_subset(cu, 24)
    batch = at.PersistentBatch(cu, tiles)
    tn = cu.Tune(16, 16, 64, 8)
    batch.render(engine, tn)
    cu.synchronize()
    want = (batch.counts.download(np.uint32), batch.smooth.download(np.float64), batch.iters.total())
    g = batch.capture(engine, tn)
    other = at.PersistentBatch(cu, light_subset(cu, 5))
    try:
        for k in range(4):
            batch.counts.fill_u32(0)
            other.render(engine, cu.Tune(8, 32, 256, 16))  # eager work between replays
            cu.synchronize()
            g.replay()
            assert g.last_ms() > 0.0
            got = (batch.counts.download(np.uint32), batch.smooth.download(np.float64),
                   batch.iters.total())
            assert np.array_equal(got[0], want[0])
            assert np.array_equal(got[1], want[1])
            assert got[2] == want[2]
    finally:
        g.close()
