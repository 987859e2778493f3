"""Autotuner for batched tile renders: picks the engine and its launch shape by
measuring them on the batch itself.

Candidates:
  * the straggler-compaction frame pipeline (cu_frame): pass-1 wave shape (the
    pixel block a wave covers: 8x8 .. 64x1 pixels at 1 spp), unchecked block
    length, hand-off threshold (lanes still running when a wave hands its
    stragglers to the queue), number of tail passes, and whether one persistent
    drain launch runs what the tail passes leave instead of the last of them
    running it to the end (its idle-lane refill threshold), or the whole queue
    with no tail pass;
  * the persistent lane-refill kernel (cu_escape_tiles_persistent): block
    length, resident waves per CU, pixels taken per queue grab and the number
    of idle lanes that trigger a refill.

Both engines are searched in the precision asked for (FP64 or FP32).  The
search is coordinate descent from a default point, one knob at a time in
the order above, each candidate timed with HIP events on its stream (median of
`repeats` after `warmup` renders) right after the incumbent is timed again, so
that the two are compared at the same GPU clock.  Every candidate renders the same pixels;
the caller can check their outputs against each other.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

PIPELINE_SPACE = {
    "wave_shape": (0, 1, 2, 3),
    "block_steps": (16, 8),
    "handoff_lanes": (8, 16, 24, 32),
    "tail_passes": (3, 5, 7),
    "drain_refill": (0, 16, 32),
}
PIPELINE_START = {"block_steps": 16, "handoff_lanes": 16, "tail_passes": 5, "wave_shape": 0,
                  "drain_refill": 0}
DRAIN_ONLY_REFILL = (24, 32, 40, 48)  # tried after the descent with tail_passes 0

PERSISTENT_SPACE = {
    "block_steps": (16, 8, 32),
    "waves_per_cu": (32, 16, 24),
    "chunk": (256, 64, 1024),
    "refill_lanes": (16, 8, 32),
}
PERSISTENT_START = {"block_steps": 16, "waves_per_cu": 32, "chunk": 256, "refill_lanes": 16}


def frame_tune(cu, p):
    """cu.FrameTune from a dict of the pipeline knobs."""
    return cu.FrameTune(p["block_steps"], p["handoff_lanes"], p["tail_passes"], p["wave_shape"],
                        p.get("drain_refill", 0))


def _median_ms(fn, repeats, warmup):
    for _ in range(warmup):
        fn()
    ts = [fn() for _ in range(repeats)]
    return float(np.median(ts))


def _challenge(best, best_ms, cand, measure, table, log):
    """Time `cand` right after timing the incumbent again, so that both run at the
    same clock (the GPU's clock drifts over a search, and a process's first renders
    run while it ramps up); returns the winner and its time."""
    best_now = measure(best)
    ms = measure(cand)
    table.append((cand, ms))
    if log:
        log(f"[autotune] {cand} {ms:.3f} ms (incumbent {best_now:.3f} ms)")
    if ms < best_now:
        return cand, ms
    return best, min(best_ms, best_now)


def _descend(space, start, measure, log):
    best = dict(start)
    best_ms = measure(best)
    table = [(dict(best), best_ms)]
    for knob, values in space.items():
        for v in values:
            if v == best[knob]:
                continue
            cand = dict(best)
            cand[knob] = v
            best, best_ms = _challenge(best, best_ms, cand, measure, table, log)
    return best, best_ms, table


def search_pipeline(measure, log=None):
    """Coordinate descent over PIPELINE_SPACE, then the drain over the whole
    pass-1 queue (tail_passes 0), which one knob at a time cannot reach: tail
    passes 0 needs a drain.  `measure(knobs) -> ms`; returns (best, ms, table)."""
    best, best_ms, table = _descend(PIPELINE_SPACE, PIPELINE_START, measure, log)
    for refill in DRAIN_ONLY_REFILL:
        cand = dict(best, tail_passes=0, drain_refill=refill)
        best, best_ms = _challenge(best, best_ms, cand, measure, table, log)
    return best, best_ms, table


class PersistentBatch:
    """Device copy of a tile batch laid out back to back for the persistent kernel."""

    def __init__(self, cu, tiles, precision=64):
        self.cu = cu
        self.precision = precision
        arr = (cu.TileDesc * len(tiles))()
        offset = 0
        for k, t in enumerate(tiles):
            C.memmove(C.byref(arr[k]), C.byref(t), C.sizeof(t))
            arr[k].work_offset = offset
            arr[k].out_offset = offset
            arr[k].out_pitch = t.width
            offset += t.width * t.height
        self.n, self.total = len(tiles), offset
        self.d_tiles = cu.DeviceBuffer(C.sizeof(arr))
        cu.abi.call("cu_memcpy_htod", self.d_tiles.ptr, C.cast(arr, C.c_void_p), C.sizeof(arr))
        self.counts = cu.DeviceBuffer(offset * 4)
        self.smooth = cu.DeviceBuffer(offset * 8)
        self.iters = cu.IterationCounter()
        self.queue = cu.DeviceBuffer(8)
        self.variant = 1 if all(cu.tile_fused_safe(t, precision) for t in tiles) else 0

    def render(self, engine, tune) -> float:
        self.iters.clear()
        engine.escape_tiles_persistent(self.precision, self.variant, self.d_tiles, self.n, self.total,
                                       self.counts, self.smooth, self.iters, self.queue, tune)
        return engine.last_kernel_ms()

    def capture(self, engine, tune):
        """The persistent launch (clears included) as one replayable HIP graph."""
        g = C.c_void_p()
        self.cu.abi.call("cu_escape_tiles_persistent_capture", engine.handle, self.precision, self.variant,
                         self.d_tiles.ptr, self.n, C.c_uint64(self.total), self.counts.ptr,
                         self.smooth.ptr, self.iters.ptr, self.queue.ptr, C.byref(tune), C.byref(g))
        return CapturedLaunch(self.cu, g)


class CapturedLaunch:
    """A captured launch sequence (cu_graph)."""

    def __init__(self, cu, handle):
        self.cu, self.handle = cu, handle

    def replay(self, stream=None) -> None:
        self.cu.abi.call("cu_graph_launch", self.handle, stream.handle if stream else None)

    def last_ms(self) -> float:
        ms = C.c_float(0)
        self.cu.abi.call("cu_graph_last_ms", self.handle, C.byref(ms))
        return ms.value

    def close(self) -> None:
        if self.handle and self.handle.value:
            self.cu.abi.call("cu_graph_destroy", self.handle)
            self.handle = C.c_void_p()


def tune_batch(cu, engine, tiles, precision=64, repeats=3, warmup=1, persistent=True, log=None):
    """Measure both engines on `tiles`; returns a dict with the winner and the tables."""
    frame = cu.Frame(engine, tiles, precision, -1)

    def measure_pipeline(p):
        tn = frame_tune(cu, p)

        def one():
            frame.render(tn)
            return frame.last_ms()[1]
        return _median_ms(one, repeats, warmup)

    pbest, pms, ptable = search_pipeline(measure_pipeline, log)
    result = {"pipeline": {"best": pbest, "ms": pms, "table": ptable}}
    frame.close()
    best = ("pipeline", pbest, pms)
    if persistent:
        batch = PersistentBatch(cu, tiles, precision)

        def measure_persistent(p):
            tn = cu.Tune(p["block_steps"], p["waves_per_cu"], p["chunk"], p["refill_lanes"])
            return _median_ms(lambda: batch.render(engine, tn), repeats, warmup)

        qbest, qms, qtable = _descend(PERSISTENT_SPACE, PERSISTENT_START, measure_persistent, log)
        result["persistent"] = {"best": qbest, "ms": qms, "table": qtable}
        if qms < pms:
            best = ("persistent", qbest, qms)
    result["engine"], result["tune"], result["ms"] = best
    return result
