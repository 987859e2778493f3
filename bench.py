"""Benchmark of the escape-time render path on one MI355X.

Workload (one step): the 4096x4096 plain-Mandelbrot frame at centre
(-0.743643887037151, 0.131825904205330), real-axis width 1e-4, 4 seeded-jitter
samples per pixel, maxIter 5000, bailout |z|^2 >= 256, rendered in FP64 with the
reference's rounding (bit-exact escape counts) by the persistent lane-refill
kernel, then composed to RGBA8 through a 256-entry palette.  Inputs are resident
on the device before timing.  value = Giga-iterations/s (iterations executed,
an interior sample counts maxIter) over the timed steps.

Also reported: the FP32 variant of the same frame, the thread-per-pixel kernel,
the roofline of the dominant kernel (algorithmic FP64 FLOPs = 8 per iteration
against the 78.6 TFLOP/s vector peak; its add/mul-only ceiling is half that) and
the oracle's CPU rate on a bounded row sample of the same frame, on the host's cores.

python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Giga-iter/s (FP32 perturb vs FP64 direct, % vector peak); p50/p99 frame ms"
CENTER = (-0.743643887037151, 0.131825904205330)
W = H = 4096
SPP = 4
MAX_ITER = 5000
BAILOUT_SQ = 256.0
SEED = 20261202
FLOPS_PER_ITER = 8
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def frame_tile(cu):
    seg = cu.plane_segment_from_center(CENTER[0], CENTER[1], 1e-4, W, H)
    return seg, cu.make_tile(*seg, W, H, MAX_ITER, BAILOUT_SQ, spp=SPP, pattern=cu.SAMPLE_JITTER,
                             seed=SEED)


class Frame:
    """Device buffers of one C2 frame and its launches."""

    def __init__(self, cu, engine, tile, precision, tune):
        import ctypes as C
        self.cu, self.engine, self.tile, self.precision, self.tune = cu, engine, tile, precision, tune
        n = tile.width * tile.height
        self.counts = cu.DeviceBuffer(n * 4)
        self.smooth = cu.DeviceBuffer(n * 8)
        self.rgba = cu.DeviceBuffer(n * 4)
        self.iters = cu.IterationCounter()
        self.queue = cu.DeviceBuffer(8)
        pal = (np.arange(256, dtype=np.uint32) * 0x00010305) | 0xFF000000
        self.palette = cu.DeviceBuffer.from_array(pal.astype(np.uint32))
        self.d_tile = cu.DeviceBuffer(128)
        cu.abi.call("cu_memcpy_htod", self.d_tile.ptr, C.byref(tile), C.c_size_t(128))
        self.variant = 1 if cu.tile_fused_safe(tile, precision) else 0
        self.pipeline = cu.Frame(engine, [tile], precision, -1)
        self.frame_tune = cu.FrameTune(16, 16, 3, 0)
        self.render_ms = []   # dominant kernel per step (HIP events on its stream)
        self.total_ms = []    # whole render per step
        self.pipe_iters = 0

    def step(self, kernel: str):
        cu, e, t = self.cu, self.engine, self.tile
        if kernel == "pipeline":
            self.pipeline.render(self.frame_tune, self.palette, 256, self.rgba)
            p1, tot = self.pipeline.last_ms()
            self.render_ms.append(p1)
            self.total_ms.append(tot)
            return
        if kernel == "persistent":
            e.escape_tiles_persistent(self.precision, self.variant, self.d_tile, 1,
                                      t.width * t.height, self.counts, self.smooth, self.iters,
                                      self.queue, self.tune)
        else:
            e.escape_tile(self.precision, t, self.counts, self.smooth, self.iters, self.variant)
        self.render_ms.append(e.last_kernel_ms())
        self.total_ms.append(self.render_ms[-1])
        e.compose_smooth(self.smooth, t.width, self.rgba, t.width, t.width, t.height, self.palette, 256)

    def iterations(self, kernel: str, steps: int) -> int:
        if kernel == "pipeline":  # the pipeline clears its counter per render
            return self.pipeline.iterations() * steps
        return self.iters.total()


def timed(cu, frame: Frame, kernel: str, steps: int, warmup: int):
    for _ in range(warmup):
        frame.step(kernel)
    cu.synchronize()
    frame.iters.clear()
    frame.render_ms = []
    frame.total_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        frame.step(kernel)
    cu.synchronize()
    dt = time.perf_counter() - t0
    return dt, frame.iterations(kernel, steps)


def cpu_baseline(cu, tile, seconds_target=15.0):
    """The oracle (a scalar C++ restatement of the reference iteration, OpenMP over rows)
    on a bounded row band of the same frame."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    # probe 4 rows to size the sample, then time a band of rows around the centre
    r0 = H // 2
    t0 = time.perf_counter()
    _oracle.escape_tile(tile, 64, rows=(r0, r0 + 4), threads=threads)
    probe = max(time.perf_counter() - t0, 1e-3)
    rows = int(max(4, min(H, 4 * seconds_target / probe)))
    rows = min(rows, H)
    a = max(0, r0 - rows // 2)
    t0 = time.perf_counter()
    _, _, it = _oracle.escape_tile(tile, 64, rows=(a, a + rows), threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(it / dt / 1e9, 4), "unit": "Giga-iter/s", "cores": threads,
            "kind": "port",
            "sample": f"rows {a}..{a + rows - 1} of the 4096x4096 FP64 frame (4 spp, maxIter 5000), "
                      f"{it} iterations in {dt:.1f} s, g++ -O2 -ffp-contract=off, OpenMP {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kernel", choices=["pipeline", "persistent", "grid"], default="pipeline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-variants", action="store_true", help="skip the FP32 / grid variants")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after timing, copy this many MiB device-to-device once (a plain copy "
                         "kernel of known size that scales the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    cu = importlib.import_module("chaos-ultra_amd")
    cu.init(0)
    engine = cu.Engine()
    seg, tile = frame_tile(cu)
    tune = cu.Tune(16, 32, 256, 16)

    log(f"[bench] FP64 frame {W}x{H} spp={SPP} maxIter={MAX_ITER} kernel={args.kernel}")
    f64 = Frame(cu, engine, tile, 64, tune)
    dt, iters = timed(cu, f64, args.kernel, args.steps, args.warmup)
    value = iters / dt / 1e9
    ms_step = dt / args.steps * 1e3
    kern_ms = float(np.mean(f64.render_ms))
    per_launch_iters = iters / args.steps
    achieved_tflops = per_launch_iters * FLOPS_PER_ITER / (kern_ms * 1e-3) / 1e12
    log(f"[bench] FP64 {value:.1f} G-iter/s, {ms_step:.2f} ms/step, render kernel {kern_ms:.2f} ms")

    variants = {}
    if not args.no_variants:
        for other in ("pipeline", "grid", "persistent"):
            if other == args.kernel:
                continue
            dt2, it2 = timed(cu, f64, other, max(2, args.steps // 2), 1)
            variants[f"fp64_{other}"] = {"giga_iter_s": round(it2 / dt2 / 1e9, 2),
                                         "render_ms": round(float(np.mean(f64.render_ms)), 3)}
            log(f"[bench] FP64 {other}: {variants[f'fp64_{other}']}")
        for ft in ((8, 16, 3), (16, 8, 3), (16, 24, 3), (16, 16, 5)):
            f64.frame_tune = cu.FrameTune(*ft, 0)
            dt4, it4 = timed(cu, f64, "pipeline", max(2, args.steps // 2), 1)
            key = "fp64_pipeline_K%d_h%d_t%d" % ft
            variants[key] = {"giga_iter_s": round(it4 / dt4 / 1e9, 2),
                             "pass1_ms": round(float(np.mean(f64.render_ms)), 3),
                             "render_ms": round(float(np.mean(f64.total_ms)), 3)}
            log(f"[bench] {key}: {variants[key]}")
        f64.frame_tune = cu.FrameTune(16, 16, 3, 0)
        f32 = Frame(cu, engine, tile, 32, tune)
        dt3, it3 = timed(cu, f32, args.kernel, max(2, args.steps // 2), 1)
        variants["fp32_plain"] = {"giga_iter_s": round(it3 / dt3 / 1e9, 2),
                                  "render_ms": round(float(np.mean(f32.render_ms)), 3),
                                  "note": "FP32 c-grid at this zoom is coarser than the pixel "
                                          "(pixel 2.4e-8 < ulp_f32 5.96e-8)"}
        log(f"[bench] FP32: {variants['fp32_plain']}")

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Giga-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain Mandelbrot FP64 4096x4096, 4 jittered spp, centre "
                               "(-0.743643887037151, 0.131825904205330), width 1e-4, maxIter 5000, "
                               "bailout |z|^2>=256, smooth value + RGBA8 compose (256-entry palette)",
                   "kernel": args.kernel, "variant": "fused" if f64.variant else "exact",
                   "tune": {"block_steps": tune.block_steps, "waves_per_cu": tune.waves_per_cu,
                            "chunk": tune.chunk, "refill_lanes": tune.refill_lanes},
                   "iterations_per_step": int(per_launch_iters)},
        "roofline": {"bound": "valu", "achieved": round(achieved_tflops, 2),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved_tflops / FP64_PEAK_TFLOPS, 4),
                     "frac_of_add_mul_ceiling": round(achieved_tflops / (FP64_PEAK_TFLOPS / 2), 4),
                     "kernel_ms": round(kern_ms, 3), "flops_per_iteration": FLOPS_PER_ITER,
                     "traffic": None},
        "variants": variants,
    }
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle, bounded sample)")
        result["cpu_baseline"] = cpu_baseline(cu, tile)
        log(f"[bench] CPU {result['cpu_baseline']['value']} G-iter/s")
    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib << 20
        src, dst = cu.DeviceBuffer(n), cu.DeviceBuffer(n)
        src.fill_u32(1)
        cu.abi.call("cu_memcpy_dtod", dst.ptr, src.ptr, n)
        cu.synchronize()
        log(f"[bench] copied {args.calibrate_copy_mib} MiB device-to-device for counter calibration")
    engine.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
