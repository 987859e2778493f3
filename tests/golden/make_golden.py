"""Write the golden fixtures under tests/golden/ from the reference's own code.

Run here (where /root/reference exists): python tests/golden/make_golden.py

Sources of truth:
  * computeFractal of src/main/cuda/fractals/mandelbrot.cu and julia.cu, compiled
    for the host with FP contraction off by oracle/Makefile (target `ref`, output
    oracle/_ref/libref.so).  Only input/output vectors are written: full-frame
    SHA-256 digests of the uint32 escape counts plus sampled (px, py, count) rows,
    and per-point outputs.
  * palettes/*.png of the reference (image data): copied to tests/golden/palettes/
    as inputs of the palette loader, and their first rows as ARGB words the way
    java.awt.image.BufferedImage.getRGB returns them (decoded by PIL).
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFERENCE = os.environ.get("REFERENCE", "/root/reference")
sys.path.insert(0, os.path.dirname(HERE))

import _oracle  # noqa: E402


def build_ref():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), f"REFERENCE={REFERENCE}",
                    "ref"], check=True)
    lib = _oracle.ref_lib()
    if lib is None:
        raise SystemExit("oracle/_ref/libref.so did not build")
    return lib


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u4").tobytes()).hexdigest()


def samples(counts: np.ndarray, rng, n=2048):
    h, w = counts.shape
    pts = set()
    for x in (0, w - 1):
        for y in (0, h - 1):
            pts.add((x, y))
    while len(pts) < n:
        pts.add((int(rng.integers(w)), int(rng.integers(h))))
    return [[x, y, int(counts[y, x])] for x, y in sorted(pts)]


def frame(fn, w, h, left, top, pw, ph, real):
    counts = np.zeros((h, w), dtype=np.uint32)
    for py in range(h):
        for px in range(w):
            if real is np.float32:
                x = np.float32(left) + (np.float32(px) + np.float32(0.5)) * np.float32(pw)
                y = np.float32(top) - (np.float32(py) + np.float32(0.5)) * np.float32(ph)
            else:
                x = left + (px + 0.5) * pw
                y = top - (py + 0.5) * ph
            counts[py, px] = int(fn(float(x), float(y)))
    return counts


def main():
    ref = build_ref()
    rng = np.random.default_rng(20261202)
    out = {}

    # C1: plain Mandelbrot, 1024x1024 pixel centres over [-2,1]x[-1.5,1.5], maxIter 1000
    ps = 3.0 / 1024
    for name, fn, real in (("fp64", ref.ref_mandelbrot_d, np.float64),
                           ("fp32", ref.ref_mandelbrot_f, np.float32)):
        counts = frame(lambda x, y: fn(1000, x, y), 1024, 1024, -2.0, 1.5, ps, ps, real)
        out[f"c1_mandelbrot_{name}"] = {
            "width": 1024, "height": 1024, "left": -2.0, "top": 1.5, "pixel": ps,
            "max_iter": 1000, "sampling": "pixel centres", "count_convention": "computeFractal",
            "sha256_counts_u32le": sha(counts), "samples": samples(counts, rng),
            "sum": int(counts.astype(np.uint64).sum()),
        }
        print(name, "c1 sum", out[f"c1_mandelbrot_{name}"]["sum"], file=sys.stderr)

    # Julia module default c (ModuleJulia.java:41), 512x512 over [-1.5,1.5]^2, maxIter 900
    jc = (-0.4, 0.6)
    ps = 3.0 / 512
    for name, fn, real in (("fp64", ref.ref_julia_d, np.float64), ("fp32", ref.ref_julia_f, np.float32)):
        counts = frame(lambda x, y: fn(900, x, y, jc[0], jc[1]), 512, 512, -1.5, 1.5, ps, ps, real)
        out[f"julia_{name}"] = {
            "width": 512, "height": 512, "left": -1.5, "top": 1.5, "pixel": ps, "max_iter": 900,
            "julia_c": list(jc), "sampling": "pixel centres",
            "sha256_counts_u32le": sha(counts), "samples": samples(counts, rng, 1024),
        }

    # random points, both modules, both precisions
    pts = []
    for _ in range(2000):
        cx, cy = float(rng.uniform(-2.2, 0.8)), float(rng.uniform(-1.3, 1.3))
        pts.append([cx, cy,
                    float(ref.ref_mandelbrot_d(2000, cx, cy)),
                    float(ref.ref_mandelbrot_f(2000, cx, cy)),
                    float(ref.ref_julia_d(2000, 0.8 * cx, cy, -0.8, 0.156)),
                    float(ref.ref_julia_f(2000, 0.8 * cx, cy, -0.8, 0.156))])
    out["points"] = {"max_iter": 2000, "julia_c": [-0.8, 0.156], "julia_z_scale_x": 0.8,
                     "columns": ["cx", "cy", "mandel_d", "mandel_f", "julia_d", "julia_f"],
                     "rows": pts}

    with open(os.path.join(HERE, "reference_iteration.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))

    # palettes (image data of the reference): the PNG files as loader inputs, and
    # their first rows decoded by PIL as the expected ARGB words
    from PIL import Image
    import shutil
    pal = {}
    os.makedirs(os.path.join(HERE, "palettes"), exist_ok=True)
    for fn in sorted(os.listdir(os.path.join(REFERENCE, "palettes"))):
        shutil.copyfile(os.path.join(REFERENCE, "palettes", fn), os.path.join(HERE, "palettes", fn))
        im = Image.open(os.path.join(REFERENCE, "palettes", fn)).convert("RGB")
        w, _ = im.size
        argb = [(0xFF << 24) | (r << 16) | (g << 8) | b for r, g, b in (im.getpixel((x, 0)) for x in range(w))]
        pal[fn] = argb
    with open(os.path.join(HERE, "palettes_argb.json"), "w") as f:
        json.dump(pal, f, separators=(",", ":"))
    print("wrote fixtures", file=sys.stderr)


if __name__ == "__main__":
    main()
