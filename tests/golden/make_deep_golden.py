"""Write tests/golden/deep_dd_subset.npz: the double-double direct oracle on a
seeded 65 536-pixel subset of the 1920x1080 deep-zoom frames.

Frames: Mandelbrot, centre (-0.743643887037151, 0.131825904205330), real-axis
widths 1e-6, 1e-10 and 1e-20 over 1920 px (square pixels), maxIter 50000,
bailout |z|^2 >= 256.  Pixel (px, py) samples
centre + ((px + 0.5 - 960) * pw, (540 - (py + 0.5)) * pw).

The subset is numpy.random.default_rng(SEED).choice(1920*1080, 65536, replace=False),
sorted; the fixture stores the indices, the escape counts (Mandelbrot convention:
maxIter -> 0) and the smooth value n + 1 - log2(ln|z|^2 / ln 256) as float64.

Run: python tests/golden/make_deep_golden.py [out.npz]   (about 15 s on 8 cores)
"""
from __future__ import annotations

import importlib
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import _oracle  # noqa: E402

CENTRE = (-0.743643887037151, 0.131825904205330)
W, H = 1920, 1080
MAX_ITER = 50000
BAILOUT_SQ = 256.0
WIDTHS = (1e-6, 1e-10, 1e-20)
SEED = 20261202
N = 65536
OUT = os.path.join(HERE, "deep_dd_subset.npz")


def subset():
    idx = np.random.default_rng(SEED).choice(W * H, N, replace=False)
    return np.sort(idx).astype(np.uint32)


def main():
    cu = importlib.import_module("chaos-ultra_amd")
    idx = subset()
    px, py = idx % W, idx // W
    out = {"index": idx}
    for w in WIDTHS:
        d = cu.make_deep(CENTRE[0], CENTRE[1], w, W, H, MAX_ITER, BAILOUT_SQ)
        counts, mag, state, _ = _oracle.deep_pixels(_oracle.DEEP_DIRECT_DD, px, py, d)
        lb = math.log(BAILOUT_SQ)
        smooth = np.where(state == 1,
                          counts + 1.0 - np.log2(np.log(np.where(state == 1, mag, BAILOUT_SQ)) / lb),
                          0.0)
        key = f"{w:.0e}"
        out[f"counts_{key}"] = counts.astype(np.uint16)
        out[f"smooth_{key}"] = smooth.astype(np.float64)
        print(f"width {w}: mean count {counts.mean():.1f}, interior {(state == 2).mean():.3f}")
    dest = sys.argv[1] if len(sys.argv) > 1 else OUT
    np.savez_compressed(dest, **out)
    print("wrote", dest, os.path.getsize(dest), "bytes")


if __name__ == "__main__":
    main()
