"""Seeded synthetic workloads of the render path (host-side input generation).

uneven_tiles(): a batch of tiles of very uneven size, depth and kind — 70 %
Mandelbrot, 30 % Julia with c uniform in a disk of radius 0.8; square tiles of
side 64 * 16**u px (log-uniform in [64, 1024]); plane width 3 * (1e-6/3)**u
(log-uniform in [1e-6, 3]); maxIter 256 * 256**u (log-uniform in [256, 65536]);
each centred on a point within 1e-3 of the set's boundary, found by bisection
on a segment parallel to the real axis between a sample that stays bounded
for BOUNDED_TEST iterations and one that escapes.

Every draw comes from numpy.random.default_rng(seed) in a fixed order (all tile
parameters first, then segment rows round by round), and the membership tests
are float64 numpy, so a seed always gives the same batch.
"""
from __future__ import annotations

import math

import numpy as np

BOUNDED_TEST = 1024      # iterations that count a point as inside
SEGMENT_SAMPLES = 1024   # samples along a segment before bisection
BISECTIONS = 48          # halvings of the bracketing interval (2.6e-3 / 2**48)
ROUNDS_PER_C = 8         # segment rows tried before a Julia tile redraws its c


def bounded(zx, zy, cx, cy, iters=BOUNDED_TEST):
    """Vectorised: True where |z| <= 2 for `iters` steps of z <- z^2 + c."""
    zx = np.array(zx, dtype=np.float64)
    zy = np.array(zy, dtype=np.float64)
    cx = np.broadcast_to(np.asarray(cx, dtype=np.float64), zx.shape)
    cy = np.broadcast_to(np.asarray(cy, dtype=np.float64), zx.shape)
    alive = np.ones(zx.shape, dtype=bool)
    for _ in range(iters):
        x2, y2 = zx * zx, zy * zy
        alive &= (x2 + y2) <= 4.0
        if not alive.any():
            break
        nx = np.where(alive, x2 - y2 + cx, zx)
        zy = np.where(alive, 2.0 * zx * zy + cy, zy)
        zx = nx
    return alive


def _member(julia, x, y, jcx, jcy):
    """Membership of points (x, y) (rows = tiles); julia/jcx/jcy per row."""
    j = julia[:, None]
    zx = np.where(j, x, 0.0)
    zy = np.where(j, y, 0.0)
    cx = np.where(j, jcx[:, None], x)
    cy = np.where(j, jcy[:, None], y)
    return bounded(zx, zy, cx, cy)


def _julia_c(rng):
    r, t = 0.8 * math.sqrt(rng.random()), 2 * math.pi * rng.random()
    return r * math.cos(t), r * math.sin(t)


def uneven_tiles(cu, n=256, seed=20261202, bailout_sq=4.0, spp=1):
    """The uneven-tile batch as TileDesc objects, plus a description per tile."""
    rng = np.random.default_rng(seed)
    side = np.empty(n, np.int64)
    width = np.empty(n)
    max_iter = np.empty(n, np.int64)
    julia = np.zeros(n, bool)
    jcx, jcy = np.zeros(n), np.zeros(n)
    for k in range(n):
        side[k] = int(round(64 * 16 ** rng.random()))
        width[k] = 3.0 * (1e-6 / 3.0) ** rng.random()
        max_iter[k] = int(round(256 * 256 ** rng.random()))
        julia[k] = rng.random() < 0.3
        if julia[k]:
            jcx[k], jcy[k] = _julia_c(rng)
    lo = np.where(julia, -2.0, -2.1)
    hi = np.where(julia, 2.0, 0.6)
    yr = np.where(julia, 1.5, 1.15)
    t = np.linspace(0.0, 1.0, SEGMENT_SAMPLES)
    cx_out, cy_out, gap = np.full(n, np.nan), np.zeros(n), np.zeros(n)
    a, b, ia, yk = np.zeros(n), np.zeros(n), np.zeros(n, bool), np.zeros(n)
    tries = np.zeros(n, np.int64)
    todo = np.arange(n)
    while todo.size:
        for k in todo:  # a new row (and after ROUNDS_PER_C misses, a new Julia c)
            if julia[k] and tries[k] and tries[k] % ROUNDS_PER_C == 0:
                jcx[k], jcy[k] = _julia_c(rng)
            yk[k] = rng.uniform(-yr[k], yr[k])
            tries[k] += 1
        xs = lo[todo, None] + (hi - lo)[todo, None] * t[None, :]
        ins = _member(julia[todo], xs, np.broadcast_to(yk[todo, None], xs.shape), jcx[todo], jcy[todo])
        found = []
        for r, k in enumerate(todo):
            edges = np.nonzero(ins[r, :-1] != ins[r, 1:])[0]
            if edges.size:
                e = int(edges[rng.integers(edges.size)])
                a[k], b[k], ia[k] = xs[r, e], xs[r, e + 1], ins[r, e]
                found.append(k)
        found = np.array(found, dtype=np.int64)
        if found.size:
            for _ in range(BISECTIONS):  # all tiles of this round at once
                m = 0.5 * (a[found] + b[found])
                im = _member(julia[found], m[:, None], yk[found, None], jcx[found], jcy[found])[:, 0]
                same = im == ia[found]
                a[found] = np.where(same, m, a[found])
                b[found] = np.where(same, b[found], m)
            cx_out[found] = 0.5 * (a[found] + b[found])
            cy_out[found] = yk[found]
            gap[found] = np.abs(b[found] - a[found])
        todo = np.array([k for k in todo if np.isnan(cx_out[k])], dtype=np.int64)
    tiles, info = [], []
    for k in range(n):
        h = width[k] / 2
        cx, cy = cx_out[k], cy_out[k]
        jc = (float(jcx[k]), float(jcy[k]))
        tiles.append(cu.make_tile(cx - h, cy - h, cx + h, cy + h, int(side[k]), int(side[k]),
                                  int(max_iter[k]), bailout_sq,
                                  kind=cu.FRACTAL_JULIA if julia[k] else cu.FRACTAL_MANDELBROT,
                                  julia_c=jc, spp=spp))
        info.append({"side": int(side[k]), "width": float(width[k]), "max_iter": int(max_iter[k]),
                     "kind": "julia" if julia[k] else "mandelbrot", "julia_c": jc,
                     "centre": (float(cx), float(cy)), "bisection_gap": float(gap[k]),
                     "bracket": (float(a[k]), float(b[k])),
                     "segment_rows": int(tries[k])})
    return tiles, info
