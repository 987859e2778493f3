"""Checks that pin the oracle's restatements to what the reference's source and
mathematics fix, written against a loaded oracle build `L` (tests/_oracle.bind).

tests/test_oracle_pins.py runs every check on oracle/liboracle.so, where each
must pass; tests/test_oracle_mutants.py builds copies of chaos_oracle.cpp with
one planted mistake each and shows that some check here fails on every one.

None of these restates the oracle's code path: the iteration is brute-forced in
numpy's IEEE float32/float64 arithmetic (a different engine), palette slots and
grey levels are known answers written out by hand, the under-sampled kernel is
checked for which pixels it writes, and foveation against a float64 closed form.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

PIXEL_INFO = np.dtype([("value", "<f4"), ("weight", "<f4"), ("is_reused", "u1"),
                       ("pad", "u1", (3,)), ("weight_of_new_samples", "<f4")])


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# computeFractal<Real> (mandelbrot.cu:11-25, julia.cu:5-17) brute-forced in numpy
# ---------------------------------------------------------------------------
def brute_escape(dtype, cx, cy, max_iter, zx=None, zy=None, bailout_sq=4.0):
    """Vectorised mandelbrot.cu:16-21 in numpy `dtype` arithmetic (each operation
    rounds to dtype, no contraction).  Returns (iterations executed, escaped)."""
    cx = np.asarray(cx, dtype)
    cy = np.asarray(cy, dtype)
    x = np.zeros_like(cx) if zx is None else np.asarray(zx, dtype).copy()
    y = np.zeros_like(cy) if zy is None else np.asarray(zy, dtype).copy()
    two = dtype(2)
    bsq = dtype(bailout_sq)
    n = np.zeros(cx.shape, np.uint32)
    running = np.ones(cx.shape, bool)
    escaped = np.zeros(cx.shape, bool)
    with np.errstate(over="ignore", invalid="ignore"):
        for _ in range(max_iter):
            inside = (x * x + y * y) < bsq          # :16  x*x + y*y < 4
            escaped |= running & ~inside
            running &= inside
            if not running.any():
                break
            xn = x * x - y * y + cx                 # :17  (x*x - y*y) + cx
            y = np.where(running, two * x * y + cy, y)  # :18  ((2*x)*y) + cy
            x = np.where(running, xn, x)
            n += running
    return n, escaped


def _grid(n=48):
    xs = np.linspace(-2.1, 0.7, n)
    ys = np.linspace(-1.25, 1.25, n)
    gx, gy = np.meshgrid(xs, ys)
    # points close to the boundary, where rounding decides the count
    rng = np.random.default_rng(2024)
    t = rng.uniform(0, 2 * np.pi, 600)
    bx = 0.5 * np.cos(t) - 0.25 * np.cos(2 * t) + rng.normal(0, 2e-3, t.size)
    by = 0.5 * np.sin(t) - 0.25 * np.sin(2 * t) + rng.normal(0, 2e-3, t.size)
    return np.concatenate([gx.ravel(), bx]), np.concatenate([gy.ravel(), by])


def check_mandelbrot_bruteforce(L, precision: int, max_iter: int = 400) -> int:
    """or_mandelbrot_{f,d} equals the numpy brute force of the reference loop at
    every grid and near-boundary point, interior -> 0 (mandelbrot.cu:22-23).
    Returns how many points the FP32 and FP64 brute forces disagree on (the
    check only discriminates the two precisions when that is > 0)."""
    gx, gy = _grid()
    dt = np.float32 if precision == 32 else np.float64
    cxs, cys = gx.astype(dt), gy.astype(dt)
    n, esc = brute_escape(dt, cxs, cys, max_iter)
    want = np.where(esc, n, 0)
    f = L.or_mandelbrot_f if precision == 32 else L.or_mandelbrot_d
    got = np.array([f(max_iter, float(a), float(b)) for a, b in zip(cxs, cys)])
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (f"FP{precision} computeFractal differs from the brute force at "
                           f"{bad.size} points, first c={cxs[bad[0]]!r},{cys[bad[0]]!r}: "
                           f"{got[bad[0]]} vs {want[bad[0]]}")
    n32, e32 = brute_escape(np.float32, gx.astype(np.float32), gy.astype(np.float32), max_iter)
    n64, e64 = brute_escape(np.float64, gx, gy, max_iter)
    return int((np.where(e32, n32, 0) != np.where(e64, n64, 0)).sum())


def check_julia_bruteforce(L, precision: int, max_iter: int = 300) -> None:
    """or_julia_{f,d}: z0 = pixel, c = (Real)julia_c, interior -> maxIter (julia.cu:5-17),
    against the brute force; an asymmetric c, so a swapped component shows."""
    gx, gy = _grid(40)
    dt = np.float32 if precision == 32 else np.float64
    jc = (-0.7269, 0.1889)
    zx, zy = (gx * 0.9).astype(dt), gy.astype(dt)
    cx = np.full(zx.shape, jc[0], dt)
    cy = np.full(zy.shape, jc[1], dt)
    n, _ = brute_escape(dt, cx, cy, max_iter, zx=zx, zy=zy)
    f = L.or_julia_f if precision == 32 else L.or_julia_d
    got = np.array([f(max_iter, float(a), float(b), jc[0], jc[1]) for a, b in zip(zx, zy)])
    assert np.array_equal(got, n), f"FP{precision} Julia differs at {(got != n).sum()} points"


def check_special_points(L) -> None:
    """Orbits that are exact in both precisions: c = -2 stops at |z|^2 = 4 exactly
    (the test is a strict <); c = 1, 2i, 1+i and 1/2 leave after 2, 1, 2 and 5
    steps through dyadic values; c = 0, -1, i and 1/4 stay bounded (interior -> 0).
    Julia: z0 = 0 with c = 0 is interior (-> maxIter), a z0 outside takes no step."""
    for f in (L.or_mandelbrot_d, L.or_mandelbrot_f):
        assert f(1000, -2.0, 0.0) == 1      # 0 -> -2: |z|^2 = 4 is not < 4
        assert f(1000, 1.0, 0.0) == 2       # 0 -> 1 -> 2
        assert f(1000, 0.0, 2.0) == 1       # 0 -> 2i: |z|^2 = 4
        assert f(1000, 2.0, 0.0) == 1
        assert f(1000, 0.0, 0.0) == 0       # fixed point
        assert f(1000, -1.0, 0.0) == 0      # period 2: 0, -1, 0, ...
        assert f(1000, 0.0, 1.0) == 0       # 0, i, -1+i, -i, -1+i, ...
        assert f(1000, 0.25, 0.0) == 0      # the cusp converges to 1/2 from below
        assert f(1000, 1.0, 1.0) == 2       # 0 -> 1+i -> 1+3i
        assert f(1000, 0.5, 0.0) == 5       # 0, 1/2, 3/4, 17/16, 417/256, 3.15...: all dyadic
    for f in (L.or_julia_d, L.or_julia_f):
        assert f(50, 0.0, 0.0, 0.0, 0.0) == 50   # interior -> maxIter
        assert f(50, 3.0, 0.0, 0.0, 0.0) == 0    # starts outside: no step
        assert f(50, 2.0, 0.0, -2.0, 0.0) == 0   # |z0|^2 = 4 is outside
        assert f(50, 1.0, 1.0, 0.0, 0.0) == 1    # (1+i)^2 = 2i: |z|^2 = 4 after one step


def check_mirror_symmetry(L, precision: int) -> None:
    """Pixel centres y = 1.5 - (py + 0.5) * 3/1024 are dyadics with 12 bits, exact in
    float and double, so row 1023 - py is the exact negation of row py, and negation
    is exact through every step: the C1 frame's counts are mirror-symmetric."""
    f = L.or_mandelbrot_f if precision == 32 else L.or_mandelbrot_d
    for py in (0, 100, 383, 470, 511):
        y = 1.5 - (py + 0.5) * 3.0 / 1024
        for px in range(0, 1024, 13):
            x = -2.0 + (px + 0.5) * 3.0 / 1024
            assert f(300, x, y) == f(300, x, -y)


# ---------------------------------------------------------------------------
# planar tiles: FP32 sample positions and the FP32 tile path
# ---------------------------------------------------------------------------
def check_tile_points_f32(L, cu) -> None:
    """or_tile_point_f on the C1 mapping is exact: the pixel centres
    -2 + (px + 0.5) * 3/1024 and 1.5 - (py + 0.5) * 3/1024 are dyadics with at most
    13 significant bits.  Corner samples lie on the pixel grid itself.  Jittered
    samples lie inside their pixel, within one float ulp."""
    x, y = C.c_float(), C.c_float()
    t = cu.make_tile(-2.0, -1.5, 1.0, 1.5, 1024, 1024, 10, 4.0)
    for px, py in [(0, 0), (1023, 1023), (511, 3), (7, 640), (300, 1000)]:
        L.or_tile_point_f(C.byref(t), px, py, 0, C.byref(x), C.byref(y))
        assert x.value == -2.0 + (px + 0.5) * 3 / 1024, (px, x.value)
        assert y.value == 1.5 - (py + 0.5) * 3 / 1024, (py, y.value)
    t = cu.make_tile(-2.0, -1.5, 1.0, 1.5, 1024, 1024, 10, 4.0, pattern=cu.SAMPLE_CORNER)
    for px, py in [(0, 0), (5, 9), (1000, 17)]:
        L.or_tile_point_f(C.byref(t), px, py, 0, C.byref(x), C.byref(y))
        assert (x.value, y.value) == (-2.0 + px * 3 / 1024, 1.5 - py * 3 / 1024)
    t = cu.make_tile(-0.75, 0.05, -0.74, 0.06, 64, 64, 10, 4.0, spp=4, pattern=cu.SAMPLE_JITTER,
                     seed=9)
    ulp = np.spacing(np.float32(0.75))
    pw = 0.01 / 64
    for px, py, s in [(0, 0, 0), (63, 63, 3), (10, 40, 2)]:
        L.or_tile_point_f(C.byref(t), px, py, s, C.byref(x), C.byref(y))
        assert -0.75 + px * pw - 2 * ulp <= x.value <= -0.75 + (px + 1) * pw + 2 * ulp
        assert 0.06 - (py + 1) * pw - 2 * ulp <= y.value <= 0.06 - py * pw + 2 * ulp
    # A segment with no dyadic structure: the position is evaluated in float32, each
    # operation rounded, in the order leftTop + (1,-1) * (pixel + delta) * pixelSize
    # (fractalRendererGeneric.cu:119-124) from the float-cast segment
    # (cudarenderer/KernelMainFloat.java:19-21).  numpy's float32 is the engine here.
    f32 = np.float32
    dxd, dyd = C.c_double(), C.c_double()
    for pattern in (cu.SAMPLE_CENTRE, cu.SAMPLE_JITTER):
        t = cu.make_tile(-0.7436939, 0.1317759, -0.7435939, 0.1318759, 333, 333, 10, 4.0, spp=2,
                         pattern=pattern, seed=77)
        for px, py, s in [(0, 0, 0), (332, 332, 1), (17, 250, 1), (199, 3, 0), (101, 101, 1)]:
            if pattern == cu.SAMPLE_CENTRE:
                dx = dy = 0.5
            else:
                L.or_jitter(77, (py * 333 + px) * 2 + s, C.byref(dxd), C.byref(dyd))
                dx, dy = dxd.value, dyd.value
            L.or_tile_point_f(C.byref(t), px, py, s, C.byref(x), C.byref(y))
            wx = f32(t.left_f) + (f32(px) + f32(dx)) * f32(t.pixel_w_f)
            wy = f32(t.top_f) - (f32(py) + f32(dy)) * f32(t.pixel_h_f)
            assert (x.value, y.value) == (float(wx), float(wy)), (pattern, px, py, s)


def check_tile_rows_f32(L, cu) -> None:
    """The FP32 tile path (pixel centres, bailout 256, 2 x 2 of the C1 view) equals
    the float32 brute force at the exact dyadic centres."""
    w = h = 96
    t = cu.make_tile(-2.0, -1.5, 1.0, 1.5, w, h, 250, 256.0)
    counts = np.zeros((h, w), np.uint32)
    smooth = np.zeros((h, w), np.float64)
    L.or_escape_tile_rows(C.byref(t), 32, 0, h, _ptr(counts), _ptr(smooth), 1)
    px, py = np.meshgrid(np.arange(w), np.arange(h))
    cx = (-2.0 + (px + 0.5) * 3.0 / w).astype(np.float32)   # exact: w = 96 = 3 * 32
    cy = (1.5 - (py + 0.5) * 3.0 / h).astype(np.float32)
    n, esc = brute_escape(np.float32, cx, cy, 250, bailout_sq=256.0)
    assert np.array_equal(counts, np.where(esc, n, 0))


# ---------------------------------------------------------------------------
# compose / colorize / colorizeSampleCount (fractalRendererGeneric.cu:455-475,
# 67-77; mandelbrot.cu:27-35)
# ---------------------------------------------------------------------------
# (value, palette slot for L = 256): L - (round(v) mod L) - 1, round half away from zero
COMPOSE_KNOWN = [
    (0.0, 255), (1.0, 254), (2.4999, 253), (2.5, 252), (3.5, 251), (254.6, 0), (255.0, 0),
    (256.0, 255), (256.5, 254), (515.0, 252), (1000000.5, 190), (-0.3, 255), (-3.0, 255),
]


def _palette_with_guard(L_len: int) -> np.ndarray:
    # entry k holds 0x10000 + k; entries past L hold a guard so an off-by-one read is seen
    pal = np.full(L_len + 8, 0xDEAD0000, np.uint32)
    pal[:L_len] = 0x10000 + np.arange(L_len, dtype=np.uint32)
    return pal


def check_compose_palette(L) -> None:
    pal = _palette_with_guard(256)
    vals = np.array([v for v, _ in COMPOSE_KNOWN], np.float32)
    pix = np.zeros((1, vals.size), PIXEL_INFO)
    pix["value"][0] = vals
    pix["weight"][0] = 1.0
    out = np.zeros((1, vals.size), np.uint32)
    L.or_compose(_ptr(pix), vals.size * 16, _ptr(out), vals.size, vals.size, 1, _ptr(pal), 256,
                 1.0, 0)
    want = np.array([0x10000 + k for _, k in COMPOSE_KNOWN], np.uint32)
    assert np.array_equal(out[0], want), list(zip(vals, out[0] - 0x10000, want - 0x10000))


def check_compose_smooth_palette(L) -> None:
    pal = _palette_with_guard(256)
    vals = np.array([v for v, _ in COMPOSE_KNOWN] + [-1e9], np.float64)
    out = np.zeros((1, vals.size), np.uint32)
    sm = vals.reshape(1, -1).copy()
    L.or_compose_smooth(_ptr(sm), vals.size, _ptr(out), vals.size, vals.size, 1, _ptr(pal), 256)
    want = np.array([0x10000 + k for _, k in COMPOSE_KNOWN] + [0x10000 + 255], np.uint32)
    assert np.array_equal(out[0], want), list(zip(vals, out[0] - 0x10000, want - 0x10000))
    # a palette of 7 entries: the modulus is the palette length
    pal7 = _palette_with_guard(7)
    sm = np.array([[0.0, 6.0, 7.0, 13.49, 13.5]], np.float64)
    out = np.zeros((1, 5), np.uint32)
    L.or_compose_smooth(_ptr(sm), 5, _ptr(out), 5, 5, 1, _ptr(pal7), 7)
    assert list(out[0] - 0x10000) == [6, 0, 6, 0, 6]


def check_colorize_sample_count(L) -> None:
    """Grey level = (int)(255 * min(n, full) / full) with full = max(1, maxSS) as an
    integer; a reused pixel shows its weight of new samples."""
    cases = [  # (weight, is_reused, weight_of_new, max_ss, grey)
        (64.0, 0, 0.0, 64.0, 255), (32.0, 0, 0.0, 64.0, 127), (16.0, 0, 0.0, 64.0, 63),
        (100.0, 0, 0.0, 64.0, 255), (0.0, 0, 0.0, 64.0, 0), (3.9, 0, 0.0, 8.0, 95),
        (8.0, 1, 2.0, 8.0, 63), (1.0, 0, 0.0, 0.5, 255), (0.5, 0, 0.0, 0.5, 0),
    ]
    for weight, reused, won, max_ss, grey in cases:
        pix = np.zeros((1, 1), PIXEL_INFO)
        pix["weight"] = weight
        pix["is_reused"] = reused
        pix["weight_of_new_samples"] = won
        out = np.zeros((1, 1), np.uint32)
        pal = np.zeros(4, np.uint32)
        L.or_compose(_ptr(pix), 16, _ptr(out), 1, 1, 1, _ptr(pal), 1, max_ss, 1)
        want = grey | (grey << 8) | (grey << 16) | (0xFF << 24)
        assert int(out[0, 0]) == want, (weight, reused, won, max_ss, hex(int(out[0, 0])), grey)


# ---------------------------------------------------------------------------
# fractalRenderUnderSampled (fractalRendererGeneric.cu:478-503)
# ---------------------------------------------------------------------------
def check_undersampled(L) -> None:
    """Only pixels (k*level, j*level) with k*level < W - level and j*level < H - level
    are written; each holds computeFractal<float> at its pixel corner and weight
    1/level.  The corners are evaluated in float32, the reference's order
    (left_bottom_x + idx_x * pixelWidth, right_top_y - idx_y * pixelHeight); the
    segment is not symmetric about the real axis, so a row counted from the bottom
    shows."""
    for W, H, level in [(64, 32, 4), (23, 17, 5), (40, 40, 3), (9, 9, 8)]:
        seg = (-2.0, -1.25, 1.0, 1.75)
        out = np.zeros((H, W), PIXEL_INFO)
        out["value"] = -7.0        # sentinel
        out["weight"] = -7.0
        L.or_render_undersampled(_ptr(out), W * 16, W, H, *seg, 200, level)
        pw = (np.float32(seg[2]) - np.float32(seg[0])) / np.float32(W)
        ph = (np.float32(seg[3]) - np.float32(seg[1])) / np.float32(H)
        for y in range(H):
            for x in range(W):
                written = x % level == 0 and y % level == 0 and x < W - level and y < H - level
                if not written:
                    assert out["value"][y, x] == -7.0, (W, H, level, x, y)
                    continue
                cx = np.float32(seg[0]) + np.float32(x) * pw
                cy = np.float32(seg[3]) - np.float32(y) * ph
                assert out["value"][y, x] == L.or_mandelbrot_f(200, float(cx), float(cy))
                assert out["weight"][y, x] == np.float32(1) / np.float32(level)


# ---------------------------------------------------------------------------
# getFoveationAdvisedSampleCount + linearMapping (fractalRendererGeneric.cu:209-257)
# ---------------------------------------------------------------------------
def foveation_closed_form(px, py, fx, fy, max_ss):
    """Float64 closed form: the 8x4 tile corner's distance to the focus, times
    0.02652 cm per pixel, seen from 60 cm, in degrees; full budget up to the foveal
    angle (5.5 degrees, scaled by maxSS below 1), falling linearly to 0 at 60
    degrees and below 0 beyond; at least 1 inside the fovea."""
    nx, ny = px - px % 8, py - py % 4
    d_cm = math.hypot(fx - nx, fy - ny) * 0.02652
    angle = math.degrees(math.atan(d_cm / 60.0))
    foveal = 5.5 if max_ss >= 1 else 5.5 * max_ss
    rel = min(1.0, (60.0 - angle) / (60.0 - foveal))
    adv = max_ss * rel
    inside = angle <= foveal
    if inside:
        adv = max(1.0, adv)
    return adv, inside


def check_foveation(L) -> None:
    adv, ins = C.c_float(), C.c_int()
    focus = (1000, 700)
    offsets = [(0, 0), (3, 2), (7, 3), (150, 0), (0, 260), (400, 300), (1500, 0), (2300, 1700),
               (3000, 900), (5000, 0)]
    seen_negative = False
    for max_ss in (64.0, 8.0, 1.0, 0.5):
        for ox, oy in offsets:
            px, py = focus[0] + ox, focus[1] + oy
            L.or_foveation(px, py, focus[0], focus[1], max_ss, C.byref(adv), C.byref(ins))
            want, want_in = foveation_closed_form(px, py, focus[0], focus[1], max_ss)
            assert bool(ins.value) == want_in, (max_ss, ox, oy)
            assert abs(adv.value - want) <= 2e-5 * max(1.0, max_ss), (max_ss, ox, oy, adv.value, want)
            seen_negative |= want < 0
    assert seen_negative  # beyond 60 degrees the budget is negative, as in the reference


# ---------------------------------------------------------------------------
# the C1 frame in full
# ---------------------------------------------------------------------------
def brute_escape_compact(dtype, cx, cy, max_iter, bailout_sq=4.0):
    """brute_escape for large inputs: the running points are gathered into dense
    arrays every 32 steps, so the cost follows the points still iterating."""
    cx = np.asarray(cx, dtype).ravel()
    cy = np.asarray(cy, dtype).ravel()
    idx = np.arange(cx.size)
    x = np.zeros_like(cx)
    y = np.zeros_like(cy)
    n = np.zeros(cx.size, np.uint32)
    escaped = np.zeros(cx.size, bool)
    two, bsq = dtype(2), dtype(bailout_sq)
    step = 0
    with np.errstate(over="ignore", invalid="ignore"):
        while idx.size and step < max_iter:
            for _ in range(min(32, max_iter - step)):
                inside = (x * x + y * y) < bsq
                gone = ~inside & ~escaped[idx]
                escaped[idx[gone]] = True
                live = inside & ~escaped[idx]
                xn = x * x - y * y + cx
                y = np.where(live, two * x * y + cy, y)
                x = np.where(live, xn, x)
                n[idx[live]] += 1
                step += 1
            keep = ~escaped[idx]
            idx, x, y, cx, cy = idx[keep], x[keep], y[keep], cx[keep], cy[keep]
    # points still running after max_iter steps are the interior: the loop ends on
    # i == maxIter without another test (mandelbrot.cu:16)
    return n, escaped


def check_c1_frame(L, cu, precision: int) -> None:
    """The C1 frame (1024 x 1024 pixel centres of [-2, 1] x [-1.5, 1.5], maxIter
    1000) through the oracle's tile path equals the numpy brute force of the
    reference loop at every pixel, with the interior (count maxIter) shown as 0."""
    w = h = 1024
    t = cu.make_tile(-2.0, -1.5, 1.0, 1.5, w, h, 1000, 4.0)
    counts = np.zeros((h, w), np.uint32)
    L.or_escape_tile_rows(C.byref(t), precision, 0, h, _ptr(counts), None, 0)
    px, py = np.meshgrid(np.arange(w), np.arange(h))
    dt = np.float32 if precision == 32 else np.float64
    cx = (-2.0 + (px + 0.5) * 3.0 / w).astype(dt)   # exact dyadics in both precisions
    cy = (1.5 - (py + 0.5) * 3.0 / h).astype(dt)
    n, esc = brute_escape_compact(dt, cx, cy, 1000)
    want = np.where(esc, n, 0).reshape(h, w)
    bad = np.argwhere(counts != want)
    assert bad.size == 0, f"FP{precision} C1 frame differs at {len(bad)} pixels, first {bad[0]}"
    assert np.array_equal(counts, counts[::-1])     # mirror symmetry of the dyadic centres


# ---------------------------------------------------------------------------
# sampleTheFractal's sample counts (fractalRendererGeneric.cu:85-155, 24-34)
# ---------------------------------------------------------------------------
# Regions where every sample of every pixel has the same count, so each vote
# group is unanimous and the stop rules give known sample counts:
#   R1 all counts 1 (|c|^2 >= 4), R2 all counts 2 (|c|^2 < 4 <= |c^2 + c|^2),
#   R0 all inside the main cardioid (computeFractal -> 0, so the mean is 0 and
#   the index of dispersion 0/0 is NaN: no vote passes).
REGIONS = {"R1": (3.0, 3.0, 3.5, 3.5), "R2": (1.5, 0.0, 1.6, 0.1), "R0": (-0.15, 0.05, -0.05, 0.15)}
# (region, maxSuperSampling, adaptive, value, weight)
VOTE_KNOWN = [
    ("R2", 16.0, False, 2, 9),   # check at i == n/2 = 8 (the && / || precedence): dispersion 0 stops
    ("R2", 16.0, True, 2, 2),    # i == 1: samples 0 and 1 equal in the whole group
    ("R2", 4.0, False, 2, 3),    # i == 2 == n/2
    ("R2", 2.6, False, 2, 2),    # n = round(2.6) = 3, n/2 = 1: the i == 1 rule runs without adaptive SS
    ("R2", 1.0, False, 2, 1),
    ("R1", 16.0, False, 1, 9),
    ("R0", 16.0, False, 0, 16),  # NaN dispersion: every sample taken
    ("R0", 16.0, True, 0, 2),    # |0 - 0| < FLT_EPSILON at i == 1
    ("R0", 4.0, False, 0, 4),
    ("R0", 3.6, False, 0, 4),    # n = round(3.6) = 4, not 3
    ("R0", 0.5, False, 0, 0),    # sampleCountF < 1: no sample, weight 0
]


def check_render_votes(L) -> None:
    w, h = 16, 8  # four 8x4 vote groups
    for region, max_ss, adaptive, value, weight in VOTE_KNOWN:
        for precision in (64, 32):
            out = np.zeros((h, w), PIXEL_INFO)
            L.or_render_main(precision, 0, 0.0, 0.0, _ptr(out), w * 16, w, h, *REGIONS[region], 200,
                             max_ss, 1 if adaptive else 0)
            assert np.all(out["value"] == value), (region, max_ss, adaptive, precision,
                                                   np.unique(out["value"]))
            assert np.all(out["weight"] == weight), (region, max_ss, adaptive, precision,
                                                     np.unique(out["weight"]))
