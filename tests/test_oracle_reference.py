"""The oracle against mathematics (CPU): invariants, special cases, closed forms,
asymptotics and exact-arithmetic orbits, chosen so that a dropped term, a wrong
sign/index or a swapped operand in the oracle fails at least one of them.  The
brute-force and known-answer pins (and the planted-mistake runs that show they
bite) are in tests/test_oracle_pins.py and tests/test_oracle_mutants.py.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN


def test_special_points(oracle):
    L = oracle.lib()
    assert L.or_mandelbrot_d(1000, 0.0, 0.0) == 0       # interior -> 0
    assert L.or_mandelbrot_d(1000, 1.0, 0.0) == 2       # 0, 1, 2: |2|^2 = 4 stops
    assert L.or_mandelbrot_d(1000, -2.0, 0.0) == 1      # z1 = -2, |z|^2 = 4 not < 4
    assert L.or_mandelbrot_d(1000, 0.25, 0.0) == 0      # cusp, bounded
    assert L.or_mandelbrot_d(1000, 2.0, 0.0) == 1
    assert L.or_julia_d(50, 0.0, 0.0, 0.0, 0.0) == 50   # Julia interior -> maxIter
    assert L.or_julia_d(50, 3.0, 0.0, 0.0, 0.0) == 0    # starts outside


def test_mirror_symmetry_c1(cu, oracle):
    # pixel centres y = 1.5 - (py+0.5)*3/1024 are exact dyadics, so row 1023-py is the
    # exact negation of row py and negation is exact through every step
    tile = cu.make_tile(-2.0, -1.5, 1.0, 1.5, 256, 256, 500, 4.0)
    counts, smooth, _ = oracle.escape_tile(tile, 64)
    assert np.array_equal(counts, counts[::-1])
    assert np.array_equal(smooth, smooth[::-1])


def test_julia_point_symmetry(cu, oracle):
    tile = cu.make_tile(-1.5, -1.5, 1.5, 1.5, 128, 128, 300, 4.0, kind=cu.FRACTAL_JULIA,
                        julia_c=(-0.8, 0.156))
    counts, _, _ = oracle.escape_tile(tile, 64)
    assert np.array_equal(counts, counts[::-1, ::-1])


def test_julia_is_mandelbrot_shifted(oracle):
    # z0 = c in the Julia loop is the Mandelbrot state after one step
    L = oracle.lib()
    rng = np.random.default_rng(5)
    for _ in range(2000):
        cx, cy = rng.uniform(-2, 0.6), rng.uniform(-1.2, 1.2)
        m = L.or_mandelbrot_d(400, cx, cy)
        j = L.or_julia_d(399, cx, cy, cx, cy)
        if m != 0:
            assert j == m - 1
        else:
            assert j == 399


@pytest.mark.parametrize("eps", [1e-3, 1e-4, 1e-5])
def test_seahorse_asymptotic(oracle, eps):
    # c = -3/4 + i*eps escapes after N ~ pi/eps steps (N*eps -> pi)
    n = oracle.lib().or_mandelbrot_d(10 ** 7, -0.75, eps)
    assert abs(n * eps - math.pi) < 20 * eps


@pytest.mark.parametrize("eps", [1e-4, 1e-6, 1e-8])
def test_cusp_asymptotic(oracle, eps):
    # c = 1/4 + eps escapes after N ~ pi/sqrt(eps) steps
    n = oracle.lib().or_mandelbrot_d(10 ** 7, 0.25 + eps, 0.0)
    assert abs(n * math.sqrt(eps) - math.pi) < 0.05


@pytest.mark.parametrize("bailout_sq", [4.0, 256.0, 1e6])
def test_escape_bailout_against_exact_orbit(oracle, bailout_sq):
    """The generalised bailout against an exact-arithmetic orbit (mpmath, 300 digits)."""
    import ctypes as C
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 300
    L = oracle.lib()
    rng = np.random.default_rng(int(bailout_sq))
    checked = 0
    for _ in range(300):
        cx, cy = float(rng.uniform(-2.2, 0.8)), float(rng.uniform(-1.3, 1.3))
        m2, esc = C.c_double(), C.c_int()
        n = L.or_escape_d(60, 0.0, 0.0, cx, cy, bailout_sq, C.byref(m2), C.byref(esc))
        x, y = mpmath.mpf(0), mpmath.mpf(0)
        k, near, exact_m2 = 0, False, None
        while k < 60:
            mag = x * x + y * y
            if abs(mag - bailout_sq) < mpmath.mpf(10) ** -6 * bailout_sq:
                near = True
            if mag >= bailout_sq:
                exact_m2 = mag
                break
            x, y = x * x - y * y + cx, 2 * x * y + cy
            k += 1
        if near or k > 40:  # rounding could legitimately move the step (chaotic growth)
            continue
        checked += 1
        assert n == k
        assert bool(esc.value) == (exact_m2 is not None)
        if exact_m2 is not None:
            assert abs(m2.value - float(exact_m2)) <= 1e-9 * float(exact_m2)
    assert checked > 200


def test_bailout4_matches_compute_fractal(oracle):
    import ctypes as C
    L = oracle.lib()
    rng = np.random.default_rng(11)
    for _ in range(3000):
        cx, cy = rng.uniform(-2.1, 0.7), rng.uniform(-1.2, 1.2)
        m2, esc = C.c_double(), C.c_int()
        n = L.or_escape_d(800, 0.0, 0.0, cx, cy, 4.0, C.byref(m2), C.byref(esc))
        assert (n if esc.value else 0) == L.or_mandelbrot_d(800, cx, cy)


def test_smooth_value_mpmath(oracle):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 50
    L = oracle.lib()
    for n, m2, b in [(1, 4.5, 4.0), (17, 300.0, 256.0), (999, 65535.9, 256.0), (3, 1e6 + 1, 1e6),
                     (42, 17.0, 4.0)]:
        want = mpmath.mpf(n) + 1 - mpmath.log(mpmath.log(m2) / mpmath.log(b), 2)
        got = L.or_smooth(n, m2, b)
        assert abs(got - float(want)) <= 1e-14 * abs(float(want))


def test_smooth_value_is_continuous_across_bands(cu, oracle):
    # along a ray leaving the set the smooth value varies continuously while the integer
    # count jumps: consecutive samples differ by far less than one band
    L = oracle.lib()
    import ctypes as C
    prev = None
    for k in range(400):
        cx = 0.3 + k * 1e-4
        m2, esc = C.c_double(), C.c_int()
        n = L.or_escape_d(10000, 0.0, 0.0, cx, 0.0, 256.0, C.byref(m2), C.byref(esc))
        assert esc.value
        s = L.or_smooth(n, m2.value, 256.0)
        if prev is not None:
            assert abs(s - prev) < 0.35
        prev = s


def test_splitmix64_published_vectors(oracle):
    # SplitMix64 (Vigna, splitmix64.c) seeded with 1234567: the first outputs of next()
    L = oracle.lib()
    golden = 0x9E3779B97F4A7C15
    state = 1234567
    outs = []
    for _ in range(3):
        outs.append(L.or_splitmix64(state))
        state = (state + golden) % 2 ** 64
    assert outs == [6457827717110365317, 3203168211198807973, 9817491932198370423]


def test_jitter_indexes_the_splitmix_sequence(oracle):
    # sample j takes output j of SplitMix64(seed): high word -> dx, low word -> dy
    import ctypes as C
    L = oracle.lib()
    dx, dy = C.c_double(), C.c_double()
    for j, h in enumerate([6457827717110365317, 3203168211198807973, 9817491932198370423]):
        L.or_jitter(1234567, j, C.byref(dx), C.byref(dy))
        assert dx.value == (h >> 32) / 2 ** 32
        assert dy.value == (h & 0xFFFFFFFF) / 2 ** 32


def test_jitter_is_uniform_and_seeded(oracle):
    import ctypes as C
    L = oracle.lib()
    dx, dy = C.c_double(), C.c_double()
    xs = []
    for i in range(20000):
        L.or_jitter(7, i, C.byref(dx), C.byref(dy))
        assert 0.0 <= dx.value < 1.0 and 0.0 <= dy.value < 1.0
        xs.append((dx.value, dy.value))
    a = np.array(xs)
    assert abs(a.mean() - 0.5) < 0.01
    assert abs(np.corrcoef(a[:, 0], a[:, 1])[0, 1]) < 0.03
    L.or_jitter(8, 0, C.byref(dx), C.byref(dy))
    assert (dx.value, dy.value) != xs[0]


def test_default_palette_matches_reference_png(oracle):
    # palettes/defaultPalette.png holds the 1536-entry default palette
    with open(os.path.join(GOLDEN, "palettes_argb.json")) as f:
        pal = json.load(f)
    argb = np.array(pal["defaultPalette.png"], dtype=np.uint32)
    rgba = np.array([oracle.lib().or_argb_to_rgba(int(v)) for v in argb], dtype=np.uint32)
    assert np.array_equal(rgba, oracle.default_palette())


def test_argb_to_rgba_byte_order(oracle):
    L = oracle.lib()
    assert L.or_argb_to_rgba(0xFF112233) == 0xFF332211      # R lands in the low byte
    assert L.or_argb_to_rgba(0x00000080) == 0xFF800000      # sign of blue fills the top byte
    assert L.or_argb_to_rgba(0x7F000001) == 0x7F010000


def test_render_main_single_sample_is_corner_sample(cu, oracle):
    """maxSS = 1 takes the one sample at the pixel corner (delta = 0)."""
    w, h = 40, 24
    seg = (-2.0, -1.2, 1.0, 1.2)
    pix = oracle.render_main(64, cu.FRACTAL_MANDELBROT, (0, 0), w, h, seg, 300, 1.0, 0)
    L = oracle.lib()
    psx, psy = (seg[2] - seg[0]) / w, (seg[3] - seg[1]) / h
    for y in range(h):
        for x in range(w):
            c = (seg[0] + (x + 0.0) * psx, seg[3] - (y + 0.0) * psy)
            assert pix["value"][y, x] == L.or_mandelbrot_d(300, *c)
            assert pix["weight"][y, x] == 1.0


def test_render_main_non_adaptive_mean(cu, oracle):
    """Without adaptive sampling and n = 3 the pixel holds the integer mean of the
    samples at deltas 0, 1/3, 2/3 (the vote at i == n/2 cannot stop it early unless
    the whole 8x4 group is uniform, which then gives the same mean)."""
    w, h = 16, 8
    seg = (-0.8, 0.0, -0.7, 0.1)
    pix = oracle.render_main(64, cu.FRACTAL_MANDELBROT, (0, 0), w, h, seg, 400, 3.0, 0)
    L = oracle.lib()
    psx, psy = (seg[2] - seg[0]) / w, (seg[3] - seg[1]) / h
    for y in range(h):
        for x in range(w):
            s = []
            for i in range(3):
                d = i / 3.0
                s.append(int(L.or_mandelbrot_d(400, seg[0] + (x + d) * psx, seg[3] - (y + d) * psy)))
            n = int(pix["weight"][y, x])
            assert n in (2, 3)
            assert pix["value"][y, x] == sum(s[:n]) // n


def test_render_advanced_identity_reuse_copies(cu, oracle):
    """Reusing an identical rectangle reprojects every pixel onto itself (alpha = beta = 0)."""
    w, h = 64, 48
    seg = (-2.0, -1.2, 1.0, 1.2)
    prev = oracle.render_main(64, cu.FRACTAL_MANDELBROT, (0, 0), w, h, seg, 200, 2.0,
                              cu.FLAG_ADAPTIVE_SS)
    out = oracle.render_advanced(64, cu.FRACTAL_MANDELBROT, (0, 0), w, h, seg, 200, 2.0,
                                 cu.FLAG_SAMPLE_REUSE, seg, prev, (0, 0))
    # interior of the image (origin within [2, size-2)) is reused
    inner = out[3:h - 3, 2:w - 3]
    assert np.all(inner["is_reused"] == 1)
    # pixel (x, y) maps to (x, h - (h - y)) = (x, y) exactly
    assert np.array_equal(inner["value"], prev[3:h - 3, 2:w - 3]["value"])


def test_foveation_profile(oracle):
    import ctypes as C
    L = oracle.lib()
    adv, ins = C.c_float(), C.c_int()
    L.or_foveation(100, 100, 100, 100, 8.0, C.byref(adv), C.byref(ins))
    assert ins.value == 1 and adv.value == 8.0          # at the focus: full budget
    L.or_foveation(3000, 100, 0, 100, 8.0, C.byref(adv), C.byref(ins))
    assert ins.value == 0 and adv.value < 8.0           # ~53 degrees away: reduced
    L.or_foveation(100, 100, 100, 100, 0.5, C.byref(adv), C.byref(ins))
    assert ins.value == 1 and adv.value == 1.0          # fovea gets at least one sample
