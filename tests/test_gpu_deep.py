"""The deep-zoom form on the GPU against the oracle.

Bit-exact: the double-double reference orbit, every pixel's count of the FP32 and
FP64 perturbation (the kernel's unchecked blocks replay to the oracle's checked
loop), the glitch flags without rebasing, the glitch round around a secondary
reference, and the direct FP64 form; smooth values within 1e-12 relative.
Accuracy at the full 1920x1080 frames is measured against the committed
double-double fixture (seeded 65 536-pixel subset).
"""
import math
import os

import numpy as np
import pytest

import _oracle as O
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

CENTRE = (-0.743643887037151, 0.131825904205330)


def deep(cu, width_plane, w=192, h=108, max_iter=50000, rebase=True, max_references=8):
    return cu.make_deep(CENTRE[0], CENTRE[1], width_plane, w, h, max_iter, 256.0, rebase=rebase,
                        max_references=max_references)


def render(cu, engine, desc, method):
    f = cu.DeepFrame(engine, desc, method)
    try:
        f.render()
        cu.synchronize()
        return f.counts(), f.smooth(), f.glitched(), f.report(), (f.orbit() if method != 1 else None)
    finally:
        f.close()


def smooth_of(counts, mag, state):
    lb = math.log(256.0)
    return np.where(state == 1, counts + 1.0 - np.log2(np.log(np.where(state == 1, mag, 256.0)) / lb),
                    0.0)


def assert_smooth_close(got, want):
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("method,precision", [(32, 32), (64, 64)])
def test_orbit_bit_exact(gpu, engine, method, precision):
    cu = gpu
    *_, rep, orbit = render(cu, engine, deep(cu, 1e-10, 32, 18), method)
    want, M = O.reference_orbit(CENTRE[0], CENTRE[1], 50000, 256.0, precision)
    assert rep.reference_length == M == 3087
    assert orbit.shape == want.shape and np.array_equal(orbit, want)


# Centres with low parts, on the real axis (y stays zero: signs of zero), inside
# the set (the orbit runs to max_iter) and escaping after a few steps; every
# entry compared bit for bit with the oracle's.
@pytest.mark.parametrize("centre,lo", [
    ((-0.743643887037151, 0.131825904205330), (1.3e-17, -2.1e-17)),
    ((-1.0, 0.0), (0.0, 0.0)),
    ((-1.25, 0.0), (0.0, 0.0)),
    ((0.3, -0.5), (0.0, 0.0)),
    ((-0.12, 0.74), (0.0, 0.0)),
    ((0.26, 1e-3), (-4e-20, 3e-20)),
])
@pytest.mark.parametrize("method,precision", [(32, 32), (64, 64)])
def test_orbit_bits_across_centres(gpu, engine, centre, lo, method, precision):
    cu = gpu
    d = cu.make_deep(centre[0], centre[1], 1e-12, 8, 8, 20000, centre_lo=lo)
    f = cu.DeepFrame(engine, d, method)
    try:
        f.render()
        got, rep = f.orbit(), f.report()
    finally:
        f.close()
    want, M = O.reference_orbit(centre[0], centre[1], 20000, 256.0, precision, lo)
    assert rep.reference_length == M
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


# 1.0: a shallow frame around one reference, where rebases come every few steps
# and most blocks stop early at a rebase
@pytest.mark.parametrize("width", [1e-6, 1e-10, 1e-20, 1.0])
@pytest.mark.parametrize("method", [32, 64])
def test_perturbation_matches_oracle(gpu, engine, width, method):
    cu = gpu
    d = deep(cu, width, max_iter={1e-6: 20000, 1.0: 2000}.get(width, 50000))
    counts, smooth, glitch, rep, orbit = render(cu, engine, d, method)
    ys, xs = np.mgrid[0:d.height, 0:d.width]
    om = O.DEEP_PERTURB_F32 if method == 32 else O.DEEP_PERTURB_F64
    wc, wm, ws, wit = O.deep_pixels(om, xs.ravel(), ys.ravel(), d, orbit, rep.reference_length)
    assert np.array_equal(counts.ravel(), wc)
    assert rep.iterations == wit
    assert not glitch.any() and rep.glitched == 0 and rep.references == 1
    assert_smooth_close(smooth.ravel(), smooth_of(wc, wm, ws))


# frame sizes that leave partial 16x16 / 32x16 blocks and partial 8x8 waves in
# both directions, and budgets shorter than one 16-step block (checked steps only)
# or just past it
@pytest.mark.parametrize("w,h,max_iter", [(37, 23, 50000), (1, 1, 50000), (130, 9, 15),
                                          (65, 70, 17)])
@pytest.mark.parametrize("method", [32, 64])
def test_perturbation_ragged_frames_and_short_budgets(gpu, engine, w, h, max_iter, method):
    cu = gpu
    d = deep(cu, 1e-6, w=w, h=h, max_iter=max_iter)
    counts, smooth, glitch, rep, orbit = render(cu, engine, d, method)
    assert counts.shape == (h, w)
    ys, xs = np.mgrid[0:h, 0:w]
    om = O.DEEP_PERTURB_F32 if method == 32 else O.DEEP_PERTURB_F64
    wc, wm, ws, wit = O.deep_pixels(om, xs.ravel(), ys.ravel(), d, orbit, rep.reference_length)
    assert np.array_equal(counts.ravel(), wc)
    assert rep.iterations == wit
    assert_smooth_close(smooth.ravel(), smooth_of(wc, wm, ws))


def test_direct_matches_oracle(gpu, engine):
    cu = gpu
    d = deep(cu, 1e-10)
    counts, smooth, _, rep, _ = render(cu, engine, d, 1)
    ys, xs = np.mgrid[0:d.height, 0:d.width]
    wc, wm, ws, wit = O.deep_pixels(O.DEEP_DIRECT_F64, xs.ravel(), ys.ravel(), d)
    assert np.array_equal(counts.ravel(), wc)
    assert rep.iterations == wit
    assert_smooth_close(smooth.ravel(), smooth_of(wc, wm, ws))


@pytest.mark.parametrize("method", [32, 64])
def test_glitch_detection_and_secondary_round(gpu, engine, method):
    cu = gpu
    om = O.DEEP_PERTURB_F32 if method == 32 else O.DEEP_PERTURB_F64
    # no rounds: the glitch flags are the oracle's
    d0 = deep(cu, 1e-10, rebase=False, max_references=0)
    counts, _, glitch, rep, orbit = render(cu, engine, d0, method)
    ys, xs = np.mgrid[0:d0.height, 0:d0.width]
    px, py = xs.ravel(), ys.ravel()
    wc, _, ws, _ = O.deep_pixels(om, px, py, d0, orbit, rep.reference_length)
    assert np.array_equal(glitch.ravel() == 1, ws == 3)
    assert np.array_equal(counts.ravel(), wc)
    assert rep.glitched == (ws == 3).sum() > 100
    # one round: the glitched pixels again, around the glitched pixel nearest their centroid
    d1 = deep(cu, 1e-10, rebase=False, max_references=1)
    c1, _, g1, rep1, _ = render(cu, engine, d1, method)
    gl = ws == 3
    gx, gy = px[gl], py[gl]
    k = np.argmin((gx - gx.mean()) ** 2 + (gy - gy.mean()) ** 2)
    ox, oy = O.C.c_double(), O.C.c_double()
    O.lib().or_deep_offset(int(gx[k]), int(gy[k]), d1.width, d1.height, d1.pixel_w, d1.pixel_h,
                           O.C.byref(ox), O.C.byref(oy))
    rx, ry = O.dd_add(CENTRE[0], 0.0, ox.value), O.dd_add(CENTRE[1], 0.0, oy.value)
    orbit2, M2 = O.reference_orbit(rx[0], ry[0], d1.max_iter, 256.0, method, lo=(rx[1], ry[1]))
    c2, _, s2, _ = O.deep_pixels(om, gx, gy, d1, orbit2, M2, ref=(ox.value, oy.value))
    want = wc.copy()
    want[gl] = c2
    assert np.array_equal(c1.ravel(), want)
    assert rep1.references == 2 and rep1.glitched == (s2 == 3).sum()
    assert np.array_equal(g1.ravel()[gl] == 1, s2 == 3)


def test_deep_rejects(gpu, engine):
    cu = gpu
    bad = deep(cu, 1e-10)
    bad.bailout_sq = 2.0
    with pytest.raises(cu.IllegalArgumentError, match="bailout_sq"):
        cu.DeepFrame(engine, bad)
    far = cu.make_deep(1.9, 0.9, 1e-3, 64, 64, 100)
    with pytest.raises(cu.IllegalArgumentError, match=r"\|c\| <= 2"):
        cu.DeepFrame(engine, far)
    with pytest.raises(cu.IllegalArgumentError, match="method"):
        cu.DeepFrame(engine, deep(cu, 1e-10), 16)
    f = cu.DeepFrame(engine, deep(cu, 1e-10))
    try:
        with pytest.raises(cu.IllegalStateError):
            f.report()
    finally:
        f.close()


# Share of the 65 536-pixel subset whose count differs from the double-double
# oracle, per width: the bound each method is held to.  FP32 deltas are far less
# accurate at this centre (measured 6.48 % at 1e-6 and 12.95 % at 1e-10, equal to
# the oracle's FP32 perturbation, which the kernel matches bit for bit); direct
# FP64 is held to what it measures (1.80 % and 15.21 %).
MISMATCH_BOUND = {1e-6: {64: 0.005, 32: 0.070, 1: 0.020},
                  1e-10: {64: 0.005, 32: 0.135, 1: 0.160},
                  1e-20: {64: 0.005, 32: 0.005}}


@pytest.mark.parametrize("width", [1e-6, 1e-10, 1e-20])
def test_c4_frames_against_double_double(gpu, engine, width):
    """1920x1080 at maxIter 50000: each method's share of mismatching counts on the
    seeded subset against the double-double oracle stays within MISMATCH_BOUND
    (FP64 deltas 0.1-0.3 %, FP32 deltas 6.5 % / 13.0 %, direct FP64 1.8 % / 15.2 %)."""
    cu = gpu
    g = np.load(os.path.join(GOLDEN, "deep_dd_subset.npz"))
    idx = g["index"]
    want = g[f"counts_{width:.0e}"].astype(np.uint32)
    rates = {}
    for method in (64, 32, 1):
        if method == 1 and width == 1e-20:
            continue  # below the FP64 ulp of the centre: every pixel samples the same c
        c, s, _, rep, _ = render(cu, engine, deep(cu, width, 1920, 1080), method)
        got = c.ravel()[idx]
        rates[method] = float(np.mean(got != want))
        print(f"width {width:g} method {method}: mismatch vs double-double "
              f"{100 * rates[method]:.3f} %, {rep.iterations} iterations, {rep.total_ms:.2f} ms")
    for method, bound in MISMATCH_BOUND[width].items():
        assert rates[method] <= bound, (width, method, rates[method])


@pytest.mark.parametrize("centre,width", [(CENTRE, 1e-10), ((-0.5, 0.5), 0.3)])
@pytest.mark.parametrize("method", [32, 64])
def test_kept_orbit_gives_fresh_results(gpu, engine, method, centre, width):
    """With CU_DEEP_REUSE_ORBIT a zoom step toward the same centre keeps the orbit
    (computed to the creation max_iter) and renders what a fresh orbit renders.
    The second centre is interior: its orbit never escapes, so the kept orbit is
    longer than a fresh one for the frame's max_iter."""
    cu = gpu

    def desc(w, mi, flags):
        d = cu.make_deep(centre[0], centre[1], w, 192, 108, mi, 256.0)
        d.flags = flags
        return d

    keep = cu.DEEP_REBASE | cu.DEEP_REUSE_ORBIT
    f = cu.DeepFrame(engine, desc(width, 4000, keep), method)
    try:
        f.render()
        assert f.report().references == 1
        f.retarget(desc(width * 0.97, 3200, keep))
        f.render()
        rep = f.report()
        assert rep.references == 0
        kept = f.counts(), f.smooth()
    finally:
        f.close()
    counts, smooth, _, rep2, _ = render(cu, engine, desc(width * 0.97, 3200, cu.DEEP_REBASE), method)
    assert rep2.references == 1
    assert np.array_equal(kept[0], counts)
    assert np.array_equal(kept[1], smooth)
    assert counts.any()
