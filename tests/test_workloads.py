"""The seeded uneven-tile generator (chaos-ultra_amd/workloads.py), without a GPU:
reproducible, inside its stated ranges, and every centre within 1e-3 of a
bounded/escaping transition along its real-axis row."""
import importlib
import math

import numpy as np
import pytest

cu = importlib.import_module("chaos-ultra_amd")
wl = importlib.import_module("chaos-ultra_amd.workloads")


@pytest.fixture(scope="module")
def batch():
    return wl.uneven_tiles(cu, 64, seed=7)


def test_reproducible(batch):
    tiles, info = batch
    again, _ = wl.uneven_tiles(cu, 64, seed=7)
    assert all(bytes(a) == bytes(b) for a, b in zip(tiles, again))
    other, _ = wl.uneven_tiles(cu, 8, seed=8)
    assert bytes(other[0]) != bytes(tiles[0])


def test_ranges(batch):
    tiles, info = batch
    for t, i in zip(tiles, info):
        assert 64 <= i["side"] <= 1024 and t.width == t.height == i["side"]
        assert 1e-6 <= i["width"] <= 3.0
        assert 256 <= i["max_iter"] <= 65536 and t.max_iter == i["max_iter"]
        assert t.pixel_w == pytest.approx(i["width"] / i["side"], rel=1e-12)
        if i["kind"] == "julia":
            assert t.kind == cu.FRACTAL_JULIA and math.hypot(*i["julia_c"]) <= 0.8
        else:
            assert t.kind == cu.FRACTAL_MANDELBROT
    frac = np.mean([i["kind"] == "julia" for i in info])
    assert 0.15 <= frac <= 0.45


def member(i, x, y):
    x, y = np.atleast_1d(x).astype(float), np.atleast_1d(y).astype(float)
    if i["kind"] == "julia":
        return wl.bounded(x, y, i["julia_c"][0], i["julia_c"][1])
    return wl.bounded(np.zeros_like(x), np.zeros_like(y), x, y)


def test_centres_on_the_boundary(batch):
    """The centre is the midpoint of a final bisection bracket whose ends test
    differently (one bounded for 1024 steps, one escaping): a transition of the
    set lies within the bracket, far closer than 1e-3."""
    _, info = batch
    for i in info:
        cx, cy = i["centre"]
        a, b = i["bracket"]
        assert abs(b - a) == i["bisection_gap"] < 1e-12
        assert min(a, b) <= cx <= max(a, b)
        ia, ib = member(i, [a, b], [cy, cy])
        assert ia != ib
