"""The oracle's FP32 iteration, palette compose, under-sampled kernel and
foveation, pinned by brute force, known answers and closed forms (CPU).

The checks live in tests/oracle_pins.py so that tests/test_oracle_mutants.py can
run the same checks against oracle builds with planted mistakes.
"""
import pytest

import oracle_pins as P


@pytest.mark.parametrize("precision", [32, 64])
def test_mandelbrot_bruteforce(oracle, precision):
    differ = P.check_mandelbrot_bruteforce(oracle.lib(), precision)
    # the float32 and float64 brute forces disagree at some of these points, so
    # agreeing with one of them everywhere says which precision the oracle runs in
    assert differ > 0


@pytest.mark.parametrize("precision", [32, 64])
def test_julia_bruteforce(oracle, precision):
    P.check_julia_bruteforce(oracle.lib(), precision)


def test_special_points_both_precisions(oracle):
    P.check_special_points(oracle.lib())


@pytest.mark.parametrize("precision", [32, 64])
def test_mirror_symmetry_dyadic_centres(oracle, precision):
    P.check_mirror_symmetry(oracle.lib(), precision)


def test_tile_points_f32(cu, oracle):
    P.check_tile_points_f32(oracle.lib(), cu)


def test_tile_rows_f32_bruteforce(cu, oracle):
    P.check_tile_rows_f32(oracle.lib(), cu)


def test_compose_palette_known_answers(oracle):
    P.check_compose_palette(oracle.lib())


def test_compose_smooth_known_answers(oracle):
    P.check_compose_smooth_palette(oracle.lib())


def test_colorize_sample_count_known_answers(oracle):
    P.check_colorize_sample_count(oracle.lib())


def test_undersampled_pixels_and_values(oracle):
    P.check_undersampled(oracle.lib())


def test_foveation_closed_form(oracle):
    P.check_foveation(oracle.lib())


@pytest.mark.parametrize("precision", [64, 32])
def test_c1_frame_bruteforce(cu, oracle, precision):
    P.check_c1_frame(oracle.lib(), cu, precision)


def test_render_vote_sample_counts(oracle):
    P.check_render_votes(oracle.lib())
