"""Each oracle pin catches the mistakes it is there for (CPU).

Every case below copies oracle/chaos_oracle.cpp, plants one plausible mistake
(a wrong precision, operation order, comparison, rounding mode, index, constant
or sign; each keeps every memory access in bounds), builds the copy with the
oracle's own flags into a temporary directory, and runs the pins of
tests/oracle_pins.py against it: at least one of them must fail.  The real oracle passes all of them (tests/test_oracle_pins.py).
"""
from __future__ import annotations

import os
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

import _oracle
import oracle_pins as P

ORACLE_DIR = os.path.join(_oracle.ROOT, "oracle")
SOURCE = os.path.join(ORACLE_DIR, "chaos_oracle.cpp")

# (name, text in chaos_oracle.cpp, replacement)
MUTANTS = [
    # --- computeFractal (mandelbrot.cu:16-23, julia.cu:5-17)
    ("bailout_le", "if (!(m < bailout_sq)) {", "if (!(m <= bailout_sq)) {"),
    ("x_order", "Real xn = x * x - y * y + cx;", "Real xn = x * x + (cx - y * y);"),
    ("y_in_double", "y = 2 * x * y + cy;            // ((2*x)*y) + cy",
     "y = (Real)(2 * (double)x * y + cy);"),
    ("x_in_double", "Real xn = x * x - y * y + cx;", "Real xn = (Real)((double)x * x - (double)y * y + cx);"),
    ("y_sign", "y = 2 * x * y + cy;            // ((2*x)*y) + cy", "y = 2 * x * y - cy;"),
    ("interior_max", "if (i == max_iter) return 0;  // mandelbrot.cu:22-23",
     "if (i == max_iter) return (float)max_iter;"),
    ("julia_c_swapped", "escape_loop<Real>(max_iter, px, py, (Real)jcx, (Real)jcy, (Real)4, &m, &esc);",
     "escape_loop<Real>(max_iter, px, py, (Real)jcy, (Real)jcx, (Real)4, &m, &esc);"),
    ("julia_interior_zero", "        return (float)i;\n    }\n    uint32_t i = escape_loop",
     "        return i == max_iter ? 0.0f : (float)i;\n    }\n    uint32_t i = escape_loop"),
    # --- FP32 sample positions of a tile
    ("tile_f_sign", "*y = t.top_f - ((float)py + (float)dy) * t.pixel_h_f;",
     "*y = t.top_f + ((float)py + (float)dy) * t.pixel_h_f;"),
    ("tile_f_in_double", "*x = t.left_f + ((float)px + (float)dx) * t.pixel_w_f;",
     "*x = (float)(t.left + ((double)px + dx) * t.pixel_w);"),
    ("tile_f_swap", "*x = t.left_f + ((float)px + (float)dx) * t.pixel_w_f;",
     "*x = t.left_f + ((float)px + (float)dy) * t.pixel_w_f;"),
    # --- colorize / compose (mandelbrot.cu:27-35, fractalRendererGeneric.cu:455-475)
    ("colorize_rint", "uint32_t vi = value > 0.0f ? (uint32_t)std::round(value) : 0u;",
     "uint32_t vi = value > 0.0f ? (uint32_t)std::rint(value) : 0u;"),
    ("colorize_floor", "uint32_t vi = value > 0.0f ? (uint32_t)std::round(value) : 0u;",
     "uint32_t vi = value > 0.0f ? (uint32_t)value : 0u;"),
    ("colorize_off_by_one", "return palette[length - (vi % length) - 1];",
     "return palette[length - (vi % length)];"),
    ("colorize_no_reverse", "return palette[length - (vi % length) - 1];",
     "return palette[vi % length];"),
    ("slot_rint", "uint32_t vi = v > 0.0 ? (uint32_t)std::round(v) : 0u;",
     "uint32_t vi = v > 0.0 ? (uint32_t)std::rint(v) : 0u;"),
    ("slot_off_by_one", "    return length - (vi % length) - 1;\n}",
     "    return (length - (vi % length)) % length;\n}"),
    ("slot_negative", "uint32_t vi = v > 0.0 ? (uint32_t)std::round(v) : 0u;",
     "uint32_t vi = (uint32_t)(int64_t)std::round(v);"),
    # --- colorizeSampleCount (fractalRendererGeneric.cu:67-77)
    ("grey_rounded", "uint32_t v = (uint32_t)(int)(255 * rel) & 0xffu;",
     "uint32_t v = (uint32_t)(int)std::round(255 * rel) & 0xffu;"),
    ("grey_no_clamp", "sample_count = std::min(sample_count, full);", "(void)0;"),
    ("grey_reused_weight",
     "if (p.is_reused) color = colorize_sample_count((uint32_t)p.weight_of_new_samples, full);",
     "(void)0;"),
    # --- fractalRenderUnderSampled (fractalRendererGeneric.cu:478-503)
    ("under_row_bound", "for (uint64_t y = 0; (uint32_t)y < H - level; y += level)",
     "for (uint64_t y = 0; (uint32_t)y <= H - level; y += level)"),
    ("under_weight", "p.weight = 1 / (float)level;", "p.weight = (float)level;"),
    ("under_centre", "float cx = lbx + (uint32_t)x * pw;", "float cx = lbx + ((uint32_t)x + 0.5f) * pw;"),
    ("under_y_from_bottom", "float cy = rty - (uint32_t)y * ph;", "float cy = lby + (uint32_t)y * ph;"),
    # --- sampleTheFractal / computeDispersion (fractalRendererGeneric.cu:85-155, 24-34)
    ("vote_precedence",
     "if (l->in_loop && ((adaptive && (i > 0 && i < kAdaptiveThreshold)) ||\n                               (i == l->sample_count / 2)))",
     "if (l->in_loop && (adaptive && ((i > 0 && i < kAdaptiveThreshold) ||\n                               (i == l->sample_count / 2))))"),
    ("vote1_count", "if (i == 1 && vote1) {\n                for (Lane *l : part) l->sample_count = i + 1;",
     "if (i == 1 && vote1) {\n                for (Lane *l : part) l->sample_count = i + 2;"),
    ("dispersion_nan_guard", "    return variance / mean;", "    return mean == 0 ? 0 : variance / mean;"),
    ("count_truncated", "l->sample_count = std::min(kMaxSuperSampling, (uint32_t)std::round(l->sample_count_f));",
     "l->sample_count = std::min(kMaxSuperSampling, (uint32_t)l->sample_count_f);"),
    ("half_rounded_up", "(i == l->sample_count / 2)))", "(i == (l->sample_count + 1) / 2)))"),
    ("no_sample_weight", "            l->sample_count_f = 0;\n            l->result = 0;\n            l->in_loop = false;\n            continue;",
     "            l->sample_count_f = 1;\n            l->result = 0;\n            l->in_loop = false;\n            continue;"),
    # --- getFoveationAdvisedSampleCount (fractalRendererGeneric.cu:209-257)
    ("fov_pixel_cm", "const float pixel_width_cm = 0.02652;", "const float pixel_width_cm = 0.0265;"),
    ("fov_tile_transposed", "uint32_t nx = px - px % 8, ny = py - py % 4;",
     "uint32_t nx = px - px % 4, ny = py - py % 8;"),
    ("fov_peripheral", "const float peripheral = 60;", "const float peripheral = 50;"),
    ("fov_screen", "const float screen_distance = 60;", "const float screen_distance = 50;"),
    ("fov_foveal_scaled", "const float foveal = max_ss >= 1 ? 5.5 : 5.5 * max_ss;",
     "const float foveal = 5.5;"),
    ("fov_clamp_low", "float rel = (float)std::fmin(1.0, (double)lm);",
     "float rel = (float)std::fmax(0.0, std::fmin(1.0, (double)lm));"),
    ("fov_min_one", "adv = (float)std::fmax(1.0, (double)adv);", "(void)0;"),
]


def _pins(L, cu):
    """Run every pin; return the names of those that fail."""
    checks = {
        "mandelbrot32": lambda: P.check_mandelbrot_bruteforce(L, 32),
        "mandelbrot64": lambda: P.check_mandelbrot_bruteforce(L, 64),
        "julia32": lambda: P.check_julia_bruteforce(L, 32),
        "julia64": lambda: P.check_julia_bruteforce(L, 64),
        "special": lambda: P.check_special_points(L),
        "mirror32": lambda: P.check_mirror_symmetry(L, 32),
        "tile_points_f32": lambda: P.check_tile_points_f32(L, cu),
        "tile_rows_f32": lambda: P.check_tile_rows_f32(L, cu),
        "compose": lambda: P.check_compose_palette(L),
        "compose_smooth": lambda: P.check_compose_smooth_palette(L),
        "grey": lambda: P.check_colorize_sample_count(L),
        "undersampled": lambda: P.check_undersampled(L),
        "foveation": lambda: P.check_foveation(L),
        "votes": lambda: P.check_render_votes(L),
    }
    failed = []
    for name, fn in checks.items():
        try:
            fn()
        except AssertionError:
            failed.append(name)
    return failed


@pytest.fixture(scope="module")
def mutant_libs():
    src = open(SOURCE).read()
    tmp = tempfile.mkdtemp(prefix="oracle_mutants_")

    def build(m):
        name, old, new = m
        assert src.count(old) == 1, f"mutant {name}: the text to replace is not unique/present"
        cpp = os.path.join(tmp, f"{name}.cpp")
        with open(cpp, "w") as f:
            f.write(src.replace(old, new))
        so = os.path.join(tmp, f"lib{name}.so")
        # the oracle's flags (oracle/Makefile); -Bsymbolic keeps each copy's calls inside it
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-w", "-ffp-contract=off", "-fno-fast-math",
               "-fopenmp", f"-I{ORACLE_DIR}", "-shared", "-Wl,-Bsymbolic", "-o", so, cpp]
        p = subprocess.run(cmd, capture_output=True, text=True)
        assert p.returncode == 0, p.stderr
        return name, so

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        libs = dict(ex.map(build, MUTANTS))
    yield libs
    for so in libs.values():
        os.remove(so)
    for m in MUTANTS:
        os.remove(os.path.join(tmp, f"{m[0]}.cpp"))
    os.rmdir(tmp)


def test_unmutated_copy_passes(cu, oracle):
    assert _pins(oracle.lib(), cu) == []


@pytest.mark.parametrize("name", [m[0] for m in MUTANTS])
def test_mutant_is_caught(cu, mutant_libs, name):
    L = _oracle.bind(mutant_libs[name])
    failed = _pins(L, cu)
    assert failed, f"no pin catches the planted mistake {name}"
