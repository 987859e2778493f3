"""The native host replay of the reference's call stacks on the GPU.

chaosultra_replay (csrc/harness/replay_calls.cpp) drives the C ABI through the
reference host's sequences: (A) provider + module load + initializeRendering,
(B) renderQuality, (C) three renderFast zoom frames with reuse and foveation,
(D) forceReload + launchDebugKernel + fractal switch, (E) Julia custom
parameters.  Every rendered step's pixel_info_t buffer and RGBA8 image must equal
the oracle's restatement bit for bit, the reload must reproduce (B), and the
expected error codes (missing code object, closed renderer, malformed
parameters, double free) must come back as the reference's exceptions would.
"""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def load_step(cu, out, rec):
    w, h = rec["width"], rec["height"]
    info = np.fromfile(os.path.join(out, rec["step"] + ".info"), dtype=cu.PIXEL_INFO_DTYPE)
    rgba = np.fromfile(os.path.join(out, rec["step"] + ".rgba"), dtype="<u4")
    return info.reshape(h, w), rgba.reshape(h, w)


def test_replay_call_stacks(gpu, oracle, tmp_path):
    cu = gpu
    exe = os.path.join(os.path.dirname(cu.abi.LIB_PATH), "chaosultra_replay")
    out = str(tmp_path)
    p = subprocess.run([exe, cu.KERNELS_DIR, out], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout
    assert p.stdout.count("as expected") == 4
    assert "hello from mandelbrot" in p.stdout  # the debug kernel's printf (launchDebugKernel)
    pal = np.fromfile(os.path.join(out, "palette.rgba"), dtype="<u4")
    assert np.array_equal(pal, cu.default_palette())
    with open(os.path.join(out, "manifest.jsonl")) as f:
        steps = {r["step"]: r for r in map(json.loads, f)}
    assert list(steps) == ["B", "C1", "C2", "C3", "D", "E"]
    got = {}
    for name, rec in steps.items():
        info, rgba = load_step(cu, out, rec)
        got[name] = info
        prec = 32 if rec["precision"] == cu.PRECISION_SINGLE else 64
        kind = cu.FRACTAL_JULIA if rec["fractal"] == "julia" else cu.FRACTAL_MANDELBROT
        jc = tuple(float(v) for v in rec["julia_c"].split(";")) if rec["julia_c"] else (0.0, 0.0)
        args = (prec, kind, jc, rec["width"], rec["height"], tuple(rec["segment"]),
                rec["max_iter"], rec["max_ss"], rec["flags"])
        if "prev_step" in rec:
            want = oracle.render_advanced(*args, tuple(rec["prev_segment"]), got[rec["prev_step"]],
                                          tuple(rec["focus"]))
        else:
            want = oracle.render_main(*args)
        for f in ("value", "weight", "is_reused", "weight_of_new_samples"):
            assert np.array_equal(info[f], want[f]), f"step {name}: {f} differs"
        assert np.array_equal(rgba, oracle.compose(want, pal, rec["max_ss"], False)), name
    # the reloaded module renders (B) again
    assert np.array_equal(got["D"], got["B"])
    assert steps["C3"]["flags"] & cu.FLAG_SAMPLE_REUSE and got["C3"]["is_reused"].any()
