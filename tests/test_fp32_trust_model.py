"""Whether a precision-escalation round (FP32 deltas first, untrusted pixels again
with FP64 deltas) can pay on the deep-zoom frames: a CPU model (CPU).

tests/emulate/fp32_trust.cpp runs the oracle's FP32-delta loop on seeded pixels
of the 1920x1080 frames with a first-order rounding-error bound carried along,
flags the escaping pixels whose escape test or the test before it lies within
that bound of the bailout radius, and compares with FP64 deltas.  At 1e-6 the
bound flags about a quarter of the pixels and catches about 99 % of those whose
FP32 count is wrong; at 1e-10 every pixel rebases with its dc below the FP32
resolution of the rebased delta and the bound flags all of them.  With FP32
issuing at only twice the FP64 rate on this GPU (and the bound adding about half
again to the FP32 step), the escalated frame costs more than FP64 deltas alone,
so the deep-zoom form offers FP32 and FP64 deltas without an escalation round.
"""
import os
import subprocess

import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    exe = tmp_path_factory.mktemp("trust") / "fp32_trust"
    oracle = os.path.join(ROOT, "oracle")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-Wextra",
           os.path.join(ROOT, "tests", "emulate", "fp32_trust.cpp"), "-o", str(exe),
           f"-L{oracle}", "-loracle", f"-Wl,-rpath,{oracle}"]
    build = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if build.returncode != 0:
        pytest.fail(build.stdout[-4000:])
    return str(exe)


def run(model, width, pixels, seed=7):
    out = subprocess.run([model, width, "50000", str(pixels), str(seed)], stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:]
    toks = out.stdout.split()
    return {k: float(v) for k, v in zip(toks[::2], toks[1::2])}


def test_trust_bound_at_1e6(model):
    r = run(model, "1e-6", 2048)
    assert r["loop_mismatch"] == 0               # the bound does not change the iteration
    assert 0.04 <= r["differ"] <= 0.09           # FP32 counts off FP64-delta counts (6.5 %)
    assert 0.15 <= r["flagged"] <= 0.35          # a quarter of the frame would be re-rendered
    assert r["caught"] >= 0.95 and r["after"] <= 0.005


def test_trust_bound_at_1e10(model):
    r = run(model, "1e-10", 2048)
    assert r["loop_mismatch"] == 0
    assert 0.09 <= r["differ"] <= 0.16           # 13 %
    assert r["flagged"] >= 0.99                  # every pixel: escalation = FP64 deltas + FP32
