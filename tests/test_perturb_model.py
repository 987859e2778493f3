"""The perturbation kernel's schedule, modelled on the CPU, against the oracle.

tests/emulate/perturb_waves.cpp runs perturb_pixel's wave control flow (32-step
unchecked blocks for the whole wave, a rerun up to the first step where a rebase
test fires, escape tested once per block and replayed with checked steps) on
8x8-pixel waves of 1920x1080 deep frames, and compares every pixel's count, state
and |z|^2 with the oracle's checked loop (or_perturb_pixel_f/_d).  It pins the
claim that the schedule is bit-identical to the checked loop at every kind of
frame: deep (few rebases, long uniform runs), 1e-10 (frequent rebases), 1e-20,
shallow (a rebase every few steps), and a reference that escapes in 7 steps (no
blocks at all).  The GPU tests check the kernel itself against the same oracle.
"""
import os
import subprocess

import pytest

from conftest import ROOT

CENTRE = ("-0.743643887037151", "0.131825904205330")
CASES = [
    # precision, centre, width, max_iter, waves, seed
    (32, CENTRE, "1e-6", 20000, 16, 1),
    (64, CENTRE, "1e-6", 20000, 8, 1),
    (32, CENTRE, "1e-10", 50000, 16, 2),
    (64, CENTRE, "1e-10", 50000, 16, 2),
    (32, CENTRE, "1e-20", 50000, 8, 3),
    (32, CENTRE, "1.0", 2000, 16, 4),
    (64, CENTRE, "1.0", 2000, 16, 4),
    (32, ("0.5", "0.0"), "1e-3", 2000, 16, 6),
]


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    exe = tmp_path_factory.mktemp("model") / "perturb_waves"
    oracle = os.path.join(ROOT, "oracle")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-Wextra",
           os.path.join(ROOT, "tests", "emulate", "perturb_waves.cpp"), "-o", str(exe),
           f"-L{oracle}", "-loracle", f"-Wl,-rpath,{oracle}"]
    build = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if build.returncode != 0:
        pytest.fail(build.stdout[-4000:])
    return str(exe)


@pytest.mark.parametrize("prec,centre,width,max_iter,waves,seed", CASES)
def test_schedule_matches_checked_loop(model, prec, centre, width, max_iter, waves, seed):
    run = subprocess.run([model, str(prec), centre[0], centre[1], width, str(max_iter), str(waves),
                          str(seed)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=300)
    assert run.returncode == 0, run.stdout[-2000:]
    fields = dict(zip(run.stdout.split()[::2], run.stdout.split()[1::2]))
    assert fields["mismatches"] == "0" and int(fields["pixels"]) == 64 * waves
    if centre == CENTRE and width in ("1e-6", "1.0"):
        # the cases exercise both halves of the schedule
        assert int(fields["blocks"]) > 0 and int(fields["reruns"]) > 0
