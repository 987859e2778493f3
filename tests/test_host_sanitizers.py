"""Host code of the C ABI under AddressSanitizer + UndefinedBehaviorSanitizer.

tests/sanitize/host_sanitize.cpp is compiled with g++ together with the
library's host sources (csrc/host/*.cpp) and run on the CPU: the palette PNG
reader on every truncation and single-byte corruption of the reference's palette
files and on headers of every colour type, depth and size; PNG output round
trips; palettes, plane mapping, precision switch, launch geometry and tile
set-up at edge arguments.  No device code is built or run; the test passes when
that program exits 0 with no sanitizer report."""
import glob
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HOST = os.path.join(ROOT, "chaos-ultra_amd", "csrc", "host")


def test_host_code_under_asan_ubsan(tmp_path):
    exe = tmp_path / "host_sanitize"
    sources = sorted(glob.glob(os.path.join(HOST, "*.cpp")))
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "tests", "sanitize", "host_sanitize.cpp"), *sources, "-o", str(exe),
           f"-L{ROCM}/lib", "-lamdhip64", "-lz", f"-Wl,-rpath,{ROCM}/lib"]
    build = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if build.returncode != 0:
        pytest.fail(build.stdout[-4000:])
    pngs = sorted(glob.glob(os.path.join(GOLDEN, "palettes", "*.png")))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe), *pngs], cwd=tmp_path, stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True, env=env, timeout=600)
    assert run.returncode == 0, run.stdout[-4000:]
    assert "host_sanitize: ok" in run.stdout
    assert "runtime error" not in run.stdout and "AddressSanitizer" not in run.stdout
