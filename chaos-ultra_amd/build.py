"""Build the gfx950 code objects and the C-ABI library in-tree.

Outputs (git-ignored, shipped with the tree):
  kernels/<fractal>.hsaco       one module per fractal (render_generic.hip + fractals/<fractal>.hip),
                                the counterpart of src/main/cuda/compile.sh's one PTX per fractal
  kernels/escape_family.hsaco   planar / persistent escape-time kernels and the smooth compose
  lib/libchaosultra_hip.so      the C ABI (include/chaosultra.h), host C++ over the HIP runtime,
                                plus the allocation-free deep render's kernels (deep_static.hip)

Run as a script (`python chaos-ultra_amd/build.py`) or through __graft_entry__.build().
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
KERNELS = os.path.join(PKG, "csrc", "kernels")
HOST = os.path.join(PKG, "csrc", "host")
OUT_KERNELS = os.path.join(PKG, "kernels")
OUT_LIB = os.path.join(PKG, "lib")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

FRACTALS = ("mandelbrot", "julia")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",  # the reference's rounding: no fused multiply-add unless written
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]
HOST_FLAGS = [
    "-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", "-ffp-contract=off",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
]
HOST_SOURCES = ("driver.cpp", "renderer.cpp", "engine.cpp", "frame.cpp", "deep.cpp", "serve.cpp",
                "png_io.cpp")
# host sources with device code built into the library itself (compiled by hipcc)
HOST_HIP_SOURCES = ("deep_static.hip",)


def _run(cmd: list[str], cwd: str | None = None) -> None:
    p = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}\n{out}")


def _hipcc() -> str:
    return os.path.join(ROCM, "bin", "hipcc")


def _module_cmd(name: str) -> list[str]:
    out = os.path.join(OUT_KERNELS, f"{name}.hsaco")
    if name == "escape_family":
        return [_hipcc(), *HIP_FLAGS, "--genco", "escape_family.hip", "-o", out]
    return [_hipcc(), *HIP_FLAGS, "--genco", f'-DCU_FRACTAL_HEADER="fractals/{name}.hip"',
            "render_generic.hip", "-o", out]


def build_modules(jobs: int = 4) -> list[str]:
    os.makedirs(OUT_KERNELS, exist_ok=True)
    names = list(FRACTALS) + ["escape_family"]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda n: _run(_module_cmd(n), cwd=KERNELS), names))
    return [os.path.join(OUT_KERNELS, f"{n}.hsaco") for n in names]


def build_library() -> str:
    os.makedirs(OUT_LIB, exist_ok=True)
    objs = []
    with ThreadPoolExecutor(max_workers=len(HOST_SOURCES) + len(HOST_HIP_SOURCES)) as ex:
        def compile_one(src: str) -> str:
            obj = os.path.join(OUT_LIB, src.replace(".cpp", ".o"))
            _run(["g++", *HOST_FLAGS, "-c", os.path.join(HOST, src), "-o", obj])
            return obj

        def compile_hip(src: str) -> str:
            obj = os.path.join(OUT_LIB, src.replace(".hip", ".o"))
            _run([_hipcc(), *HIP_FLAGS, "-fPIC", "-c", os.path.join(HOST, src), "-o", obj])
            return obj
        objs = list(ex.map(compile_one, HOST_SOURCES)) + list(ex.map(compile_hip, HOST_HIP_SOURCES))
    lib = os.path.join(OUT_LIB, "libchaosultra_hip.so")
    tmp = lib + ".tmp"
    _run(["g++", "-shared", "-o", tmp, *objs, f"-L{ROCM}/lib", "-lamdhip64", "-lz",
          f"-Wl,-rpath,{ROCM}/lib", "-Wl,--no-undefined"])
    os.replace(tmp, lib)
    for o in objs:
        os.remove(o)
    return lib


def build_replay() -> str:
    """chaosultra_replay: the native host that replays the reference's call stacks
    through the C ABI (csrc/harness/replay_calls.cpp)."""
    exe = os.path.join(OUT_LIB, "chaosultra_replay")
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra",
          os.path.join(ROOT, "chaos-ultra_amd", "csrc", "harness", "replay_calls.cpp"),
          "-o", exe, f"-L{OUT_LIB}", "-lchaosultra_hip", "-Wl,-rpath,$ORIGIN"])
    return exe


def build_oracle() -> None:
    """The CPU oracle (test infrastructure) and, where the reference tree is present,
    the reference's own iteration functions compiled for the host."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all() -> None:
    with ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(build_modules), ex.submit(build_library), ex.submit(build_oracle)]
        for f in futs:
            f.result()
    build_replay()


if __name__ == "__main__":
    build_all()
    print("built:", OUT_KERNELS, OUT_LIB, file=sys.stderr)
