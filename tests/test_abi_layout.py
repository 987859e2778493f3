"""The ctypes mirrors of the C ABI's structs (chaos-ultra_amd/abi.py) against the
header itself: a probe compiled with gcc from include/chaosultra.h prints
sizeof and every field's offsetof, which must equal the ctypes layout.  The Java
shim's ModelMarshal offsets are pinned to the same numbers (test_java_shim.py)."""
import ctypes as C
import os
import subprocess

import pytest

from conftest import ROOT

STRUCTS = {  # C struct -> ctypes class name in abi.py
    "cu_rendering_model": "RenderingModel",
    "cu_tile_desc": "TileDesc",
    "cu_tune": "Tune",
    "cu_frame_tune": "FrameTune",
    "cu_deep_desc": "DeepDesc",
    "cu_deep_report": "DeepReport",
    "cu_serve_config": "ServeConfig",
    "cu_frame_request": "FrameRequest",
    "cu_frame_done": "FrameDone",
    "cu_serve_report": "ServeReport",
}


@pytest.fixture(scope="module")
def c_layout(cu, tmp_path_factory):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "chaosultra.h"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        py = getattr(cu.abi, pyname)
        lines.append(f'    printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'    printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["    return 0;", "}"]
    d = tmp_path_factory.mktemp("layout")
    src, exe = d / "probe.c", d / "probe"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE, text=True).stdout
    return {tuple(l.split()[:2]): int(l.split()[2]) for l in out.splitlines()}


@pytest.mark.parametrize("cname", sorted(STRUCTS))
def test_ctypes_mirror_matches_header(cu, c_layout, cname):
    py = getattr(cu.abi, STRUCTS[cname])
    assert c_layout[(cname, "sizeof")] == C.sizeof(py)
    for f, _ in py._fields_:
        assert c_layout[(cname, f)] == getattr(py, f).offset, f
