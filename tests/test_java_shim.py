"""Consistency of the Java/JNI shim (chaos-ultra_amd/java/), which this image
cannot compile (no JDK): every native method of HipNative has its JNI function
in chaosultra_jni.c with the matching name, every C-ABI call the glue makes is
declared in include/chaosultra.h, and ModelMarshal's field offsets equal the
offsets of cu_rendering_model (checked through the ctypes mirror, itself checked
against the header's layout by the library)."""
import ctypes as C
import os
import re

from conftest import ROOT

JAVA = os.path.join(ROOT, "chaos-ultra_amd", "java")
PKG = os.path.join(JAVA, "src", "main", "java", "cz", "cuni", "mff", "cgg", "teichmaa", "chaosultra",
                   "hiprenderer")


def read(*p):
    with open(os.path.join(*p)) as f:
        return f.read()


def test_native_methods_have_jni_functions():
    java = read(PKG, "HipNative.java")
    natives = set(re.findall(r"static native \S+ (\w+)\(", java))
    assert len(natives) == 19
    glue = read(JAVA, "jni", "chaosultra_jni.c")
    jni = set(re.findall(r"^NATIVE\(\w+, (\w+)\)", glue, re.M))
    assert natives == jni


def test_glue_calls_only_declared_entry_points():
    header = read(ROOT, "include", "chaosultra.h")
    declared = set(re.findall(r"\b(cu_\w+)\s*\(", header))
    glue = read(JAVA, "jni", "chaosultra_jni.c")
    called = set(re.findall(r"\b(cu_\w+)\s*\(", glue))
    assert called and called <= declared, called - declared


def test_model_marshal_offsets_match_the_struct(cu):
    java = read(PKG, "ModelMarshal.java")
    consts = {k: int(v) for k, v in re.findall(r"(\w+) = (\d+)", java)}
    M = cu.RenderingModel
    fields = {"LBX": "left_bottom_x", "LBY": "left_bottom_y", "RTX": "right_top_x",
              "RTY": "right_top_y", "MAX_ITER": "max_iterations", "MAX_SS": "max_super_sampling",
              "ADAPTIVE": "use_adaptive_super_sampling", "VISUALISE": "visualise_sample_count",
              "FOVEATION": "use_foveated_rendering", "FOCUS_X": "mouse_focus_x",
              "FOCUS_Y": "mouse_focus_y", "ZOOMING": "zooming", "ZOOMING_IN": "zooming_in",
              "REUSE": "use_sample_reuse", "REUSE_DIRTY": "sample_reuse_cache_dirty",
              "PRECISION": "floating_point_precision", "ERROR_COUNT": "error_count",
              "LAST_ERROR": "last_error", "CUSTOM_PARAMS": "fractal_custom_params"}
    for const, field in fields.items():
        assert consts[const] == getattr(M, field).offset, const
    assert consts["TEXT_BYTES"] == M.last_error.size == M.fractal_custom_params.size
    native = read(PKG, "HipNative.java")
    assert int(re.search(r"MODEL_BYTES = (\d+)", native).group(1)) == C.sizeof(M)
