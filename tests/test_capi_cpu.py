"""The C-ABI libraries load and export every declared entry point (CPU; no kernel runs)."""
import ctypes
import os
import re

import numpy as np
import pytest

import pdflow
from pdflow import lib as plib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pd-flow_amd", "pdflow")


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(pdf_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [
    ("pdflow_hip.h", "libpdflow_hip.so"),
    ("pdflow_cnn.h", "libpdflow_hip.so"),
    ("pdflow_png.h", "libpdflow_png.so"),
    ("pdflow_synth.h", "libpdflow_synth.so"),
    ("pdflow_cudalib.h", "libpdflow_csf.so"),
])
def test_every_declared_symbol_is_exported(header, lib):
    names = declared(header)
    assert names, header
    so = ctypes.CDLL(os.path.join(PKG, lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_csf_dropin_exports_reference_bridges():
    """The drop-in library exports the reference's bridge functions and CSF_cuda methods
    (C++ names, pdflow_cudalib.h:125-174)."""
    so = os.path.join(PKG, "libpdflow_csf.so")
    syms = os.popen(f"nm -D --defined-only {so} | c++filt").read()
    for name in ("ObjectToDevice(CSF_cuda*)", "GaussianPyramidBridge(CSF_cuda*, unsigned int, unsigned int)",
                 "AssignZerosBridge", "UpsampleBridge", "ImageGradientsBridge", "RijBridge", "WarpingBridge",
                 "MuAndStepSizesBridge", "DualVariablesBridge", "PrimalVariablesBridge", "DivergenceBridge",
                 "GradientBridge", "FilterBridge", "MotionFieldBridge", "DebugBridge", "BridgeBack",
                 "CSF_cuda::readParameters", "CSF_cuda::allocateDevMemory", "CSF_cuda::allocateMemoryNewLevel",
                 "CSF_cuda::copyNewFrames", "CSF_cuda::freeDeviceMemory", "CSF_cuda::freeLevelVariables",
                 "CSF_cuda::copyAllSolutions", "CSF_cuda::copyMotionField"):
        assert name in syms, name


def test_dropin_header_compiles_without_gpu_headers(tmp_path):
    """include/pdflow_cudalib.h is plain C++: a reference-style app compiles with g++ alone."""
    src = tmp_path / "app.cpp"
    src.write_text('#include "pdflow_cudalib.h"\n'
                   'struct App { CSF_cuda csf_host, *csf_device; };\n'
                   'int main() { App a; a.csf_device = ObjectToDevice(&a.csf_host);\n'
                   '  AssignZerosBridge(a.csf_device); BridgeBack(&a.csf_host, a.csf_device);\n'
                   '  return a.csf_host.lastStatus() == 2 ? 0 : 1; }\n')
    inc = os.path.join(ROOT, "include")
    rc = os.system(f"g++ -std=c++11 -fsyntax-only -I{inc} {src}")
    assert rc == 0


def test_create_rejects_bad_arguments_before_touching_the_device():
    lib = plib.hip_library()
    ctx = ctypes.c_void_p()
    mask = pdflow.gaussian_mask()
    fp = mask.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    bad = [
        (120, 160, 0),        # cam_mode 0
        (120, 160, 7),        # cam_mode does not divide 640x480
        (480, 640, 2),        # solve larger than the 320x240 input
        (100, 133, 4),        # does not halve ctf_levels-1 times
    ]
    for rows, cols, cam in bad:
        levels = 4
        rc = lib.pdf_create(ctypes.byref(ctx), rows, cols, 0.04, 0.35, 75.0, fp, levels, cam, 1.0, 0.8)
        assert rc == 1, (rows, cols, cam)
        assert lib.pdf_last_create_error()
    assert lib.pdf_create(None, 120, 160, 0.04, 0.35, 75.0, fp, 4, 4, 1.0, 0.8) == 1
    # every stage rejects a null context
    for st in ("assign_zeros", "rij", "warping", "filter", "motion_field"):
        assert getattr(lib, "pdf_" + st)(None) == 1


def test_batch_entry_points_reject_null_handles_before_touching_the_device():
    lib = plib.hip_library()
    pairs = (ctypes.c_uint * 2)(0, 1)
    e = (ctypes.c_float * 2)()
    r = (ctypes.c_float * 2)()
    assert lib.pdf_batch_continue(None, pairs, 2, 16) == 1
    assert lib.pdf_batch_metrics_pairs(None, pairs, 2, e, r) == 1
    assert lib.pdf_batch_resolve_plain(None, pairs, 2) == 1
    assert lib.pdf_batch_metrics(None, e, r) == 1
    n, first = ctypes.c_uint64(), ctypes.c_uint32()
    assert lib.pdf_check_exact_sequence(5, 0, 1, ctypes.byref(n), ctypes.byref(first)) == 1
    assert lib.pdf_check_exact_sequence(0, 2, 1, ctypes.byref(n), ctypes.byref(first)) == 1
    assert lib.pdf_check_exact_sequence(0, 0, 1, None, ctypes.byref(first)) == 1


def test_field_table_matches_header():
    text = open(os.path.join(ROOT, "include", "pdflow_fields.h")).read()
    body = text[text.index("enum pdf_field_id"):text.index("PDF_NUM_FIELDS")]
    ids = re.findall(r"PDF_([A-Z0-9_]+)", body)
    assert [x.lower() for x in ids] == list(plib.FIELDS.keys())


def test_arena_limit_fails_creates_with_nomem_before_touching_the_device():
    """Allocator-failure injection from the host: with an arena cap below what a solver or a batch
    needs, pdf_create and pdf_batch_create return PDF_ERR_NOMEM (4) before any device call, so this
    runs without a GPU.  Lifting the cap restores the normal path (no GPU here: argument errors
    still come first, and the allocation itself is not attempted by this test)."""
    lib = plib.hip_library()
    mask = pdflow.gaussian_mask()
    fp = mask.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    ctx = ctypes.c_void_p()
    try:
        assert lib.pdf_set_arena_limit(1 << 20) == 0                 # 1 MiB
        rc = lib.pdf_create(ctypes.byref(ctx), 120, 160, 0.04, 0.35, 75.0, fp, 4, 4, 1.0, 0.8)
        assert rc == 4 and not ctx.value
        msg = lib.pdf_last_create_error().decode()
        assert "limit" in msg and "1048576" in msg, msg
        rows = (ctypes.c_uint * 2)(240, 120)
        b = ctypes.c_void_p()
        rc = lib.pdf_batch_create(ctypes.byref(b), 2, rows, None, 0.04, 0.35, 75.0, fp, 1.0, None)
        assert rc == 4 and not b.value
        assert "limit" in lib.pdf_batch_last_error(None).decode()
        # a bad argument is still reported as such under the cap
        assert lib.pdf_create(ctypes.byref(ctx), 120, 160, 0.04, 0.35, 75.0, fp, 4, 7, 1.0, 0.8) == 1
    finally:
        lib.pdf_set_arena_limit(0)
