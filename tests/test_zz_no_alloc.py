"""The allocation-free pair solve (pdf_solve_pair_noalloc) against pdf_solve_pair_device.

Both run the same kernels on the same inputs; the allocation-free one keeps its working planes in
the library's per-GPU device array instead of a context's arena.  Two calls in a row, the second
with smaller frames than the first, must each give the flow of the context path bit for bit, so
nothing that a call leaves in the shared array can reach the next one.
"""
import ctypes
import sys

import numpy as np
import pytest

from pdflow import PDFlowError, Solver, synth_pair_float
from pdflow import lib as plib

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def upload_pair(seed, cam_mode):
    frames = synth_pair_float(seed, 480 // cam_mode, 640 // cam_mode)
    bufs = []
    for a in frames:
        b = plib.DeviceBuffer(a.nbytes)
        b.upload(a)
        bufs.append(b)
    return bufs


def context_flow(bufs, rows, cam_mode, iters=None):
    """pdf_solve_pair_device on a fresh context: the entry point the allocation-free one mirrors"""
    s = Solver(rows, cam_mode)
    s.solve_pair_device(bufs, iters=iters)
    s.synchronize()
    ptrs = [ctypes.c_void_p() for _ in range(3)]
    s._chk(s.lib.pdf_motion_field_device(s.ctx, *[ctypes.byref(p) for p in ptrs]))
    out = []
    for p in ptrs:
        a = np.empty(rows * (rows * 4 // 3), np.float32)
        if s.lib.pdf_memcpy_d2h(a.ctypes.data, p, a.nbytes):
            raise PDFlowError(3, "device-to-host copy")
        out.append(a)
    s.close()
    return out


def noalloc_flow(bufs, rows, cam_mode, iters=None):
    n = rows * (rows * 4 // 3)
    outs = [plib.DeviceBuffer(4 * n) for _ in range(3)]
    plib.solve_pair_noalloc(bufs, outs, rows=rows, cam_mode=cam_mode, iters=iters)
    # a blocking copy on the null stream follows the solve queued there
    return [o.download(np.empty(n, np.float32)) for o in outs]


def test_noalloc_equals_the_context_solve_twice_in_a_row(gpu_available):
    # first a VGA pair (480x640 frames, 6 levels), then a smaller one: 320x240 frames solved at
    # 120x160 (4 levels, the finest level one pyramid level below the frames) with another
    # schedule, after the larger call has filled the shared planes
    cases = [(480, 1, 7001, None), (120, 2, 7002, [20, 30, 40, 50])]
    for rows, cam_mode, seed, iters in cases:
        say(f"allocation-free solve rows={rows} cam_mode={cam_mode} seed={seed} iters={iters}")
        bufs = upload_pair(seed, cam_mode)
        got = noalloc_flow(bufs, rows, cam_mode, iters)
        want = context_flow(bufs, rows, cam_mode, iters)
        for name, a, b in zip(("dx", "dy", "dz"), got, want):
            assert np.isfinite(a).all(), f"{name}: non-finite values"
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (
                f"rows={rows}: {name} differs from pdf_solve_pair_device in {np.sum(a != b)} of {a.size} values")
        assert np.abs(got[0]).max() > 0, "the pair moves, so the flow is not all zero"


def test_noalloc_rejects_bad_arguments(gpu_available):
    bufs = upload_pair(7003, 4)
    outs = [plib.DeviceBuffer(4 * 120 * 160) for _ in range(3)]
    with pytest.raises(PDFlowError) as e:
        plib.solve_pair_noalloc(bufs, outs, rows=120, cam_mode=3)
    assert e.value.code == 1 and "cam_mode" in str(e.value)
    with pytest.raises(PDFlowError) as e:
        plib.solve_pair_noalloc(bufs, outs, rows=120, cam_mode=4, cols=150)
    assert e.value.code == 1
    lib = plib.hip_library()
    mask = np.ones(25, np.float32) / 25
    rc = lib.pdf_solve_pair_noalloc(120, 160, 0.04, 0.35, 75.0, plib._fptr(mask), 4, 4, 1.0,
                                    bufs[0].ptr, None, bufs[2].ptr, bufs[3].ptr, None,
                                    outs[0].ptr, outs[1].ptr, outs[2].ptr, None)
    assert rc == 1 and lib.pdf_last_create_error().decode() == "null argument"


def test_noalloc_solve_captured_into_a_graph(gpu_available):
    """Code under HIP graph capture is one of the callers the entry point is for: its launches,
    captured on a stream, form a graph whose every replay gives the eager call's flow (the outputs
    are overwritten with NaN by eager copies between the replays)."""
    hip = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p
    for name, args in (("hipStreamCreate", [ctypes.POINTER(vp)]), ("hipStreamBeginCapture", [vp, ctypes.c_int]),
                       ("hipStreamEndCapture", [vp, ctypes.POINTER(vp)]),
                       ("hipGraphInstantiate", [ctypes.POINTER(vp), vp, vp, ctypes.c_char_p, ctypes.c_size_t]),
                       ("hipGraphLaunch", [vp, vp]), ("hipStreamSynchronize", [vp]), ("hipGraphExecDestroy", [vp]),
                       ("hipGraphDestroy", [vp]), ("hipStreamDestroy", [vp])):
        getattr(hip, name).argtypes = args
        getattr(hip, name).restype = ctypes.c_int
    rows, cam_mode = 240, 2
    say(f"allocation-free solve under graph capture rows={rows} cam_mode={cam_mode}")
    bufs = upload_pair(7004, cam_mode)
    want = noalloc_flow(bufs, rows, cam_mode)      # eager; also resolves the device array's address
    n = rows * (rows * 4 // 3)
    outs = [plib.DeviceBuffer(4 * n) for _ in range(3)]
    stream, graph, exe = vp(), vp(), vp()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        assert hip.hipStreamBeginCapture(stream, 1) == 0        # hipStreamCaptureModeThreadLocal
        try:
            plib.solve_pair_noalloc(bufs, outs, rows=rows, cam_mode=cam_mode, stream=stream)
        finally:
            end = hip.hipStreamEndCapture(stream, ctypes.byref(graph))
        assert end == 0 and graph.value
        assert hip.hipGraphInstantiate(ctypes.byref(exe), graph, None, None, 0) == 0
        for rep in range(3):
            for o in outs:
                o.upload(np.full(n, np.nan, np.float32))
            assert hip.hipGraphLaunch(exe, stream) == 0
            assert hip.hipStreamSynchronize(stream) == 0
            got = [o.download(np.empty(n, np.float32)) for o in outs]
            for name, a, b in zip(("dx", "dy", "dz"), got, want):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"replay {rep}: {name} differs"
    finally:
        if exe.value:
            hip.hipGraphExecDestroy(exe)
        if graph.value:
            hip.hipGraphDestroy(graph)
        hip.hipStreamDestroy(stream)
