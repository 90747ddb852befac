"""Warm-start injection (the learned initialisation's entry point) against the oracle."""
import sys

import numpy as np
import pytest

from oracle import Oracle
from pdflow import Batch, Solver, level_table, synth_pair_float
from pdflow.lib import DeviceBuffer

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(7)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_warm(rows, cam, frames, start, warm, iters):
    """The oracle driven stage by stage: the start level is zeroed, then du_prev, dv_prev,
    dw_new and dw_acc are overwritten with the warm planes; finer levels as usual."""
    o = Oracle(rows, cam, threads=8)
    I1, Z1, I2, Z2 = frames
    o.copy_new_frames(I1, Z1)
    o.gaussian_pyramid()
    o.copy_new_frames(I2, Z2)
    o.gaussian_pyramid()
    for (i, r, c, limg, _) in level_table(rows, cam):
        if i < start:
            continue
        o.allocate_level(r, c, i, limg)
        o.stage("assign_zeros")
        if i == start:
            P = r * c
            o.set_field("du_prev", warm[:P])
            o.set_field("dv_prev", warm[P:2 * P])
            o.set_field("dw_new", warm[2 * P:])
            o.set_field("dw_acc", warm[2 * P:])
        else:
            o.stage("upsample")
        for st in ("rij", "image_gradients", "warping", "mu_and_step_sizes"):
            o.stage(st)
        o.pd_iterations(iters[i])
        o.stage("filter")
        o.stage("motion_field")
    return o.motion_field()


@pytest.mark.parametrize("rows,start", [(240, 2), (120, 1), (480, 2)])
def test_warm_start_matches_oracle_and_batch(gpu_available, rows, start):
    cam = 480 // rows
    say(f"warm start rows={rows} start={start}")
    frames = synth_pair_float(900 + rows, 480 // cam, 640 // cam)
    t = level_table(rows, cam)
    r0, c0 = t[start][1], t[start][2]
    P = r0 * c0
    warm = np.concatenate([RNG.normal(0, 0.6, P), RNG.normal(0, 0.6, P), RNG.normal(0, 0.02, P)]).astype(np.float32)
    iters = [max(1, x[4] // 4) + 1 for x in t]           # reduced, odd and even counts
    g = Solver(rows, cam)
    g.copy_new_frames(frames[0], frames[1])
    g.gaussian_pyramid()
    g.copy_new_frames(frames[2], frames[3])
    g.gaussian_pyramid()
    buf = DeviceBuffer(warm.nbytes)
    buf.upload(warm)
    g.solve_warm(start, buf.ptr.value, iters)
    got = g.motion_field()
    want = oracle_warm(rows, cam, frames, start, warm, iters)
    for a, b, n in zip(got, want, ("dx", "dy", "dz")):
        np.testing.assert_array_equal(a, b, err_msg=n)
    # the batch, with one warm pair between two cold ones
    b = Batch([rows, rows, rows], iters=[iters, iters, iters])
    for i in range(3):
        b.set_inputs(i, *frames)
    ptr = b.set_start_level(1, start)
    from pdflow.lib import hip_library
    import ctypes
    hip_library().pdf_memcpy_h2d(ctypes.c_void_p(ptr), warm.ctypes.data, warm.nbytes)
    b.solve(graph=True)
    b.synchronize()
    for a, w in zip(b.flow(1), got):
        np.testing.assert_array_equal(a, w)
    cold = Solver(rows, cam).solve_pair(*frames, iters=iters)
    for a, w in zip(b.flow(0), cold):
        np.testing.assert_array_equal(a, w)
