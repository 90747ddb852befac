"""The benchmark's batch configurations at their real sizes, checked pair by pair.

Mixed plain batch: 128 pairs, 64 of 480x640, 38 of 240x320 and 26 of 120x160 in one batch;
sampled pairs of every size equal their single-pair solves bit for bit.

Learned batch: 256 pairs of 480x640 with the committed network at the 60x80 level and every
level's iterations divided by 2.5.  Sampled pairs equal the explicit warm start (the network's
inputs gathered from the pyramids, one forward pass, its prediction injected at 60x80), one
sampled pair's warm planes give the oracle's staged warm solve (the per-pair sequence of
scene_flow_impair.cpp:95-168 started at 60x80, injection point pdflow_cudalib.cu:556-568), the
energy, residual and end-point-error reductions equal the NumPy restatements, and a pair the
fallback rule rejects ends with its plain solve, bit for bit.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import metrics as mt
from pdflow import Batch, Net, Solver, default_params, synth_pair_float, synth_pairs_float
from pdflow.lib import DeviceBuffer, hip_library
from pdflow.train import cut_schedule

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pd-flow_amd",
                       "pdflow", "weights", "warmstart_qvga.bin")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_mixed_128_pair_batch_matches_single_pairs(gpu_available):
    rows_list = [480] * 64 + [240] * 38 + [120] * 26
    say("mixed batch 64x480 + 38x240 + 26x120")
    frames = {}
    b = Batch(rows_list)
    for i, r in enumerate(rows_list):
        cam = 480 // r
        frames[i] = synth_pair_float(6100 + i, 480 // cam, 640 // cam)
        b.set_inputs(i, *frames[i])
    b.solve(graph=True)
    b.synchronize()
    for i in (0, 37, 63, 64, 90, 101, 102, 115, 127):
        r = rows_list[i]
        want = Solver(r, 480 // r).solve_pair(*frames[i])
        for c, (a, w) in enumerate(zip(b.flow(i), want)):
            assert np.array_equal(a, w), (i, r, c)


def oracle_warm_pair(frames, warm, iters):
    """One VGA pair through the oracle, started at the 60x80 level (coarse-first index 2) from the
    warm planes (du, dv in pixels of that level, dw in metres) with zero duals"""
    from oracle import Oracle
    from pdflow import level_table
    o = Oracle(480, 1, threads=8)
    o.copy_new_frames(frames[0], frames[1])
    o.gaussian_pyramid()
    o.copy_new_frames(frames[2], frames[3])
    o.gaussian_pyramid()
    for (i, r, c, limg, _) in level_table(480, 1):
        if i < 2:
            continue
        o.allocate_level(r, c, i, limg)
        o.stage("assign_zeros")
        if i == 2:
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


def test_learned_256_vga_batch(gpu_available):
    n, rows, cols = 256, 480, 640
    say(f"learned batch of {n} pairs of {rows}x{cols}")
    planes = synth_pairs_float(8800, n, rows, cols)
    cut = cut_schedule(rows, 2.5)
    net = Net(max_batch=n, h=60, w=80)
    net.load(WEIGHTS)
    lb = Batch([rows] * n, iters=[cut] * n)
    pb = Batch([rows] * n)
    for i in range(n):
        frames = [p[i] for p in planes]
        lb.set_inputs(i, *frames)
        pb.set_inputs(i, *frames)
    lb.set_learned(net)
    lb.solve(graph=True)
    pb.solve(graph=True)
    lb.synchronize()
    pb.synchronize()
    sampled = (0, 97, 200, 255)
    learned = {i: lb.flow(i) for i in sampled}
    plain = {i: pb.flow(i) for i in sampled}

    # the explicit path: the same network inputs gathered from the solved batch's pyramids, one
    # forward pass of all 256, the prediction injected into a small batch of the sampled pairs
    h, w = 60, 80
    bx, bt, bm = DeviceBuffer(n * h * w * 16), DeviceBuffer(n * h * w * 12), DeviceBuffer(n * h * w * 4)
    lb.learn_samples(h, w, bx.ptr.value, bt.ptr.value, bm.ptr.value)
    x = bx.download(np.empty((n, h, w, 4), np.float32))
    y = net.forward(x)
    assert np.isfinite(y).all()
    wb = Batch([rows] * len(sampled), iters=[cut] * len(sampled))
    warm = {}
    for j, i in enumerate(sampled):
        wb.set_inputs(j, *(p[i] for p in planes))
        ptr = wb.set_start_level(j, 2)
        warm[i] = np.ascontiguousarray(np.concatenate(
            [y[i, ..., 0].T.ravel(), y[i, ..., 1].T.ravel(), (y[i, ..., 2] / np.float32(10)).T.ravel()]), np.float32)
        assert hip_library().pdf_memcpy_h2d(ctypes.c_void_p(ptr), warm[i].ctypes.data, warm[i].nbytes) == 0
    wb.solve(graph=False)
    wb.synchronize()
    for j, i in enumerate(sampled):
        for c, (a, e) in enumerate(zip(learned[i], wb.flow(j))):
            assert np.array_equal(a, e), ("explicit warm start", i, c)

    # one sampled pair's warm planes through the oracle's staged warm solve
    i = 97
    say(f"oracle warm solve of pair {i}")
    want = oracle_warm_pair([p[i] for p in planes], warm[i], cut)
    for c, (a, e) in enumerate(zip(learned[i], want)):
        np.testing.assert_array_equal(a, e, err_msg=f"oracle warm solve, component {c}")

    # the rule's reductions on the sampled pairs against the NumPy restatements
    e_all, r_all = lb.metrics()
    pe, pr = pb.metrics()
    epe = lb.epe(pb)
    p = default_params(rows, 1)
    names = ["dct", "dcu", "dcv", "ddt", "ddu", "ddv", "mu_uv", "ri", "rj", "du_prev", "dv_prev",
             "du_new", "dv_new", "dw_new"]
    for i in sampled:
        f = {k: lb.field(i, k) for k in names}
        want_e = mt.energy(f, rows, cols, p["lambda_i"], p["lambda_d"])
        I1, _, _ = lb.pyramid(i, "colour_old", 0)
        I2, _, _ = lb.pyramid(i, "colour", 0)
        Z1, _, _ = lb.pyramid(i, "depth_old", 0)
        want_r = mt.residual(I1, I2, Z1, lb.field(i, "du_l"), lb.field(i, "dv_l"), rows, cols)
        assert abs(e_all[i] - want_e) <= 1e-4 * want_e, (i, e_all[i], want_e)
        assert abs(r_all[i] - want_r) <= 1e-4 * want_r, (i, r_all[i], want_r)
        want_epe = mt.epe(learned[i], plain[i], Z1)
        assert abs(epe[i] - want_epe) <= 1e-4 * want_epe + 1e-9, (i, epe[i], want_epe)

    # the fallback rule with the plain solver's references: rejected pairs end with their plain
    # solve, accepted ones keep the learned flow
    flags = lb.fallback(pe, pr, 0.05)
    expect = (e_all > 1.05 * pe) | (r_all > 1.05 * pr)
    assert np.array_equal(flags.astype(bool), expect)
    rejected = [i for i in range(n) if flags[i]]
    accepted = [i for i in sampled if not flags[i]]
    say(f"fallback: {len(rejected)} of {n} pairs re-solved")
    assert rejected, "the cut schedule leaves some pairs above the rule"
    for i in rejected[:3] + rejected[-1:]:
        pf = pb.flow(i)
        for c, (a, e) in enumerate(zip(lb.flow(i), pf)):
            assert np.array_equal(a, e), ("fallback", i, c)
    for i in accepted:
        for c, (a, e) in enumerate(zip(lb.flow(i), learned[i])):
            assert np.array_equal(a, e), ("accepted", i, c)
