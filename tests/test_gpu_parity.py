"""GPU parity: the HIP library against the CPU oracle on the same seeded inputs.

Both sides evaluate the reference's per-pixel arithmetic in the same order without FMA
contraction, so stage outputs are expected to be bit-identical; the contract bar is the
tolerance 1e-4 m absolute + 1e-3 relative per flow component (FP32).
"""
import ctypes
import sys

import numpy as np
import pytest

from oracle import Oracle
from pdflow import PDFlowError, Solver, level_table, synth_pair_float, synth_static_float
from pdflow import lib as plib

pytestmark = pytest.mark.gpu

ATOL, RTOL = 1e-4, 1e-3

SETUP_STAGES = ["rij", "image_gradients", "warping", "mu_and_step_sizes"]
STAGE_FIELDS = {
    "assign_zeros": ["du_prev", "dv_prev", "du_new", "pd", "puu", "du_acc", "dw_acc"],
    "upsample": ["du_prev", "dv_prev", "dw_new", "pd", "puu", "puv", "pvu", "pvv", "pwu", "pwv", "dw_acc"],
    "rij": ["ri", "rj", "ri_2", "rj_2"],
    "image_gradients": ["dcu_aux", "dcv_aux", "ddu_aux", "ddv_aux"],
    "warping": ["dct", "dcu", "dcv", "ddt", "ddu", "ddv"],
    "mu_and_step_sizes": ["mu_uv", "sigma_pd", "sigma_puvx", "sigma_puvy", "sigma_pwx", "sigma_pwy",
                          "tau_u", "tau_v", "tau_w"],
    "iterate": ["du_new", "dv_new", "dw_new", "du_acc", "dv_acc", "dw_acc", "pd", "puu", "puv",
                "pvu", "pvv", "pwu", "pwv"],
    "filter": ["du_new", "dv_new", "dw_new", "du_l", "dv_l", "dw_l", "pd_l", "puu_l", "pwv_l"],
    "motion_field": ["dx", "dy", "dz"],
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def both(rows, cam_mode, seed):
    I1, Z1, I2, Z2 = synth_pair_float(seed, 480 // cam_mode, 640 // cam_mode)
    g, o = Solver(rows, cam_mode), Oracle(rows, cam_mode, threads=8)
    for s in (g, o):
        s.copy_new_frames(I1, Z1)
        s.gaussian_pyramid()
        s.copy_new_frames(I2, Z2)
        s.gaussian_pyramid()
    return g, o, (I1, Z1, I2, Z2)


def assert_same(a, b, what):
    if np.array_equal(a, b):
        return
    d = np.abs(a.astype(np.float64) - b)
    bad = d > ATOL + RTOL * np.abs(b)
    assert not bad.any(), f"{what}: {bad.sum()} of {a.size} outside tolerance, max diff {d.max()}"
    say(f"  {what}: within tolerance, not bit-identical (max diff {d.max():.3g}, "
        f"{np.mean(a != b):.2%} differ)")


@pytest.mark.parametrize("rows,cam_mode", [(120, 4), (480, 1)])
def test_pyramid_parity(gpu_available, rows, cam_mode):
    say(f"pyramid parity rows={rows} cam_mode={cam_mode}")
    g, o, _ = both(rows, cam_mode, seed=41)
    for lvl in range(g.pyr_levels):
        for kind in plib.PYR_KINDS:
            assert_same(g.get_pyramid(kind, lvl), o.get_pyramid(kind, lvl), f"{kind}[{lvl}]")


@pytest.mark.parametrize("fused", [True, False])
def test_stage_by_stage_parity_120(gpu_available, fused):
    """Every stage of every level of a 120x160 solve, compared plane by plane in lock-step,
    including after 1 and 2 iterations and after the full loop."""
    say(f"lock-step parity 120x160 fused={fused}")
    g, o, _ = both(120, 4, seed=42)
    for (i, rows_i, cols_i, limg, iters) in level_table(120, 4):
        for s in (g, o):
            s.allocate_level(rows_i, cols_i, i, limg)
        plan = ["assign_zeros"] + (["upsample"] if i > 0 else []) + SETUP_STAGES
        for st in plan:
            g.stage(st)
            o.stage(st)
            for f in STAGE_FIELDS[st]:
                assert_same(g.get_field(f), o.get_field(f), f"level {i} {st} {f}")
        done = 0
        for n in (1, 1, iters - 2):
            if fused:
                g.pd_iterations(n)
            else:
                for _ in range(n):
                    for st in ("gradient", "dual_variables", "divergence", "primal_variables"):
                        g.stage(st)
            o.pd_iterations(n)
            done += n
            for f in STAGE_FIELDS["iterate"]:
                assert_same(g.get_field(f), o.get_field(f), f"level {i} after {done} iterations {f}")
        for st in ("filter", "motion_field"):
            g.stage(st)
            o.stage(st)
            for f in STAGE_FIELDS[st]:
                assert_same(g.get_field(f), o.get_field(f), f"level {i} {st} {f}")


@pytest.mark.parametrize("rows,cam_mode,seed", [(120, 4, 7), (240, 2, 8), (480, 1, 9), (240, 1, 10)])
def test_end_to_end_flow_parity(gpu_available, rows, cam_mode, seed):
    say(f"end-to-end parity rows={rows} cam_mode={cam_mode} seed={seed}")
    g, o, _ = both(rows, cam_mode, seed)
    g.solve()
    o.solve()
    flow_g, flow_o = g.motion_field(), o.motion_field()
    exact = np.mean([np.mean(a == b) for a, b in zip(flow_g, flow_o)])
    say(f"  bit-identical fraction {exact:.6f}")
    for name, a, b in zip(("dx", "dy", "dz"), flow_g, flow_o):
        assert_same(a, b, name)


def test_zero_motion_is_exactly_zero_at_vga(gpu_available):
    say("zero motion VGA")
    I, Z = synth_static_float(77, 480, 640)
    g = Solver(480, 1)
    dx, dy, dz = g.solve_pair(I, Z, I, Z)
    assert np.all(dx == 0) and np.all(dy == 0) and np.all(dz == 0)


def test_graph_replays_match_eager(gpu_available):
    """The captured solve equals the eager one bit for bit on every replay, with eager work
    and synchronisations in between."""
    say("graph replay vs eager 240x320")
    I1, Z1, I2, Z2 = synth_pair_float(5, 240, 320)
    g = Solver(240, 2)
    want = g.solve_pair(I1, Z1, I2, Z2)
    for rep in range(3):
        got = g.solve_pair(I1, Z1, I2, Z2, graph=True)
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b)
        eager = g.solve_pair(I2, Z2, I1, Z1)          # different work in between
        assert not np.array_equal(eager[0], want[0])
    # odd iteration counts leave the ping-pong planes swapped; replays must still agree
    odd = [3, 5, 7, 9, 11]
    e = g.solve_pair(I1, Z1, I2, Z2, iters=odd)
    for rep in range(2):
        got = g.solve_pair(I1, Z1, I2, Z2, iters=odd, graph=True)
        for a, b in zip(got, e):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("fused", [1, 0])
def test_csf_dropin_matches_oracle(gpu_available, fused):
    """The reference app's own call sequence, written against include/pdflow_cudalib.h."""
    say(f"CSF_cuda drop-in fused={fused}")
    import os
    so = ctypes.CDLL(os.path.join(os.path.dirname(plib.lib_path()), "libpdflow_csf.so"))
    fp = ctypes.POINTER(ctypes.c_float)
    so.pdf_csf_evaluate_pair.argtypes = [ctypes.c_uint, ctypes.c_uint] + [fp] * 7 + [ctypes.c_int]
    rows, cam = 120, 4
    I1, Z1, I2, Z2 = synth_pair_float(13, 120, 160)
    out = [np.zeros(120 * 160, np.float32) for _ in range(3)]
    P = lambda a: a.ctypes.data_as(fp)
    rc = so.pdf_csf_evaluate_pair(rows, cam, P(I1), P(Z1), P(I2), P(Z2), *[P(a) for a in out], fused)
    assert rc == 0
    o = Oracle(rows, cam, threads=8)
    want = o.solve_pair(I1, Z1, I2, Z2)
    for a, b, n in zip(out, want, ("dx", "dy", "dz")):
        assert_same(a, b, n)


def test_error_paths(gpu_available):
    say("error paths")
    g = Solver(120, 4)
    with pytest.raises(PDFlowError) as e:
        g.pd_iterations(3)                       # no level selected
    assert e.value.code == 2
    with pytest.raises(PDFlowError) as e:
        g.allocate_level(100, 160, 0, 0)        # size does not match the pyramid slot
    assert e.value.code == 1
    g.allocate_level(15, 20, 0, 3)
    lib = plib.hip_library()
    buf = np.zeros(300, np.float32)
    assert lib.pdf_get_field(g.ctx, 999, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 1
    with pytest.raises(PDFlowError):
        g.stage("upsample")                      # level 0 has nothing to upsample
    with pytest.raises(PDFlowError):
        g.gaussian_pyramid(12)


def test_nan_and_empty_depth_inputs_do_not_break_the_solve(gpu_available):
    """All-zero depth (no geometry at all) is accepted and gives zero flow everywhere (K15
    zeroes pixels without depth); NaN intensities are accepted (the call succeeds, the NaNs
    propagate through the TV coupling as they do in the reference)."""
    say("degenerate inputs")
    I1, Z1, I2, Z2 = synth_pair_float(3, 120, 160)
    g = Solver(120, 4)
    zero = np.zeros_like(Z1)
    dx, dy, dz = g.solve_pair(I1, zero, I2, zero)
    assert np.all(dx == 0) and np.all(dy == 0) and np.all(dz == 0)
    In = I2.copy()
    In[::97] = np.nan
    g.solve_pair(I1, Z1, In, Z2)
    # and the context is still usable afterwards
    a = g.solve_pair(I1, Z1, I2, Z2)
    b = Oracle(120, 4, threads=8).solve_pair(I1, Z1, I2, Z2)
    for x, y, n in zip(a, b, ("dx", "dy", "dz")):
        assert_same(x, y, n)


@pytest.mark.parametrize("k,shape", [(2, 0), (3, 0), (4, 0), (6, 0), (8, 0), (2, 1), (4, 1), (8, 1)])
def test_temporal_blocking_matches_oracle(gpu_available, k, shape):
    """k iterations per launch on every level (min_pixels=0), including remainders, in both
    region shapes, against the oracle's one-stage-at-a-time iterations: bit-identical expected."""
    say(f"temporal blocking k={k} shape={shape}")
    g, o, _ = both(240, 2, seed=31 + k)
    g.set_iteration_kernel(k, 0)
    g.set_tb_shape(shape)
    for (i, rows_i, cols_i, limg, iters) in level_table(240, 2):
        for s in (g, o):
            s.allocate_level(rows_i, cols_i, i, limg)
            for st in ["assign_zeros"] + (["upsample"] if i > 0 else []) + SETUP_STAGES:
                s.stage(st)
        for n in (k + 1, iters - k - 1):
            g.pd_iterations(n)
            o.pd_iterations(n)
            for f in STAGE_FIELDS["iterate"]:
                assert_same(g.get_field(f), o.get_field(f), f"k={k} level {i} {f}")
        for s in (g, o):
            s.stage("filter")
            s.stage("motion_field")
    for a, b, n in zip(g.motion_field(), o.motion_field(), ("dx", "dy", "dz")):
        assert_same(a, b, n)


@pytest.mark.parametrize("fused_setup,level_kernel", [(True, True), (False, True), (True, False), (False, False)])
def test_solve_options_agree_bit_for_bit(gpu_available, fused_setup, level_kernel):
    """Fused set-up kernels and the whole-level kernel change launch structure only."""
    say(f"solve options fused_setup={fused_setup} level_kernel={level_kernel}")
    I1, Z1, I2, Z2 = synth_pair_float(55, 240, 320)
    ref = Solver(240, 2)
    ref.set_solve_options(False, False)
    ref.set_iteration_kernel(1, 0)
    want = ref.solve_pair(I1, Z1, I2, Z2)
    g = Solver(240, 2)
    g.set_solve_options(fused_setup, level_kernel)
    got = g.solve_pair(I1, Z1, I2, Z2)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
    # carry planes of the finest level agree too
    for f in ("du_l", "dv_l", "dw_l", "pd_l", "puu_l", "pwv_l"):
        np.testing.assert_array_equal(g.get_field(f), ref.get_field(f))


def test_device_pair_path_matches_host_path(gpu_available):
    """pdf_solve_pair_device (two-frame pyramid from device planes) equals the reference
    ingest sequence from host planes, with and without the graph."""
    say("device pair path")
    from pdflow.lib import DeviceBuffer
    I1, Z1, I2, Z2 = synth_pair_float(66, 480, 640)
    g = Solver(480, 1)
    want = g.solve_pair(I1, Z1, I2, Z2)
    bufs = []
    for arr in (I1, Z1, I2, Z2):
        b = DeviceBuffer(arr.nbytes)
        b.upload(arr)
        bufs.append(b)
    for graph in (False, True, True):
        g.solve_pair_device(bufs, graph=graph)
        for a, b in zip(g.motion_field(), want):
            np.testing.assert_array_equal(a, b)


def test_streaming_matches_oracle_pairs(gpu_available):
    """Streaming (one pyramid per new frame, old/new swap) gives, for every new frame, the
    flow the oracle computes for (previous frame, this frame)."""
    say("streaming 240x320, 4 frames")
    from pdflow import synth_sequence_float
    from pdflow.lib import DeviceBuffer
    g = Solver(240, 2)
    o = Oracle(240, 2, threads=8)
    prev = None
    for t in range(4):
        I, Z = synth_sequence_float(1234, 240, 320, t)
        bI, bZ = DeviceBuffer(I.nbytes), DeviceBuffer(Z.nbytes)
        bI.upload(I)
        bZ.upload(Z)
        solved = g.stream_frame_device(bI, bZ, graph=(t % 2 == 1))
        o.copy_new_frames(I, Z)
        o.gaussian_pyramid()
        assert solved == (t > 0)
        if t > 0:
            o.solve()
            for a, b, n in zip(g.motion_field(), o.motion_field(), ("dx", "dy", "dz")):
                assert_same(a, b, f"frame {t} {n}")
        prev = (bI, bZ)


def test_copy_all_solutions_per_level(gpu_available):
    """copyAllSolutions after each level (the live viewer's per-level export,
    scene_flow_visualization.cpp:197-199), including the coarse levels' dx, which use the
    finest cols in inv_f (.cu:1152)."""
    say("copyAllSolutions per level")
    g, o, _ = both(120, 4, seed=77)
    for (i, r, c, limg, iters) in level_table(120, 4):
        for s in (g, o):
            s.allocate_level(r, c, i, limg)
            for st in ["assign_zeros"] + (["upsample"] if i > 0 else []) + SETUP_STAGES:
                s.stage(st)
            s.pd_iterations(iters)
            s.stage("filter")
            s.stage("motion_field")
        got = g.copy_all_solutions()
        for n in ("dx", "dy", "dz"):
            assert_same(got[n], o.get_field(n), f"level {i} {n}")
        for n, kind in (("depth", "depth"), ("depth_old", "depth_old"), ("colour", "colour"),
                        ("xx_old", "xx_old"), ("yy", "yy")):
            np.testing.assert_array_equal(got[n], o.get_pyramid(kind, limg))


def test_non_default_parameters_match_the_oracle(gpu_available):
    """Every parameter changed from the pair evaluator's (lambda_i, lambda_d, mu, a lopsided 5x5
    mask whose rows and columns differ, fields of view): the single-pair solve equals the
    oracle's bit for bit, and so does a batch of two pairs against the single-pair solves."""
    from pdflow import Batch
    mask = np.outer([1.0, 2.0, 6.0, 3.0, 1.0], [2.0, 3.0, 5.0, 4.0, 1.0]).astype(np.float32)
    mask /= mask.sum()
    over = dict(lambda_i=0.07, lambda_d=0.2, mu=40.0, g_mask=mask, fovh=1.1, fovv=0.85)
    rows, cam = 240, 2
    say("non-default parameters: 240x320 pair against the oracle, then a batch of two")
    pairs = [synth_pair_float(s, 240, 320) for s in (77, 78)]
    o = Oracle(rows, cam, threads=8, **over)
    want = o.solve_pair(*pairs[0])
    g = Solver(rows, cam, **over)
    single = [g.solve_pair(*p) for p in pairs]
    for a, b, n in zip(single[0], want, ("dx", "dy", "dz")):
        np.testing.assert_array_equal(a, b, err_msg=n)
    b = Batch([rows, rows], **over)
    for i, p in enumerate(pairs):
        b.set_inputs(i, *p)
    b.solve(graph=True)
    b.synchronize()
    for i in range(2):
        for a, w in zip(b.flow(i), single[i]):
            np.testing.assert_array_equal(a, w, err_msg=f"batch pair {i}")


def _hard_pairs():
    """(name, I1, Z1, I2, Z2) at 120x160: a large depth hole plus striped holes, constant and
    equal intensities (zero image gradients everywhere), sharp depth steps from 0.3 to 10 m"""
    I1, Z1, I2, Z2 = synth_pair_float(91, 120, 160)
    out = []
    h1, h2 = Z1.copy().reshape(160, 120), Z2.copy().reshape(160, 120)   # column-major planes
    h1[40:90, 20:70] = 0.0
    h2[50:100, 30:80] = 0.0
    h1[:, ::7] = 0.0
    h2[::11, :] = 0.0
    out.append(("holes", I1, h1.ravel().copy(), I2, h2.ravel().copy()))
    flat = np.full_like(I1, 128.0)
    out.append(("flat intensity", flat, Z1, flat.copy(), Z2))
    s1, s2 = Z1.copy().reshape(160, 120), Z2.copy().reshape(160, 120)
    s1[::2, :] = np.where(s1[::2, :] > 0, 0.3, 0.0)
    s1[1::2, :] = np.where(s1[1::2, :] > 0, 10.0, 0.0)
    s2[:, ::3] = np.where(s2[:, ::3] > 0, 9.5, 0.0)
    out.append(("depth steps", I1, s1.ravel().copy(), I2, s2.ravel().copy()))
    return out


def test_hard_inputs_match_the_oracle(gpu_available):
    """Holes, textureless images and depth steps take the border and rejection branches of the
    pyramid, gradients, warp, prox and median filter: the flow equals the oracle's bit for bit,
    single and batched."""
    from pdflow import Batch
    pairs = _hard_pairs()
    g, o = Solver(120, 4), Oracle(120, 4, threads=8)
    b = Batch([120] * len(pairs))
    for i, (name, *p) in enumerate(pairs):
        say(f"hard input: {name}")
        got = g.solve_pair(*p)
        want = o.solve_pair(*p)
        for a, w, n in zip(got, want, ("dx", "dy", "dz")):
            np.testing.assert_array_equal(a, w, err_msg=f"{name} {n}")
        b.set_inputs(i, *p)
    b.solve(graph=True)
    b.synchronize()
    for i, (name, *p) in enumerate(pairs):
        for a, w, n in zip(b.flow(i), o.solve_pair(*p), ("dx", "dy", "dz")):
            np.testing.assert_array_equal(a, w, err_msg=f"batched {name} {n}")
