"""GPU tests of the learned warm start inside the batched solver and of the per-pair quality
reductions and fallback rule (include/pdflow_hip.h, pdf_batch_set_learned ... learn_samples)."""
import numpy as np
import pytest

import metrics as mt

pytestmark = pytest.mark.gpu

ROWS = [240, 120, 240]          # QVGA (cam_mode 2), 120x160 (cam_mode 4), QVGA
SEED = 9100


@pytest.fixture(scope="module")
def pdflow(gpu_available):
    import pdflow
    return pdflow


def _fill(pdflow, b, rows, seed=SEED):
    for i, r in enumerate(rows):
        b.set_inputs(i, *pdflow.synth_pair_float(seed + i, r, r * 4 // 3))


def _plain(pdflow, rows=ROWS, seed=SEED):
    b = pdflow.Batch(rows)
    _fill(pdflow, b, rows, seed)
    b.solve(graph=True)
    b.synchronize()
    return b


def _cut(pdflow, rows, factor=4):
    """the reference schedule of each pair with every level's count divided by factor"""
    return [[max(1, n // factor) for n in pdflow.num_max_iter(pdflow.ctf_levels_for_rows(r))] for r in rows]


def _flows(b, n):
    return [b.flow(i) for i in range(n)]


def test_metrics_match_numpy(pdflow):
    b = _plain(pdflow)
    e, r = b.metrics()
    p = pdflow.default_params(480, 1)
    for i, rows in enumerate(ROWS):
        cols = rows * 4 // 3
        names = ["dct", "dcu", "dcv", "ddt", "ddu", "ddv", "mu_uv", "ri", "rj", "du_prev", "dv_prev",
                 "du_new", "dv_new", "dw_new"]
        f = {n: b.field(i, n) for n in names}
        want_e = mt.energy(f, rows, cols, p["lambda_i"], p["lambda_d"])
        I1, _, _ = b.pyramid(i, "colour_old", 0)
        I2, _, _ = b.pyramid(i, "colour", 0)
        Z1, _, _ = b.pyramid(i, "depth_old", 0)
        want_r = mt.residual(I1, I2, Z1, b.field(i, "du_l"), b.field(i, "dv_l"), rows, cols)
        assert abs(e[i] - want_e) <= 1e-4 * want_e, (i, e[i], want_e)
        assert abs(r[i] - want_r) <= 1e-4 * want_r, (i, r[i], want_r)
        assert want_r < 20.0          # a converged solve explains most of the intensity change


def test_quality_and_bounds_rule(pdflow):
    """pdf_batch_quality: the energy and residual of pdf_batch_metrics plus the zero-motion
    residual; pdf_batch_fallback_bounds: the calibrated rule re-solves exactly the pairs outside
    its bounds, and with the committed calibration it rejects a corrupted network's pairs"""
    import json
    import os
    b = _plain(pdflow)
    e, r = b.metrics()
    qe, qr, q0 = b.quality()
    assert np.array_equal(e, qe) and np.array_equal(r, qr)
    for i, rows in enumerate(ROWS):
        I1, _, _ = b.pyramid(i, "colour_old", 0)
        I2, _, _ = b.pyramid(i, "colour", 0)
        Z1, _, _ = b.pyramid(i, "depth_old", 0)
        zero = np.zeros(rows * rows * 4 // 3)
        want = mt.residual(I1, I2, Z1, zero, zero, rows, rows * 4 // 3)
        assert abs(q0[i] - want) <= 1e-4 * want, (i, q0[i], want)
        assert qr[i] < q0[i]                 # the flow explains part of the intensity change
    sub = b.quality([2, 0])
    assert sub[0][0] == qe[2] and sub[2][1] == q0[0]
    want = _flows(b, len(ROWS))
    n = len(ROWS)
    lb = pdflow.Batch(ROWS, iters=_cut(pdflow, ROWS))
    _fill(pdflow, lb, ROWS)
    lb.set_learned(_net(pdflow, n))
    lb.solve(graph=True)
    lb.synchronize()
    learned = _flows(lb, n)
    assert not lb.fallback_bounds(np.inf, np.inf).any()
    for i in range(n):
        assert np.array_equal(lb.flow(i)[0], learned[i][0])
    assert lb.fallback_bounds(1e-6, 1e-6).all()
    for i in range(n):
        for c in range(3):
            assert np.array_equal(lb.flow(i)[c], want[i][c]), (i, c)
    rule = json.load(open(os.path.join(os.path.dirname(pdflow.__file__), "weights", "warmstart_qvga_rule.json")))
    bad = pdflow.Batch(ROWS, iters=_cut(pdflow, ROWS))
    _fill(pdflow, bad, ROWS)
    bad.set_learned(_net(pdflow, n, seed=3, scale=3.0))
    bad.solve(graph=True)
    bad.synchronize()
    flags = bad.fallback_bounds(rule["energy_per_pixel_max"], rule["residual_ratio_max"])
    assert flags.all()
    for i in range(n):
        assert np.array_equal(bad.flow(i)[2], want[i][2])


def test_epe_matches_numpy(pdflow):
    a = _plain(pdflow)
    assert np.all(a.epe(a) == 0)
    c = pdflow.Batch(ROWS, iters=_cut(pdflow, ROWS))
    _fill(pdflow, c, ROWS)
    c.solve(graph=False)
    c.synchronize()
    got = c.epe(a)
    for i, rows in enumerate(ROWS):
        Z1, _, _ = a.pyramid(i, "depth_old", 0)
        want = mt.epe(c.flow(i), a.flow(i), Z1)
        assert abs(got[i] - want) <= 1e-4 * want + 1e-9, (i, got[i], want)
        assert want > 0


def test_learn_samples(pdflow):
    from pdflow.lib import DeviceBuffer
    b = _plain(pdflow)
    n, h, w = len(ROWS), 60, 80
    bx, bt, bm = DeviceBuffer(n * h * w * 16), DeviceBuffer(n * h * w * 12), DeviceBuffer(n * h * w * 4)
    b.learn_samples(h, w, bx.ptr.value, bt.ptr.value, bm.ptr.value)
    x = bx.download(np.empty((n, h, w, 4), np.float32))
    t = bt.download(np.empty((n, h, w, 3), np.float32))
    m = bm.download(np.empty((n, h, w), np.float32))
    assert np.all(m == 1)
    for i, rows in enumerate(ROWS):
        level = {240: 2, 120: 1}[rows]         # pyramid image of the 60x80 level
        planes = [b.pyramid(i, k, level)[0].reshape(w, h).T for k in ("colour_old", "depth_old", "colour", "depth")]
        want = np.stack([planes[0] / np.float32(255) - np.float32(0.5), planes[1] * np.float32(0.25),
                         planes[2] / np.float32(255) - np.float32(0.5), planes[3] * np.float32(0.25)], axis=-1)
        np.testing.assert_allclose(x[i], want, rtol=0, atol=1e-6)
        s = rows // h
        for c, name in enumerate(("du_l", "dv_l", "dw_l")):
            fine = b.field(i, name).reshape(rows * 4 // 3, rows).T.astype(np.float64)
            blk = fine.reshape(h, s, w, s).mean(axis=(1, 3))
            want_t = blk / s if c < 2 else blk * 10
            np.testing.assert_allclose(t[i, ..., c], want_t, rtol=1e-5, atol=1e-6)


def _net(pdflow, n, seed=1, scale=None):
    net = pdflow.Net(max_batch=n, h=60, w=80, seed=seed)
    if scale is not None:
        net.set_params(net.params() * np.float32(scale))
    return net


def test_learned_equals_explicit_warm_start(pdflow):
    """network inside the solve == gather -> forward -> inject -> warm start done by hand"""
    from pdflow.lib import DeviceBuffer
    n = len(ROWS)
    net = _net(pdflow, n)
    sched = _cut(pdflow, ROWS)
    lb = pdflow.Batch(ROWS, iters=sched)
    _fill(pdflow, lb, ROWS)
    lb.set_learned(net)
    lb.solve(graph=False)
    lb.synchronize()
    got = _flows(lb, n)
    # by hand: network inputs of a plain batch's pyramids, forward, then warm planes
    pb = _plain(pdflow)
    h, w = 60, 80
    bx, bt, bm = DeviceBuffer(n * h * w * 16), DeviceBuffer(n * h * w * 12), DeviceBuffer(n * h * w * 4)
    pb.learn_samples(h, w, bx.ptr.value, bt.ptr.value, bm.ptr.value)
    x = bx.download(np.empty((n, h, w, 4), np.float32))
    y = net.forward(x)
    assert np.isfinite(y).all() and np.abs(y).max() > 0.1
    wb = pdflow.Batch(ROWS, iters=sched)
    _fill(pdflow, wb, ROWS)
    for i in range(n):
        ptr = wb.set_start_level(i, 2)           # the 60x80 level of 120- and 240-row pairs
        planes = np.concatenate([y[i, ..., 0].T.ravel(), y[i, ..., 1].T.ravel(), (y[i, ..., 2] / np.float32(10)).T.ravel()])
        rc = wb.lib.pdf_memcpy_h2d(ptr, np.ascontiguousarray(planes, np.float32).ctypes.data, planes.nbytes)
        assert rc == 0
    wb.solve(graph=False)
    wb.synchronize()
    want = _flows(wb, n)
    for i in range(n):
        for c in range(3):
            assert np.array_equal(got[i][c], want[i][c]), (i, c)
    # the captured graph replays the same solve
    lb.solve(graph=True)
    lb.synchronize()
    lb.solve(graph=True)
    lb.synchronize()
    for i in range(n):
        for c in range(3):
            assert np.array_equal(lb.flow(i)[c], want[i][c]), (i, c)


def test_resolve_plain_and_fallback(pdflow):
    n = len(ROWS)
    plain = _plain(pdflow)
    pe, pr = plain.metrics()
    want = _flows(plain, n)
    lb = pdflow.Batch(ROWS, iters=_cut(pdflow, ROWS))
    _fill(pdflow, lb, ROWS)
    lb.set_learned(_net(pdflow, n))
    lb.solve(graph=True)
    lb.synchronize()
    learned = _flows(lb, n)
    # nothing exceeds an infinite reference: no fallback, flow unchanged
    flags = lb.fallback(np.full(n, np.inf), np.full(n, np.inf))
    assert not flags.any()
    for i in range(n):
        assert np.array_equal(lb.flow(i)[0], learned[i][0])
    # references far below: every pair falls back and then equals its plain solve bit for bit
    flags = lb.fallback(pe * 0.01, pr * 0.01)
    assert flags.all()
    for i in range(n):
        for c in range(3):
            assert np.array_equal(lb.flow(i)[c], want[i][c]), (i, c)
    # a corrupted network output (weights scaled up 3x per layer) starts every pair far off:
    # energy or residual end above the plain solver's and every pair falls back
    bad = pdflow.Batch(ROWS, iters=_cut(pdflow, ROWS))
    _fill(pdflow, bad, ROWS)
    bad.set_learned(_net(pdflow, n, seed=3, scale=3.0))
    bad.solve(graph=True)
    bad.synchronize()
    be, br = bad.metrics()
    assert np.all((be > 1.05 * pe) | (br > 1.05 * pr))
    flags = bad.fallback(pe, pr)
    assert flags.all()
    for i in range(n):
        assert np.array_equal(bad.flow(i)[2], want[i][2])
    # the next solve of the batch is the learned one again (plan state restored)
    lb.solve(graph=True)
    lb.synchronize()
    for i in range(n):
        assert np.array_equal(lb.flow(i)[1], learned[i][1])


def test_set_learned_rejects_bad_pairs(pdflow):
    net = _net(pdflow, 1)
    b = pdflow.Batch([240, 240])
    with pytest.raises(pdflow.PDFlowError):
        b.set_learned(net)                      # two pairs, max_batch 1
    with pytest.raises(pdflow.PDFlowError):
        b.set_learned(net, [5])
    small = pdflow.Batch([30])                  # 30x40: no 60x80 level
    with pytest.raises(pdflow.PDFlowError):
        small.set_learned(net)
    b.set_learned(net, [1])
    b.set_learned(None)


def test_set_learned_error_leaves_every_pair_cold(pdflow):
    """A rejected pdf_batch_set_learned (a pair listed twice, an index out of range, too many
    pairs) changes no pair: the next solve is the plain solve, bit for bit."""
    rows = [240, 240]
    want = _flows(_plain(pdflow, rows), 2)
    net = _net(pdflow, 2)
    b = pdflow.Batch(rows)
    _fill(pdflow, b, rows)
    for bad in ([0, 0], [1, 0, 1], [0, 7]):
        with pytest.raises(pdflow.PDFlowError):
            b.set_learned(net, bad)
        b.solve(graph=False)
        b.synchronize()
        for i in range(2):
            for c in range(3):
                assert np.array_equal(b.flow(i)[c], want[i][c]), (bad, i, c)
    # a learned set-up followed by a rejected one also leaves the batch plain
    b.set_learned(net, [1])
    with pytest.raises(pdflow.PDFlowError):
        b.set_learned(net, [1, 1])
    b.solve(graph=True)
    b.synchronize()
    for i in range(2):
        for c in range(3):
            assert np.array_equal(b.flow(i)[c], want[i][c]), ("after learned", i, c)


def test_trained_network_beats_the_ablation(pdflow):
    """The committed network (trained on other seeds) against the same 4x iteration cut without
    it, on held-out QVGA pairs: end-point error against the rendered ground truth."""
    import os
    from pdflow.train import EVAL_SEED0, cut_schedule
    weights = os.path.join(os.path.dirname(pdflow.__file__), "weights", "warmstart_qvga.bin")
    n, rows = 32, 240
    *planes, gt = pdflow.synth_pairs_float(EVAL_SEED0 + 5000, n, rows, 320, with_gt=True)
    net = pdflow.Net(max_batch=n, h=60, w=80)
    net.load(weights)
    cut = [cut_schedule(rows, 4)] * n

    def solve(iters, learned):
        b = pdflow.Batch([rows] * n, iters=iters)
        for i in range(n):
            b.set_inputs(i, *(p[i] for p in planes))
        if learned:
            b.set_learned(net)
        b.solve(graph=True)
        b.synchronize()
        return b

    def gt_epe(b):
        out = []
        for i in range(n):
            ok = planes[1][i] > 0
            out.append(np.sqrt(sum((f.astype(np.float64) - g) ** 2 for f, g in zip(b.flow(i), gt[i])))[ok].mean())
        return np.asarray(out)

    plain, learned, ablation = solve(None, False), solve(cut, True), solve(cut, False)
    e_plain, e_learned, e_abl = gt_epe(plain).mean(), gt_epe(learned).mean(), gt_epe(ablation).mean()
    # the network's gain over the cut schedule alone, and accuracy within 2 mm of the plain solve's
    # (against the rendered ground truth).  Against the plain solution itself the learned start is
    # not closer than the ablation (about 12.6 against 11.5 mm on held-out QVGA pairs), so that
    # is not asserted.
    assert e_learned < e_abl - 0.003, (e_learned, e_abl)
    assert e_learned <= e_plain + 0.002, (e_learned, e_plain)


def test_two_learned_batches_in_flight_match_one_at_a_time(pdflow):
    """The benchmark's pipeline: two learned batches, each with its own network instance and
    stream, one solving while the other continues some pairs and re-solves another. Every flow
    equals that of the same batch run alone."""
    import os
    from pdflow.train import cut_schedule
    weights = os.path.join(os.path.dirname(pdflow.__file__), "weights", "warmstart_qvga.bin")
    n, rows = 8, 240
    cut = [cut_schedule(rows, 2.5)] * n
    inputs = {s: pdflow.synth_pairs_float(s, n, rows, 320) for s in (9400, 9500)}

    def make(seed):
        net = pdflow.Net(max_batch=n, h=60, w=80)
        net.load(weights)
        b = pdflow.Batch([rows] * n, iters=cut)
        for i in range(n):
            b.set_inputs(i, *(p[i] for p in inputs[seed][:4]))
        b.set_learned(net)
        return b, net

    def finish(b):
        b.continue_finest([0, 3, 5], 16)
        b.resolve_plain([1])

    want = {}
    for seed in (9400, 9500):
        b, _ = make(seed)
        b.solve(graph=True)
        finish(b)
        b.synchronize()
        want[seed] = _flows(b, n)
    (a, net_a), (c, net_c) = make(9400), make(9500)
    for rep in range(2):
        a.solve(graph=True)
        c.solve(graph=True)          # in flight while a finishes
        finish(a)
        finish(c)
        a.synchronize()
        c.synchronize()
        for seed, b in ((9400, a), (9500, c)):
            for i, (got, exp) in enumerate(zip(_flows(b, n), want[seed])):
                for k in range(3):
                    assert np.array_equal(got[k], exp[k]), (rep, seed, i, k)
