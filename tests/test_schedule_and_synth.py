"""Schedule table, mask and synthetic generator (CPU)."""
import json
import os

import numpy as np
import pytest

from pdflow import schedule, synth_hash, synth_pair, synth_pair_float, synth_static_float
import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("rows,levels,iters,pix_iters", [
    # the per-size workload of the reference schedule (scene_flow_impair.cpp:40, :46-52)
    (120, 4, [55, 70, 85, 100], 2_428_500),
    (240, 5, [40, 55, 70, 85, 100], 9_726_000),
    (480, 6, [25, 40, 55, 70, 85, 100], 38_911_500),
    (15, 1, [100], 30_000),
    (60, 3, [70, 85, 100], 60 * 80 * 100 + 30 * 40 * 85 + 15 * 20 * 70),
])
def test_schedule_table(rows, levels, iters, pix_iters):
    assert schedule.ctf_levels_for_rows(rows) == levels
    assert schedule.num_max_iter(levels) == iters
    cam_mode = 480 // rows
    assert schedule.pixel_iterations(rows, cam_mode) == pix_iters
    # the C oracle agrees
    lib = oracle.load()
    assert lib.orc_ctf_levels(rows) == levels
    out = (oracle.ctypes.c_uint * 6)()
    lib.orc_num_max_iter(levels, out)
    assert list(out)[:levels] == iters
    # the coarsest level is always 15x20
    t = schedule.level_table(rows, cam_mode)
    assert (t[0][1], t[0][2]) == (15, 20)
    assert (t[-1][1], t[-1][2]) == (rows, rows * 4 // 3)


def test_level_image_when_input_is_larger_than_solve():
    # rows=240 solved from 640x480 input: pyramid slot 1 is the finest solve level
    t = schedule.level_table(240, 1)
    assert [x[3] for x in t] == [5, 4, 3, 2, 1]


def test_gaussian_mask():
    m = schedule.gaussian_mask()
    assert m.sum() == np.float32(1.0)
    b = np.array([1, 4, 6, 4, 1], np.float64)
    np.testing.assert_array_equal(m.reshape(5, 5).T, (np.outer(b, b) / 256).astype(np.float32))
    c = np.zeros(25, np.float32)
    oracle.load().orc_gaussian_mask(oracle._fp(c))
    np.testing.assert_array_equal(c, m)


def test_synth_deterministic_and_shaped():
    a = synth_pair(11, 120, 160)
    b = synth_pair(11, 120, 160)
    c = synth_pair(12, 120, 160)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert any(not np.array_equal(x, y) for x, y in zip(a, c))
    i1, z1, i2, z2 = a
    for z in (z1, z2):
        hole = np.mean(z == 0)
        assert 0.04 < hole < 0.075, hole
        zm = z[z > 0] / 5000.0
        assert zm.min() >= 0.5 and zm.max() <= 4.5
    assert i1.dtype == np.uint8 and z1.dtype == np.uint16
    # frame 2 differs (motion + noise), intensity textured
    assert np.mean(i1 != i2) > 0.5
    assert i1.std() > 15


def test_synth_float_conversion_and_gt():
    i1, z1, i2, z2 = synth_pair(5, 60, 80)
    I1, Z1, I2, Z2, gt = synth_pair_float(5, 60, 80, with_gt=True)
    np.testing.assert_array_equal(I1, i1.astype(np.float32))
    np.testing.assert_array_equal(Z1, (z1.astype(np.float64) / 5000.0).astype(np.float32))
    m = Z1 > 0
    mag = np.linalg.norm(gt[:, m], axis=0)
    assert np.all(gt[:, ~m] == 0)
    # |t| <= 5 cm plus a <= 2 degree rotation about the patch centre
    assert mag.max() < 0.2 and mag.mean() > 0.005


def test_synth_static_equals_frame1():
    I, Z = synth_static_float(9, 60, 80)
    I1, Z1, _, _ = synth_pair_float(9, 60, 80)
    np.testing.assert_array_equal(I, I1)
    np.testing.assert_array_equal(Z, Z1)


def test_synth_golden_hashes():
    """Pins the generator's output bytes (tests/golden/make_golden.py wrote them)."""
    with open(os.path.join(GOLDEN, "synth_hashes.json")) as f:
        golden = json.load(f)
    for key, want in golden.items():
        seed, rows, cols = (int(x) for x in key.split("_"))
        got = [synth_hash(a) for a in synth_pair(seed, rows, cols)]
        assert got == want, key


def test_oracle_regression_fixture():
    """The oracle's 60x80 end-to-end flow still matches the committed fixture bit for bit."""
    d = np.load(os.path.join(GOLDEN, "oracle_60x80.npz"))
    o = oracle.Oracle(60, 8)
    got = o.solve_pair(d["I1"], d["Z1"], d["I2"], d["Z2"])
    for name, g in zip(("dx", "dy", "dz"), got):
        np.testing.assert_array_equal(g, d[name])


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asking for more exits non-zero with a message, before any GPU
    work."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU only" in (r.stdout + r.stderr)
    assert '"metric"' not in r.stdout
