"""Pins the NumPy restatement of the quality figures (oracle/metrics.py) on the CPU: against the
C oracle's bilinear warp, brute-force loops and closed forms."""
import numpy as np

import metrics as mt
import oracle as orc


def test_bilinear_matches_c_oracle():
    rng = np.random.default_rng(0)
    rows, cols = 7, 9
    mat = rng.uniform(0, 255, rows * cols).astype(np.float32)
    iu = np.concatenate([rng.uniform(-2, cols + 1, 200), np.arange(cols, dtype=float), [np.nan, 3.0, 2.5]])
    iv = np.concatenate([rng.uniform(-2, rows + 1, 200), np.full(cols, 2.0), [1.5, np.nan, 4.0]])
    iu, iv = iu.astype(np.float32), iv.astype(np.float32)
    got = mt.bilinear(mat, rows, cols, iu, iv)
    want = np.array([orc.interpolate(mat, rows, cols, float(a), float(b)) for a, b in zip(iu, iv)])
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-3)


def test_grad_r_brute_force():
    rng = np.random.default_rng(1)
    rows, cols = 4, 5
    f = rng.standard_normal(rows * cols)
    ri = rng.uniform(0.5, 2, rows * cols)
    rj = rng.uniform(0.5, 2, rows * cols)
    g1, g2 = mt.grad_r(f, ri, rj, rows, cols)
    for u in range(cols):
        for v in range(rows):
            i = v + u * rows
            assert g1[i] == (ri[i] * (f[i + rows] - f[i]) if u < cols - 1 else 0.0)
            assert g2[i] == (rj[i] * (f[i + 1] - f[i]) if v < rows - 1 else 0.0)


def _fields(rng, P):
    names = ["dct", "dcu", "dcv", "ddt", "ddu", "ddv", "mu_uv", "ri", "rj", "du_prev", "dv_prev",
             "du_new", "dv_new", "dw_new"]
    return {n: rng.standard_normal(P) for n in names}


def test_energy_closed_forms():
    rng = np.random.default_rng(2)
    rows, cols = 6, 8
    P = rows * cols
    f = _fields(rng, P)
    # zero increment on a constant total flow: only the data terms at du = dv = dw = 0
    f["du_new"][:] = 0
    f["dv_new"][:] = 0
    f["dw_new"][:] = 0
    f["du_prev"][:] = 1.5
    f["dv_prev"][:] = -0.5
    want = np.sum(np.abs(f["dct"]) + f["mu_uv"] * np.abs(f["ddt"]))
    assert abs(mt.energy(f, rows, cols, 0.04, 0.35) - want) < 1e-9 * max(1, abs(want))
    # a ramp du = a*u with unit weights: TV of u is lambda_i * |a| on all but the last column
    f2 = dict(f)
    for k in ("dct", "ddt", "dcu", "dcv", "ddu", "ddv", "du_prev", "dv_prev"):
        f2[k] = np.zeros(P)
    f2["ri"] = np.ones(P)
    f2["rj"] = np.ones(P)
    a = 0.25
    f2["du_new"] = a * (np.arange(P) // rows).astype(float)
    want = 0.04 * abs(a) * rows * (cols - 1)
    assert abs(mt.energy(f2, rows, cols, 0.04, 0.35) - want) < 1e-12


def test_residual_and_epe_closed_forms():
    rows, cols = 5, 7
    idx = np.arange(rows * cols)
    v, u = idx % rows, idx // rows
    I1 = (3.0 * u + 2.0 * v).astype(np.float32)
    Z = np.ones(rows * cols, np.float32)
    Z[3] = 0
    # zero flow on identical frames
    assert mt.residual(I1, I1, Z, np.zeros_like(I1), np.zeros_like(I1), rows, cols) == 0.0
    # frame 2 shifted one column right: flow +1 in u is exact except where clamped (last col)
    I2 = (3.0 * (u - 1) + 2.0 * v).astype(np.float32)
    r = mt.residual(I1, I2, Z, np.ones_like(I1), np.zeros_like(I1), rows, cols)
    clamped = (u == cols - 1) & (Z > 0)
    assert abs(r - 3.0 * clamped.sum() / (Z > 0).sum()) < 1e-12
    a = [np.zeros(rows * cols)] * 3
    b = [np.full(rows * cols, 0.3), np.full(rows * cols, 0.4), np.zeros(rows * cols)]
    assert abs(mt.epe(a, b, Z) - 0.5) < 1e-12


def test_zero_motion_residual_is_the_mean_intensity_change():
    """with zero flow the bilinear warp reads I2 at integer pixels: the residual is the mean of
    |I2 - I1| over the pixels with frame-1 depth"""
    rng = np.random.default_rng(5)
    rows, cols = 9, 12
    I1 = rng.uniform(0, 255, rows * cols).astype(np.float32)
    I2 = rng.uniform(0, 255, rows * cols).astype(np.float32)
    Z1 = np.where(rng.uniform(size=rows * cols) < 0.2, 0.0, 2.0).astype(np.float32)
    zero = np.zeros(rows * cols)
    want = np.abs(I2.astype(np.float64) - I1)[Z1 > 0].mean()
    assert abs(mt.residual(I1, I2, Z1, zero, zero, rows, cols) - want) < 1e-9 * want
