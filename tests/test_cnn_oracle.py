"""Pins the NumPy reference of the warm-start network (oracle/cnn_oracle.py) on the CPU:
convolution against a brute-force loop, backward against central finite differences."""
import numpy as np
import pytest

import cnn_oracle as co


def test_param_count_and_flat_round_trip():
    n = co.num_params()
    assert n == sum(c * 9 * i + c for _, i, c, _ in co.ARCH)
    assert 400_000 < n < 600_000
    flat = np.random.default_rng(0).standard_normal(n).astype(np.float32)
    assert np.array_equal(co.flatten(co.unflatten(flat)), flat)


def test_conv3x3_matches_brute_force():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 5, 6, 3))
    w = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    got = co.conv3x3(x, w, b)
    want = np.zeros((2, 5, 6, 4))
    for n in range(2):
        for y in range(5):
            for xx in range(6):
                for o in range(4):
                    s = b[o]
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            if 0 <= y + dy < 5 and 0 <= xx + dx < 6:
                                s += np.dot(x[n, y + dy, xx + dx], w[o, dy + 1, dx + 1])
                    want[n, y, xx, o] = s
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_conv_grads_adjoint():
    # <conv(x), dz> = <x, dx(dz)> for the bias-free convolution (dx is the adjoint)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, 4, 5, 3))
    w = rng.standard_normal((2, 3, 3, 3))
    dz = rng.standard_normal((1, 4, 5, 2))
    dx, dw, db = co.conv3x3_grads(x, w, dz)
    lhs = np.sum(co.conv3x3(x, w, np.zeros(2)) * dz)
    assert abs(lhs - np.sum(x * dx)) < 1e-10 * max(1, abs(lhs))
    assert abs(lhs - np.sum(w * dw)) < 1e-10 * max(1, abs(lhs))
    np.testing.assert_allclose(db, dz.sum(axis=(0, 1, 2)))


def test_pool_up_adjoints():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 4, 6, 2))
    y = rng.standard_normal((1, 2, 3, 2))
    assert abs(np.sum(co.pool(x) * y) - np.sum(x * co.pool_grad(y))) < 1e-12
    assert abs(np.sum(co.up(y) * x) - np.sum(y * co.up_grad(x))) < 1e-12


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0], np.float32)
    r = co.bf16(v)
    # ties to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6 (even), 1 + 2^-9 -> 1
    assert list(r) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 0.0]


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    flat = (rng.standard_normal(co.num_params()) * 0.2).astype(np.float64)
    x = rng.standard_normal((1, 8, 8, 4))
    dout = rng.standard_normal((1, 8, 8, 3))

    def loss(p):
        out, _ = co.forward(p, x, emulate_bf16=False, dtype=np.float64)
        return np.sum(out * dout)

    out, cache = co.forward(flat, x, emulate_bf16=False, dtype=np.float64)
    g = co.backward(cache, dout)
    assert g.shape == flat.shape
    # a few parameters of every layer (weights and bias)
    idx, off = [], 0
    for _, ci, c, _ in co.ARCH:
        nw = c * 9 * ci
        idx += list(off + rng.choice(nw, 3, replace=False)) + [off + nw + int(rng.integers(c))]
        off += nw + c
    # forward casts parameters through float32, so step on a float32-representable grid
    h = 2.0 ** -10
    for i in idx:
        p1, p2 = flat.copy(), flat.copy()
        p1[i] += h
        p2[i] -= h
        fd = (loss(p1) - loss(p2)) / (2 * h)
        assert abs(fd - g[i]) <= 2e-3 * max(1.0, abs(g[i])), (i, fd, g[i])
