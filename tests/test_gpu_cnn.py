"""GPU checks of the warm-start network's HIP/MFMA kernels (include/pdflow_cnn.h) against the
NumPy reference in oracle/cnn_oracle.py."""

import numpy as np
import pytest

import cnn_oracle as co

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pdflow(gpu_available):
    import pdflow
    return pdflow


def _bf16_bits(x):
    return (co.bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def _inputs(rng, n, h, w):
    x = np.empty((n, h, w, 4), np.float32)
    x[..., 0] = rng.uniform(-0.5, 0.5, (n, h, w))
    x[..., 1] = rng.uniform(0.0, 1.0, (n, h, w))
    x[..., 2] = rng.uniform(-0.5, 0.5, (n, h, w))
    x[..., 3] = rng.uniform(0.0, 1.0, (n, h, w))
    return x


def test_mfma_operand_layout(pdflow):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((16, 16)).astype(np.float32)
    b = rng.standard_normal((16, 16)).astype(np.float32)
    c = pdflow.mfma_probe(_bf16_bits(a), _bf16_bits(b))
    want = co.bf16(a).astype(np.float64) @ co.bf16(b).astype(np.float64)
    np.testing.assert_allclose(c, want, rtol=1e-5, atol=1e-5)
    # a permutation matrix picks rows/columns: catches a transposed operand exactly
    p = np.eye(16, dtype=np.float32)[rng.permutation(16)]
    c = pdflow.mfma_probe(_bf16_bits(p), _bf16_bits(b))
    assert np.array_equal(c, p @ co.bf16(b))


def test_params_deterministic_and_round_trip(pdflow, tmp_path):
    n1 = pdflow.Net(max_batch=2, h=16, w=24, seed=7)
    n2 = pdflow.Net(max_batch=2, h=16, w=24, seed=7)
    p = n1.params()
    assert p.size == co.num_params()
    assert np.array_equal(p, n2.params())
    # biases start at zero, weights He-normal
    layers = co.unflatten(p)
    assert all(np.all(b == 0) for _, b in layers)
    w0 = layers[1][0]
    assert abs(w0.std() - np.sqrt(2.0 / (9 * 32))) < 0.1 * np.sqrt(2.0 / (9 * 32))
    q = np.random.default_rng(1).standard_normal(p.size).astype(np.float32)
    n1.set_params(q)
    assert np.array_equal(n1.params(), q)
    path = tmp_path / "w.bin"
    n1.save(path)
    n2.load(path)
    assert np.array_equal(n2.params(), q)
    with open(path, "r+b") as f:
        f.write(b"XXXXXXXX")
    with pytest.raises(pdflow.PDFlowError):
        n2.load(path)


def test_bad_arguments(pdflow):
    with pytest.raises(pdflow.PDFlowError):
        pdflow.Net(max_batch=1, h=30, w=40)          # h not a multiple of 4
    net = pdflow.Net(max_batch=2, h=16, w=24)
    with pytest.raises(pdflow.PDFlowError):
        net.forward(np.zeros((3, 16, 24, 4), np.float32))


@pytest.mark.parametrize("n,h,w", [(2, 60, 80), (1, 16, 24)])
def test_forward_matches_reference(pdflow, n, h, w):
    rng = np.random.default_rng(10 + h)
    net = pdflow.Net(max_batch=n, h=h, w=w, seed=3)
    flat = net.params()
    # non-zero biases so the bias path is exercised
    layers = co.unflatten(flat.copy())
    for _, b in layers:
        b[:] = rng.uniform(-0.05, 0.05, b.shape)
    flat = co.flatten(layers).astype(np.float32)
    net.set_params(flat)
    x = _inputs(rng, n, h, w)
    got = net.forward(x)
    want, _ = co.forward(flat, x, emulate_bf16=True)
    # BF16 activations: the GPU and the reference round the same FP32 values, which differ
    # only in summation order, so a few activations differ by one BF16 ulp
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < 1e-2, err
    assert np.max(np.abs(got - want)) < 5e-2 * np.max(np.abs(want))


def test_backward_matches_reference(pdflow):
    rng = np.random.default_rng(20)
    n, h, w = 2, 32, 40
    net = pdflow.Net(max_batch=n, h=h, w=w, seed=5)
    flat = net.params()
    x = _inputs(rng, n, h, w)
    dy = rng.standard_normal((n, h, w, 3)).astype(np.float32) * 1e-2
    got = net.backward_from(x, dy)
    _, cache = co.forward(flat, x, emulate_bf16=True, dtype=np.float64)
    want = co.backward(cache, co.bf16(dy).astype(np.float64))
    # per layer: BF16 activation gradients give percent-level element errors; the direction
    # of every layer's gradient must agree closely
    off = 0
    for name, ci, c, _ in co.ARCH:
        sz = c * 9 * ci + c
        g, r = got[off:off + sz], want[off:off + sz]
        cos = np.dot(g, r) / (np.linalg.norm(g) * np.linalg.norm(r))
        rel = np.linalg.norm(g - r) / np.linalg.norm(r)
        assert cos > 0.998 and rel < 0.06, (name, cos, rel)
        off += sz


def test_training_reduces_loss(pdflow):
    # a fixed batch: a few Adam steps on the L1 loss must bring it down
    from pdflow.lib import DeviceBuffer
    rng = np.random.default_rng(30)
    n, h, w = 4, 32, 40
    net = pdflow.Net(max_batch=n, h=h, w=w, seed=9)
    x = _inputs(rng, n, h, w)
    t = np.stack([x[..., 0] - x[..., 2], 0.5 * x[..., 1], x[..., 3] - x[..., 1]], axis=-1).astype(np.float32)
    m = (rng.uniform(size=(n, h, w)) > 0.1).astype(np.float32)
    bx, bt, bm = DeviceBuffer(x.nbytes), DeviceBuffer(t.nbytes), DeviceBuffer(m.nbytes)
    bx.upload(x)
    bt.upload(t)
    bm.upload(m)
    losses = [net.train_step_ptr(bx.ptr.value, bt.ptr.value, bm.ptr.value, n, 2e-3, want_loss=True)
              for _ in range(60)]
    # the first loss is the L1 loss of the initial network
    out0, _ = co.forward(pdflow.Net(max_batch=n, h=h, w=w, seed=9).params(), x)
    l0 = np.sum(m[..., None] * np.abs(out0 - t)) / (3 * n * h * w)
    assert abs(losses[0] - l0) < 2e-2 * l0
    assert losses[-1] < 0.5 * losses[0], losses[::10]
