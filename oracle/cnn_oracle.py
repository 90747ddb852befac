"""NumPy reference of the warm-start network -- TEST INFRASTRUCTURE ONLY (the checker of the
HIP/MFMA kernels in pd-flow_amd/csrc/pdflow_cnn*.hip).  No counterpart in the reference repo.

Architecture and flat parameter order as in include/pdflow_cnn.h.  ``forward`` can round to
BF16 at the same points as the GPU (input, weights, every activation) so the two agree to FP32
accumulation-order differences; ``backward`` is the exact float64 gradient of the same graph
(pinned on the CPU by finite differences).
"""
import numpy as np

ARCH = [("e1a", 4, 32, True), ("e1b", 32, 32, True), ("e2a", 32, 64, True), ("e2b", 64, 64, True),
        ("b1", 64, 128, True), ("b2", 128, 128, True), ("d2a", 192, 64, True), ("d2b", 64, 64, True),
        ("d1a", 96, 32, True), ("d1b", 32, 32, True), ("head", 32, 3, False)]


def num_params():
    return sum(co * 9 * ci + co for _, ci, co, _ in ARCH)


def unflatten(flat):
    """flat vector -> list of (w [cout, 3, 3, cin], b [cout]) in layer order"""
    out, off = [], 0
    for _, ci, co, _ in ARCH:
        w = flat[off:off + co * 9 * ci].reshape(co, 3, 3, ci)
        off += co * 9 * ci
        b = flat[off:off + co]
        off += co
        out.append((w, b))
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in layers])


def bf16(x):
    """round to nearest even BF16, returned as float32"""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def conv3x3(x, w, b):
    """same-padded 3x3 convolution, x [N,H,W,Cin], w [Cout,3,3,Cin] -> [N,H,W,Cout]"""
    N, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((N, H, W, w.shape[0]), dtype=np.result_type(x, w))
    for dy in range(3):
        for dx in range(3):
            out += np.einsum("nhwc,oc->nhwo", xp[:, dy:dy + H, dx:dx + W, :], w[:, dy, dx, :])
    return out + b


def conv3x3_grads(x, w, dz):
    """gradients of conv3x3 w.r.t. x, w, b given dz = dL/d(out)"""
    N, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for dy in range(3):
        for dx in range(3):
            dw[:, dy, dx, :] = np.einsum("nhwo,nhwc->oc", dz, xp[:, dy:dy + H, dx:dx + W, :])
            dxp[:, dy:dy + H, dx:dx + W, :] += np.einsum("nhwo,oc->nhwc", dz, w[:, dy, dx, :])
    return dxp[:, 1:-1, 1:-1, :], dw, dz.sum(axis=(0, 1, 2))


def pool(x):
    N, H, W, C = x.shape
    return x.reshape(N, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_grad(dy):
    return np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) / 4.0


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_grad(dy):
    N, H, W, C = dy.shape
    return dy.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def forward(flat, x, emulate_bf16=True, dtype=np.float32):
    """returns (out [N,H,W,3], cache)"""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype))
    layers = [(r(w).astype(dtype), b.astype(dtype)) for w, b in unflatten(np.asarray(flat, np.float32))]
    c = {}

    def conv(i, inp):
        w, b = layers[i]
        z = conv3x3(inp, w, b)
        if ARCH[i][3]:
            z = np.maximum(z, 0)
        c[ARCH[i][0]] = (inp, z)
        return z if i == len(ARCH) - 1 else r(z)

    x0 = r(x).astype(dtype)
    e1a = conv(0, x0)
    e1b = conv(1, e1a)
    p1 = r(pool(e1b))
    e2a = conv(2, p1)
    e2b = conv(3, e2a)
    p2 = r(pool(e2b))
    b1 = conv(4, p2)
    b2 = conv(5, b1)
    cat2 = np.concatenate([up(b2), e2b], axis=3)
    d2a = conv(6, cat2)
    d2b = conv(7, d2a)
    cat1 = np.concatenate([up(d2b), e1b], axis=3)
    d1a = conv(8, cat1)
    d1b = conv(9, d1a)
    out = conv(10, d1b)
    c["layers"] = layers
    return out, c


def backward(cache, dout):
    """exact gradients (in the cache's dtype) of sum(out * dout) w.r.t. the flat parameters"""
    layers = cache["layers"]
    grads = [None] * len(ARCH)

    def back(i, dy):
        inp, z = cache[ARCH[i][0]]
        dz = dy * (z > 0) if ARCH[i][3] else dy
        dx, dw, db = conv3x3_grads(inp, layers[i][0], dz)
        grads[i] = (dw, db)
        return dx

    g_d1b = back(10, dout)
    g_d1a = back(9, g_d1b)
    g_cat1 = back(8, g_d1a)
    g_d2b = up_grad(g_cat1[..., :64])
    g_e1b = g_cat1[..., 64:].copy()
    g_d2a = back(7, g_d2b)
    g_cat2 = back(6, g_d2a)
    g_b2 = up_grad(g_cat2[..., :128])
    g_e2b = g_cat2[..., 128:].copy()
    g_b1 = back(5, g_b2)
    g_p2 = back(4, g_b1)
    g_e2b = g_e2b + pool_grad(g_p2)
    g_e2a = back(3, g_e2b)
    g_p1 = back(2, g_e2a)
    g_e1b = g_e1b + pool_grad(g_p1)
    g_e1a = back(1, g_e1b)
    back(0, g_e1a)
    return flatten(grads)
