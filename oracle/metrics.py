"""NumPy restatement of the per-pair quality figures the batched solver reduces on the GPU
(pdf_batch_metrics / pdf_batch_epe) -- TEST INFRASTRUCTURE ONLY.

The reference never evaluates these; they follow from its updates:
  energy   E = sum |dct + dcu du + dcv dv|                      intensity prox, pdflow_cudalib.cu:906-923
             + mu_uv |ddt + ddu du + ddv dv - dw|                depth dual, .cu:836-846
             + lambda_i (|grad_r(du+du_prev)| + |grad_r(dv+dv_prev)|) + lambda_d |grad_r dw|
                                                                 TV duals .cu:849-885, grad_r .cu:982-1013
  residual mean over pixels with frame-1 depth of |I2(x + u) - I1(x)|, bilinear warp .cu:654-682
  EPE      mean over pixels with frame-1 depth of |(dx, dy, dz) - (dx', dy', dz')|
All planes column-major (index v + u*rows).
"""
import numpy as np


def grad_r(f, ri, rj, rows, cols):
    """r-weighted forward differences, zero at the last column / row (.cu:982-1013)"""
    F = f.reshape(cols, rows)
    g1 = np.zeros_like(F)
    g2 = np.zeros_like(F)
    g1[:-1, :] = F[1:, :] - F[:-1, :]
    g2[:, :-1] = F[:, 1:] - F[:, :-1]
    return ri * g1.ravel(), rj * g2.ravel()


def energy(fields, rows, cols, lambda_i, lambda_d):
    f = {k: np.asarray(v, np.float64) for k, v in fields.items()}
    du, dv, dw = f["du_new"], f["dv_new"], f["dw_new"]
    data_i = np.abs(f["dct"] + f["dcu"] * du + f["dcv"] * dv)
    data_d = f["mu_uv"] * np.abs(f["ddt"] + f["ddu"] * du + f["ddv"] * dv - dw)
    gu = grad_r(du + f["du_prev"], f["ri"], f["rj"], rows, cols)
    gv = grad_r(dv + f["dv_prev"], f["ri"], f["rj"], rows, cols)
    gw = grad_r(dw, f["ri"], f["rj"], rows, cols)
    tv = lambda_i * (np.hypot(*gu) + np.hypot(*gv)) + lambda_d * np.hypot(*gw)
    return float(np.sum(data_i + data_d + tv))


def bilinear(mat, rows, cols, iu, iv):
    """vectorised interpolatePixel (.cu:654-682): clamp, then the four ceil/floor cases"""
    iu = np.nan_to_num(np.asarray(iu, np.float64), nan=0.0)
    iv = np.nan_to_num(np.asarray(iv, np.float64), nan=0.0)
    iu = np.clip(iu, 0.0, cols - 1.0)
    iv = np.clip(iv, 0.0, rows - 1.0)
    su, fu = np.ceil(iu).astype(np.int64), np.floor(iu).astype(np.int64)
    sv, fv = np.ceil(iv).astype(np.int64), np.floor(iv).astype(np.int64)
    m = np.asarray(mat, np.float64)
    top = (su - iu) * m[sv + rows * fu] + (iu - fu) * m[sv + rows * su]
    bot = (su - iu) * m[fv + rows * fu] + (iu - fu) * m[fv + rows * su]
    out = (sv - iv) * bot + (iv - fv) * top
    cu = np.rint(iu).astype(np.int64)
    cv = np.rint(iv).astype(np.int64)
    eu, ev = su == fu, sv == fv
    only_u = eu & ~ev
    out[only_u] = ((sv - iv) * m[fv + rows * cu] + (iv - fv) * m[sv + rows * cu])[only_u]
    only_v = ev & ~eu
    out[only_v] = ((su - iu) * m[cv + rows * fu] + (iu - fu) * m[cv + rows * su])[only_v]
    both = eu & ev
    out[both] = m[np.rint(iv + rows * iu).astype(np.int64)][both]
    return out


def residual(I1, I2, Zold, du_l, dv_l, rows, cols):
    idx = np.arange(rows * cols)
    v, u = idx % rows, idx // rows
    w = bilinear(I2, rows, cols, u + np.asarray(du_l, np.float64), v + np.asarray(dv_l, np.float64))
    ok = np.asarray(Zold) > 0
    r = np.abs(w - np.asarray(I1, np.float64))[ok]
    return float(r.mean()) if r.size else 0.0


def epe(a, b, zold):
    ok = np.asarray(zold) > 0
    d = np.sqrt(sum((np.asarray(x, np.float64) - np.asarray(y, np.float64)) ** 2 for x, y in zip(a, b)))[ok]
    return float(d.mean()) if d.size else 0.0
