"""Pins the CPU oracle against what the mathematics fixes (CPU only).

The reference ships no tests, fixtures or golden outputs, and cannot be compiled here (it needs
nvcc and cuda_runtime.h), so each restated stage is checked against closed forms, invariants and
brute force computed independently in float64, chosen so that a dropped term, a wrong sign or
index, or a transposed operand fails at least one of them.
"""
import numpy as np
import pytest

import oracle
from oracle import Oracle
from pdflow import synth_pair_float, synth_static_float

RNG = np.random.default_rng(1234)


def tiny(rows=15, cam_mode=32):
    """A one-level oracle whose level 0 is (480/cam_mode) x (640/cam_mode)."""
    o = Oracle(rows, cam_mode)
    o.allocate_level(o.height, o.width, 0, 0)
    return o


def cm(a, rows):
    """column-major flat -> (rows, cols) view"""
    return a.reshape(-1, rows).T


# ------------------------------------------------------------------ pyramid (K1)

def test_pyramid_of_constants():
    o = Oracle(60, 8)
    I = np.full(o.width * o.height, 137.0, np.float32)
    Z = np.full(o.width * o.height, 1.7, np.float32)
    o.copy_new_frames(I, Z)
    o.gaussian_pyramid()
    tan_half = np.float64(o.tan_half)
    for lvl in range(o.pyr_levels):
        rows, cols = o.height >> lvl, o.width >> lvl
        col = o.get_pyramid("colour", lvl)
        dep = o.get_pyramid("depth", lvl)
        # binomial weights are dyadic: colour stays exactly constant, interior and border
        assert np.all(col == np.float32(137.0))
        np.testing.assert_allclose(dep, 1.7, rtol=2e-7)
        # pinhole back-projection x = (u - cx) Z / f, f = cols / (2 tan(fovh/2))
        u = np.repeat(np.arange(cols), rows).astype(np.float64)
        v = np.tile(np.arange(rows), cols).astype(np.float64)
        f = cols / (2 * tan_half)
        np.testing.assert_allclose(o.get_pyramid("xx", lvl), (u - (cols - 1) / 2) * dep / f,
                                   rtol=2e-6, atol=1e-7)
        np.testing.assert_allclose(o.get_pyramid("yy", lvl), (v - (rows - 1) / 2) * dep / f,
                                   rtol=2e-6, atol=1e-7)


def test_pyramid_colour_is_binomial_downsampling():
    o = Oracle(60, 8)
    rows0, cols0 = o.height, o.width
    I = RNG.uniform(0, 255, rows0 * cols0).astype(np.float32)
    Z = np.full(rows0 * cols0, 2.0, np.float32)
    o.copy_new_frames(I, Z)
    o.gaussian_pyramid(2)
    fine = cm(I, rows0).astype(np.float64)
    coarse = cm(o.get_pyramid("colour", 1), rows0 // 2)
    b = np.array([1, 4, 6, 4, 1], np.float64) / 16
    for v in range(rows0 // 2):
        for u in range(cols0 // 2):
            num = den = 0.0
            for k in range(-2, 3):
                for l in range(-2, 3):
                    vv, uu = 2 * v + k, 2 * u + l
                    if 0 <= vv < rows0 and 0 <= uu < cols0:
                        num += b[k + 2] * b[l + 2] * fine[vv, uu]
                        den += b[k + 2] * b[l + 2]
            assert abs(coarse[v, u] - num / den) < 2e-4, (v, u)


def test_pyramid_depth_edge_rules():
    o = Oracle(60, 8)
    rows0, cols0 = o.height, o.width
    Zf = np.ones((rows0, cols0), np.float32)
    Zf[:, cols0 // 2:] = 2.0                      # a 1 m depth step
    Zf[10, 10] = 0.0                              # a hole at an even (centre) site
    Zf[30:40, 60:70] = 0.0                        # a hole region
    I = np.zeros(rows0 * cols0, np.float32)
    o.copy_new_frames(I, Zf.T.ravel())
    o.gaussian_pyramid(2)
    Zc = cm(o.get_pyramid("depth", 1), rows0 // 2)
    # taps more than 0.1 m from the centre are excluded: both sides of the step stay exact
    assert Zc[5, cols0 // 4 - 1] == np.float32(1.0)
    assert Zc[20, cols0 // 4] == np.float32(2.0)
    # a hole at the centre tap makes the coarse pixel a hole (no tap is within 0.1 m of 0)
    assert Zc[5, 5] == 0.0
    # a fully empty neighbourhood gives 0
    assert Zc[17, 32] == 0.0


# ------------------------------------------------------------------ Rij (K5), gradients (K6)

def test_rij_of_fronto_parallel_plane():
    o = tiny()
    rows, cols = o.height, o.width
    Zp = 2.5
    o.copy_new_frames(np.zeros(rows * cols, np.float32), np.full(rows * cols, Zp, np.float32))
    o.gaussian_pyramid()
    o.copy_new_frames(np.zeros(rows * cols, np.float32), np.full(rows * cols, Zp, np.float32))
    o.gaussian_pyramid()
    o.stage("rij")
    ri, rj = cm(o.get_field("ri"), rows), cm(o.get_field("rj"), rows)
    inv_f = 2 * np.float64(o.tan_half) / cols
    # neighbouring points are Z*inv_f apart on a plane facing the camera: r = 2/(Z inv_f)
    np.testing.assert_allclose(ri[:, :-1], 2 / (Zp * inv_f), rtol=1e-5)
    np.testing.assert_allclose(rj[:-1, :], 2 / (Zp * inv_f), rtol=1e-5)
    assert np.all(ri[:, -1] == 1.0) and np.all(rj[-1, :] == 1.0)   # zero differences give 1


def test_rij_of_two_planes_at_different_depths():
    """Frame 1 (old) faces the camera at Z1, frame 2 (new) at Z2: ri, rj follow frame 1 and
    ri_2, rj_2 follow frame 2, each 2/(Z inv_f) from its own frame's point spacing."""
    o = tiny()
    rows, cols = o.height, o.width
    Z1, Z2 = 2.5, 1.25
    for Zp in (Z1, Z2):
        o.copy_new_frames(np.zeros(rows * cols, np.float32), np.full(rows * cols, Zp, np.float32))
        o.gaussian_pyramid()
    o.stage("rij")
    inv_f = 2 * np.float64(o.tan_half) / cols
    for name, Zp in (("ri", Z1), ("rj", Z1), ("ri_2", Z2), ("rj_2", Z2)):
        r = cm(o.get_field(name), rows)
        inner = r[:, :-1] if name.startswith("ri") else r[:-1, :]
        np.testing.assert_allclose(inner, 2 / (Zp * inv_f), rtol=1e-5, err_msg=name)


@pytest.mark.parametrize("axis", ["u", "v"])
def test_image_gradients_follow_the_connected_neighbour(axis):
    """The central differences are weighted by frame 2's connectivity: across a depth step of
    frame 2 the difference towards the far side counts less, so the gradient of a parabola at
    the step lies near the one-sided difference on the connected side.  Frame 1 is flat, so
    weights taken from it would give the plain mean of the two one-sided differences."""
    o = tiny()
    rows, cols = o.height, o.width
    u = np.repeat(np.arange(cols), rows).astype(np.float64)
    v = np.tile(np.arange(rows), cols).astype(np.float64)
    t = u if axis == "u" else v
    I = (t * t).astype(np.float32)
    step = 9 if axis == "u" else 6                 # the step lies between t = step and step + 1
    Z2 = np.where(t > step, 3.5, 2.0).astype(np.float32)
    o.copy_new_frames(I, np.full(rows * cols, 2.0, np.float32))
    o.gaussian_pyramid()
    o.copy_new_frames(I, Z2)
    o.gaussian_pyramid()
    o.stage("rij")
    o.stage("image_gradients")
    g = cm(o.get_field("dcu_aux" if axis == "u" else "dcv_aux"), rows)
    back, fwd = 2 * step - 1, 2 * step + 1        # the one-sided differences of t^2 at t = step
    at = g[:, step] if axis == "u" else g[step, :]
    assert np.all(at > back) and np.all(at < fwd)                 # a convex combination ...
    assert np.all(at - back < 0.25 * (fwd - back)), at            # ... weighted to the near side
    # one pixel beyond the step the far side is the backward one
    at1 = g[:, step + 1] if axis == "u" else g[step + 1, :]
    assert np.all((2 * step + 3) - at1 < 0.25 * 2), at1
    # away from the step both sides are connected equally: the centred difference 2t
    away = g[:, 3] if axis == "u" else g[3, :]
    np.testing.assert_allclose(away, 6.0, rtol=1e-6)


def test_image_gradients_of_a_ramp():
    o = tiny()
    rows, cols = o.height, o.width
    u = np.repeat(np.arange(cols), rows).astype(np.float32)
    v = np.tile(np.arange(rows), cols).astype(np.float32)
    I = (3 * u + 2 * v).astype(np.float32)
    Z = np.full(rows * cols, 2.0, np.float32)
    Z[7 + 5 * rows] = 0.0
    for _ in range(2):
        o.copy_new_frames(I, Z)
        o.gaussian_pyramid()
    o.stage("rij")
    o.stage("image_gradients")
    np.testing.assert_allclose(o.get_field("dcu_aux"), 3.0, rtol=1e-6)
    np.testing.assert_allclose(o.get_field("dcv_aux"), 2.0, rtol=1e-6)
    ddu = cm(o.get_field("ddu_aux"), rows)
    assert ddu[7, 5] == 0.0                      # interior depth gradient is 0 on a hole


# ------------------------------------------------------------------ warping (K7)

def _bilinear64(mat, rows, cols, iu, iv):
    m = cm(mat, rows).astype(np.float64)
    iu = min(max(iu, 0.0), cols - 1.0)
    iv = min(max(iv, 0.0), rows - 1.0)
    u0, v0 = int(np.floor(iu)), int(np.floor(iv))
    u1, v1 = min(u0 + 1, cols - 1), min(v0 + 1, rows - 1)
    a, b = iu - u0, iv - v0
    return ((1 - a) * (1 - b) * m[v0, u0] + a * (1 - b) * m[v0, u1] +
            (1 - a) * b * m[v1, u0] + a * b * m[v1, u1])


def test_bilinear_brute_force():
    rows, cols = 7, 9
    mat = RNG.uniform(-50, 50, rows * cols).astype(np.float32)
    pts = [(RNG.uniform(-2, cols + 1), RNG.uniform(-2, rows + 1)) for _ in range(400)]
    pts += [(3.0, 4.0), (3.0, 2.5), (4.5, 2.0), (0.0, 0.0), (cols - 1.0, rows - 1.0), (8.7, -3.0)]
    for iu, iv in pts:
        iu, iv = np.float32(iu), np.float32(iv)
        got = oracle.interpolate(mat, rows, cols, iu, iv)
        assert abs(got - _bilinear64(mat, rows, cols, float(iu), float(iv))) < 1e-4, (iu, iv)


def test_depth_interpolation_hole_fallback():
    rows, cols = 6, 8
    mat = RNG.uniform(1, 3, rows * cols).astype(np.float32)
    m2 = cm(mat, rows)
    # no zero tap: bilinear
    assert abs(oracle.interpolate(mat, rows, cols, 2.25, 1.75, depth=True) -
               _bilinear64(mat, rows, cols, 2.25, 1.75)) < 1e-5
    # a zero tap: nearest neighbour with round-half-even on both axes
    hole = mat.copy()
    cm(hole, rows)[1, 3] = 0.0
    got = oracle.interpolate(hole, rows, cols, 2.5, 1.5, depth=True)   # rint -> (v=2, u=2)
    assert got == m2[2, 2]
    got = oracle.interpolate(hole, rows, cols, 3.5, 0.5, depth=True)   # rint -> (v=0, u=4)
    assert got == m2[0, 4]


def test_warping_zero_flow_and_integer_shift():
    o = tiny()
    rows, cols = o.height, o.width
    I1, Z1, I2, Z2 = synth_pair_float(3, rows, cols)
    o.copy_new_frames(I1, Z1)
    o.gaussian_pyramid()
    o.copy_new_frames(I2, Z2)
    o.gaussian_pyramid()
    o.stage("assign_zeros")
    o.stage("rij")
    o.stage("image_gradients")
    o.stage("warping")
    np.testing.assert_array_equal(o.get_field("dct"), I2 - I1)
    np.testing.assert_array_equal(o.get_field("dcu"), o.get_field("dcu_aux"))
    np.testing.assert_array_equal(o.get_field("ddv"), o.get_field("ddv_aux"))
    np.testing.assert_array_equal(o.get_field("ddt"), np.where(Z2 > 0, Z2 - Z1, 0).astype(np.float32))
    # shift by one column: compare with the column to the right (clamped at the border)
    o.set_field("du_prev", np.ones(rows * cols, np.float32))
    o.stage("warping")
    shifted = cm(I2, rows)[:, np.minimum(np.arange(cols) + 1, cols - 1)]
    np.testing.assert_array_equal(cm(o.get_field("dct"), rows), shifted - cm(I1, rows))


# ------------------------------------------------------------------ step sizes (K8)

def _level_after_setup(seed=21, rows=15, cam_mode=32):
    o = tiny(rows, cam_mode)
    r, c = o.height, o.width
    I1, Z1, I2, Z2 = synth_pair_float(seed, r, c)
    o.copy_new_frames(I1, Z1)
    o.gaussian_pyramid()
    o.copy_new_frames(I2, Z2)
    o.gaussian_pyramid()
    for st in ("assign_zeros", "rij", "image_gradients", "warping", "mu_and_step_sizes"):
        o.stage(st)
    return o


def test_step_sizes_are_diagonal_preconditioning():
    """sigma/tau are the inverse row/column abs-sums of the linear operator K of the
    primal-dual problem (Pock & Chambolle 2011, alpha = 1), built here densely in float64
    from its definition: depth data term mu (ddu du + ddv dv - dw) and the r-weighted
    forward differences of du, dv (lambda_i) and dw (lambda_d)."""
    o = _level_after_setup()
    rows, cols = o.level_dims()
    P = rows * cols
    g = {n: o.get_field(n).astype(np.float64) for n in
         ("ri", "rj", "ddu", "ddv", "mu_uv", "sigma_pd", "sigma_puvx", "sigma_puvy",
          "sigma_pwx", "sigma_pwy", "tau_u", "tau_v", "tau_w")}
    li, ld = float(o.p["lambda_i"]), float(o.p["lambda_d"])
    # mu_uv = mu / (1 + 1000 |grad Z|^2)
    ddt = o.get_field("ddt").astype(np.float64)
    np.testing.assert_allclose(g["mu_uv"], 75.0 / (1 + 1000 * (g["ddu"] ** 2 + g["ddv"] ** 2 + ddt ** 2)), rtol=1e-6)
    K = np.zeros((7 * P, 3 * P))
    for x in range(P):
        v, u = x % rows, x // rows
        K[x, x] = g["mu_uv"][x] * g["ddu"][x]
        K[x, P + x] = g["mu_uv"][x] * g["ddv"][x]
        K[x, 2 * P + x] = -g["mu_uv"][x]
        for comp, lam in ((0, li), (1, li), (2, ld)):
            if u < cols - 1:
                row = P * (1 + 2 * comp) + x
                K[row, comp * P + x + rows] += lam * g["ri"][x]
                K[row, comp * P + x] -= lam * g["ri"][x]
            if v < rows - 1:
                row = P * (2 + 2 * comp) + x
                K[row, comp * P + x + 1] += lam * g["rj"][x]
                K[row, comp * P + x] -= lam * g["rj"][x]
    rs = np.abs(K).sum(axis=1)
    cs = np.abs(K).sum(axis=0)
    inner = np.array([(x % rows) < rows - 1 and (x // rows) < cols - 1 for x in range(P)])
    np.testing.assert_allclose(g["sigma_pd"], 1 / (rs[:P] + 1e-10), rtol=1e-6)
    np.testing.assert_allclose(g["sigma_puvx"][inner], 1 / rs[P:2 * P][inner], rtol=1e-6)
    np.testing.assert_allclose(g["sigma_puvy"][inner], 1 / rs[2 * P:3 * P][inner], rtol=1e-6)
    np.testing.assert_allclose(g["sigma_pwx"][inner], 1 / rs[5 * P:6 * P][inner], rtol=1e-6)
    np.testing.assert_allclose(g["sigma_pwy"][inner], 1 / rs[6 * P:7 * P][inner], rtol=1e-6)
    np.testing.assert_allclose(g["tau_u"][inner], 1 / cs[:P][inner], rtol=1e-6)
    np.testing.assert_allclose(g["tau_v"][inner], 1 / cs[P:2 * P][inner], rtol=1e-6)
    np.testing.assert_allclose(g["tau_w"][inner], 1 / cs[2 * P:][inner], rtol=1e-6)
    # reference quirk (.cu:801): at the last column the weight of the (zero) gradient row
    # is still counted, so tau is smaller there than the exact column sum implies
    last_col = np.array([(x // rows) == cols - 1 and (x % rows) < rows - 1 for x in range(P)])
    np.testing.assert_allclose(1 / g["tau_u"][last_col] - cs[:P][last_col],
                               li * g["ri"][last_col], rtol=1e-5)


# ------------------------------------------------------------------ iteration (K9-K12)

def _random_level(o):
    rows, cols = o.level_dims()
    P = rows * cols
    fields = {}
    for n in ("du_acc", "dv_acc", "dw_acc", "du_prev", "dv_prev", "puu", "puv", "pvu", "pvv", "pwu", "pwv"):
        fields[n] = RNG.normal(0, 1, P).astype(np.float32)
    for n in ("ri", "rj"):
        fields[n] = RNG.uniform(0.5, 40, P).astype(np.float32)
    for n, a in fields.items():
        o.set_field(n, a)
    return fields


def test_gradient_of_a_ramp():
    o = tiny()
    rows, cols = o.level_dims()
    f = _random_level(o)
    u = np.repeat(np.arange(cols), rows).astype(np.float32)
    v = np.tile(np.arange(rows), cols).astype(np.float32)
    o.set_field("du_acc", 0.5 * u + 0.25 * v)
    o.set_field("du_prev", np.zeros(rows * cols, np.float32))
    o.set_field("dw_acc", -2 * u + 4 * v)
    o.stage("gradient")
    g1, g2 = cm(o.get_field("gradu1"), rows), cm(o.get_field("gradu2"), rows)
    ri, rj = cm(f["ri"], rows), cm(f["rj"], rows)
    np.testing.assert_allclose(g1[:, :-1], 0.5 * ri[:, :-1], rtol=1e-6)
    np.testing.assert_allclose(g2[:-1, :], 0.25 * rj[:-1, :], rtol=1e-6)
    assert np.all(g1[:, -1] == 0) and np.all(g2[-1, :] == 0)
    np.testing.assert_allclose(cm(o.get_field("gradw1"), rows)[:, :-1], -2 * ri[:, :-1], rtol=1e-6)
    np.testing.assert_allclose(cm(o.get_field("gradw2"), rows)[:-1, :], 4 * rj[:-1, :], rtol=1e-6)


def test_divergence_is_negative_adjoint_of_gradient():
    o = tiny(30, 16)
    f = _random_level(o)
    o.stage("gradient")
    o.stage("divergence")
    g = {n: o.get_field(n).astype(np.float64) for n in
         ("gradu1", "gradu2", "gradv1", "gradv2", "gradw1", "gradw2", "divpu", "divpv", "divpw")}
    ub = f["du_acc"].astype(np.float64) + f["du_prev"]
    vb = f["dv_acc"].astype(np.float64) + f["dv_prev"]
    wb = f["dw_acc"].astype(np.float64)
    for bar, g1, g2, p1, p2, div in ((ub, "gradu1", "gradu2", "puu", "puv", "divpu"),
                                     (vb, "gradv1", "gradv2", "pvu", "pvv", "divpv"),
                                     (wb, "gradw1", "gradw2", "pwu", "pwv", "divpw")):
        lhs = np.sum(g[g1] * f[p1] + g[g2] * f[p2])
        rhs = -np.sum(bar * g[div])
        scale = np.sum(np.abs(g[g1] * f[p1]) + np.abs(g[g2] * f[p2]))
        assert abs(lhs - rhs) < 1e-5 * scale, (div, lhs, rhs)


def test_dual_projection_invariants():
    o = tiny()
    rows, cols = o.level_dims()
    P = rows * cols
    f = _random_level(o)
    for n in ("sigma_pd", "mu_uv", "ddu", "ddv", "ddt", "sigma_puvx", "sigma_puvy", "sigma_pwx", "sigma_pwy"):
        o.set_field(n, RNG.uniform(0.1, 3, P).astype(np.float32))
    for n in ("gradu1", "gradu2", "gradv1", "gradv2", "gradw1", "gradw2"):
        o.set_field(n, RNG.normal(0, 20, P).astype(np.float32))
    o.set_field("pd", RNG.normal(0, 2, P).astype(np.float32))
    o.stage("dual_variables")
    assert np.all(np.abs(o.get_field("pd")) <= 1)
    for a, b in (("puu", "puv"), ("pvu", "pvv"), ("pwu", "pwv")):
        assert np.all(np.hypot(o.get_field(a), o.get_field(b)) <= 1 + 1e-6)
    # inside the disc nothing changes: zero gradients and small duals
    small = (RNG.uniform(-0.5, 0.5, P) * 0.7).astype(np.float32)
    for n in ("gradu1", "gradu2", "gradv1", "gradv2", "gradw1", "gradw2"):
        o.set_field(n, np.zeros(P, np.float32))
    for n in ("puu", "puv", "pvu", "pvv", "pwu", "pwv"):
        o.set_field(n, small)
    o.stage("dual_variables")
    for n in ("puu", "puv", "pvu", "pvv", "pwu", "pwv"):
        np.testing.assert_array_equal(o.get_field(n), small)


def test_primal_prox_is_the_l1_proximal_map():
    """With zero divergence and pd, the primal step is the proximal map of
    |dct + dcu du + dcv dv| in the tau-weighted norm; checked by brute force along the
    only direction the minimiser can move, x(t) = x0 - t (tau_u dcu, tau_v dcv), t in [-1,1]."""
    o = tiny()
    rows, cols = o.level_dims()
    P = rows * cols
    x0u = RNG.normal(0, 1, P).astype(np.float32)
    x0v = RNG.normal(0, 1, P).astype(np.float32)
    tu = RNG.uniform(0.05, 1, P).astype(np.float32)
    tv = RNG.uniform(0.05, 1, P).astype(np.float32)
    dcu = RNG.normal(0, 2, P).astype(np.float32)
    dcv = RNG.normal(0, 2, P).astype(np.float32)
    dct = RNG.normal(0, 3, P).astype(np.float32)
    for n, a in (("du_new", x0u), ("dv_new", x0v), ("dw_new", np.zeros(P, np.float32)),
                 ("tau_u", tu), ("tau_v", tv), ("tau_w", np.ones(P, np.float32)), ("dcu", dcu),
                 ("dcv", dcv), ("dct", dct), ("pd", np.zeros(P, np.float32)),
                 ("divpu", np.zeros(P, np.float32)), ("divpv", np.zeros(P, np.float32)),
                 ("divpw", np.zeros(P, np.float32)), ("mu_uv", np.ones(P, np.float32)),
                 ("ddu", np.zeros(P, np.float32)), ("ddv", np.zeros(P, np.float32))):
        o.set_field(n, a)
    o.stage("primal_variables")
    du, dv = o.get_field("du_new"), o.get_field("dv_new")
    t = np.linspace(-1, 1, 200001)
    for x in range(0, P, 7):
        a, b = float(tu[x] * dcu[x]), float(tv[x] * dcv[x])
        cu_, cv_ = x0u[x] - t * a, x0v[x] - t * b
        obj = (0.5 * (cu_ - x0u[x]) ** 2 / tu[x] + 0.5 * (cv_ - x0v[x]) ** 2 / tv[x] +
               np.abs(dct[x] + dcu[x] * cu_ + dcv[x] * cv_))
        k = np.argmin(obj)
        assert abs(du[x] - cu_[k]) < 1e-4 * (1 + abs(a)) and abs(dv[x] - cv_[k]) < 1e-4 * (1 + abs(b)), x
    np.testing.assert_array_equal(o.get_field("du_acc"), (2 * du - x0u).astype(np.float32))


def test_zero_motion_gives_exactly_zero_flow():
    I, Z = synth_static_float(17, 60, 80)
    o = Oracle(60, 8)
    dx, dy, dz = o.solve_pair(I, Z, I, Z)
    assert np.all(dx == 0) and np.all(dy == 0) and np.all(dz == 0)


# ------------------------------------------------------------------ filter and motion field

def _median64(f, w):
    order = np.argsort(f, kind="stable")
    fs, ws = np.asarray(f, np.float64)[order], np.asarray(w, np.float64)[order]
    cum = np.cumsum(ws)
    return np.interp(0.5 * cum[-1], cum, fs)


@pytest.mark.parametrize("n", [9, 6, 4])
def test_weighted_median_brute_force(n):
    for _ in range(200):
        fu, fv, fw = (RNG.normal(0, 1, n).astype(np.float32) for _ in range(3))
        w = RNG.uniform(0.05, 1, n).astype(np.float32)
        mu, mv, mw = oracle.weighted_median3(fu, fv, fw, w)
        for m, f in ((mu, fu), (mv, fv), (mw, fw)):
            assert abs(m - _median64(f, w)) < 1e-5


def test_weighted_median_uniform_is_mean_of_middle_pair():
    f = np.array([5, 1, 9, 3, 7, 2, 8, 4, 6], np.float32)
    w = np.full(9, 0.25, np.float32)
    mu, _, _ = oracle.weighted_median3(f, f, f, w)
    assert mu == np.float32(4.5)          # mean of the 4th and 5th smallest, not the 5th


@pytest.mark.parametrize("via", ["depth", "ddt"])
def test_median_filter_weights(via):
    """The median filter's tap weight is 1 / (1 + 5 dz^2 + 10 dt^2) with dz the depth difference to
    the centre and dt the tap's ddt (.cu:1075).  Four taps of value 0 at the centre's depth and five
    of value 1 at dz = 0.2 (or at dt = 0.1) move the half-weight point into the first 1 by a
    fraction fixed by those weights: (4 + 5 w) / 2 - 4 = w * m with w = 1/1.2 (1/1.1) gives m = 0.1
    (0.3), worked out by hand."""
    o = _level_after_setup()
    rows, cols = o.level_dims()
    P = rows * cols
    v0, u0 = 7, 10
    depth = np.ones(P, np.float32)
    ddt = np.zeros(P, np.float32)
    du = np.zeros(P, np.float32)
    ones = [(-1, -1), (0, -1), (1, 0), (1, 1), (-1, 1)]       # five taps; the other four stay 0
    for k, l in ones:
        i = (v0 + k) + (u0 + l) * rows
        du[i] = 1.0
        if via == "depth":
            depth[i] = 1.2
        else:
            ddt[i] = 0.1
    o.set_pyramid("depth_old", 0, depth)
    for name, val in (("du_new", du), ("dv_new", du), ("dw_new", du), ("ddt", ddt),
                      ("du_prev", np.zeros(P, np.float32)), ("dv_prev", np.zeros(P, np.float32))):
        o.set_field(name, val)
    o.stage("filter")
    want = 0.1 if via == "depth" else 0.3
    for name in ("du_l", "dv_l", "dw_l"):
        assert abs(float(o.get_field(name)[v0 + u0 * rows]) - want) < 1e-5, name


def test_upsampling_of_a_constant_field():
    o = Oracle(30, 16)                    # levels 15x20 and 30x40
    rows, cols = 30, 40
    o.allocate_level(15, 20, 0, 1)
    c = np.float32(0.3)
    for n in ("du_l", "dv_l", "dw_l", "pd_l", "puu_l", "puv_l", "pvu_l", "pvv_l", "pwu_l", "pwv_l"):
        o.set_field(n, np.full(300, c, np.float32))
    o.allocate_level(rows, cols, 1, 0)
    o.stage("assign_zeros")
    o.stage("upsample")
    b = np.array([1, 4, 6, 4, 1], np.float64) / 16
    g = np.outer(b, b)
    z = np.zeros((rows, cols))
    z[::2, ::2] = float(c)
    for name, scale in (("du_prev", 2), ("dv_prev", 2), ("dw_new", 1), ("pd", 1), ("pwv", 1), ("dw_acc", 1)):
        got = cm(o.get_field(name), rows)
        for v in range(rows):
            for u in range(cols):
                num, out = 0.0, 0.0
                for k in range(-2, 3):
                    for l in range(-2, 3):
                        vv, uu = v + k, u + l
                        if 0 <= vv < rows and 0 <= uu < cols:
                            num += 4 * g[k + 2, l + 2] * z[vv, uu]
                        else:
                            out += g[k + 2, l + 2]
                want = scale * num / (1 - out)
                if 2 <= v <= rows - 3 and 2 <= u <= cols - 3:
                    assert abs(want - scale * float(c)) < 1e-12   # interior: the constant itself
                assert abs(got[v, u] - want) < 1e-6, (name, v, u)


def test_motion_field_is_linearised_pinhole_motion():
    o = tiny(30, 16)
    rows, cols = o.level_dims()
    P = rows * cols
    I1, Z1, I2, Z2 = synth_pair_float(8, rows, cols)
    o.copy_new_frames(I1, Z1)
    o.gaussian_pyramid()
    o.copy_new_frames(I2, Z2)
    o.gaussian_pyramid()
    du = RNG.normal(0, 0.7, P).astype(np.float32)
    dv = RNG.normal(0, 0.7, P).astype(np.float32)
    dw = RNG.normal(0, 0.03, P).astype(np.float32)
    o.set_field("du_l", du)
    o.set_field("dv_l", dv)
    o.set_field("dw_l", dw)
    o.stage("motion_field")
    dx, dy, dz = o.motion_field()
    f = cols / (2 * np.float64(o.tan_half))
    u = np.repeat(np.arange(cols), rows) - (cols - 1) / 2
    v = np.tile(np.arange(rows), cols) - (rows - 1) / 2
    Z = Z1.astype(np.float64)
    m = Z > 0
    # exact displacement of the back-projected point: ((u+du)(Z+dw) - u Z) / f
    X = ((u + du) * (Z + dw) - u * Z) / f
    Y = ((v + dv) * (Z + dw) - v * Z) / f
    so = np.abs(du * dw / f) + 1e-6        # the formula drops the second-order du*dw/f term
    assert np.all(np.abs(dx[m] - dw[m]) == 0)
    assert np.all(np.abs(dy[m] - X[m]) <= so[m] + 1e-6 * np.abs(X[m]))
    assert np.all(np.abs(dz[m] - Y[m]) <= np.abs(dv * dw / f)[m] + 1e-6 + 1e-6 * np.abs(Y[m]))
    assert np.all(dx[~m] == 0) and np.all(dy[~m] == 0) and np.all(dz[~m] == 0)


def test_saturation():
    o = _level_after_setup()
    rows, cols = o.level_dims()
    P = rows * cols
    du = RNG.normal(0, 2, P).astype(np.float32)
    prev = RNG.normal(0, 1, P).astype(np.float32)
    o.set_field("du_new", du)
    o.set_field("dv_new", du)
    o.set_field("dw_new", np.ones(P, np.float32))
    o.set_field("du_prev", prev)
    o.set_field("dv_prev", prev)
    o.stage("filter")
    zold = o.get_pyramid("depth_old", 0)
    np.testing.assert_array_equal(o.get_field("du_new"), np.clip(du, -1, 1) + prev)
    np.testing.assert_array_equal(o.get_field("dw_new"), np.where(zold == 0, 0, 1).astype(np.float32))
    # holes pass through the median unchanged
    h = zold == 0
    np.testing.assert_array_equal(o.get_field("du_l")[h], o.get_field("du_new")[h])


def test_pyramid_edge_aware_depth_brute_force():
    """Depth downsampling weights each tap by g * (0.1 - |d - d_centre|) over taps with
    d > 0 and |d - d_centre| < 0.1 m (the edge-aware rule), recomputed in float64."""
    o = Oracle(60, 8)
    rows0, cols0 = o.height, o.width
    Zf = (1.5 + RNG.uniform(-0.12, 0.12, (rows0, cols0))).astype(np.float32)
    Zf[RNG.uniform(size=Zf.shape) < 0.1] = 0.0
    o.copy_new_frames(np.zeros(rows0 * cols0, np.float32), Zf.T.ravel())
    o.gaussian_pyramid(2)
    Zc = cm(o.get_pyramid("depth", 1), rows0 // 2)
    b = np.array([1, 4, 6, 4, 1], np.float64) / 16
    z64 = Zf.astype(np.float64)
    for v in range(rows0 // 2):
        for u in range(cols0 // 2):
            dc = z64[2 * v, 2 * u]
            num = den = 0.0
            for k in range(-2, 3):
                for l in range(-2, 3):
                    vv, uu = 2 * v + k, 2 * u + l
                    if 0 <= vv < rows0 and 0 <= uu < cols0:
                        d = z64[vv, uu]
                        if d > 0 and abs(d - dc) < 0.1:
                            w = b[k + 2] * b[l + 2] * (0.1 - abs(d - dc))
                            num += w * d
                            den += w
            want = num / den if num > 0 else 0.0
            assert abs(Zc[v, u] - want) < 2e-6 * max(1.0, want), (v, u)


def test_depth_dual_step_is_l1_dual_ascent():
    """pd is the dual of the depth term |ddt + ddu du + ddv dv - dw| (new depth = old depth
    + dw along the ray): with a very large step it jumps to the sign of that residual."""
    o = tiny()
    rows, cols = o.level_dims()
    P = rows * cols
    vals = {n: RNG.normal(0, 1, P).astype(np.float32) for n in ("du_acc", "dv_acc", "dw_acc", "ddu", "ddv", "ddt")}
    for n, a in vals.items():
        o.set_field(n, a)
    o.set_field("pd", np.zeros(P, np.float32))
    o.set_field("mu_uv", np.ones(P, np.float32))
    o.set_field("sigma_pd", np.full(P, 1e6, np.float32))
    for n in ("gradu1", "gradu2", "gradv1", "gradv2", "gradw1", "gradw2"):
        o.set_field(n, np.zeros(P, np.float32))
    o.stage("dual_variables")
    res = (vals["ddt"].astype(np.float64) + vals["ddu"] * vals["du_acc"].astype(np.float64) +
           vals["ddv"] * vals["dv_acc"].astype(np.float64) - vals["dw_acc"])
    np.testing.assert_array_equal(o.get_field("pd"), np.sign(res).astype(np.float32))


@pytest.mark.parametrize("rows,cam_mode", [(60, 8), (120, 4)])
def test_solver_recovers_ground_truth_motion(rows, cam_mode):
    """End to end against physics: on rendered pairs the estimated flow is far closer to the
    generator's true 3-D motion than zero flow is (signs, axes and units all matter here)."""
    errs, mags = [], []
    for seed in (101, 102, 103):
        I1, Z1, I2, Z2, gt = synth_pair_float(seed, 480 // cam_mode, 640 // cam_mode, with_gt=True)
        o = Oracle(rows, cam_mode, threads=4)
        est = np.stack(o.solve_pair(I1, Z1, I2, Z2))
        m = Z1 > 0
        errs.append(np.median(np.linalg.norm(est[:, m] - gt[:, m], axis=0)))
        mags.append(np.median(np.linalg.norm(gt[:, m], axis=0)))
    assert np.mean(errs) < 0.35 * np.mean(mags), (errs, mags)
