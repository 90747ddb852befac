"""Solver geometry and schedule rules of the reference pair evaluator.

Pure host arithmetic, mirroring scene_flow_impair.cpp:38-58 and :103-108.
"""
import math

import numpy as np

# reference defaults (scene_flow_impair.cpp:41-42, :67-69)
LAMBDA_I = np.float32(0.04)
LAMBDA_D = np.float32(0.35)
MU = np.float32(75.0)
# M_PI*62.5f/180.f is evaluated in double and stored as float
FOVH = np.float32(math.pi * 62.5 / 180.0)
FOVV = np.float32(math.pi * 48.5 / 180.0)


def ctf_levels_for_rows(rows):
    """ctf_levels = int(log2(float(rows/15))) + 1 (scene_flow_impair.cpp:40)."""
    return int(math.log2(float(np.float32(rows // 15)))) + 1


def num_max_iter(ctf_levels):
    """Iterations per level, coarse first: 100 at the finest, 15 fewer per coarser level
    (scene_flow_impair.cpp:46-52)."""
    out = [0] * 6
    for i in range(5, -1, -1):
        out[i] = 100 if i >= ctf_levels - 1 else out[i + 1] - 15
    return out[:ctf_levels]


def gaussian_mask():
    """25-float binomial mask, m[i + 5 j] = b_i b_j / 256 (scene_flow_impair.cpp:55-58)."""
    b = [1, 4, 6, 4, 1]
    m = np.zeros(25, np.float32)
    for i in range(5):
        for j in range(5):
            m[i + 5 * j] = np.float32(b[i] * b[j]) / np.float32(256.0)
    return m


def level_table(rows, cam_mode):
    """(level i, rows_i, cols_i, level_image, iterations) for each solve level, coarse first
    (scene_flow_impair.cpp:103-108)."""
    cols = rows * 320 // 240
    width = 640 // cam_mode
    L = ctf_levels_for_rows(rows)
    lw = int(math.log2(width // cols))
    its = num_max_iter(L)
    out = []
    for i in range(L):
        s = 2 ** (L - (i + 1))
        out.append((i, rows // s, cols // s, L - i + lw - 1, its[i]))
    return out


def pixel_iterations(rows, cam_mode, iters=None):
    """Sum over levels of P_l * N_l."""
    t = level_table(rows, cam_mode)
    if iters is None:
        iters = [x[4] for x in t]
    return sum(r * c * n for (_, r, c, _, _), n in zip(t, iters))
