"""ctypes binding of the seeded synthetic RGB-D pair generator (include/pdflow_synth.h)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libpdflow_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(path)
        fp = ctypes.POINTER(ctypes.c_float)
        u8 = ctypes.POINTER(ctypes.c_uint8)
        u16 = ctypes.POINTER(ctypes.c_uint16)
        lib.pdf_synth_pair.argtypes = [ctypes.c_uint64, ctypes.c_uint, ctypes.c_uint, u8, u16, u8, u16, fp]
        lib.pdf_synth_pair_float.argtypes = [ctypes.c_uint64, ctypes.c_uint, ctypes.c_uint, fp, fp, fp, fp, fp]
        lib.pdf_synth_static_float.argtypes = [ctypes.c_uint64, ctypes.c_uint, ctypes.c_uint, fp, fp]
        lib.pdf_synth_sequence_float.argtypes = [ctypes.c_uint64, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, fp, fp]
        lib.pdf_synth_pairs_float.argtypes = [ctypes.c_uint64, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                              fp, fp, fp, fp, fp, ctypes.c_int]
        lib.pdf_synth_hash.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        lib.pdf_synth_hash.restype = ctypes.c_uint64
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def synth_pair(seed, rows, cols, with_gt=False):
    """uint8 intensity and uint16 depth (metres*5000) of both frames, column-major flat
    arrays of rows*cols; plus the (3, rows*cols) ground-truth motion if asked."""
    P = rows * cols
    i1 = np.zeros(P, np.uint8)
    i2 = np.zeros(P, np.uint8)
    z1 = np.zeros(P, np.uint16)
    z2 = np.zeros(P, np.uint16)
    gt = np.zeros(3 * P, np.float32) if with_gt else None
    rc = _load().pdf_synth_pair(seed, rows, cols, _p(i1, ctypes.c_uint8), _p(z1, ctypes.c_uint16),
                                _p(i2, ctypes.c_uint8), _p(z2, ctypes.c_uint16),
                                _p(gt, ctypes.c_float) if with_gt else None)
    if rc:
        raise RuntimeError(f"pdf_synth_pair failed ({rc})")
    if with_gt:
        return i1, z1, i2, z2, gt.reshape(3, P)
    return i1, z1, i2, z2


def synth_pair_float(seed, rows, cols, with_gt=False):
    """Solver input planes I = float(u8), Z = float(u16)/5000 of both frames."""
    P = rows * cols
    arrs = [np.zeros(P, np.float32) for _ in range(4)]
    gt = np.zeros(3 * P, np.float32) if with_gt else None
    rc = _load().pdf_synth_pair_float(seed, rows, cols, *[_p(a, ctypes.c_float) for a in arrs],
                                      _p(gt, ctypes.c_float) if with_gt else None)
    if rc:
        raise RuntimeError(f"pdf_synth_pair_float failed ({rc})")
    if with_gt:
        return (*arrs, gt.reshape(3, P))
    return tuple(arrs)


def synth_pairs_float(seed0, n, rows, cols, threads=None, with_gt=False):
    """n pairs (seeds seed0..seed0+n-1) as four (n, rows*cols) float arrays (plus the
    (n, 3, rows*cols) ground-truth motion if asked), rendered on `threads` host threads
    (default: up to 16)."""
    P = rows * cols
    arrs = [np.empty((n, P), np.float32) for _ in range(4)]
    gt = np.empty((n, 3, P), np.float32) if with_gt else None
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    rc = _load().pdf_synth_pairs_float(seed0, n, rows, cols, *[_p(a, ctypes.c_float) for a in arrs],
                                       _p(gt, ctypes.c_float) if with_gt else None, threads)
    if rc:
        raise RuntimeError(f"pdf_synth_pairs_float failed ({rc})")
    return (*arrs, gt) if with_gt else tuple(arrs)


def synth_static_float(seed, rows, cols):
    P = rows * cols
    I = np.zeros(P, np.float32)
    Z = np.zeros(P, np.float32)
    rc = _load().pdf_synth_static_float(seed, rows, cols, _p(I, ctypes.c_float), _p(Z, ctypes.c_float))
    if rc:
        raise RuntimeError(f"pdf_synth_static_float failed ({rc})")
    return I, Z


def synth_sequence_float(seed, rows, cols, t):
    """Frame t of the seeded moving scene (I, Z float planes)."""
    P = rows * cols
    I = np.zeros(P, np.float32)
    Z = np.zeros(P, np.float32)
    rc = _load().pdf_synth_sequence_float(seed, rows, cols, t, _p(I, ctypes.c_float), _p(Z, ctypes.c_float))
    if rc:
        raise RuntimeError(f"pdf_synth_sequence_float failed ({rc})")
    return I, Z


def synth_hash(arr):
    a = np.ascontiguousarray(arr)
    return int(_load().pdf_synth_hash(a.ctypes.data, a.nbytes))
