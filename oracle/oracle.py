"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only the tests, the smoke check and the CPU-baseline leg of the benchmark use this module;
it is the checker, never the thing measured or shipped.  ``Oracle`` mirrors
``pdflow.Solver`` stage by stage so the two can be stepped in lock-step.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_libm = ctypes.CDLL("libm.so.6")
_libm.tanf.argtypes = [ctypes.c_float]
_libm.tanf.restype = ctypes.c_float


def tanf(x):
    """libm tanf, the same routine the HIP library's host code calls."""
    return np.float32(_libm.tanf(np.float32(x)))


def lib_path():
    # PDF_ORACLE_LIB: another build of the oracle (tests/test_oracle_mutations.py loads builds
    # with planted mistakes to show that the known-answer tests reject them)
    return os.environ.get("PDF_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")


def load():
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run `make -C oracle`")
    lib = ctypes.CDLL(path)
    vp, c_uint, c_int, c_float = ctypes.c_void_p, ctypes.c_uint, ctypes.c_int, ctypes.c_float
    fp = ctypes.POINTER(ctypes.c_float)
    up = ctypes.POINTER(ctypes.c_uint)
    lib.orc_create.argtypes = [c_uint, c_uint, c_float, c_float, c_float, fp, c_uint, c_uint,
                               c_float, c_float, c_float]
    lib.orc_create.restype = vp
    lib.orc_destroy.argtypes = [vp]
    lib.orc_set_threads.argtypes = [vp, c_int]
    lib.orc_copy_new_frames.argtypes = [vp, fp, fp]
    lib.orc_gaussian_pyramid.argtypes = [vp, c_uint, c_uint]
    lib.orc_allocate_level.argtypes = [vp, c_uint, c_uint, c_uint, c_uint]
    for st in ("assign_zeros", "upsample", "rij", "image_gradients", "warping", "mu_and_step_sizes",
               "gradient", "dual_variables", "divergence", "primal_variables", "filter",
               "motion_field"):
        getattr(lib, "orc_" + st).argtypes = [vp]
    lib.orc_pd_iterations.argtypes = [vp, c_uint]
    lib.orc_solve.argtypes = [vp, up]
    lib.orc_get_field.argtypes = [vp, c_int, fp]
    lib.orc_set_field.argtypes = [vp, c_int, fp]
    lib.orc_get_pyramid.argtypes = [vp, c_int, c_uint, fp]
    lib.orc_set_pyramid.argtypes = [vp, c_int, c_uint, fp]
    lib.orc_rows_i.argtypes = [vp]
    lib.orc_rows_i.restype = c_uint
    lib.orc_cols_i.argtypes = [vp]
    lib.orc_cols_i.restype = c_uint
    lib.orc_ctf_levels.argtypes = [c_uint]
    lib.orc_ctf_levels.restype = c_uint
    lib.orc_num_max_iter.argtypes = [c_uint, up]
    lib.orc_gaussian_mask.argtypes = [fp]
    lib.orc_interpolate.argtypes = [fp, c_uint, c_uint, c_float, c_float]
    lib.orc_interpolate.restype = c_float
    lib.orc_interpolate_depth.argtypes = [fp, c_uint, c_uint, c_float, c_float]
    lib.orc_interpolate_depth.restype = c_float
    lib.orc_weighted_median3.argtypes = [c_int, fp, fp, fp, fp, fp, fp, fp]
    _lib = lib
    return lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


# field and pyramid ids: same order as include/pdflow_fields.h (shared with the library
# binding, which is the single place the names are listed)
def _ids():
    import sys
    pkg = os.path.join(os.path.dirname(_HERE), "pd-flow_amd")
    if pkg not in sys.path:
        sys.path.insert(0, pkg)
    from pdflow.lib import FIELDS, PYR_KINDS
    return FIELDS, PYR_KINDS


class Oracle:
    """CPU restatement of one CSF_cuda solve (see pdflow_oracle.h)."""

    def __init__(self, rows=120, cam_mode=4, threads=1, **overrides):
        from pdflow.lib import default_params
        p = default_params(rows, cam_mode)
        p.update(overrides)
        self.p = p
        self.lib = load()
        self.FIELDS, self.PYR = _ids()
        self.width, self.height = 640 // p["cam_mode"], 480 // p["cam_mode"]
        mask = np.ascontiguousarray(p["g_mask"], np.float32)
        self.tan_half = tanf(np.float32(0.5) * np.float32(p["fovh"]))
        self.ctx = self.lib.orc_create(p["rows"], p["cols"], p["lambda_i"], p["lambda_d"], p["mu"],
                                       _fp(mask), p["ctf_levels"], p["cam_mode"], p["fovh"],
                                       p["fovv"], self.tan_half)
        if not self.ctx:
            raise ValueError("orc_create rejected the parameters")
        self.lib.orc_set_threads(self.ctx, threads)
        self.pyr_levels = int(round(np.log2(self.width // p["cols"]))) + p["ctf_levels"]

    def close(self):
        if self.ctx:
            self.lib.orc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_threads(self, n):
        self.lib.orc_set_threads(self.ctx, n)

    def copy_new_frames(self, I, Z):
        I = np.ascontiguousarray(I, np.float32)
        Z = np.ascontiguousarray(Z, np.float32)
        self.lib.orc_copy_new_frames(self.ctx, _fp(I), _fp(Z))

    def gaussian_pyramid(self, levels=None):
        self.lib.orc_gaussian_pyramid(self.ctx, levels or self.pyr_levels, self.p["cam_mode"])

    def allocate_level(self, rows_i, cols_i, level, level_image):
        self.lib.orc_allocate_level(self.ctx, rows_i, cols_i, level, level_image)

    def stage(self, name):
        getattr(self.lib, "orc_" + name)(self.ctx)

    def pd_iterations(self, n):
        self.lib.orc_pd_iterations(self.ctx, n)

    def solve(self, iters=None):
        if iters is None:
            from pdflow.schedule import num_max_iter
            iters = num_max_iter(self.p["ctf_levels"])
        arr = (ctypes.c_uint * len(iters))(*iters)
        self.lib.orc_solve(self.ctx, arr)

    def level_dims(self):
        return self.lib.orc_rows_i(self.ctx), self.lib.orc_cols_i(self.ctx)

    def get_field(self, name):
        r, c = self.level_dims()
        out = np.empty(r * c, np.float32)
        self.lib.orc_get_field(self.ctx, self.FIELDS[name], _fp(out))
        return out

    def set_field(self, name, values):
        v = np.ascontiguousarray(values, np.float32)
        self.lib.orc_set_field(self.ctx, self.FIELDS[name], _fp(v))

    def get_pyramid(self, kind, level):
        n = (self.width * self.height) >> (2 * level)
        out = np.empty(n, np.float32)
        self.lib.orc_get_pyramid(self.ctx, self.PYR[kind], level, _fp(out))
        return out

    def set_pyramid(self, kind, level, values):
        v = np.ascontiguousarray(values, np.float32)
        assert v.size == (self.width * self.height) >> (2 * level)
        self.lib.orc_set_pyramid(self.ctx, self.PYR[kind], level, _fp(v))

    def motion_field(self):
        return self.get_field("dx"), self.get_field("dy"), self.get_field("dz")

    def solve_pair(self, I1, Z1, I2, Z2, iters=None):
        self.copy_new_frames(I1, Z1)
        self.gaussian_pyramid()
        self.copy_new_frames(I2, Z2)
        self.gaussian_pyramid()
        self.solve(iters)
        return self.motion_field()


def interpolate(mat, rows, cols, iu, iv, depth=False):
    m = np.ascontiguousarray(mat, np.float32)
    fn = load().orc_interpolate_depth if depth else load().orc_interpolate
    return np.float32(fn(_fp(m), rows, cols, np.float32(iu), np.float32(iv)))


def weighted_median3(fu, fv, fw, pres):
    n = len(fu)
    a = [np.ascontiguousarray(x, np.float32) for x in (fu, fv, fw, pres)]
    outs = [ctypes.c_float() for _ in range(3)]
    load().orc_weighted_median3(n, *[_fp(x) for x in a],
                                *[ctypes.cast(ctypes.byref(o), ctypes.POINTER(ctypes.c_float)) for o in outs])
    return tuple(np.float32(o.value) for o in outs)
