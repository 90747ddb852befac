"""ctypes binding of libpdflow_hip.so (C-ABI: include/pdflow_hip.h).

``Solver`` is one solver context, the counterpart of one reference ``CSF_cuda`` object
(pdflow_cudalib.h:38-154); its methods carry the reference names of the bridges and host
methods they replace.
"""
import ctypes
import os

import numpy as np

from . import schedule

_HERE = os.path.dirname(os.path.abspath(__file__))

# plane ids, in the order of include/pdflow_fields.h
_FIELD_NAMES = [
    "dct", "dcu", "dcv", "ddt", "ddu", "ddv",
    "dcu_aux", "dcv_aux", "ddu_aux", "ddv_aux",
    "gradu1", "gradu2", "gradv1", "gradv2", "gradw1", "gradw2",
    "divpu", "divpv", "divpw",
    "sigma_pd", "sigma_puvx", "sigma_puvy", "sigma_pwx", "sigma_pwy",
    "tau_u", "tau_v", "tau_w",
    "mu_uv", "ri", "rj", "ri_2", "rj_2",
    "du_acc", "dv_acc", "dw_acc", "du_prev", "dv_prev",
    "du_upsamp", "dv_upsamp", "dw_upsamp", "pd_upsamp",
    "puu_upsamp", "puv_upsamp", "pvu_upsamp", "pvv_upsamp", "pwu_upsamp", "pwv_upsamp",
    "du_new", "dv_new", "dw_new", "pd", "puu", "puv", "pvu", "pvv", "pwu", "pwv",
    "dx", "dy", "dz",
    "du_l", "dv_l", "dw_l", "pd_l", "puu_l", "puv_l", "pvu_l", "pvv_l", "pwu_l", "pwv_l",
]
FIELDS = {n: i for i, n in enumerate(_FIELD_NAMES)}
PYR_KINDS = {n: i for i, n in enumerate(
    ["colour", "colour_old", "depth", "depth_old", "xx", "xx_old", "yy", "yy_old"])}

STATUS = {0: "ok", 1: "bad argument", 2: "out of sequence", 3: "HIP error", 4: "out of device memory"}


class PDFlowError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib_path():
    return os.path.join(_HERE, "libpdflow_hip.so")


def hip_library():
    """Loads libpdflow_hip.so; raises if it is not built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: the HIP library must be built "
                           "(__graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(path)
    c_uint, c_int, c_float = ctypes.c_uint, ctypes.c_int, ctypes.c_float
    vp = ctypes.c_void_p
    fp = ctypes.POINTER(ctypes.c_float)
    up = ctypes.POINTER(ctypes.c_uint)
    sigs = {
        "pdf_create": ([ctypes.POINTER(vp), c_uint, c_uint, c_float, c_float, c_float, fp, c_uint,
                        c_uint, c_float, c_float], c_int),
        "pdf_destroy": ([vp], c_int),
        "pdf_set_arena_limit": ([ctypes.c_size_t], c_int),
        "pdf_set_stream": ([vp, vp], c_int),
        "pdf_synchronize": ([vp], c_int),
        "pdf_last_error": ([vp], ctypes.c_char_p),
        "pdf_last_create_error": ([], ctypes.c_char_p),
        "pdf_copy_new_frames": ([vp, fp, fp], c_int),
        "pdf_copy_new_frames_device": ([vp, vp, vp], c_int),
        "pdf_gaussian_pyramid": ([vp, c_uint, c_uint], c_int),
        "pdf_allocate_level": ([vp, c_uint, c_uint, c_uint, c_uint], c_int),
        "pdf_free_level": ([vp], c_int),
        "pdf_pd_iterations": ([vp, c_uint], c_int),
        "pdf_set_iteration_kernel": ([vp, c_int, c_uint], c_int),
        "pdf_set_solve_options": ([vp, c_int, c_int], c_int),
        "pdf_copy_motion_field": ([vp, fp, fp, fp], c_int),
        "pdf_copy_all_solutions": ([vp] + [fp] * 11, c_int),
        "pdf_motion_field_device": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)], c_int),
        "pdf_solve": ([vp, up], c_int),
        "pdf_solve_graph": ([vp, up], c_int),
        "pdf_solve_pair_device": ([vp, vp, vp, vp, vp, up, c_int], c_int),
        "pdf_solve_pair_noalloc": ([c_uint, c_uint, c_float, c_float, c_float, fp, c_uint, c_uint, c_float,
                                    vp, vp, vp, vp, up, vp, vp, vp, vp], c_int),
        "pdf_get_field": ([vp, c_int, fp], c_int),
        "pdf_set_field": ([vp, c_int, fp], c_int),
        "pdf_get_pyramid": ([vp, c_int, c_uint, fp], c_int),
        "pdf_level_dims": ([vp, up, up], c_int),
        "pdf_tan_half_fovh": ([vp], c_float),
        "pdf_enable_kernel_timing": ([vp, c_int], c_int),
        "pdf_kernel_timing": ([vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)], c_int),
        "pdf_device_alloc": ([ctypes.POINTER(vp), ctypes.c_size_t], c_int),
        "pdf_device_free": ([vp], c_int),
        "pdf_memcpy_h2d": ([vp, vp, ctypes.c_size_t], c_int),
        "pdf_memcpy_d2h": ([vp, vp, ctypes.c_size_t], c_int),
        "pdf_host_alloc": ([ctypes.POINTER(vp), ctypes.c_size_t], c_int),
        "pdf_host_free": ([vp], c_int),
        "pdf_device_count": ([ctypes.POINTER(c_int)], c_int),
    }
    up_ = ctypes.POINTER(ctypes.c_uint)
    fpp = ctypes.POINTER(ctypes.POINTER(ctypes.c_float))
    sigs.update({
        "pdf_batch_create": ([ctypes.POINTER(vp), c_uint, up_, up_, c_float, c_float, c_float, fp, c_float, up_], c_int),
        "pdf_batch_destroy": ([vp], c_int),
        "pdf_batch_size": ([vp, up_], c_int),
        "pdf_batch_last_error": ([vp], ctypes.c_char_p),
        "pdf_batch_set_iteration_kernel": ([vp, c_int, c_uint, c_int], c_int),
        "pdf_batch_inputs_device": ([vp, c_uint, fpp, fpp, fpp, fpp], c_int),
        "pdf_batch_set_inputs": ([vp, c_uint, fp, fp, fp, fp], c_int),
        "pdf_batch_solve": ([vp, c_int], c_int),
        "pdf_batch_synchronize": ([vp], c_int),
        "pdf_batch_wait": ([vp, vp], c_int),
        "pdf_batch_flow_device": ([vp, c_uint, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), up_, up_], c_int),
        "pdf_batch_copy_flow": ([vp, c_uint, fp, fp, fp], c_int),
        "pdf_batch_upload_inputs_async": ([vp, fp, ctypes.c_size_t], c_int),
        "pdf_batch_set_setup_tile": ([vp, c_int], c_int),
        "pdf_batch_setup_tile": ([vp, ctypes.POINTER(c_int)], c_int),
        "pdf_batch_download_flows_async": ([vp, fp, ctypes.c_size_t], c_int),
        "pdf_batch_launch_count": ([vp, up_], c_int),
        "pdf_batch_set_start_level": ([vp, c_uint, c_uint, fpp], c_int),
        "pdf_batch_set_schedule": ([vp, c_uint, up_], c_int),
        "pdf_solve_warm": ([vp, up, c_uint, vp], c_int),
        "pdf_stream_frame_device": ([vp, vp, vp, up, c_int, ctypes.POINTER(c_int)], c_int),
        "pdf_batch_enable_kernel_timing": ([vp, c_int], c_int),
        "pdf_device_copy": ([vp, vp, ctypes.c_size_t], c_int),
        "pdf_batch_continue": ([vp, ctypes.POINTER(c_uint), c_uint, c_uint], c_int),
        "pdf_batch_metrics_pairs": ([vp, ctypes.POINTER(c_uint), c_uint, fp, fp], c_int),
        "pdf_check_exact_sequence": ([c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_uint32)], c_int),
        "pdf_set_tb_shape": ([vp, c_int], c_int),
        "pdf_batch_set_tb_shape": ([vp, c_int], c_int),
        "pdf_batch_autotune": ([vp, ctypes.c_char_p, ctypes.POINTER(c_int), ctypes.POINTER(c_int), fp], c_int),
        "pdf_batch_field_device": ([vp, c_uint, c_int, ctypes.POINTER(vp)], c_int),
        "pdf_batch_pyramid_device": ([vp, c_uint, c_int, c_uint, ctypes.POINTER(vp), up_, up_], c_int),
        "pdf_batch_set_learned": ([vp, vp, up_, c_uint], c_int),
        "pdf_batch_metrics": ([vp, fp, fp], c_int),
        "pdf_batch_epe": ([vp, vp, fp], c_int),
        "pdf_batch_quality": ([vp, up_, c_uint, fp, fp, fp], c_int),
        "pdf_batch_fallback_bounds": ([vp, c_float, c_float, ctypes.POINTER(ctypes.c_ubyte), up_], c_int),
        "pdf_batch_resolve_plain": ([vp, up_, c_uint], c_int),
        "pdf_batch_fallback": ([vp, fp, fp, c_float, ctypes.POINTER(ctypes.c_ubyte), up_], c_int),
        "pdf_batch_learn_samples": ([vp, c_uint, c_uint, vp, vp, vp], c_int),
        "pdf_batch_kernel_timing": ([vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)], c_int),
    })
    for stage in ("assign_zeros", "upsample", "rij", "image_gradients", "warping",
                  "mu_and_step_sizes", "gradient", "dual_variables", "divergence",
                  "primal_variables", "filter", "motion_field"):
        sigs["pdf_" + stage] = ([vp], c_int)
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def device_count():
    n = ctypes.c_int(0)
    hip_library().pdf_device_count(ctypes.byref(n))
    return n.value


EXACT_SQRT, EXACT_RCP, EXACT_INV_SQRT, EXACT_ONE_OVER, EXACT_HALF_OVER = 0, 1, 2, 3, 4


def check_exact_sequence(which, lo, hi):
    """Runs one of the iteration kernels' short exact sequences and the compiler's IEEE operation
    on every float whose bit pattern lies in [lo, hi]; returns (mismatches, first mismatching bit
    pattern or None)."""
    n, first = ctypes.c_uint64(0), ctypes.c_uint32(0)
    st = hip_library().pdf_check_exact_sequence(which, lo, hi, ctypes.byref(n), ctypes.byref(first))
    if st != 0:
        raise PDFlowError(st, "pdf_check_exact_sequence")
    return n.value, (None if n.value == 0 else first.value)


def default_params(rows, cam_mode):
    """Reference parameters of the pair evaluator for a solve of `rows` rows
    (scene_flow_impair.cpp:38-69)."""
    return dict(rows=rows, cols=rows * 320 // 240, lambda_i=schedule.LAMBDA_I,
                lambda_d=schedule.LAMBDA_D, mu=schedule.MU, g_mask=schedule.gaussian_mask(),
                ctf_levels=schedule.ctf_levels_for_rows(rows), cam_mode=cam_mode,
                fovh=schedule.FOVH, fovv=schedule.FOVV)


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class DeviceBuffer:
    """A raw device allocation made through the library (no torch needed)."""

    def __init__(self, nbytes):
        self.lib = hip_library()
        self.ptr = ctypes.c_void_p()
        rc = self.lib.pdf_device_alloc(ctypes.byref(self.ptr), nbytes)
        if rc:
            raise PDFlowError(rc, f"device allocation of {nbytes} bytes")
        self.nbytes = nbytes

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        rc = self.lib.pdf_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes)
        if rc:
            raise PDFlowError(rc, "host-to-device copy")

    def download(self, arr):
        rc = self.lib.pdf_memcpy_d2h(arr.ctypes.data, self.ptr, arr.nbytes)
        if rc:
            raise PDFlowError(rc, "device-to-host copy")
        return arr

    def free(self):
        if self.ptr:
            self.lib.pdf_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedBuffer:
    """Page-locked host memory (pdf_host_alloc) seen as a float32 NumPy array of `floats`
    elements; copies from and to it can overlap kernels."""

    def __init__(self, floats):
        self.lib = hip_library()
        self.ptr = ctypes.c_void_p()
        rc = self.lib.pdf_host_alloc(ctypes.byref(self.ptr), floats * 4)
        if rc:
            raise PDFlowError(rc, f"page-locked allocation of {floats * 4} bytes")
        self.array = np.ctypeslib.as_array(ctypes.cast(self.ptr, ctypes.POINTER(ctypes.c_float)), shape=(floats,))

    def free(self):
        if self.ptr:
            self.array = None
            self.lib.pdf_host_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Solver:
    """One solver context (one reference CSF_cuda object)."""

    def __init__(self, rows=120, cam_mode=4, **overrides):
        p = default_params(rows, cam_mode)
        p.update(overrides)
        self.p = p
        self.lib = hip_library()
        self.width, self.height = 640 // p["cam_mode"], 480 // p["cam_mode"]
        self.ctx = ctypes.c_void_p()
        mask = np.ascontiguousarray(p["g_mask"], np.float32)
        rc = self.lib.pdf_create(ctypes.byref(self.ctx), p["rows"], p["cols"], p["lambda_i"],
                                 p["lambda_d"], p["mu"], _fptr(mask), p["ctf_levels"],
                                 p["cam_mode"], p["fovh"], p["fovv"])
        if rc:
            raise PDFlowError(rc, self.lib.pdf_last_create_error().decode())
        self.pyr_levels = int(round(np.log2(self.width // p["cols"]))) + p["ctf_levels"]

    # -- plumbing
    def _chk(self, rc):
        if rc:
            raise PDFlowError(rc, self.lib.pdf_last_error(self.ctx).decode())

    def close(self):
        if self.ctx:
            self.lib.pdf_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def tan_half(self):
        return np.float32(self.lib.pdf_tan_half_fovh(self.ctx))

    # -- reference API, stage by stage
    def copy_new_frames(self, I, Z):
        I = np.ascontiguousarray(I, np.float32)
        Z = np.ascontiguousarray(Z, np.float32)
        assert I.size == self.width * self.height and Z.size == I.size
        self._chk(self.lib.pdf_copy_new_frames(self.ctx, _fptr(I), _fptr(Z)))

    def gaussian_pyramid(self, levels=None):
        self._chk(self.lib.pdf_gaussian_pyramid(self.ctx, levels or self.pyr_levels, self.p["cam_mode"]))

    def allocate_level(self, rows_i, cols_i, level, level_image):
        self._chk(self.lib.pdf_allocate_level(self.ctx, rows_i, cols_i, level, level_image))

    def stage(self, name):
        self._chk(getattr(self.lib, "pdf_" + name)(self.ctx))

    def pd_iterations(self, n):
        self._chk(self.lib.pdf_pd_iterations(self.ctx, n))

    def set_iteration_kernel(self, k, min_pixels=0):
        """k iterations per launch (temporal blocking) on levels of >= min_pixels pixels."""
        self._chk(self.lib.pdf_set_iteration_kernel(self.ctx, k, min_pixels))

    def set_tb_shape(self, shape):
        """region shape of the temporally blocked kernel: 0 = 64x32 (1024 threads), 1 = 32x32"""
        self._chk(self.lib.pdf_set_tb_shape(self.ctx, shape))

    def set_solve_options(self, fused_setup=True, level_kernel=True):
        self._chk(self.lib.pdf_set_solve_options(self.ctx, int(fused_setup), int(level_kernel)))

    def solve(self, iters=None, graph=False):
        arr = None if iters is None else (ctypes.c_uint * len(iters))(*iters)
        fn = self.lib.pdf_solve_graph if graph else self.lib.pdf_solve
        self._chk(fn(self.ctx, arr))

    def synchronize(self):
        self._chk(self.lib.pdf_synchronize(self.ctx))

    def level_dims(self):
        r, c = ctypes.c_uint(), ctypes.c_uint()
        self._chk(self.lib.pdf_level_dims(self.ctx, ctypes.byref(r), ctypes.byref(c)))
        return r.value, c.value

    def get_field(self, name):
        r, c = self.level_dims()
        out = np.empty(r * c, np.float32)
        self._chk(self.lib.pdf_get_field(self.ctx, FIELDS[name], _fptr(out)))
        return out

    def set_field(self, name, values):
        v = np.ascontiguousarray(values, np.float32)
        r, c = self.level_dims()
        assert v.size == r * c
        self._chk(self.lib.pdf_set_field(self.ctx, FIELDS[name], _fptr(v)))

    def get_pyramid(self, kind, level):
        n = (self.width * self.height) >> (2 * level)
        out = np.empty(n, np.float32)
        self._chk(self.lib.pdf_get_pyramid(self.ctx, PYR_KINDS[kind], level, _fptr(out)))
        return out

    def motion_field(self):
        r, c = self.level_dims()
        dx, dy, dz = (np.empty(r * c, np.float32) for _ in range(3))
        self._chk(self.lib.pdf_copy_motion_field(self.ctx, _fptr(dx), _fptr(dy), _fptr(dz)))
        return dx, dy, dz

    def stream_frame_device(self, dI, dZ, iters=None, graph=True):
        """Streaming: one new frame (device buffers); returns True when a solve ran."""
        arr = None if iters is None else (ctypes.c_uint * len(iters))(*iters)
        solved = ctypes.c_int(0)
        self._chk(self.lib.pdf_stream_frame_device(self.ctx, dI.ptr, dZ.ptr, arr, 1 if graph else 0,
                                                   ctypes.byref(solved)))
        return bool(solved.value)

    def solve_warm(self, start_level, warm_device_ptr, iters=None):
        """Solve from `start_level` initialised by the 3 device planes at warm_device_ptr."""
        arr = None if iters is None else (ctypes.c_uint * len(iters))(*iters)
        self._chk(self.lib.pdf_solve_warm(self.ctx, arr, start_level, ctypes.c_void_p(warm_device_ptr)))

    def solve_pair_device(self, bufs, iters=None, graph=False):
        """One pair from four device buffers (I1, Z1, I2, Z2); flow stays on the device."""
        arr = None if iters is None else (ctypes.c_uint * len(iters))(*iters)
        self._chk(self.lib.pdf_solve_pair_device(self.ctx, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                                                 bufs[3].ptr, arr, 1 if graph else 0))

    def copy_all_solutions(self):
        """copyAllSolutions (pdflow_cudalib.cu:224-237): the level's flow and both frames'
        depth, colour, xx, yy planes, as a dict of host arrays."""
        r, c = self.level_dims()
        names = ("dx", "dy", "dz", "depth", "depth_old", "colour", "colour_old", "xx", "xx_old", "yy", "yy_old")
        out = {n: np.empty(r * c, np.float32) for n in names}
        self._chk(self.lib.pdf_copy_all_solutions(self.ctx, *[_fptr(out[n]) for n in names]))
        return out

    def solve_pair(self, I1, Z1, I2, Z2, iters=None, graph=False):
        """The pair evaluator end to end from host planes: returns (dx, dy, dz)."""
        self.copy_new_frames(I1, Z1)
        self.gaussian_pyramid()
        self.copy_new_frames(I2, Z2)
        self.gaussian_pyramid()
        self.solve(iters, graph=graph)
        return self.motion_field()

    def enable_kernel_timing(self, on=True):
        self._chk(self.lib.pdf_enable_kernel_timing(self.ctx, 1 if on else 0))

    def kernel_timing(self):
        """(launches, total ms, pixel-iterations, pixels covered) of the timed iteration launches"""
        n, ms, pix, pixl = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
        self._chk(self.lib.pdf_kernel_timing(self.ctx, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(pix),
                                             ctypes.byref(pixl)))
        return n.value, ms.value, pix.value, pixl.value


def solve_pair_noalloc(bufs, outs, rows=120, cam_mode=4, iters=None, stream=None, **overrides):
    """One pair through pdf_solve_pair_noalloc: four device buffers in (I1, Z1, I2, Z2), the flow
    written to three device buffers of rows x cols floats (dx, dy, dz); no device allocation,
    working planes in the library's per-GPU array.  Asynchronous on `stream` (None: the null
    stream)."""
    p = default_params(rows, cam_mode)
    p.update(overrides)
    lib = hip_library()
    mask = np.ascontiguousarray(p["g_mask"], np.float32)
    arr = None if iters is None else (ctypes.c_uint * len(iters))(*iters)
    rc = lib.pdf_solve_pair_noalloc(p["rows"], p["cols"], p["lambda_i"], p["lambda_d"], p["mu"], _fptr(mask),
                                    p["ctf_levels"], p["cam_mode"], p["fovh"],
                                    *[b.ptr for b in bufs], arr, *[o.ptr for o in outs], stream)
    if rc:
        raise PDFlowError(rc, lib.pdf_last_create_error().decode())


class Batch:
    """A batched ragged solve: many pairs of mixed size solved by shared launches."""

    def __init__(self, rows_list, iters=None, cam_modes=None, **overrides):
        self.lib = hip_library()
        p = default_params(480, 1)
        p.update(overrides)
        self.rows = list(rows_list)
        n = len(self.rows)
        rows_arr = (ctypes.c_uint * n)(*self.rows)
        cam_arr = None if cam_modes is None else (ctypes.c_uint * n)(*cam_modes)
        self.cam_modes = list(cam_modes) if cam_modes is not None else [480 // r for r in self.rows]
        it_arr = None
        if iters is not None:
            if len(iters) != n:
                raise ValueError(f"one schedule per pair expected ({n}), got {len(iters)}")
            flat = []
            for sched in iters:
                flat += list(sched) + [0] * (6 - len(sched))
            it_arr = (ctypes.c_uint * len(flat))(*flat)
        mask = np.ascontiguousarray(p["g_mask"], np.float32)
        self.ctx = ctypes.c_void_p()
        rc = self.lib.pdf_batch_create(ctypes.byref(self.ctx), n, rows_arr, cam_arr, p["lambda_i"],
                                       p["lambda_d"], p["mu"], _fptr(mask), p["fovh"], it_arr)
        if rc:
            raise PDFlowError(rc, self.lib.pdf_batch_last_error(None).decode())

    def _chk(self, rc):
        if rc:
            raise PDFlowError(rc, self.lib.pdf_batch_last_error(self.ctx).decode())

    def close(self):
        if self.ctx:
            self.lib.pdf_batch_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_iteration_kernel(self, k, min_pixels=0, level_kernel=True):
        self._chk(self.lib.pdf_batch_set_iteration_kernel(self.ctx, k, min_pixels, int(level_kernel)))

    def set_tb_shape(self, shape):
        self._chk(self.lib.pdf_batch_set_tb_shape(self.ctx, shape))

    SETUP_TILES = ((32, 8), (16, 16), (8, 32), (32, 4), (16, 8), (8, 16), (8, 8))

    def set_setup_tile(self, tile):
        """tile of the LDS-staged set-up stencils (level-init, median filter): an index into
        SETUP_TILES (rows x columns, one pixel per thread)"""
        self._chk(self.lib.pdf_batch_set_setup_tile(self.ctx, tile))

    def setup_tile(self):
        t = ctypes.c_int()
        self._chk(self.lib.pdf_batch_setup_tile(self.ctx, ctypes.byref(t)))
        return t.value

    def autotune(self, cache_path=None):
        """picks the fastest region shape and iterations per launch for this batch, then the
        fastest set-up tile (see pdf_batch_autotune; setup_tile() gives it); returns (shape, k,
        ms per solve)"""
        shape, k, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        self._chk(self.lib.pdf_batch_autotune(self.ctx, str(cache_path).encode() if cache_path else None,
                                              ctypes.byref(shape), ctypes.byref(k), ctypes.byref(ms)))
        return shape.value, k.value, ms.value

    def set_inputs(self, pair, I1, Z1, I2, Z2):
        a = [np.ascontiguousarray(x, np.float32) for x in (I1, Z1, I2, Z2)]
        self._chk(self.lib.pdf_batch_set_inputs(self.ctx, pair, *[_fptr(x) for x in a]))

    def solve(self, graph=True):
        self._chk(self.lib.pdf_batch_solve(self.ctx, 1 if graph else 0))

    def synchronize(self):
        self._chk(self.lib.pdf_batch_synchronize(self.ctx))

    def wait_for(self, other):
        """work queued on this batch from now on starts only once everything queued on `other`
        so far has finished (no host wait)"""
        self._chk(self.lib.pdf_batch_wait(self.ctx, other.ctx))

    def upload_inputs_async(self, host):
        """queues the copy of every pair's I1, Z1, I2, Z2 (pair after pair) from the float32 array
        `host` (best a PinnedBuffer's) on the batch's stream; `host` must stay unchanged until the
        stream has passed the copy"""
        self._chk(self.lib.pdf_batch_upload_inputs_async(self.ctx, _fptr(host), host.size))

    def download_flows_async(self, host):
        """queues the copy of every pair's dx, dy, dz (pair after pair) into the float32 array
        `host` on the batch's stream"""
        self._chk(self.lib.pdf_batch_download_flows_async(self.ctx, _fptr(host), host.size))

    def flow(self, pair):
        r = self.rows[pair]
        c = r * 4 // 3
        out = [np.empty(r * c, np.float32) for _ in range(3)]
        self._chk(self.lib.pdf_batch_copy_flow(self.ctx, pair, *[_fptr(x) for x in out]))
        return tuple(out)

    def enable_kernel_timing(self, on=True, tb_only=False):
        self._chk(self.lib.pdf_batch_enable_kernel_timing(self.ctx, (2 if tb_only else 1) if on else 0))

    def kernel_timing(self):
        """(launches, total ms, pixel-iterations, pixels covered) of the timed iteration launches"""
        n, ms, pix, pixl = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
        self._chk(self.lib.pdf_batch_kernel_timing(self.ctx, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(pix),
                                                   ctypes.byref(pixl)))
        return n.value, ms.value, pix.value, pixl.value

    def inputs_device(self, pair):
        ptrs = [ctypes.POINTER(ctypes.c_float)() for _ in range(4)]
        self._chk(self.lib.pdf_batch_inputs_device(self.ctx, pair, *[ctypes.byref(x) for x in ptrs]))
        return [ctypes.cast(x, ctypes.c_void_p).value for x in ptrs]

    def set_start_level(self, pair, level):
        """Warm-start pair at `level`; returns the device address of its du, dv, dw planes."""
        ptr = ctypes.POINTER(ctypes.c_float)()
        self._chk(self.lib.pdf_batch_set_start_level(self.ctx, pair, level, ctypes.byref(ptr)))
        return ctypes.cast(ptr, ctypes.c_void_p).value

    def set_schedule(self, pair, iters):
        arr = (ctypes.c_uint * 6)(*(list(iters) + [0] * (6 - len(iters))))
        self._chk(self.lib.pdf_batch_set_schedule(self.ctx, pair, arr))

    def launch_count(self):
        n = ctypes.c_uint()
        self._chk(self.lib.pdf_batch_launch_count(self.ctx, ctypes.byref(n)))
        return n.value

    def field_device(self, pair, name):
        ptr = ctypes.c_void_p()
        self._chk(self.lib.pdf_batch_field_device(self.ctx, pair, FIELDS[name], ctypes.byref(ptr)))
        return ptr.value

    def field(self, pair, name):
        """a finest-level plane of pair's last solve, copied to the host"""
        r = self.rows[pair]
        out = np.empty(r * (r * 4 // 3), np.float32)
        self.synchronize()
        rc = self.lib.pdf_memcpy_d2h(out.ctypes.data, self.field_device(pair, name), out.nbytes)
        if rc:
            raise PDFlowError(rc, "device-to-host copy")
        return out

    def pyramid(self, pair, kind, level):
        ptr, r, c = ctypes.c_void_p(), ctypes.c_uint(), ctypes.c_uint()
        self._chk(self.lib.pdf_batch_pyramid_device(self.ctx, pair, PYR_KINDS[kind], level, ctypes.byref(ptr),
                                                    ctypes.byref(r), ctypes.byref(c)))
        out = np.empty(r.value * c.value, np.float32)
        self.synchronize()
        rc = self.lib.pdf_memcpy_d2h(out.ctypes.data, ptr, out.nbytes)
        if rc:
            raise PDFlowError(rc, "device-to-host copy")
        return out, r.value, c.value

    def set_learned(self, net, pairs=None):
        """learned warm start for `pairs` (all if None) from the network `net` (None: off)"""
        arr = None
        n = 0
        if pairs is not None:
            n = len(pairs)
            arr = (ctypes.c_uint * max(n, 1))(*pairs)
        self._learned_net = net          # keeps the network alive while the batch uses it
        self._chk(self.lib.pdf_batch_set_learned(self.ctx, net.ctx if net is not None else None, arr, n))

    def metrics(self):
        n = len(self.rows)
        e = np.empty(n, np.float32)
        r = np.empty(n, np.float32)
        self._chk(self.lib.pdf_batch_metrics(self.ctx, _fptr(e), _fptr(r)))
        return e, r

    def quality(self, pairs=None):
        """(energy, residual, zero-motion residual) of the listed pairs (all if None), in order
        (pdf_batch_quality)"""
        idx = list(range(len(self.rows))) if pairs is None else [int(i) for i in pairs]
        n = len(idx)
        e, r, r0 = (np.zeros(n, np.float32) for _ in range(3))
        if n:
            arr = None if pairs is None else (ctypes.c_uint * n)(*idx)
            self._chk(self.lib.pdf_batch_quality(self.ctx, arr, 0 if pairs is None else n, _fptr(e), _fptr(r),
                                                 _fptr(r0)))
        return e, r, r0

    def fallback_bounds(self, energy_per_pixel_max, residual_ratio_max):
        """re-solves the pairs outside the bounds (pdf_batch_fallback_bounds); returns their flags"""
        flags = np.zeros(len(self.rows), np.uint8)
        cnt = ctypes.c_uint()
        self._chk(self.lib.pdf_batch_fallback_bounds(self.ctx, energy_per_pixel_max, residual_ratio_max,
                                                     flags.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                                                     ctypes.byref(cnt)))
        return flags

    def epe(self, ref):
        out = np.empty(len(self.rows), np.float32)
        self._chk(self.lib.pdf_batch_epe(self.ctx, ref.ctx, _fptr(out)))
        return out

    def resolve_plain(self, pairs):
        n = len(pairs)
        arr = (ctypes.c_uint * max(n, 1))(*pairs)
        self._chk(self.lib.pdf_batch_resolve_plain(self.ctx, arr, n))

    def metrics_pairs(self, pairs):
        """(energy, residual) of the listed pairs only (pdf_batch_metrics_pairs)"""
        pairs = [int(i) for i in pairs]
        n = len(pairs)
        e = np.zeros(n, np.float32)
        r = np.zeros(n, np.float32)
        if n:
            arr = (ctypes.c_uint * n)(*pairs)
            self._chk(self.lib.pdf_batch_metrics_pairs(self.ctx, arr, n, _fptr(e), _fptr(r)))
        return e, r

    def continue_finest(self, pairs, extra_iters):
        """extra_iters more iterations of the listed pairs at their finest level, continuing from
        their current state, then the level's filter and motion field (pdf_batch_continue)"""
        pairs = [int(i) for i in pairs]
        n = len(pairs)
        arr = (ctypes.c_uint * max(n, 1))(*pairs)
        self._chk(self.lib.pdf_batch_continue(self.ctx, arr, n, int(extra_iters)))

    def fallback(self, energy_ref, residual_ref, tol=0.05):
        """re-solves the pairs failing the rule; returns their flags (uint8 per pair)"""
        n = len(self.rows)
        e = np.ascontiguousarray(energy_ref, np.float32)
        r = np.ascontiguousarray(residual_ref, np.float32)
        flags = np.zeros(n, np.uint8)
        cnt = ctypes.c_uint()
        self._chk(self.lib.pdf_batch_fallback(self.ctx, _fptr(e), _fptr(r), tol,
                                              flags.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                                              ctypes.byref(cnt)))
        return flags

    def learn_samples(self, h, w, x_ptr, t_ptr, m_ptr):
        self._chk(self.lib.pdf_batch_learn_samples(self.ctx, h, w, ctypes.c_void_p(x_ptr), ctypes.c_void_p(t_ptr),
                                                   ctypes.c_void_p(m_ptr)))
