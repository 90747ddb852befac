"""ctypes binding of the warm-start network (C-ABI: include/pdflow_cnn.h)."""
import ctypes

import numpy as np

from .lib import DeviceBuffer, PDFlowError, hip_library


def _sig(lib):
    if getattr(lib, "_net_sigs", False):
        return lib
    vp, c_int, c_uint, c_float = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_float
    fp = ctypes.POINTER(ctypes.c_float)
    u16p = ctypes.POINTER(ctypes.c_uint16)
    sigs = {
        "pdf_net_create": ([ctypes.POINTER(vp), c_uint, c_uint, c_uint, ctypes.c_uint64], c_int),
        "pdf_net_destroy": ([vp], c_int),
        "pdf_net_last_error": ([vp], ctypes.c_char_p),
        "pdf_net_num_params": ([vp, ctypes.POINTER(ctypes.c_size_t)], c_int),
        "pdf_net_get_params": ([vp, fp], c_int),
        "pdf_net_set_params": ([vp, fp], c_int),
        "pdf_net_get_grads": ([vp, fp], c_int),
        "pdf_net_forward": ([vp, vp, c_uint, vp], c_int),
        "pdf_net_output_device": ([vp, ctypes.POINTER(vp)], c_int),
        "pdf_net_loss_backward": ([vp, vp, vp, c_uint, fp], c_int),
        "pdf_net_backward_from": ([vp, vp, c_uint], c_int),
        "pdf_net_adam_step": ([vp, c_float], c_int),
        "pdf_net_train_step": ([vp, vp, vp, vp, c_uint, c_float, fp], c_int),
        "pdf_net_synchronize": ([vp], c_int),
        "pdf_net_train_step_indexed": ([vp, vp, vp, vp, ctypes.c_size_t, ctypes.POINTER(c_uint), c_uint, c_float,
                                        fp], c_int),
        "pdf_net_set_stream": ([vp, vp], c_int),
        "pdf_net_save": ([vp, ctypes.c_char_p], c_int),
        "pdf_net_load": ([vp, ctypes.c_char_p], c_int),
        "pdf_net_mfma_probe": ([u16p, u16p, fp], c_int),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    lib._net_sigs = True
    return lib


def _f(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Net:
    """The warm-start U-Net on the GPU (h x w inputs, batches up to max_batch)."""

    def __init__(self, max_batch=64, h=60, w=80, seed=0):
        self.lib = _sig(hip_library())
        self.h, self.w, self.max_batch = h, w, max_batch
        self.ctx = ctypes.c_void_p()
        rc = self.lib.pdf_net_create(ctypes.byref(self.ctx), max_batch, h, w, seed)
        if rc:
            raise PDFlowError(rc, self.lib.pdf_net_last_error(None).decode())
        n = ctypes.c_size_t()
        self._chk(self.lib.pdf_net_num_params(self.ctx, ctypes.byref(n)))
        self.num_params = n.value

    def _chk(self, rc):
        if rc:
            raise PDFlowError(rc, self.lib.pdf_net_last_error(self.ctx).decode())

    def close(self):
        if self.ctx:
            self.lib.pdf_net_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.empty(self.num_params, np.float32)
        self._chk(self.lib.pdf_net_get_params(self.ctx, _f(out)))
        return out

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        assert flat.size == self.num_params
        self._chk(self.lib.pdf_net_set_params(self.ctx, _f(flat)))

    def grads(self):
        out = np.empty(self.num_params, np.float32)
        self._chk(self.lib.pdf_net_get_grads(self.ctx, _f(out)))
        return out

    def forward_ptr(self, x_ptr, batch, y_ptr=None):
        self._chk(self.lib.pdf_net_forward(self.ctx, ctypes.c_void_p(x_ptr), batch,
                                           ctypes.c_void_p(y_ptr) if y_ptr else None))

    def output_ptr(self):
        p = ctypes.c_void_p()
        self._chk(self.lib.pdf_net_output_device(self.ctx, ctypes.byref(p)))
        return p.value

    def forward(self, x):
        """x: [N, h, w, 4] float32 host array -> [N, h, w, 3]"""
        x = np.ascontiguousarray(x, np.float32)
        n = x.shape[0]
        bx = DeviceBuffer(x.nbytes)
        bx.upload(x)
        by = DeviceBuffer(n * self.h * self.w * 3 * 4)
        self.forward_ptr(bx.ptr.value, n, by.ptr.value)
        self.synchronize()
        out = np.empty((n, self.h, self.w, 3), np.float32)
        return by.download(out)

    def backward_from(self, x, dy):
        """forward on x, then gradients for an explicit output gradient dy; returns flat grads"""
        x = np.ascontiguousarray(x, np.float32)
        dy = np.ascontiguousarray(dy, np.float32)
        n = x.shape[0]
        bx, bd = DeviceBuffer(x.nbytes), DeviceBuffer(dy.nbytes)
        bx.upload(x)
        bd.upload(dy)
        self.forward_ptr(bx.ptr.value, n)
        self._chk(self.lib.pdf_net_backward_from(self.ctx, bd.ptr, n))
        self.synchronize()
        return self.grads()

    def train_step_ptr(self, x_ptr, t_ptr, m_ptr, batch, lr, want_loss=False):
        loss = ctypes.c_float(0.0)
        self._chk(self.lib.pdf_net_train_step(self.ctx, ctypes.c_void_p(x_ptr), ctypes.c_void_p(t_ptr),
                                              ctypes.c_void_p(m_ptr), batch, lr,
                                              ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def train_step_indexed(self, x_ptr, t_ptr, m_ptr, total, idx, lr, want_loss=False):
        """one Adam step on samples idx of a device-resident data set"""
        idx = np.ascontiguousarray(idx, np.uint32)
        loss = ctypes.c_float(0.0)
        self._chk(self.lib.pdf_net_train_step_indexed(
            self.ctx, ctypes.c_void_p(x_ptr), ctypes.c_void_p(t_ptr), ctypes.c_void_p(m_ptr), total,
            idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), idx.size, lr,
            ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def synchronize(self):
        self._chk(self.lib.pdf_net_synchronize(self.ctx))

    def save(self, path):
        self._chk(self.lib.pdf_net_save(self.ctx, str(path).encode()))

    def load(self, path):
        self._chk(self.lib.pdf_net_load(self.ctx, str(path).encode()))


def mfma_probe(A, B):
    """C = A @ B for 16x16 BF16 bit patterns (uint16) through the library's MFMA layout"""
    lib = _sig(hip_library())
    A = np.ascontiguousarray(A, np.uint16)
    B = np.ascontiguousarray(B, np.uint16)
    C = np.empty((16, 16), np.float32)
    rc = lib.pdf_net_mfma_probe(A.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                B.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), _f(C))
    if rc:
        raise PDFlowError(rc, "mfma probe")
    return C
