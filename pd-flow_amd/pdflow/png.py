"""ctypes binding of the PNG codec (include/pdflow_png.h, libpdflow_png.so)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libpdflow_png.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(path)
        up = ctypes.POINTER(ctypes.c_uint)
        lib.pdf_png_read.argtypes = [ctypes.c_char_p, up, up, up, up, ctypes.POINTER(ctypes.c_void_p)]
        lib.pdf_png_free.argtypes = [ctypes.c_void_p]
        lib.pdf_png_free.restype = None
        lib.pdf_png_write.argtypes = [ctypes.c_char_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                      ctypes.c_void_p]
        fp = ctypes.POINTER(ctypes.c_float)
        lib.pdf_load_frame.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint, ctypes.c_uint, fp, fp]
        lib.pdf_png_size.argtypes = [ctypes.c_char_p, up, up]
        _lib = lib
    return _lib


def read(path):
    """decoded samples: uint8 or uint16 array (h, w) for one channel, (h, w, c) otherwise"""
    lib = _load()
    w, h, c, d = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
    ptr = ctypes.c_void_p()
    rc = lib.pdf_png_read(str(path).encode(), ctypes.byref(w), ctypes.byref(h), ctypes.byref(c), ctypes.byref(d),
                          ctypes.byref(ptr))
    if rc:
        raise OSError(f"pdf_png_read({path}) failed with code {rc}")
    dt = np.uint16 if d.value == 16 else np.uint8
    n = w.value * h.value * c.value
    arr = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint16 if d.value == 16 else ctypes.c_uint8)),
                                shape=(n,)).astype(dt).copy()
    lib.pdf_png_free(ptr)
    return arr.reshape(h.value, w.value) if c.value == 1 else arr.reshape(h.value, w.value, c.value)


def write(path, img):
    """writes a uint8/uint16 array (h, w) or (h, w, 3)"""
    a = np.ascontiguousarray(img)
    if a.dtype not in (np.uint8, np.uint16):
        raise TypeError("uint8 or uint16 samples expected")
    c = 1 if a.ndim == 2 else a.shape[2]
    rc = _load().pdf_png_write(str(path).encode(), a.shape[1], a.shape[0], c, 16 if a.dtype == np.uint16 else 8,
                               a.ctypes.data)
    if rc:
        raise OSError(f"pdf_png_write({path}) failed with code {rc}")


def load_frame(intensity_path, depth_path, width, height):
    """solver planes (column-major I, Z) as the reference's loadRGBDFrames makes them"""
    I = np.empty(width * height, np.float32)
    Z = np.empty(width * height, np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    rc = _load().pdf_load_frame(str(intensity_path).encode(), str(depth_path).encode(), width, height,
                                I.ctypes.data_as(fp), Z.ctypes.data_as(fp))
    if rc:
        raise OSError(f"pdf_load_frame failed with code {rc}")
    return I, Z
