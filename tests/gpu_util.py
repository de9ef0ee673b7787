"""Device buffers for GPU tests, managed through the library's own C ABI helpers."""
import ctypes as C

import numpy as np

from diploma_thesis_amd import capi


class Dev:
    """A device allocation holding a copy of a numpy array's shape and dtype."""

    def __init__(self, arr=None, shape=None, dtype=None):
        if arr is not None:
            arr = np.ascontiguousarray(arr)
            shape, dtype = arr.shape, arr.dtype
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        capi.check(capi.lib.arsim_malloc(C.byref(self.ptr), max(self.nbytes, 4)), "malloc")
        if arr is not None and self.nbytes:
            capi.check(capi.lib.arsim_memcpy_htod(self.ptr, capi.ptr(arr), self.nbytes, None), "htod")

    def get(self):
        out = np.zeros(self.shape, self.dtype)
        if self.nbytes:
            capi.check(capi.lib.arsim_memcpy_dtoh(capi.ptr(out), self.ptr, self.nbytes, None), "dtoh")
        return out

    def set(self, arr):
        arr = np.ascontiguousarray(arr, self.dtype)
        assert arr.nbytes == self.nbytes
        capi.check(capi.lib.arsim_memcpy_htod(self.ptr, capi.ptr(arr), self.nbytes, None), "htod")

    def free(self):
        if self.ptr:
            capi.lib.arsim_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def p(self):
        return self.ptr


def zeros(shape, dtype):
    return Dev(np.zeros(shape, dtype))


def sync():
    capi.check(capi.lib.arsim_stream_synchronize(None), "sync")


def rel_err_vs_scale(a, b, scale):
    """max |a - b| / max(|b|, scale) per row (vectors in the rows)."""
    num = np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64), axis=-1)
    den = np.maximum(np.linalg.norm(np.asarray(b, np.float64), axis=-1), scale)
    return float(np.max(num / den)) if num.size else 0.0
