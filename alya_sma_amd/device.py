"""Device arrays over the library's own HIP allocator (sma_device_malloc).

The Python host uses these instead of a framework's tensors so that the
process holds exactly one HIP runtime: the one libalya_sma_amd.so links.
"""
import ctypes as C

import numpy as np

from ._lib import check, lib


def device_count():
    n = C.c_int32()
    check(lib().sma_device_count(C.byref(n)), "sma_device_count")
    return n.value


def synchronize():
    check(lib().sma_device_synchronize(), "sma_device_synchronize")


class DeviceArray:
    """A contiguous device buffer with a numpy shape and dtype."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().sma_device_malloc(C.byref(p), self.nbytes), "sma_device_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        check(lib().sma_memcpy(C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), d.nbytes, 0),
              "sma_memcpy h2d")
        return d

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().sma_memcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, 1),
              "sma_memcpy d2h")
        return out

    def rows(self, start, stop):
        """Device pointer of rows [start, stop) of a 2-D array (no copy)."""
        return self.ptr + int(start) * int(np.prod(self.shape[1:], dtype=np.int64)) * self.dtype.itemsize

    def free(self):
        if getattr(self, "ptr", None):
            lib().sma_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
