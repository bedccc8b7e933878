"""Device buffers for tests that call C-ABI entry points taking device
pointers (hipMalloc / hipMemcpy / hipFree of the HIP runtime, via ctypes)."""
import ctypes as C
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        L = C.CDLL(path)
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipSetDevice.argtypes = [C.c_int]
        _hip = L
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A device allocation holding `count` elements of `dtype`."""

    def __init__(self, count, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = max(self.count * self.dtype.itemsize, 1)
        p = C.c_void_p()
        assert hip().hipMalloc(C.byref(p), self.nbytes) == 0, "hipMalloc failed"
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size:
            assert hip().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, H2D) == 0
        return d

    def host(self, count=None):
        n = self.count if count is None else int(count)
        out = np.zeros(n, self.dtype)
        if n:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, n * self.dtype.itemsize, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()
