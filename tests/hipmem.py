"""Minimal device-memory helper over libamdhip64 (ctypes) for tests of the device-pointer
entry points. No torch: the tests drive the C-ABI exactly as a C caller would."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        L = ctypes.CDLL("libamdhip64.so")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip = L
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A device copy of a numpy array (same shape and dtype: float64 unless the host array is
    float32 or uint16)."""

    def __init__(self, host):
        host = np.asarray(host)
        self.dtype = host.dtype if host.dtype in (np.float32, np.uint16) else np.dtype(np.float64)
        self.shape = host.shape
        self.nbytes = int(np.prod(host.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 8)) == 0
        self.ptr = p.value
        self.put(host)

    def put(self, host):
        h = np.ascontiguousarray(host, dtype=self.dtype)
        assert h.nbytes == self.nbytes
        assert hip().hipMemcpy(self.ptr, h.ctypes.data, self.nbytes, H2D) == 0

    def get(self):
        out = np.zeros(self.shape, dtype=self.dtype)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


def device_count():
    n = ctypes.c_int(0)
    if hip().hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value
