"""Minimal device-memory and event helper over libamdhip64 (ctypes), the HIP runtime the
library itself links, for callers of the device-pointer entry points (the cfd_nn_* operations,
the legacy functions): tests and the benchmark drive the C-ABI exactly as a C caller would."""
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
        L.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        L.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        L.hipEventDestroy.argtypes = [ctypes.c_void_p]
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


def time_null_stream(f, reps):
    """Mean device ms of f() (work issued on the NULL stream) over `reps` calls, between two
    HIP events recorded on the NULL stream."""
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip().hipEventCreate(ctypes.byref(e0)) == 0 and hip().hipEventCreate(ctypes.byref(e1)) == 0
    try:
        assert hip().hipEventRecord(e0, None) == 0
        for _ in range(reps):
            f()
        assert hip().hipEventRecord(e1, None) == 0
        assert hip().hipEventSynchronize(e1) == 0
        ms = ctypes.c_float(0)
        assert hip().hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        return ms.value / reps
    finally:
        hip().hipEventDestroy(e0)
        hip().hipEventDestroy(e1)
