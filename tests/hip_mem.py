"""Device buffers for tests that hand device pointers to the plugin, through
the HIP runtime's C API (ctypes; no torch import)."""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        L = C.CDLL("libamdhip64.so")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipSetDevice.argtypes = [C.c_int]
        _hip = L
    return _hip


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: hipError {rc}")


class DeviceArray:
    """nbytes of device memory on the current device, zero-filled."""

    def __init__(self, nbytes: int):
        self.nbytes = nbytes
        p = C.c_void_p()
        check(hip().hipMalloc(C.byref(p), nbytes), "hipMalloc")
        self.ptr = p.value
        check(hip().hipMemset(self.ptr, 0, nbytes), "hipMemset")
        check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")

    def to_numpy(self, dtype=np.float32) -> np.ndarray:
        check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2), "hipMemcpy D2H")   # hipMemcpyDeviceToHost
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()
