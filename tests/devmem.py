"""Device buffers for GPU tests that hand device pointers to the library,
allocated through the HIP runtime the library itself loaded
(libamdhip64.so.7), so that a test needs no second runtime (torch's ROCm
build carries its own)."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from fastearth3d_amd import lib
        lib.load()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DeviceArray:
    """A float64 array copied to device memory; .ptr(k) is the address of
    element k of the flattened array."""

    def __init__(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.shape, self.n = a.shape, a.size
        h = hip()
        p = C.c_void_p()
        if h.hipMalloc(C.byref(p), a.nbytes) != 0:
            raise MemoryError("hipMalloc of %d bytes failed" % a.nbytes)
        self._p = p.value
        if h.hipMemcpy(self._p, a.ctypes.data, a.nbytes, 1) != 0:   # hipMemcpyHostToDevice
            raise RuntimeError("hipMemcpy to the device failed")

    def ptr(self, k=0):
        return self._p + 8 * k

    def free(self):
        if self._p:
            hip().hipFree(self._p)
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
