"""Device buffers for tests that call entry points taking device pointers.

Allocated through the HIP runtime that libogsx.so itself links
(libamdhip64.so.7, loaded by ctypes under its soname, so the process keeps
one HIP runtime); test infrastructure only."""
import ctypes as C

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        import ogs5_amd
        ogs5_amd.lib()  # the library's runtime is the one loaded below
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipMalloc.restype = C.c_int
        h.hipFree.argtypes = [C.c_void_p]
        h.hipFree.restype = C.c_int
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpy.restype = C.c_int
        h.hipDeviceSynchronize.restype = C.c_int
        _hip = h
    return _hip


class DevArray:
    """A device copy of a numpy array (same dtype and length)."""

    def __init__(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self.dtype, self.size = a.dtype, a.size
        p = C.c_void_p()
        assert hip().hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0, "hipMalloc failed"
        self.ptr = p.value
        assert hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0, "hipMemcpy failed"

    def get(self) -> np.ndarray:
        out = np.empty(self.size, self.dtype)
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
