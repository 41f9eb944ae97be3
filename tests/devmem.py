"""Device buffers for tests that call device-pointer entry points of the
library, through the same HIP runtime the library is linked against
(/opt/rocm/lib/libamdhip64.so.7; a second HIP runtime in the process, such
as torch's own copy, does not see the device once the library holds it)."""
import ctypes as C

import numpy as np

_HIP = None
H2D, D2H = 1, 2


def hip():
    global _HIP
    if _HIP is None:
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        _HIP = h
    return _HIP


def _ck(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: hip error {rc}")


class DevArray:
    """A device copy of a numpy array (same dtype and size)."""

    def __init__(self, host=None, like=None, fill_byte=0):
        a = np.ascontiguousarray(host if host is not None else like)
        self.dtype, self.size, self.nbytes = a.dtype, a.size, a.nbytes
        self.ptr = C.c_void_p()
        _ck(hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 16)), "hipMalloc")
        if host is not None:
            _ck(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")
        else:
            _ck(hip().hipMemset(self.ptr, fill_byte, self.nbytes), "hipMemset")

    def get(self):
        out = np.empty(self.size, dtype=self.dtype)
        _ck(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _ck(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()
