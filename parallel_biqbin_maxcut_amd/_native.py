"""ctypes binding of libbiqbin_gpu.so (the HIP bounding library).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared library has not been built."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbiqbin_gpu.so")
_LIB = None

dptr = C.POINTER(C.c_double)
iptr = C.POINTER(C.c_int)


class NativeLibraryMissing(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc for gfx950)")
        L = C.CDLL(LIB_PATH)
        L.bqb_init.argtypes = [C.c_int]
        L.bqb_init.restype = C.c_int
        L.bqb_last_error.restype = C.c_char_p
        L.ipm_mc_pk.argtypes = [dptr, C.c_int, dptr, dptr, C.c_int]
        L.ipm_mc_pk.restype = None
        L.bqb_ipm_batch.argtypes = [C.c_int, iptr, dptr, dptr, dptr, iptr, iptr, iptr, dptr, dptr, C.c_int]
        L.bqb_ipm_batch.restype = C.c_int
        _declare_extra(L)
        _LIB = L
    return _LIB


def _declare_extra(L):
    """Signatures of the remaining entry points (filled in by api.py)."""
    from . import api
    api.declare(L)


def check(rc):
    if rc != 0:
        raise RuntimeError("libbiqbin_gpu: " + lib().bqb_last_error().decode())
    return rc
