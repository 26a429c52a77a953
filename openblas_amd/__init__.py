"""Python view of libopenblas_amd, the gfx950 (MI355X) implementation of
OpenBLAS's level-3 hot path (GEMM, TRSM, getrf/getrs/gesv, potrf, dsgesv).

The product is the C-ABI shared library ``openblas_amd/libopenblas_amd.so``
(declared in ``include/openblas_amd.h``); this module only loads it with
ctypes, declares the argument types of every exported symbol, and adds a
small ``DeviceBuffer`` helper for callers that keep matrices resident in HBM.
There is no CPU fallback: if the library is missing, ``lib()`` raises.
"""
import ctypes
import os
from ctypes import POINTER, c_char, c_char_p, c_double, c_float, c_int, c_long, c_size_t, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libopenblas_amd.so")
# the same entry points with 64-bit blasint / lapack_int (ILP64)
ILP64_LIB_PATH = os.path.join(_HERE, "libopenblas_amd_ilp64.so")

# ABI enum values (cblas.h:62-66)
RowMajor, ColMajor = 101, 102
NoTrans, Trans, ConjTrans, ConjNoTrans = 111, 112, 113, 114
Upper, Lower = 121, 122
NonUnit, Unit = 131, 132
Left, Right = 141, 142

_pi = POINTER(c_int)
_pd = c_void_p  # matrices are passed as raw addresses (host or device)

# name -> (restype, argtypes)
_SIGS = {
    "cblas_dgemm": (None, [c_int, c_int, c_int, c_int, c_int, c_int, c_double, _pd, c_int, _pd, c_int, c_double, _pd, c_int]),
    "cblas_sgemm": (None, [c_int, c_int, c_int, c_int, c_int, c_int, c_float, _pd, c_int, _pd, c_int, c_float, _pd, c_int]),
    "cblas_dtrsm": (None, [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_double, _pd, c_int, _pd, c_int]),
    "cblas_strsm": (None, [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_float, _pd, c_int, _pd, c_int]),
    "cblas_dgemm_batch": (None, [c_int, _pi, _pi, _pi, _pi, _pi, POINTER(c_double), POINTER(c_void_p), _pi,
                                 POINTER(c_void_p), _pi, POINTER(c_double), POINTER(c_void_p), _pi, c_int, _pi]),
    "cblas_sgemm_batch": (None, [c_int, _pi, _pi, _pi, _pi, _pi, POINTER(c_float), POINTER(c_void_p), _pi,
                                 POINTER(c_void_p), _pi, POINTER(c_float), POINTER(c_void_p), _pi, c_int, _pi]),
    "cblas_dgemm_batch_strided": (None, [c_int, c_int, c_int, c_int, c_int, c_int, c_double, _pd, c_int, c_long, _pd,
                                         c_int, c_long, c_double, _pd, c_int, c_long, c_int]),
    "cblas_sgemm_batch_strided": (None, [c_int, c_int, c_int, c_int, c_int, c_int, c_float, _pd, c_int, c_long, _pd,
                                         c_int, c_long, c_float, _pd, c_int, c_long, c_int]),
    "dgemm_": (None, [c_char_p, c_char_p, _pi, _pi, _pi, POINTER(c_double), _pd, _pi, _pd, _pi, POINTER(c_double), _pd, _pi]),
    "sgemm_": (None, [c_char_p, c_char_p, _pi, _pi, _pi, POINTER(c_float), _pd, _pi, _pd, _pi, POINTER(c_float), _pd, _pi]),
    "dtrsm_": (None, [c_char_p, c_char_p, c_char_p, c_char_p, _pi, _pi, POINTER(c_double), _pd, _pi, _pd, _pi]),
    "strsm_": (None, [c_char_p, c_char_p, c_char_p, c_char_p, _pi, _pi, POINTER(c_float), _pd, _pi, _pd, _pi]),
    "cblas_dsyrk": (None, [c_int, c_int, c_int, c_int, c_int, c_double, _pd, c_int, c_double, _pd, c_int]),
    "cblas_ssyrk": (None, [c_int, c_int, c_int, c_int, c_int, c_float, _pd, c_int, c_float, _pd, c_int]),
    "cblas_dgemmt": (None, [c_int, c_int, c_int, c_int, c_int, c_int, c_double, _pd, c_int, _pd, c_int, c_double, _pd, c_int]),
    "cblas_sgemmt": (None, [c_int, c_int, c_int, c_int, c_int, c_int, c_float, _pd, c_int, _pd, c_int, c_float, _pd, c_int]),
    "cblas_dsymm": (None, [c_int, c_int, c_int, c_int, c_int, c_double, _pd, c_int, _pd, c_int, c_double, _pd, c_int]),
    "cblas_ssymm": (None, [c_int, c_int, c_int, c_int, c_int, c_float, _pd, c_int, _pd, c_int, c_float, _pd, c_int]),
    "cblas_dtrmm": (None, [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_double, _pd, c_int, _pd, c_int]),
    "cblas_strmm": (None, [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_float, _pd, c_int, _pd, c_int]),
    "dsyrk_": (None, [c_char_p, c_char_p, _pi, _pi, POINTER(c_double), _pd, _pi, POINTER(c_double), _pd, _pi]),
    "ssyrk_": (None, [c_char_p, c_char_p, _pi, _pi, POINTER(c_float), _pd, _pi, POINTER(c_float), _pd, _pi]),
    "dgemmt_": (None, [c_char_p, c_char_p, c_char_p, _pi, _pi, POINTER(c_double), _pd, _pi, _pd, _pi, POINTER(c_double), _pd, _pi]),
    "sgemmt_": (None, [c_char_p, c_char_p, c_char_p, _pi, _pi, POINTER(c_float), _pd, _pi, _pd, _pi, POINTER(c_float), _pd, _pi]),
    "dsymm_": (None, [c_char_p, c_char_p, _pi, _pi, POINTER(c_double), _pd, _pi, _pd, _pi, POINTER(c_double), _pd, _pi]),
    "ssymm_": (None, [c_char_p, c_char_p, _pi, _pi, POINTER(c_float), _pd, _pi, _pd, _pi, POINTER(c_float), _pd, _pi]),
    "dtrmm_": (None, [c_char_p, c_char_p, c_char_p, c_char_p, _pi, _pi, POINTER(c_double), _pd, _pi, _pd, _pi]),
    "strmm_": (None, [c_char_p, c_char_p, c_char_p, c_char_p, _pi, _pi, POINTER(c_float), _pd, _pi, _pd, _pi]),
    "dgetrf_": (c_int, [_pi, _pi, _pd, _pi, c_void_p, _pi]),
    "sgetrf_": (c_int, [_pi, _pi, _pd, _pi, c_void_p, _pi]),
    "dgetf2_": (c_int, [_pi, _pi, _pd, _pi, c_void_p, _pi]),
    "sgetf2_": (c_int, [_pi, _pi, _pd, _pi, c_void_p, _pi]),
    "dgetrs_": (c_int, [c_char_p, _pi, _pi, _pd, _pi, c_void_p, _pd, _pi, _pi]),
    "openblas_amd_dgetrs_noalloc": (None, [c_char_p, _pi, _pi, _pd, _pi, c_void_p, _pd, _pi, _pi]),
    "sgetrs_": (c_int, [c_char_p, _pi, _pi, _pd, _pi, c_void_p, _pd, _pi, _pi]),
    "dgesv_": (c_int, [_pi, _pi, _pd, _pi, c_void_p, _pd, _pi, _pi]),
    "sgesv_": (c_int, [_pi, _pi, _pd, _pi, c_void_p, _pd, _pi, _pi]),
    "dpotrf_": (c_int, [c_char_p, _pi, _pd, _pi, _pi]),
    "spotrf_": (c_int, [c_char_p, _pi, _pd, _pi, _pi]),
    "dpotf2_": (c_int, [c_char_p, _pi, _pd, _pi, _pi]),
    "spotf2_": (c_int, [c_char_p, _pi, _pd, _pi, _pi]),
    "dlaswp_": (c_int, [_pi, _pd, _pi, _pi, _pi, c_void_p, _pi]),
    "slaswp_": (c_int, [_pi, _pd, _pi, _pi, _pi, c_void_p, _pi]),
    "dsgesv_": (c_int, [_pi, _pi, _pd, _pi, c_void_p, _pd, _pi, _pd, _pi, c_void_p, c_void_p, _pi, _pi]),
    "LAPACKE_dgesv": (c_int, [c_int, c_int, c_int, _pd, c_int, c_void_p, _pd, c_int]),
    "LAPACKE_dgesv_work": (c_int, [c_int, c_int, c_int, _pd, c_int, c_void_p, _pd, c_int]),
    "LAPACKE_sgesv": (c_int, [c_int, c_int, c_int, _pd, c_int, c_void_p, _pd, c_int]),
    "LAPACKE_sgesv_work": (c_int, [c_int, c_int, c_int, _pd, c_int, c_void_p, _pd, c_int]),
    "LAPACKE_dsgesv": (c_int, [c_int, c_int, c_int, _pd, c_int, c_void_p, _pd, c_int, _pd, c_int, _pi]),
    "LAPACKE_dsgesv_work": (c_int, [c_int, c_int, c_int, _pd, c_int, c_void_p, _pd, c_int, _pd, c_int, c_void_p,
                                    c_void_p, _pi]),
    "dsposv_": (c_int, [c_char_p, _pi, _pi, _pd, _pi, _pd, _pi, _pd, _pi, c_void_p, c_void_p, _pi, _pi]),
    "LAPACKE_dsposv": (c_int, [c_int, c_char, c_int, c_int, _pd, c_int, _pd, c_int, _pd, c_int, _pi]),
    "LAPACKE_dsposv_work": (c_int, [c_int, c_char, c_int, c_int, _pd, c_int, _pd, c_int, _pd, c_int, c_void_p,
                                    c_void_p, _pi]),
    "LAPACKE_dgetrf": (c_int, [c_int, c_int, c_int, _pd, c_int, c_void_p]),
    "LAPACKE_dgetrf_work": (c_int, [c_int, c_int, c_int, _pd, c_int, c_void_p]),
    "LAPACKE_sgetrf": (c_int, [c_int, c_int, c_int, _pd, c_int, c_void_p]),
    "LAPACKE_sgetrf_work": (c_int, [c_int, c_int, c_int, _pd, c_int, c_void_p]),
    "LAPACKE_dgetrs": (c_int, [c_int, c_char, c_int, c_int, _pd, c_int, c_void_p, _pd, c_int]),
    "LAPACKE_dgetrs_work": (c_int, [c_int, c_char, c_int, c_int, _pd, c_int, c_void_p, _pd, c_int]),
    "LAPACKE_sgetrs": (c_int, [c_int, c_char, c_int, c_int, _pd, c_int, c_void_p, _pd, c_int]),
    "LAPACKE_sgetrs_work": (c_int, [c_int, c_char, c_int, c_int, _pd, c_int, c_void_p, _pd, c_int]),
    "LAPACKE_dpotrf": (c_int, [c_int, c_char, c_int, _pd, c_int]),
    "LAPACKE_dpotrf_work": (c_int, [c_int, c_char, c_int, _pd, c_int]),
    "LAPACKE_spotrf": (c_int, [c_int, c_char, c_int, _pd, c_int]),
    "LAPACKE_spotrf_work": (c_int, [c_int, c_char, c_int, _pd, c_int]),
    "xerbla_": (c_int, [c_char_p, _pi, c_int]),
    "openblas_get_config": (c_char_p, []),
    "openblas_get_corename": (c_char_p, []),
    "openblas_set_num_threads": (None, [c_int]),
    "openblas_get_num_threads": (c_int, []),
    "openblas_get_num_procs": (c_int, []),
    "openblas_amd_dgemm_ozaki": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_double, _pd, c_int, _pd, c_int,
                                         c_double, _pd, c_int, c_int]),
    "openblas_amd_dgemm_ozaki2": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_double, _pd, c_int, _pd, c_int,
                                         c_double, _pd, c_int, c_int]),
    "openblas_amd_dgetrf_batch": (c_int, [_pi, _pi, POINTER(c_void_p), _pi, POINTER(c_void_p), _pi, c_int]),
    "openblas_amd_sgetrf_batch": (c_int, [_pi, _pi, POINTER(c_void_p), _pi, POINTER(c_void_p), _pi, c_int]),
    "openblas_amd_dgetrs_batch": (c_int, [c_char, _pi, _pi, POINTER(c_void_p), _pi, POINTER(c_void_p),
                                          POINTER(c_void_p), _pi, _pi, c_int]),
    "openblas_amd_sgetrs_batch": (c_int, [c_char, _pi, _pi, POINTER(c_void_p), _pi, POINTER(c_void_p),
                                          POINTER(c_void_p), _pi, _pi, c_int]),
    "openblas_amd_dpotrf_batch": (c_int, [c_char, _pi, POINTER(c_void_p), _pi, _pi, c_int]),
    "openblas_amd_spotrf_batch": (c_int, [c_char, _pi, POINTER(c_void_p), _pi, _pi, c_int]),
    "openblas_amd_malloc": (c_void_p, [c_size_t]),
    "openblas_amd_free": (None, [c_void_p]),
    "openblas_amd_memcpy": (c_int, [c_void_p, c_void_p, c_size_t, c_int]),
    "openblas_amd_memset": (c_int, [c_void_p, c_int, c_size_t]),
    "openblas_amd_synchronize": (c_int, []),
    "openblas_amd_set_sync": (None, [c_int]),
    "openblas_amd_device_count": (c_int, []),
    "openblas_amd_device_info": (c_int, [c_char_p, c_int, _pi, _pi]),
    "openblas_amd_timer_start": (None, []),
    "openblas_amd_timer_stop": (c_float, []),
    "openblas_amd_last_xerbla": (c_int, [c_char_p, c_int]),
    "openblas_amd_last_error": (c_int, [c_char_p, c_int]),
    "openblas_amd_ktimer_enable": (None, [c_int]),
    "openblas_amd_ktimer_read": (c_int, [c_char_p, POINTER(ctypes.c_long), POINTER(c_double), POINTER(c_double),
                                         POINTER(c_double)]),
    "openblas_amd_ktimer_reset": (None, []),
}

_lib = None


def lib():
    """Load libopenblas_amd.so (in-tree build) and declare its signatures."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(libopenblas_amd has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def ip(v):
    """ctypes pointer to a fresh int (Fortran by-reference scalar)."""
    return ctypes.byref(c_int(int(v)))


def last_xerbla():
    """(routine name, info) of the last argument error on this thread."""
    buf = ctypes.create_string_buffer(32)
    info = lib().openblas_amd_last_xerbla(buf, 32)
    return buf.value.decode().strip(), info


class DeviceBuffer:
    """An HBM allocation holding a numpy array's bytes (column-major data
    stays column-major: pass the same ld as on the host)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = lib().openblas_amd_malloc(max(self.nbytes, 1))
        if not self.ptr:
            raise MemoryError(f"openblas_amd_malloc({nbytes}) failed")

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a) if not (a.flags.f_contiguous or a.flags.c_contiguous) else a
        d = cls(a.nbytes)
        lib().openblas_amd_memcpy(d.ptr, a.ctypes.data, a.nbytes, 0)
        return d

    def to_array(self, out):
        lib().openblas_amd_memcpy(out.ctypes.data, self.ptr, out.nbytes, 1)
        return out

    def free(self):
        if self.ptr:
            lib().openblas_amd_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
