"""The reference library itself, as installed on this machine -- TEST AND
BASELINE INFRASTRUCTURE ONLY.

scipy ships OpenBLAS 0.3.28 (the reference, built by the scipy project) as
``scipy.libs/libscipy_openblas-*.so`` with ``scipy_``-prefixed LP64 symbols.
This module dlopens it and calls its exported entry points (run and timed
only, never inspected) for two purposes:
  * tests/golden/make_golden.py and the oracle tests pin the C oracle against
    the reference's own results;
  * bench.py's ``cpu_baseline`` leg times the reference CPU path on the host
    cores.
numpy's ILP64 ``libscipy_openblas64_`` (suffix ``64_``) is the fallback.
"""
import ctypes
import glob
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_float, c_int, c_void_p

import numpy as np

_lib = None
_prefix = None
_suffix = ""
_ilp64 = False


def _find():
    import scipy
    import numpy
    cands = []
    for mod, suf, ilp in ((scipy, "", False), (numpy, "64_", True)):
        d = os.path.join(os.path.dirname(os.path.dirname(mod.__file__)), mod.__name__ + ".libs")
        for p in sorted(glob.glob(os.path.join(d, "libscipy_openblas*.so"))):
            cands.append((p, suf, ilp))
    return cands


def lib():
    global _lib, _suffix, _ilp64
    if _lib is None:
        for path, suf, ilp in _find():
            if ("openblas64" in os.path.basename(path)) != ilp:
                continue
            try:
                L = ctypes.CDLL(path)
                getattr(L, "scipy_openblas_get_config" + suf)
            except (OSError, AttributeError):
                continue
            _lib, _suffix, _ilp64 = L, suf, ilp
            break
        if _lib is None:
            raise RuntimeError("no installed OpenBLAS (scipy.libs / numpy.libs) found")
    return _lib


def _fn(name):
    return getattr(lib(), "scipy_" + name + _suffix)


def _it():
    return ctypes.c_int64 if _ilp64 else c_int


def config():
    f = _fn("openblas_get_config")
    f.restype = c_char_p
    return f().decode()


def set_num_threads(n):
    _fn("openblas_set_num_threads")(_it()(n))


def _p(a):
    return c_void_p(a.ctypes.data)


def cblas_gemm(transa, transb, alpha, A, B, beta, C):
    """In place on Fortran-ordered C."""
    dt = C.dtype
    m, n = C.shape
    k = A.shape[0] if transa else A.shape[1]
    f = _fn("cblas_dgemm" if dt == np.float64 else "cblas_sgemm")
    I, F = _it(), (c_double if dt == np.float64 else c_float)
    f(I(102), I(112 if transa else 111), I(112 if transb else 111), I(m), I(n), I(k), F(alpha), _p(A),
      I(max(A.shape[0], 1)), _p(B), I(max(B.shape[0], 1)), F(beta), _p(C), I(max(m, 1)))
    return C


def cblas_trsm(side, uplo, transa, diag, alpha, A, B):
    """side 0 left; uplo 1 lower; diag 1 non-unit -- in place on B."""
    dt = B.dtype
    m, n = B.shape
    f = _fn("cblas_dtrsm" if dt == np.float64 else "cblas_strsm")
    I, F = _it(), (c_double if dt == np.float64 else c_float)
    f(I(102), I(142 if side else 141), I(122 if uplo else 121), I(112 if transa else 111), I(131 if diag else 132),
      I(m), I(n), F(alpha), _p(A), I(max(A.shape[0], 1)), _p(B), I(max(m, 1)))
    return B


def getrf(A):
    """In place; returns (ipiv int32, info)."""
    I = _it()
    m, n = A.shape
    ipiv = np.zeros(max(min(m, n), 1), np.int64 if _ilp64 else np.int32)
    info = I(0)
    _fn("dgetrf_" if A.dtype == np.float64 else "sgetrf_")(byref(I(m)), byref(I(n)), _p(A), byref(I(max(m, 1))),
                                                           _p(ipiv), byref(info))
    return ipiv[:min(m, n)].astype(np.int32), info.value


def getrs(trans, LU, ipiv, B):
    I = _it()
    n, nrhs = B.shape
    piv = np.ascontiguousarray(ipiv, dtype=np.int64 if _ilp64 else np.int32)
    info = I(0)
    _fn("dgetrs_" if B.dtype == np.float64 else "sgetrs_")(c_char_p(b"T" if trans else b"N"), byref(I(n)),
                                                           byref(I(nrhs)), _p(LU), byref(I(max(n, 1))), _p(piv), _p(B),
                                                           byref(I(max(n, 1))), byref(info))
    return info.value


def gesv(A, B):
    I = _it()
    n, nrhs = B.shape
    ipiv = np.zeros(max(n, 1), np.int64 if _ilp64 else np.int32)
    info = I(0)
    _fn("dgesv_" if A.dtype == np.float64 else "sgesv_")(byref(I(n)), byref(I(nrhs)), _p(A), byref(I(max(n, 1))),
                                                         _p(ipiv), _p(B), byref(I(max(n, 1))), byref(info))
    return ipiv[:n].astype(np.int32), info.value


def potrf(uplo, A):
    I = _it()
    n = A.shape[0]
    info = I(0)
    _fn("dpotrf_" if A.dtype == np.float64 else "spotrf_")(c_char_p(b"L" if uplo else b"U"), byref(I(n)), _p(A),
                                                           byref(I(max(n, 1))), byref(info))
    return info.value


def dsgesv(A, B):
    """Returns (X, ipiv, iter, info); A is overwritten only on fallback."""
    I = _it()
    n, nrhs = B.shape
    X = np.zeros((n, nrhs), order="F")
    ipiv = np.zeros(max(n, 1), np.int64 if _ilp64 else np.int32)
    work = np.zeros(max(n * nrhs, 1))
    swork = np.zeros(max(n * (n + nrhs), 1), np.float32)
    it, info = I(0), I(0)
    _fn("dsgesv_")(byref(I(n)), byref(I(nrhs)), _p(A), byref(I(max(n, 1))), _p(ipiv), _p(B), byref(I(max(n, 1))),
                   _p(X), byref(I(max(n, 1))), _p(work), _p(swork), byref(it), byref(info))
    return X, ipiv[:n].astype(np.int32), it.value, info.value


def dsposv(uplo, A, B):
    """uplo 1 lower / 0 upper.  Returns (X, iter, info); A is overwritten
    (by its FP64 factor) only on the fallback."""
    I = _it()
    n, nrhs = B.shape
    X = np.zeros((n, nrhs), order="F")
    work = np.zeros(max(n * nrhs, 1))
    swork = np.zeros(max(n * (n + nrhs), 1), np.float32)
    it, info = I(0), I(0)
    _fn("dsposv_")(c_char_p(b"L" if uplo else b"U"), byref(I(n)), byref(I(nrhs)), _p(A), byref(I(max(n, 1))), _p(B),
                   byref(I(max(n, 1))), _p(X), byref(I(max(n, 1))), _p(work), _p(swork), byref(it), byref(info))
    return X, it.value, info.value


def lapacke_potrf(layout, uplo, A):
    """LAPACKE_dpotrf(layout, uplo, n, a, lda) on the memory of A (a
    Fortran-ordered array; for row-major the caller passes the transpose).
    Returns the LAPACKE info."""
    I = _it()
    n = A.shape[0]
    f = _fn("LAPACKE_dpotrf")
    f.restype = I
    return int(f(c_int(layout), ctypes.c_char(uplo), I(n), _p(A), I(max(n, 1))))


def dgemm_batch(As, Bs, Cs, alpha=1.0, beta=0.0):
    """cblas_dgemm_batch with one group per problem (NN, column-major).
    The installed 0.3.28 build crashes (SIGSEGV) when a problem has
    m*n*k <= 100^3 (the small-matrix kernel route of
    interface/gemm_batch.c:300-318): call it with larger problems only."""
    I = _it()
    cnt = len(As)
    it = np.int64 if _ilp64 else np.int32
    m = np.array([c.shape[0] for c in Cs], it)
    n = np.array([c.shape[1] for c in Cs], it)
    k = np.array([a.shape[1] for a in As], it)
    ta = np.full(cnt, 111, it)
    al = np.full(cnt, alpha)
    be = np.full(cnt, beta)
    pa = (c_void_p * cnt)(*[a.ctypes.data for a in As])
    pb = (c_void_p * cnt)(*[b.ctypes.data for b in Bs])
    pc = (c_void_p * cnt)(*[c.ctypes.data for c in Cs])
    gs = np.ones(cnt, it)
    _fn("cblas_dgemm_batch")(I(102), _p(ta), _p(ta), _p(m), _p(n), _p(k), _p(al), pa, _p(m), pb, _p(k), _p(be), pc,
                             _p(m), I(cnt), _p(gs))


def _ld(M, layout):
    """Leading dimension of a 2-D array in the given CBLAS layout (its
    memory must be Fortran-ordered for 102, C-ordered for 101)."""
    return max(M.shape[0] if layout == 102 else M.shape[1], 1)


def cblas_syrk(layout, uplo, trans, alpha, A, beta, C):
    """uplo 121/122, trans 111/112; in place on C (n x n)."""
    dt = C.dtype
    I, F = _it(), (c_double if dt == np.float64 else c_float)
    n = C.shape[0]
    k = A.shape[0] if trans == 112 else A.shape[1]
    _fn("cblas_dsyrk" if dt == np.float64 else "cblas_ssyrk")(I(layout), I(uplo), I(trans), I(n), I(k), F(alpha),
                                                             _p(A), I(_ld(A, layout)), F(beta), _p(C),
                                                             I(_ld(C, layout)))
    return C


def cblas_gemmt(layout, uplo, transa, transb, alpha, A, B, beta, C):
    dt = C.dtype
    I, F = _it(), (c_double if dt == np.float64 else c_float)
    m = C.shape[0]
    k = A.shape[0] if transa == 112 else A.shape[1]
    _fn("cblas_dgemmt" if dt == np.float64 else "cblas_sgemmt")(I(layout), I(uplo), I(transa), I(transb), I(m), I(k),
                                                               F(alpha), _p(A), I(_ld(A, layout)), _p(B),
                                                               I(_ld(B, layout)), F(beta), _p(C), I(_ld(C, layout)))
    return C


def cblas_symm(layout, side, uplo, alpha, A, B, beta, C):
    dt = C.dtype
    I, F = _it(), (c_double if dt == np.float64 else c_float)
    m, n = C.shape
    _fn("cblas_dsymm" if dt == np.float64 else "cblas_ssymm")(I(layout), I(side), I(uplo), I(m), I(n), F(alpha), _p(A),
                                                             I(_ld(A, layout)), _p(B), I(_ld(B, layout)), F(beta),
                                                             _p(C), I(_ld(C, layout)))
    return C


def cblas_trmm(layout, side, uplo, transa, diag, alpha, A, B):
    dt = B.dtype
    I, F = _it(), (c_double if dt == np.float64 else c_float)
    m, n = B.shape
    _fn("cblas_dtrmm" if dt == np.float64 else "cblas_strmm")(I(layout), I(side), I(uplo), I(transa), I(diag), I(m),
                                                             I(n), F(alpha), _p(A), I(_ld(A, layout)), _p(B),
                                                             I(_ld(B, layout)))
    return B
