"""CPU oracle for the level-3 hot path -- TEST INFRASTRUCTURE ONLY.

ctypes view of oracle/liboracle.so (plain C restatement of the reference's
algorithms, see oracle.c for the file:line each function follows).  Only
tests/, the smoke check and bench.py's CPU-baseline leg use it, always as the
checker; the product (openblas_amd/) never imports it.

Each function takes and returns numpy arrays (column-major copies are made
as needed) so tests read like the reference's own testers.
"""
import ctypes
import os
import subprocess
from ctypes import c_double, c_float, c_int, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = c_void_p
        sig = {
            "or_dgemm": (None, [c_int, c_int, c_int, c_int, c_int, c_double, P, c_int, P, c_int, c_double, P, c_int]),
            "or_sgemm": (None, [c_int, c_int, c_int, c_int, c_int, c_float, P, c_int, P, c_int, c_float, P, c_int]),
            "or_dtrsm": (None, [c_int, c_int, c_int, c_int, c_int, c_int, c_double, P, c_int, P, c_int]),
            "or_strsm": (None, [c_int, c_int, c_int, c_int, c_int, c_int, c_float, P, c_int, P, c_int]),
            "or_dgetf2": (c_int, [c_int, c_int, P, c_int, P]),
            "or_sgetf2": (c_int, [c_int, c_int, P, c_int, P]),
            "or_dlaswp": (None, [c_int, P, c_int, c_int, c_int, P, c_int]),
            "or_slaswp": (None, [c_int, P, c_int, c_int, c_int, P, c_int]),
            "or_dgetrs": (None, [c_int, c_int, c_int, P, c_int, P, P, c_int]),
            "or_sgetrs": (None, [c_int, c_int, c_int, P, c_int, P, P, c_int]),
            "or_dpotf2": (c_int, [c_int, c_int, P, c_int]),
            "or_spotf2": (c_int, [c_int, c_int, P, c_int]),
            "or_dsgesv": (c_int, [c_int, c_int, P, c_int, P, P, c_int, P, c_int, P]),
            "or_dsposv": (c_int, [c_int, c_int, c_int, P, c_int, P, c_int, P, c_int, P]),
            "or_dpotrs": (None, [c_int, c_int, c_int, P, c_int, P, c_int]),
            "or_dmmch_ratio": (c_double, [c_int, c_int, c_int, c_int, c_int, c_double, P, c_int, P, c_int, c_double,
                                          P, c_int, P, c_int, c_double]),
            "or_dbeg": (c_double, [c_int]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _f(a, dt):
    return np.array(a, dtype=dt, order="F", copy=True)


def _p(a):
    return a.ctypes.data


def gemm(transa, transb, alpha, A, B, beta, C):
    """C <- alpha*op(A)*op(B) + beta*C (returns a new array)."""
    dt = np.float64 if np.asarray(C).dtype == np.float64 else np.float32
    A, B, C = _f(A, dt), _f(B, dt), _f(C, dt)
    m, n = C.shape
    k = A.shape[0] if transa else A.shape[1]
    f = lib().or_dgemm if dt == np.float64 else lib().or_sgemm
    f(int(transa), int(transb), m, n, k, alpha, _p(A), max(A.shape[0], 1), _p(B), max(B.shape[0], 1), beta, _p(C),
      max(m, 1))
    return C


def trsm(side, uplo, transa, diag, alpha, A, B):
    """side 0 left / 1 right, uplo 1 lower / 0 upper, diag 1 non-unit / 0 unit."""
    dt = np.float64 if np.asarray(B).dtype == np.float64 else np.float32
    A, B = _f(A, dt), _f(B, dt)
    m, n = B.shape
    f = lib().or_dtrsm if dt == np.float64 else lib().or_strsm
    f(side, uplo, transa, diag, m, n, alpha, _p(A), max(A.shape[0], 1), _p(B), max(m, 1))
    return B


def getf2(A):
    """Returns (LU, ipiv 1-based int32, info)."""
    dt = np.float64 if np.asarray(A).dtype == np.float64 else np.float32
    A = _f(A, dt)
    m, n = A.shape
    ipiv = np.zeros(max(min(m, n), 1), np.int32)
    f = lib().or_dgetf2 if dt == np.float64 else lib().or_sgetf2
    info = f(m, n, _p(A), max(m, 1), _p(ipiv))
    return A, ipiv[:min(m, n)], info


def laswp(A, k1, k2, ipiv, incx):
    dt = np.float64 if np.asarray(A).dtype == np.float64 else np.float32
    A = _f(A, dt)
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    f = lib().or_dlaswp if dt == np.float64 else lib().or_slaswp
    f(A.shape[1], _p(A), max(A.shape[0], 1), k1, k2, _p(ipiv), incx)
    return A


def getrs(trans, LU, ipiv, B):
    dt = np.float64 if np.asarray(B).dtype == np.float64 else np.float32
    LU, B = _f(LU, dt), _f(B, dt)
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    n, nrhs = B.shape
    f = lib().or_dgetrs if dt == np.float64 else lib().or_sgetrs
    f(int(trans), n, nrhs, _p(LU), max(n, 1), _p(ipiv), _p(B), max(n, 1))
    return B


def potf2(uplo, A):
    """uplo 1 lower / 0 upper.  Returns (factor, info)."""
    dt = np.float64 if np.asarray(A).dtype == np.float64 else np.float32
    A = _f(A, dt)
    f = lib().or_dpotf2 if dt == np.float64 else lib().or_spotf2
    info = f(uplo, A.shape[0], _p(A), max(A.shape[0], 1))
    return A, info


def dsgesv(A, B):
    """Returns (A_out, ipiv, X, iter, info)."""
    A, B = _f(A, np.float64), _f(B, np.float64)
    n, nrhs = B.shape
    X = np.zeros((n, nrhs), order="F")
    ipiv = np.zeros(max(n, 1), np.int32)
    it = c_int(0)
    info = lib().or_dsgesv(n, nrhs, _p(A), max(n, 1), _p(ipiv), _p(B), max(n, 1), _p(X), max(n, 1),
                           ctypes.addressof(it))
    return A, ipiv[:n], X, it.value, info


def dsposv(uplo, A, B):
    """uplo 1 lower / 0 upper.  Returns (A_out, X, iter, info)."""
    A, B = _f(A, np.float64), _f(B, np.float64)
    n, nrhs = B.shape
    X = np.zeros((n, nrhs), order="F")
    it = c_int(0)
    info = lib().or_dsposv(int(uplo), n, nrhs, _p(A), max(n, 1), _p(B), max(n, 1), _p(X), max(n, 1),
                           ctypes.addressof(it))
    return A, X, it.value, info


def potrs(uplo, F, B):
    """Solve with a Cholesky factor held in the uplo triangle of F (FP64)."""
    F, B = _f(F, np.float64), _f(B, np.float64)
    n, nrhs = B.shape
    lib().or_dpotrs(int(uplo), n, nrhs, _p(F), max(n, 1), _p(B), max(n, 1))
    return B


def dmmch_ratio(transa, transb, alpha, A, B, beta, C_in, C_out, eps=None):
    A, B, C_in, C_out = (_f(x, np.float64) for x in (A, B, C_in, C_out))
    m, n = C_in.shape
    k = A.shape[0] if transa else A.shape[1]
    if eps is None:
        eps = np.finfo(np.float64).eps
    return lib().or_dmmch_ratio(int(transa), int(transb), m, n, k, alpha, _p(A), max(A.shape[0], 1), _p(B),
                                max(B.shape[0], 1), beta, _p(C_in), max(m, 1), _p(C_out), max(m, 1), eps)


def dbeg(reset=False):
    return lib().or_dbeg(1 if reset else 0)


# ---- Ozaki scheme II (no reference routine; the cblas_dgemm contract) ---------
# The published scheme (Ozaki, Uchino, Imamura, "Ozaki Scheme II: a
# GEMM-oriented emulation of floating-point matrix multiplication using an
# integer modular technique", 2025) restated by its definition rather than by
# its modular arithmetic: the operands are truncated to integers, the integer
# product is exact, and the result is rounded once.  The moduli and the
# Chinese remainder theorem are only the device's way of forming that exact
# product, so the oracle forms it with Python integers.
CRT_MODULI = (256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197,
              193, 191, 181, 179, 173, 167, 163, 157, 151, 149, 139, 137, 131, 127, 113)


def crt_log2P(nm):
    """log2 of the product of the first nm moduli, as the library rounds it
    (the product rounded to FP64, then log2)."""
    P = 1
    for p in CRT_MODULI[:nm]:
        P *= p
    return float(np.log2(float(P)))


def crt_bits(nm, K, spread_a=0.0, spread_b=0.0):
    """(tA, tB) for nm moduli: the integer bits of each operand such that
    K * 2^(tA+tB) <= P/2 (the balanced range of the CRT), split by the
    spreads (equal spreads: tA = floor(total/2))."""
    import math
    kbits = math.ceil(math.log2(K)) + 1.0
    total = math.floor(crt_log2P(nm) - kbits - 1e-9)
    tA = math.floor((total + spread_a - spread_b) / 2.0)
    tA = max(1, min(total - 1, tA))
    return tA, total - tA


def ozaki2(A, B, tA, tB):
    """op(A) (M x K) times op(B) (K x N) by scheme II with tA / tB integer
    bits: a'_ik = trunc(a_ik 2^(tA - e_i)), e_i the frexp exponent of the
    row's largest |a| (0 for a zero row), b'_kj likewise per column; C' =
    A'B' exactly (Python integers); C = C' 2^(e_i + f_j - tA - tB), rounded
    once (Python's int -> float conversion rounds to nearest, ties to even).
    For small sizes (pure Python)."""
    import math
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    M, K = A.shape
    N = B.shape[1]

    def ex(v):
        m = float(np.max(np.abs(v))) if v.size else 0.0
        return math.frexp(m)[1] if m > 0 else 0

    ea = [ex(A[i]) for i in range(M)]
    fb = [ex(B[:, j]) for j in range(N)]
    from fractions import Fraction

    def trunc_scaled(x, sh):   # trunc(x 2^sh), exact (int() of a Fraction truncates toward zero)
        return int(Fraction(x) * (Fraction(2) ** sh))
    Ai = [[trunc_scaled(float(A[i, k]), tA - ea[i]) for k in range(K)] for i in range(M)]
    Bj = [[trunc_scaled(float(B[k, j]), tB - fb[j]) for k in range(K)] for j in range(N)]
    C = np.empty((M, N))
    for i in range(M):
        for j in range(N):
            s = sum(x * y for x, y in zip(Ai[i], Bj[j]))
            sh = ea[i] + fb[j] - tA - tB
            # exact scaling of an integer by 2^sh, one rounding
            C[i, j] = float(s * (1 << sh)) if sh >= 0 else _int_ratio_to_float(s, 1 << (-sh))
    return C


def _int_ratio_to_float(n, d):
    """n / d (d a power of two) correctly rounded to FP64."""
    from fractions import Fraction
    return float(Fraction(n, d))
