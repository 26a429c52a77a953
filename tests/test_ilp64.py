"""The ILP64 library (libopenblas_amd_ilp64.so: blasint and lapack_int are
64-bit, the reference's INTERFACE64=1 build, common.h:269-278 and
lapacke.h's LAPACK_ILP64).

CPU: it exports the same symbols as the LP64 library, and its argument
checks read all 64 bits of every integer: a size of 2^32 (low word 0) or
-2^32 must be rejected with the xerbla number the LP64 library gives for the
same error at a small size, where a 32-bit read of the low word would see
another error, or a valid 0 and return quietly.

GPU: the same calls through both libraries give bit-identical results,
with 64-bit pivot vectors (IPIV, ITER) in host and in device memory.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, byref, c_char, c_char_p, c_double, c_int, c_long, c_void_p

import numpy as np
import pytest

import openblas_amd as oa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIG = 1 << 32

_pl = POINTER(c_long)
_pd = c_void_p
_SIGS64 = {
    "cblas_dgemm": (None, [c_int, c_int, c_int, c_long, c_long, c_long, c_double, _pd, c_long, _pd, c_long, c_double,
                           _pd, c_long]),
    "cblas_dtrsm": (None, [c_int, c_int, c_int, c_int, c_int, c_long, c_long, c_double, _pd, c_long, _pd, c_long]),
    "cblas_dsyrk": (None, [c_int, c_int, c_int, c_long, c_long, c_double, _pd, c_long, c_double, _pd, c_long]),
    "cblas_dgemm_batch": (None, [c_int, POINTER(c_int), POINTER(c_int), _pl, _pl, _pl, POINTER(c_double),
                                 POINTER(c_void_p), _pl, POINTER(c_void_p), _pl, POINTER(c_double), POINTER(c_void_p),
                                 _pl, c_long, _pl]),
    "dgemm_": (None, [c_char_p, c_char_p, _pl, _pl, _pl, POINTER(c_double), _pd, _pl, _pd, _pl, POINTER(c_double),
                      _pd, _pl]),
    "dgetrf_": (c_int, [_pl, _pl, _pd, _pl, c_void_p, _pl]),
    "dgetrs_": (c_int, [c_char_p, _pl, _pl, _pd, _pl, c_void_p, _pd, _pl, _pl]),
    "dgesv_": (c_int, [_pl, _pl, _pd, _pl, c_void_p, _pd, _pl, _pl]),
    "dpotrf_": (c_int, [c_char_p, _pl, _pd, _pl, _pl]),
    "dlaswp_": (c_int, [_pl, _pd, _pl, _pl, _pl, c_void_p, _pl]),
    "dsgesv_": (c_int, [_pl, _pl, _pd, _pl, c_void_p, _pd, _pl, _pd, _pl, c_void_p, c_void_p, _pl, _pl]),
    "LAPACKE_dgesv": (c_long, [c_int, c_long, c_long, _pd, c_long, c_void_p, _pd, c_long]),
    "LAPACKE_dgetrf": (c_long, [c_int, c_long, c_long, _pd, c_long, c_void_p]),
    "LAPACKE_dsgesv": (c_long, [c_int, c_long, c_long, _pd, c_long, c_void_p, _pd, c_long, _pd, c_long, _pl]),
    "openblas_amd_last_xerbla": (c_int, [c_char_p, c_int]),
    "openblas_amd_device_count": (c_int, []),
}


@pytest.fixture(scope="module")
def L64():
    if not os.path.exists(oa.ILP64_LIB_PATH):
        subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "openblas_amd", "csrc")], check=True)
    L = ctypes.CDLL(oa.ILP64_LIB_PATH, mode=ctypes.RTLD_LOCAL)
    for name, (res, args) in _SIGS64.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def lp(v):
    return byref(c_long(int(v)))


def xerbla64(L):
    buf = ctypes.create_string_buffer(32)
    info = L.openblas_amd_last_xerbla(buf, 32)
    return buf.value.decode().strip(), info


def _defined(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    return {n for n in names if not n.startswith("__hip_cuid_")}   # per-object compiler ids


def test_ilp64_exports_the_lp64_symbol_set(L64):
    assert _defined(oa.ILP64_LIB_PATH) == _defined(oa.LIB_PATH)


def test_ilp64_reads_all_64_bits(L64):
    L = oa.lib()
    A = np.zeros(16)
    piv = np.zeros(4, np.int64)
    info = c_long(99)
    # the LP64 library's number for the same error at a small size
    L.dgesv_(oa.ip(3), oa.ip(1), A.ctypes.data, oa.ip(1), piv.ctypes.data, A.ctypes.data, oa.ip(3),
             byref(c_int(0)))
    want = oa.last_xerbla()
    assert want == ("DGESV", 4)
    L64.dgesv_(lp(BIG), lp(1), A.ctypes.data, lp(1), piv.ctypes.data, A.ctypes.data, lp(BIG), byref(info))
    assert info.value == -4 and xerbla64(L64) == want
    L64.dgesv_(lp(-BIG), lp(0), A.ctypes.data, lp(1), piv.ctypes.data, A.ctypes.data, lp(1), byref(info))
    assert info.value == -1 and xerbla64(L64) == ("DGESV", 1)
    L64.dgetrf_(lp(BIG), lp(0), A.ctypes.data, lp(1), piv.ctypes.data, byref(info))
    assert info.value == -4 and xerbla64(L64) == ("DGETRF", 4)
    L64.dpotrf_(b"L", lp(BIG + 3), A.ctypes.data, lp(3), byref(info))
    assert info.value == -4 and xerbla64(L64) == ("DPOTRF", 4)
    it = c_long(5)
    L64.dsgesv_(lp(2), lp(1), A.ctypes.data, lp(2), piv.ctypes.data, A.ctypes.data, lp(2), A.ctypes.data,
                lp(2 - BIG), None, None, byref(it), byref(info))
    assert info.value == -9 and it.value == 0 and xerbla64(L64) == ("DSGESV", 9)
    # getrs reports through xerbla and leaves INFO untouched
    info.value = 77
    L64.dgetrs_(b"N", lp(BIG), lp(1), A.ctypes.data, lp(1), piv.ctypes.data, A.ctypes.data, lp(BIG), byref(info))
    assert info.value == 77 and xerbla64(L64) == ("DGETRS", 5)
    # CBLAS by-value sizes: column-major ldc < M
    L.cblas_dgemm(102, 111, 111, 5, 1, 1, 1.0, A.ctypes.data, 5, A.ctypes.data, 1, 0.0, A.ctypes.data, 1)
    want = oa.last_xerbla()
    L64.cblas_dgemm(102, 111, 111, BIG, 1, 1, 1.0, A.ctypes.data, BIG, A.ctypes.data, 1, 0.0, A.ctypes.data, 1)
    assert xerbla64(L64) == want and want[0] == "DGEMM"
    L.cblas_dtrsm(102, 141, 122, 111, 131, 5, 1, 1.0, A.ctypes.data, 1, A.ctypes.data, 5)
    want = oa.last_xerbla()
    L64.cblas_dtrsm(102, 141, 122, 111, 131, BIG, 1, 1.0, A.ctypes.data, 1, A.ctypes.data, BIG)
    assert xerbla64(L64) == want and want[0] == "DTRSM"
    L.cblas_dsyrk(102, 122, 111, -1, 1, 1.0, A.ctypes.data, 1, 0.0, A.ctypes.data, 1)
    want = oa.last_xerbla()
    L64.cblas_dsyrk(102, 122, 111, -BIG, 1, 1.0, A.ctypes.data, 1, 0.0, A.ctypes.data, 1)
    assert xerbla64(L64) == want and want[0] == "DSYRK"
    # LAPACKE: 64-bit sizes and return value
    assert L64.LAPACKE_dgetrf(101, 1, BIG, A.ctypes.data, 1, piv.ctypes.data) == -5
    assert L64.LAPACKE_dgesv(101, 2, BIG + 1, A.ctypes.data, 2, piv.ctypes.data, A.ctypes.data, 1) == -8


@pytest.mark.gpu
def test_ilp64_matches_lp64_bitwise(oba, L64):
    L = oba.lib()
    assert L64.openblas_amd_device_count() >= 1
    rng = np.random.default_rng(64)
    n, nrhs = 700, 3
    A0 = np.asfortranarray(rng.standard_normal((n, n)))
    B0 = np.asfortranarray(rng.standard_normal((n, nrhs)))

    # dgesv_ with host pivots
    A32, B32, p32, i32 = A0.copy(order="F"), B0.copy(order="F"), np.zeros(n, np.int32), c_int(-9)
    L.dgesv_(oba.ip(n), oba.ip(nrhs), A32.ctypes.data, oba.ip(n), p32.ctypes.data, B32.ctypes.data, oba.ip(n),
             byref(i32))
    A64, B64, p64, i64 = A0.copy(order="F"), B0.copy(order="F"), np.zeros(n, np.int64), c_long(-9)
    L64.dgesv_(lp(n), lp(nrhs), A64.ctypes.data, lp(n), p64.ctypes.data, B64.ctypes.data, lp(n), byref(i64))
    assert i32.value == 0 and i64.value == 0
    assert np.array_equal(A32, A64) and np.array_equal(B32, B64) and np.array_equal(p32, p64)

    # dgetrf_ / dgetrs_ with 64-bit pivots in device memory
    dA = oba.DeviceBuffer.from_array(A0)
    dP = oba.DeviceBuffer(8 * n)
    L64.dgetrf_(lp(n), lp(n), dA.ptr, lp(n), dP.ptr, byref(i64))
    assert i64.value == 0
    pdev = dP.to_array(np.zeros(n, np.int64))
    assert np.array_equal(pdev, p32)
    assert np.array_equal(dA.to_array(np.empty_like(A0)), A32)
    Bt = B0.copy(order="F")
    L64.dgetrs_(b"T", lp(n), lp(nrhs), dA.ptr, lp(n), dP.ptr, Bt.ctypes.data, lp(n), byref(i64))
    Bt32 = B0.copy(order="F")
    L.dgetrs_(b"T", oba.ip(n), oba.ip(nrhs), A32.ctypes.data, oba.ip(n), p32.ctypes.data, Bt32.ctypes.data,
              oba.ip(n), byref(i32))
    assert np.array_equal(Bt, Bt32)

    # dlaswp_ with 64-bit pivots (forward and backward increments)
    for inc in (1, -1):
        W32, W64 = B0.copy(order="F"), B0.copy(order="F")
        L.dlaswp_(oba.ip(nrhs), W32.ctypes.data, oba.ip(n), oba.ip(1), oba.ip(n), p32.ctypes.data, oba.ip(inc))
        L64.dlaswp_(lp(nrhs), W64.ctypes.data, lp(n), lp(1), lp(n), p64.ctypes.data, lp(inc))
        assert np.array_equal(W32, W64), inc

    # dsgesv_ (IPIV and ITER 64-bit) and LAPACKE_dsgesv, row-major
    Ad = np.asfortranarray(A0 + np.sqrt(n) * np.eye(n))
    for lib, pt, it_t, ipf in ((L, np.int32, c_int, oba.ip), (L64, np.int64, c_long, lp)):
        Aw, X, pv, it, inf = Ad.copy(order="F"), np.zeros((n, nrhs), order="F"), np.zeros(n, pt), it_t(0), it_t(0)
        lib.dsgesv_(ipf(n), ipf(nrhs), Aw.ctypes.data, ipf(n), pv.ctypes.data, B0.ctypes.data, ipf(n), X.ctypes.data,
                    ipf(n), None, None, byref(it), byref(inf))
        if lib is L:
            ref = (X.copy(), pv.astype(np.int64), it.value, inf.value)
        else:
            assert inf.value == 0 and it.value > 0
            assert np.array_equal(X, ref[0]) and np.array_equal(pv, ref[1]) and it.value == ref[2]
    Ar, Br = np.ascontiguousarray(Ad), np.ascontiguousarray(B0)
    X32, X64 = np.zeros((n, nrhs)), np.zeros((n, nrhs))
    pv32, pv64, it32, it64 = np.zeros(n, np.int32), np.zeros(n, np.int64), c_int(0), c_long(0)
    Ar32, Ar64 = Ar.copy(), Ar.copy()
    r32 = L.LAPACKE_dsgesv(101, n, nrhs, Ar32.ctypes.data, n, pv32.ctypes.data, Br.ctypes.data, nrhs,
                           X32.ctypes.data, nrhs, byref(it32))
    r64 = L64.LAPACKE_dsgesv(101, n, nrhs, Ar64.ctypes.data, n, pv64.ctypes.data, Br.ctypes.data, nrhs,
                             X64.ctypes.data, nrhs, byref(it64))
    assert r32 == 0 and r64 == 0 and it32.value == it64.value
    assert np.array_equal(X32, X64) and np.array_equal(pv32, pv64)

    # dpotrf_
    S = np.asfortranarray(A0 @ A0.T + n * np.eye(n))
    S32, S64 = S.copy(order="F"), S.copy(order="F")
    L.dpotrf_(b"L", oba.ip(n), S32.ctypes.data, oba.ip(n), byref(i32))
    L64.dpotrf_(b"L", lp(n), S64.ctypes.data, lp(n), byref(i64))
    assert i32.value == 0 and i64.value == 0 and np.array_equal(S32, S64)

    # GEMM: CBLAS by value, Fortran by reference, and the group-batched form
    m, nn, k = 333, 517, 290
    Ag = np.asfortranarray(rng.standard_normal((m, k)))
    Bg = np.asfortranarray(rng.standard_normal((k, nn)))
    Cg = np.asfortranarray(rng.standard_normal((m, nn)))
    C32, C64 = Cg.copy(order="F"), Cg.copy(order="F")
    L.cblas_dgemm(102, 111, 111, m, nn, k, 1.5, Ag.ctypes.data, m, Bg.ctypes.data, k, -0.5, C32.ctypes.data, m)
    L64.cblas_dgemm(102, 111, 111, m, nn, k, 1.5, Ag.ctypes.data, m, Bg.ctypes.data, k, -0.5, C64.ctypes.data, m)
    assert np.array_equal(C32, C64)
    C64f = Cg.copy(order="F")
    L64.dgemm_(b"N", b"N", lp(m), lp(nn), lp(k), byref(c_double(1.5)), Ag.ctypes.data, lp(m), Bg.ctypes.data, lp(k),
               byref(c_double(-0.5)), C64f.ctypes.data, lp(m))
    assert np.array_equal(C32, C64f)
    # the group-batched form (its own engine: compared with the LP64 batch call)
    Cb, Cb32 = Cg.copy(order="F"), Cg.copy(order="F")
    i64a = lambda *v: (c_long * len(v))(*v)  # noqa: E731
    i32a = lambda *v: (c_int * len(v))(*v)  # noqa: E731
    L64.cblas_dgemm_batch(102, (c_int * 1)(111), (c_int * 1)(111), i64a(m), i64a(nn), i64a(k), (c_double * 1)(1.5),
                          (c_void_p * 1)(Ag.ctypes.data), i64a(m), (c_void_p * 1)(Bg.ctypes.data), i64a(k),
                          (c_double * 1)(-0.5), (c_void_p * 1)(Cb.ctypes.data), i64a(m), 1, i64a(1))
    L.cblas_dgemm_batch(102, i32a(111), i32a(111), i32a(m), i32a(nn), i32a(k), (c_double * 1)(1.5),
                        (c_void_p * 1)(Ag.ctypes.data), i32a(m), (c_void_p * 1)(Bg.ctypes.data), i32a(k),
                        (c_double * 1)(-0.5), (c_void_p * 1)(Cb32.ctypes.data), i32a(m), 1, i32a(1))
    assert np.array_equal(Cb32, Cb)
    assert np.abs(Cb - C32).max() <= 64 * k * np.finfo(np.float64).eps * (1.5 * np.abs(Ag) @ np.abs(Bg)
                                                                           + 0.5 * np.abs(Cg)).max()
    # TRSM
    T = np.asfortranarray(np.tril(rng.standard_normal((m, m))) + 4 * np.eye(m))
    X32, X64 = Cg.copy(order="F"), Cg.copy(order="F")
    L.cblas_dtrsm(102, 141, 122, 111, 131, m, nn, 2.0, T.ctypes.data, m, X32.ctypes.data, m)
    L64.cblas_dtrsm(102, 141, 122, 111, 131, m, nn, 2.0, T.ctypes.data, m, X64.ctypes.data, m)
    assert np.array_equal(X32, X64)
    dA.free()
    dP.free()
