"""GPU parity tests for the LAPACK entry points on the hot path (?getrf,
?getrs, ?gesv, ?potrf, ?laswp, dsgesv and their LAPACKE wrappers), called
through the C ABI and checked against the oracle and the reference
library's golden vectors.

Bars: ipiv and info bit-exact on the tie-free fixtures; factors within
c*n*eps of the reference's; solves with backward error <= c*n*eps64
(LAPACK dget02-style ratio); dsgesv ITER equal to the reference's.
"""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps}


def c_int():
    return ctypes.c_int(-99)


@pytest.mark.parametrize("tag", ["d", "s"])
def test_getrf_golden_vectors(oba, golden, tag):
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    f = L.dgetrf_ if tag == "d" else L.sgetrf_
    fs = L.dgetrs_ if tag == "d" else L.sgetrs_
    j = 0
    while f"getrf{tag}_{j}_A" in golden.files:
        key = f"getrf{tag}_{j}"
        A = np.array(golden[key + "_A"], order="F")
        m, n = A.shape
        piv = np.zeros(min(m, n), np.int32)
        info = c_int()
        f(oba.ip(m), oba.ip(n), A.ctypes.data, oba.ip(m), piv.ctypes.data, ctypes.byref(info))
        assert info.value == int(golden[key + "_info"][0]), key
        assert np.array_equal(piv, golden[key + "_ipiv"]), key
        LUr = golden[key + "_LU"]
        tol = 50 * EPS[dt] * max(m, n) * max(1.0, float(np.abs(LUr).max()))
        assert np.abs(A - LUr).max() <= tol, key
        _, opiv, oinfo = oracle.getf2(golden[key + "_A"])
        assert np.array_equal(piv, opiv) and oinfo == info.value, key
        for tr in (0, 1):
            if key + f"_rs{tr}_B" in golden.files:
                X = np.array(golden[key + f"_rs{tr}_B"], order="F")
                fs(b"T" if tr else b"N", oba.ip(n), oba.ip(X.shape[1]), A.ctypes.data, oba.ip(n), piv.ctypes.data,
                   X.ctypes.data, oba.ip(n), ctypes.byref(info))
                Xr = golden[key + f"_rs{tr}_X"]
                assert np.abs(X - Xr).max() <= 1e3 * tol * max(1, float(np.abs(Xr).max())), key
        j += 1
    assert j == 5


@pytest.mark.parametrize("n", [1, 2, 31, 32, 33, 63, 64, 65, 100, 257, 600])
def test_dgetrf_vs_oracle_sizes(oba, n):
    """Recursion boundaries around the 32-wide panel; tie-free Gaussian data."""
    L = oba.lib()
    rng = np.random.default_rng(1000 + n)
    A0 = np.asfortranarray(rng.standard_normal((n, n)))
    A = A0.copy(order="F")
    piv = np.zeros(n, np.int32)
    info = c_int()
    L.dgetrf_(oba.ip(n), oba.ip(n), A.ctypes.data, oba.ip(n), piv.ctypes.data, ctypes.byref(info))
    LUo, pivo, infoo = oracle.getf2(A0)
    assert info.value == infoo == 0
    assert np.array_equal(piv, pivo)
    assert np.abs(A - LUo).max() <= 100 * n * EPS[np.float64] * max(1, np.abs(LUo).max())
    # dget01: ||P L U - A||_1 / (n ||A||_1 eps) < 30 (dtest.in threshold)
    Lm = np.tril(A, -1) + np.eye(n); Um = np.triu(A)
    PA = oracle.laswp(A0, 1, n, piv, 1)
    assert np.linalg.norm(Lm @ Um - PA, 1) / (n * np.linalg.norm(A0, 1) * EPS[np.float64]) < 30


@pytest.mark.parametrize("m,n", [(100, 37), (37, 100), (300, 65), (65, 300)])
def test_getrf_rectangular(oba, m, n):
    L = oba.lib()
    rng = np.random.default_rng(m * 7 + n)
    for dt, f in ((np.float64, L.dgetrf_), (np.float32, L.sgetrf_)):
        A0 = np.asfortranarray(rng.standard_normal((m, n)).astype(dt))
        A = A0.copy(order="F")
        piv = np.zeros(min(m, n), np.int32)
        info = c_int()
        f(oba.ip(m), oba.ip(n), A.ctypes.data, oba.ip(m), piv.ctypes.data, ctypes.byref(info))
        LUo, pivo, infoo = oracle.getf2(A0)
        assert info.value == infoo and np.array_equal(piv, pivo), dt
        assert np.abs(A - LUo).max() <= 200 * max(m, n) * EPS[dt] * max(1, np.abs(LUo).max())


@pytest.mark.parametrize("m,n,nb", [(600, 600, 64), (700, 520, 64), (520, 700, 96), (1100, 1100, 256)])
def test_getrf_lookahead_blocks(oba, monkeypatch, m, n, nb):
    """The blocked look-ahead driver (trailing updates on a second stream,
    block interchanges deferred to the columns outside the block), forced at
    test sizes by a small block width: same pivots as the oracle, factors to
    rounding, a zero column still reported at its position."""
    L = oba.lib()
    monkeypatch.setenv("OPENBLAS_AMD_LU_NB", str(nb))
    rng = np.random.default_rng(m + 3 * n + nb)
    A0 = np.asfortranarray(rng.standard_normal((m, n)))
    A = A0.copy(order="F")
    piv = np.zeros(min(m, n), np.int32)
    info = c_int()
    L.dgetrf_(oba.ip(m), oba.ip(n), A.ctypes.data, oba.ip(m), piv.ctypes.data, ctypes.byref(info))
    LUo, pivo, infoo = oracle.getf2(A0)
    assert info.value == infoo == 0
    assert np.array_equal(piv, pivo)
    assert np.abs(A - LUo).max() <= 100 * max(m, n) * EPS[np.float64] * max(1, np.abs(LUo).max())
    # FP32: dget01 ratio and |l| <= 1 (pivots may differ from the oracle's on near-ties)
    A32 = np.asfortranarray(A0.astype(np.float32))
    p32 = np.zeros(min(m, n), np.int32)
    L.sgetrf_(oba.ip(m), oba.ip(n), A32.ctypes.data, oba.ip(m), p32.ctypes.data, ctypes.byref(info))
    assert info.value == 0
    k = min(m, n)
    Lm = (np.tril(A32, -1)[:, :k] + np.eye(m, k)).astype(np.float64)
    Um = np.triu(A32)[:k, :].astype(np.float64)
    assert np.abs(np.tril(A32, -1)).max() <= 1.0
    PA = oracle.laswp(A0.astype(np.float32).astype(np.float64), 1, k, p32, 1)
    assert np.linalg.norm(Lm @ Um - PA, 1) / (max(m, n) * np.linalg.norm(A0, 1) * EPS[np.float32]) < 30
    if m == n:
        # singular: exact zero column in the third block
        Z0 = A0.copy(order="F"); Z0[:, 2 * nb + 5] = 0.0
        Z = Z0.copy(order="F")
        L.dgetrf_(oba.ip(n), oba.ip(n), Z.ctypes.data, oba.ip(n), piv.ctypes.data, ctypes.byref(info))
        LUo, pivo, infoo = oracle.getf2(Z0)
        assert info.value == infoo == 2 * nb + 6 and np.array_equal(piv, pivo)
        # the solve through the tracked permutation
        B0 = np.asfortranarray(rng.standard_normal((n, 3)))
        A1, B = A0.copy(order="F"), B0.copy(order="F")
        L.dgesv_(oba.ip(n), oba.ip(3), A1.ctypes.data, oba.ip(n), piv.ctypes.data, B.ctypes.data, oba.ip(n),
                 ctypes.byref(info))
        assert info.value == 0
        assert np.abs(A0 @ B - B0).max() / (np.abs(A0).sum(1).max() * np.abs(B).max()) < 1e-14


def test_getrf_singular_and_ties(oba):
    L = oba.lib()
    rng = np.random.default_rng(21)
    # exact zero column 40 of 100: info = 41, factorization completed
    A0 = rng.standard_normal((100, 100)); A0[:, 40] = 0.0
    A = np.asfortranarray(A0.copy())
    piv = np.zeros(100, np.int32)
    info = c_int()
    L.dgetrf_(oba.ip(100), oba.ip(100), A.ctypes.data, oba.ip(100), piv.ctypes.data, ctypes.byref(info))
    LUo, pivo, infoo = oracle.getf2(A0)
    assert info.value == infoo == 41 and np.array_equal(piv, pivo)
    # exact ties in magnitude: the first row wins (kernel/arm/iamax.c:64-74)
    T = np.asfortranarray(np.array([[1.0, 2.0, 3.0], [-1.0, 5.0, 1.0], [1.0, 0.5, 4.0]]))
    piv = np.zeros(3, np.int32)
    L.dgetrf_(oba.ip(3), oba.ip(3), T.ctypes.data, oba.ip(3), piv.ctypes.data, ctypes.byref(info))
    assert list(piv) == list(oracle.getf2(np.array([[1.0, 2.0, 3.0], [-1.0, 5.0, 1.0], [1.0, 0.5, 4.0]]))[1])
    # all-zero matrix: no interchange, info = 1
    Z = np.zeros((70, 70), order="F")
    piv = np.zeros(70, np.int32)
    L.dgetrf_(oba.ip(70), oba.ip(70), Z.ctypes.data, oba.ip(70), piv.ctypes.data, ctypes.byref(info))
    assert info.value == 1 and list(piv) == list(range(1, 71))


@pytest.mark.parametrize("n,nrhs", [(1, 1), (50, 1), (129, 16), (513, 3)])
def test_gesv_and_lapacke(oba, n, nrhs):
    L = oba.lib()
    rng = np.random.default_rng(n + nrhs)
    A0 = np.asfortranarray(rng.standard_normal((n, n)))
    B0 = np.asfortranarray(rng.standard_normal((n, nrhs)))
    A, X = A0.copy(order="F"), B0.copy(order="F")
    piv = np.zeros(n, np.int32)
    info = c_int()
    L.dgesv_(oba.ip(n), oba.ip(nrhs), A.ctypes.data, oba.ip(n), piv.ctypes.data, X.ctypes.data, oba.ip(n),
             ctypes.byref(info))
    assert info.value == 0
    bwd = np.abs(B0 - A0 @ X).max() / (np.abs(A0).sum(1).max() * np.abs(X).max() + np.abs(B0).max())
    assert bwd < 10 * n * EPS[np.float64]
    Xo = oracle.getrs(0, oracle.getf2(A0)[0], oracle.getf2(A0)[1], B0)
    assert np.abs(X - Xo).max() <= 1e-8 * max(1, np.abs(Xo).max())
    # LAPACKE row-major: the same system stored row-major
    Ar = np.ascontiguousarray(A0); Br = np.ascontiguousarray(B0)
    piv2 = np.zeros(n, np.int32)
    assert L.LAPACKE_dgesv(101, n, nrhs, Ar.ctypes.data, n, piv2.ctypes.data, Br.ctypes.data, nrhs) == 0
    assert np.array_equal(piv2, piv)
    assert np.abs(Br - X).max() <= 1e-12 * max(1, np.abs(X).max())
    # LAPACKE NaN check: -4 for A
    An = Ar.copy(); An[0, 0] = np.nan
    assert L.LAPACKE_dgesv(101, n, nrhs, An.ctypes.data, n, piv2.ctypes.data, Br.ctypes.data, nrhs) == -4


def test_getrs_transposed_and_lapacke_getrf(oba):
    L = oba.lib()
    rng = np.random.default_rng(31)
    n = 150
    A0 = np.asfortranarray(rng.standard_normal((n, n)))
    A = A0.copy(order="F")
    piv = np.zeros(n, np.int32)
    assert L.LAPACKE_dgetrf(102, n, n, A.ctypes.data, n, piv.ctypes.data) == 0
    B0 = np.asfortranarray(rng.standard_normal((n, 4)))
    for tr, op in ((b"N", A0), (b"T", A0.T)):
        X = B0.copy(order="F")
        assert L.LAPACKE_dgetrs(102, tr, n, 4, A.ctypes.data, n, piv.ctypes.data, X.ctypes.data, n) == 0
        assert np.abs(op @ X - B0).max() < 1e-11 * np.abs(X).max() * n
    # row-major getrf factors the transpose layout correctly
    Ar = np.ascontiguousarray(A0)
    pr = np.zeros(n, np.int32)
    assert L.LAPACKE_dgetrf(101, n, n, Ar.ctypes.data, n, pr.ctypes.data) == 0
    assert np.array_equal(pr, piv) and np.abs(Ar - A).max() < 1e-12 * np.abs(A).max()


@pytest.mark.parametrize("incx", [1, -1, 2])
def test_laswp_vs_oracle(oba, incx):
    L = oba.lib()
    rng = np.random.default_rng(41)
    A0 = np.asfortranarray(rng.standard_normal((30, 9)))
    ipiv = rng.integers(1, 31, 40).astype(np.int32)
    A = A0.copy(order="F")
    L.dlaswp_(oba.ip(9), A.ctypes.data, oba.ip(30), oba.ip(3), oba.ip(12), ipiv.ctypes.data, oba.ip(incx))
    assert np.array_equal(A, oracle.laswp(A0, 3, 12, ipiv, incx))


@pytest.mark.parametrize("tag", ["d", "s"])
def test_potrf_golden_vectors(oba, golden, tag):
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    f = L.dpotrf_ if tag == "d" else L.spotrf_
    for j in range(4):
        for uplo in (0, 1):
            key = f"potrf{tag}_{j}_{uplo}"
            A = np.array(golden[key + "_A"], order="F")
            n = A.shape[0]
            A0 = A.copy()
            info = c_int()
            f(b"L" if uplo else b"U", oba.ip(n), A.ctypes.data, oba.ip(n), ctypes.byref(info))
            assert info.value == int(golden[key + "_info"][0]) == 0
            tri = np.tril if uplo else np.triu
            other = np.triu if uplo else np.tril
            assert np.array_equal(other(A, 1 if uplo else -1), other(A0, 1 if uplo else -1)), "other triangle touched"
            tol = 50 * n * EPS[dt] * np.abs(golden[key + "_F"]).max()
            assert np.abs(tri(A) - tri(golden[key + "_F"])).max() <= tol, key
    # not positive definite: info = 8 (the first failing leading minor)
    for uplo in (0, 1):
        A = np.array(golden[f"potrfd_np_{uplo}_A"], order="F")
        info = c_int()
        L.dpotrf_(b"L" if uplo else b"U", oba.ip(12), A.ctypes.data, oba.ip(12), ctypes.byref(info))
        assert info.value == int(golden[f"potrfd_np_{uplo}_info"][0]) == 8


@pytest.mark.parametrize("n", [1, 63, 64, 65, 130, 300, 1000])
def test_potrf_vs_oracle_sizes(oba, n):
    L = oba.lib()
    rng = np.random.default_rng(2000 + n)
    Bm = rng.standard_normal((n, n))
    A0 = np.asfortranarray(Bm @ Bm.T + n * np.eye(n))
    for uplo in (b"L", b"U"):
        A = A0.copy(order="F")
        info = c_int()
        L.dpotrf_(uplo, oba.ip(n), A.ctypes.data, oba.ip(n), ctypes.byref(info))
        assert info.value == 0
        F, oinfo = oracle.potf2(1 if uplo == b"L" else 0, A0)
        tri = np.tril if uplo == b"L" else np.triu
        assert np.abs(tri(A) - tri(F)).max() <= 50 * n * EPS[np.float64] * np.abs(F).max()
    # a negative diagonal entry deep in the matrix: info from a later leaf
    if n >= 64:
        A = A0.copy(order="F")
        A[n - 3, n - 3] = -1e6
        info = c_int()
        L.dpotrf_(b"L", oba.ip(n), A.ctypes.data, oba.ip(n), ctypes.byref(info))
        assert info.value == n - 2


@pytest.mark.parametrize("tag,n", [("d", 2600), ("s", 1600)])
def test_potrf_lookahead_blocks(oba, tag, n):
    """Lower Cholesky from 4 blocks of 384 columns on runs the blocked driver
    with look-ahead (trailing updates on the second stream).
    Device-resident, the strict upper triangle NaN-poisoned (never read or
    written), factor against numpy's within c*n*eps, reconstruction
    L L^T = A, and a non-positive pivot in a late block: INFO = its column
    (1-based) and the columns before it equal the factor of the leading
    minor (potf2_L.c:71-75: the failing diagonal is stored, later columns
    untouched by the factorization)."""
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    eps = EPS[dt]
    rng = np.random.default_rng(2300 + n)
    G = rng.standard_normal((n, 64))
    A0 = G @ G.T + n * np.eye(n)
    Lnp = np.linalg.cholesky(A0)
    Ap = np.asfortranarray(np.tril(A0) + np.triu(np.full((n, n), np.nan), 1)).astype(dt)
    f = L.dpotrf_ if tag == "d" else L.spotrf_
    dA = oba.DeviceBuffer.from_array(Ap)
    info = ctypes.c_int(-5)
    f(b"L", oba.ip(n), dA.ptr, oba.ip(n), ctypes.byref(info))
    got = dA.to_array(np.empty_like(Ap))
    dA.free()
    assert info.value == 0
    assert np.isnan(np.triu(got, 1)[np.triu_indices(n, 1)]).all()
    Lg = np.tril(got).astype(np.float64)
    assert np.abs(Lg - Lnp).max() <= 50 * n * eps * np.abs(Lnp).max()
    assert np.abs(Lg @ Lg.T - A0).max() <= 50 * n * eps * np.abs(A0).max()
    # a negative pivot at column jf (0-based) in the last block but one
    jf = n - 300
    Ab = Ap.copy(order="F")
    Ab[jf, jf] = -1e6
    dA = oba.DeviceBuffer.from_array(Ab)
    f(b"L", oba.ip(n), dA.ptr, oba.ip(n), ctypes.byref(info))
    got = dA.to_array(np.empty_like(Ab))
    dA.free()
    assert info.value == jf + 1
    Lm = np.linalg.cholesky(A0[:jf, :jf])
    assert np.abs(np.tril(got[:jf, :jf]).astype(np.float64) - Lm).max() <= 50 * n * eps * np.abs(Lm).max()
    assert got[jf, jf] < 0   # the failing value is stored


def test_dsgesv_golden_vectors(oba, golden):
    L = oba.lib()
    for j in range(3):
        for nrhs in (1, 2):
            key = f"dsgesv_{j}_{nrhs}"
            A = np.array(golden[key + "_A"], order="F")
            B = np.array(golden[key + "_B"], order="F")
            n = A.shape[0]
            X = np.zeros((n, nrhs), order="F")
            piv = np.zeros(n, np.int32)
            it, info = c_int(), c_int()
            L.dsgesv_(oba.ip(n), oba.ip(nrhs), A.ctypes.data, oba.ip(n), piv.ctypes.data, B.ctypes.data, oba.ip(n),
                      X.ctypes.data, oba.ip(n), None, None, ctypes.byref(it), ctypes.byref(info))
            assert info.value == int(golden[key + "_info"][0]) == 0
            assert it.value == int(golden[key + "_iter"][0]), key
            A0 = golden[key + "_A"]
            bwd = np.abs(B - A0 @ X).max() / (np.abs(A0).sum(1).max() * np.abs(X).max())
            assert bwd < 1e-15, key
            if it.value >= 0:
                assert np.array_equal(A, A0), "A must be unchanged after successful refinement"


@pytest.mark.parametrize("name", ["B_overflow", "A_overflow", "A_single_singular"])
def test_dsgesv_fallback_codes(oba, dsgesv_fallback_cases, name):
    """ITER = -2 (a value beyond FLT_MAX) and -3 (single-precision LU
    singular) take the FP64 DGETRF + DGETRS exit: ITER, INFO and pivots as
    the oracle's, A overwritten by its factors, X to the same accuracy."""
    L = oba.lib()
    A0, B0, want = dsgesv_fallback_cases[name]
    n, nrhs = B0.shape
    A, B = np.asfortranarray(A0.copy()), np.asfortranarray(B0.copy())
    X = np.zeros((n, nrhs), order="F")
    piv = np.zeros(n, np.int32)
    it, info = c_int(), c_int()
    L.dsgesv_(oba.ip(n), oba.ip(nrhs), A.ctypes.data, oba.ip(n), piv.ctypes.data, B.ctypes.data, oba.ip(n),
              X.ctypes.data, oba.ip(n), None, None, ctypes.byref(it), ctypes.byref(info))
    Ao, pivo, Xo, ito, infoo = oracle.dsgesv(A0, B0)
    assert it.value == ito == want and info.value == infoo == 0
    assert np.array_equal(piv, pivo)
    assert np.abs(A - Ao).max() <= 64 * n * EPS[np.float64] * np.abs(Ao).max()
    assert np.all(np.abs(X - Xo).max(axis=0) <= 1e-12 * np.abs(Xo).max(axis=0))
    bwd = np.abs(B0 - A0 @ X).max() / (np.abs(A0).sum(1).max() * np.abs(X).max())
    assert bwd < 1e-15


def test_lapacke_dsgesv_rowmajor(oba):
    L = oba.lib()
    rng = np.random.default_rng(51)
    n, nrhs = 200, 3
    A0 = rng.standard_normal((n, n)) + 20 * np.eye(n)
    B0 = rng.standard_normal((n, nrhs))
    Ar, Br = np.ascontiguousarray(A0), np.ascontiguousarray(B0)
    Xr = np.zeros((n, nrhs))
    piv = np.zeros(n, np.int32)
    it = ctypes.c_int(-99)
    assert L.LAPACKE_dsgesv(101, n, nrhs, Ar.ctypes.data, n, piv.ctypes.data, Br.ctypes.data, nrhs,
                            Xr.ctypes.data, nrhs, ctypes.byref(it)) == 0
    _, _, Xo, ito, _ = oracle.dsgesv(A0, B0)
    assert it.value == ito >= 0
    assert np.abs(Xr - Xo).max() < 1e-13


def test_dgesv_4096_backward_error(oba):
    """A size-independent property at a large size: normwise backward error
    of the FP64 device-resident solve, and of the mixed-precision solve."""
    L = oba.lib()
    n = 4096
    rng = np.random.default_rng(61)
    A0 = np.asfortranarray(rng.standard_normal((n, n)))
    B0 = np.asfortranarray(rng.standard_normal((n, 1)))
    dA, dB = oba.DeviceBuffer.from_array(A0), oba.DeviceBuffer.from_array(B0)
    dpiv = oba.DeviceBuffer(4 * n)
    info = c_int()
    L.dgesv_(oba.ip(n), oba.ip(1), dA.ptr, oba.ip(n), dpiv.ptr, dB.ptr, oba.ip(n), ctypes.byref(info))
    assert info.value == 0
    X = dB.to_array(np.empty((n, 1), order="F"))
    bwd = np.abs(B0 - A0 @ X).max() / (np.abs(A0).sum(1).max() * np.abs(X).max() + np.abs(B0).max())
    assert bwd < 1e-14
    dA2, dB2 = oba.DeviceBuffer.from_array(A0), oba.DeviceBuffer.from_array(B0)
    dX = oba.DeviceBuffer(8 * n)
    it = c_int()
    L.dsgesv_(oba.ip(n), oba.ip(1), dA2.ptr, oba.ip(n), dpiv.ptr, dB2.ptr, oba.ip(n), dX.ptr, oba.ip(n), None, None,
              ctypes.byref(it), ctypes.byref(info))
    X2 = dX.to_array(np.empty((n, 1), order="F"))
    assert info.value == 0 and it.value >= 0
    bwd2 = np.abs(B0 - A0 @ X2).max() / (np.abs(A0).sum(1).max() * np.abs(X2).max() + np.abs(B0).max())
    assert bwd2 < 1e-14
    for d in (dA, dB, dpiv, dA2, dB2, dX):
        d.free()
