"""GPU parity tests of the boundary pieces added after the first build:
dsposv_ / LAPACKE_dsposv[_work] (lapack-netlib/SRC/dsposv.f:197,
lapacke.h:2974,8477), LAPACKE_dsgesv_work (lapacke.h:5845), the referenced-
triangle NaN pre-check of LAPACKE_dpotrf (lapacke_dpotrf.c:43-45), dsgesv_
with NRHS = 0, pivots below the smallest normal number (getf2_k.c:103-116)
and NaN pivot columns, and the INT8-slice DGEMM's fallbacks for operands it
cannot represent.

Fixtures: tests/golden/level3_golden_extra.npz, written by the reference
library itself (tests/golden/make_golden.py --extra).
"""
import ctypes
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EPS64 = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def extra():
    return np.load(os.path.join(ROOT, "tests", "golden", "level3_golden_extra.npz"))


def c_int():
    return ctypes.c_int(-99)


def sym_full(A, lower):
    """The symmetric matrix held in the lower (upper) triangle of A."""
    T = np.tril(A) if lower else np.triu(A)
    return T + T.T - np.diag(np.diag(A))


@pytest.mark.parametrize("j", [0, 1, 2])
@pytest.mark.parametrize("uplo", [0, 1])
@pytest.mark.parametrize("nrhs", [1, 2])
def test_dsposv_golden_vectors(oba, extra, j, uplo, nrhs):
    """ITER and INFO equal to the reference library's (kappa 1e2 -> 2,
    1e6 -> 3, 1e10 -> -3: SPOTRF fails and DPOTRF/DPOTRS run), X at FP64
    backward error, A untouched unless the FP64 fallback ran (then its
    factor equals the oracle's DPOTF2)."""
    L = oba.lib()
    key = f"dsposv_{j}_{uplo}_{nrhs}"
    A0, B0 = extra[key + "_A"], extra[key + "_B"]
    n = A0.shape[0]
    # garbage in the unreferenced triangle: it must not be read
    A = np.array(A0, order="F")
    garbage = np.triu(np.full((n, n), np.nan), 1) if uplo else np.tril(np.full((n, n), np.nan), -1)
    A = np.where(np.isnan(garbage), np.nan, A)
    A = np.asfortranarray(A)
    Ain = A.copy(order="F")
    B = np.array(B0, order="F")
    X = np.zeros((n, nrhs), order="F")
    it, info = c_int(), c_int()
    L.dsposv_(b"L" if uplo else b"U", oba.ip(n), oba.ip(nrhs), A.ctypes.data, oba.ip(n), B.ctypes.data, oba.ip(n),
              X.ctypes.data, oba.ip(n), None, None, ctypes.byref(it), ctypes.byref(info))
    assert info.value == int(extra[key + "_info"][0]) == 0, key
    assert it.value == int(extra[key + "_iter"][0]), key
    S = sym_full(A0, uplo == 1)
    bwd = np.abs(B0 - S @ X).max() / (np.abs(S).sum(1).max() * np.abs(X).max())
    assert bwd < 1e-15, (key, bwd)
    Xr = extra[key + "_X"]
    # two backward-stable FP64 solutions differ by up to ~kappa*eps (forward
    # error bound), so the agreement is scaled by the condition number
    kappa = (1e2, 1e6, 1e10)[j]
    assert np.abs(X - Xr).max() <= max(1e-8, 16 * kappa * EPS64) * np.abs(Xr).max(), key
    tri = np.tril if uplo else np.triu
    if it.value >= 0:
        assert np.array_equal(np.nan_to_num(A, nan=7.0), np.nan_to_num(Ain, nan=7.0)), "A must be unchanged"
    else:
        Fo, info_o = oracle.potf2(uplo, A0)
        assert info_o == 0
        assert np.abs(tri(A) - tri(Fo)).max() <= 64 * n * EPS64 * np.abs(Fo).max()
        assert np.all(np.isnan(A[np.isnan(garbage)])), "the unreferenced triangle must stay untouched"


def test_dsposv_matches_oracle_iter(oba):
    """A second source of truth for ITER: the oracle's restatement of
    dsposv.f on fresh SPD systems of several sizes."""
    L = oba.lib()
    rng = np.random.default_rng(81)
    for n, nrhs, uplo in ((33, 1, 1), (130, 3, 0), (300, 5, 1)):
        Bm = rng.standard_normal((n, n))
        A0 = np.asfortranarray(Bm @ Bm.T + 0.1 * np.eye(n))
        B0 = np.asfortranarray(rng.standard_normal((n, nrhs)))
        A, B = A0.copy(order="F"), B0.copy(order="F")
        X = np.zeros((n, nrhs), order="F")
        it, info = c_int(), c_int()
        L.dsposv_(b"L" if uplo else b"U", oba.ip(n), oba.ip(nrhs), A.ctypes.data, oba.ip(n), B.ctypes.data,
                  oba.ip(n), X.ctypes.data, oba.ip(n), None, None, ctypes.byref(it), ctypes.byref(info))
        _, Xo, ito, infoo = oracle.dsposv(uplo, A0, B0)
        assert info.value == infoo == 0 and it.value == ito, (n, it.value, ito)
        assert np.abs(X - Xo).max() <= 1e-10 * np.abs(Xo).max()


def test_dsposv_errors_and_quick_returns(oba):
    L = oba.lib()
    A = np.eye(4, order="F")
    B = np.ones((4, 1), order="F")
    X = np.zeros((4, 1), order="F")
    it, info = c_int(), c_int()
    oba.last_xerbla()
    L.dsposv_(b"X", oba.ip(4), oba.ip(1), A.ctypes.data, oba.ip(4), B.ctypes.data, oba.ip(4), X.ctypes.data,
              oba.ip(4), None, None, ctypes.byref(it), ctypes.byref(info))
    assert info.value == -1 and oba.last_xerbla() == ("DSPOSV", 1)
    L.dsposv_(b"L", oba.ip(4), oba.ip(1), A.ctypes.data, oba.ip(3), B.ctypes.data, oba.ip(4), X.ctypes.data,
              oba.ip(4), None, None, ctypes.byref(it), ctypes.byref(info))
    assert info.value == -5 and oba.last_xerbla() == ("DSPOSV", 5)
    L.dsposv_(b"L", oba.ip(4), oba.ip(1), A.ctypes.data, oba.ip(4), B.ctypes.data, oba.ip(4), X.ctypes.data,
              oba.ip(3), None, None, ctypes.byref(it), ctypes.byref(info))
    assert info.value == -9
    L.dsposv_(b"U", oba.ip(0), oba.ip(1), A.ctypes.data, oba.ip(1), B.ctypes.data, oba.ip(1), X.ctypes.data,
              oba.ip(1), None, None, ctypes.byref(it), ctypes.byref(info))
    assert info.value == 0 and it.value == 0
    # a matrix that is not positive definite in FP64 either: ITER = -3 and
    # DPOTRF's info
    A = np.eye(6, order="F")
    A[4, 4] = -2.0
    L.dsposv_(b"L", oba.ip(6), oba.ip(1), A.ctypes.data, oba.ip(6), np.ones(6).ctypes.data, oba.ip(6),
              np.zeros(6).ctypes.data, oba.ip(6), None, None, ctypes.byref(it), ctypes.byref(info))
    assert it.value == -3 and info.value == 5


@pytest.mark.parametrize("layout", [101, 102])
@pytest.mark.parametrize("uplo", [b"L", b"U"])
def test_lapacke_dsposv_layouts(oba, layout, uplo):
    L = oba.lib()
    rng = np.random.default_rng(83)
    n, nrhs = 90, 2
    Bm = rng.standard_normal((n, n))
    S = Bm @ Bm.T + n * np.eye(n)
    B0 = rng.standard_normal((n, nrhs))
    lower = uplo == b"L"
    if layout == 101:
        a = np.ascontiguousarray(S)
        b = np.ascontiguousarray(B0)
        x = np.zeros((n, nrhs))
        lda, ldb, ldx = n, nrhs, nrhs
    else:
        a = np.asfortranarray(S)
        b = np.asfortranarray(B0)
        x = np.zeros((n, nrhs), order="F")
        lda = ldb = ldx = n
    it = c_int()
    r = L.LAPACKE_dsposv(layout, uplo, n, nrhs, a.ctypes.data, lda, b.ctypes.data, ldb, x.ctypes.data, ldx,
                         ctypes.byref(it))
    assert r == 0 and it.value >= 0
    assert np.abs(S @ x - B0).max() <= 1e-12 * np.abs(B0).max()
    # the same call through the _work variant with a NaN outside the
    # referenced triangle (the high-level wrapper's pre-check is not in _work
    # and the triangle is not read)
    S2 = S.copy()
    i, j = (1, 5) if lower else (5, 1)   # logical (row, col) in the unreferenced triangle
    S2[i, j] = np.nan
    a2 = np.ascontiguousarray(S2) if layout == 101 else np.asfortranarray(S2)
    x2 = np.zeros_like(x)
    r = L.LAPACKE_dsposv_work(layout, uplo, n, nrhs, a2.ctypes.data, lda, b.ctypes.data, ldb, x2.ctypes.data, ldx,
                              None, None, ctypes.byref(it))
    assert r == 0 and np.abs(x2 - x).max() <= 1e-13 * np.abs(x).max()
    # the high-level wrapper's NaN check: referenced triangle -> -5, B -> -7
    S3 = S.copy()
    S3[j, i] = np.nan
    a3 = np.ascontiguousarray(S3) if layout == 101 else np.asfortranarray(S3)
    assert L.LAPACKE_dsposv(layout, uplo, n, nrhs, a3.ctypes.data, lda, b.ctypes.data, ldb, x.ctypes.data, ldx,
                            ctypes.byref(it)) == -5
    b3 = b.copy()
    b3.flat[7] = np.nan
    assert L.LAPACKE_dsposv(layout, uplo, n, nrhs, a.ctypes.data, lda, b3.ctypes.data, ldb, x.ctypes.data, ldx,
                            ctypes.byref(it)) == -7
    # row-major leading dimension errors of the _work layer (-6, -8, -10)
    if layout == 101:
        assert L.LAPACKE_dsposv_work(101, uplo, n, nrhs, a.ctypes.data, n - 1, b.ctypes.data, ldb, x.ctypes.data,
                                     ldx, None, None, ctypes.byref(it)) == -6
        assert L.LAPACKE_dsposv_work(101, uplo, n, nrhs, a.ctypes.data, n, b.ctypes.data, nrhs - 1, x.ctypes.data,
                                     ldx, None, None, ctypes.byref(it)) == -8
        assert L.LAPACKE_dsposv_work(101, uplo, n, nrhs, a.ctypes.data, n, b.ctypes.data, nrhs, x.ctypes.data,
                                     nrhs - 1, None, None, ctypes.byref(it)) == -10


@pytest.mark.parametrize("layout", [101, 102])
def test_lapacke_dsgesv_work(oba, layout):
    """The _work variant: same solve as dsgesv_, no NaN pre-check, row-major
    leading dimension errors -5 / -8 / -10 (lapacke_dsgesv_work.c:56-70)."""
    L = oba.lib()
    rng = np.random.default_rng(85)
    n, nrhs = 120, 2
    A0 = rng.standard_normal((n, n)) + 15 * np.eye(n)
    B0 = rng.standard_normal((n, nrhs))
    if layout == 101:
        a, b, x = np.ascontiguousarray(A0), np.ascontiguousarray(B0), np.zeros((n, nrhs))
        lda, ldb, ldx = n, nrhs, nrhs
    else:
        a, b, x = np.asfortranarray(A0), np.asfortranarray(B0), np.zeros((n, nrhs), order="F")
        lda = ldb = ldx = n
    piv = np.zeros(n, np.int32)
    it = c_int()
    r = L.LAPACKE_dsgesv_work(layout, n, nrhs, a.ctypes.data, lda, piv.ctypes.data, b.ctypes.data, ldb,
                              x.ctypes.data, ldx, None, None, ctypes.byref(it))
    _, pivo, Xo, ito, _ = oracle.dsgesv(A0, B0)
    assert r == 0 and it.value == ito >= 0
    assert np.array_equal(piv, pivo)
    assert np.abs(x - Xo).max() < 1e-13
    if layout == 101:
        for args, want in (((n - 1, ldb, ldx), -5), ((n, nrhs - 1, ldx), -8), ((n, ldb, nrhs - 1), -10)):
            assert L.LAPACKE_dsgesv_work(101, n, nrhs, a.ctypes.data, args[0], piv.ctypes.data, b.ctypes.data,
                                         args[1], x.ctypes.data, args[2], None, None, ctypes.byref(it)) == want


@pytest.mark.parametrize("layout", [101, 102])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("where", ["ref", "other"])
def test_lapacke_dpotrf_nan_check(oba, extra, layout, uplo, where):
    """LAPACKE_dpotrf returns -4 for a NaN in the referenced triangle only,
    exactly as the reference library does (fixtures), in both layouts."""
    L = oba.lib()
    key = f"lapacke_potrf_{layout}_{uplo}_{where}"
    A = extra[key + "_A"]
    n = A.shape[0]
    buf = np.ascontiguousarray(A) if layout == 101 else np.asfortranarray(A)
    r = L.LAPACKE_dpotrf(layout, uplo.encode(), n, buf.ctypes.data, n)
    assert r == int(extra[key + "_info"][0]), key


def test_dsgesv_nrhs0_factorizes(oba):
    """NRHS = 0 with N > 0: dsgesv.f returns early only for N = 0, so SGETRF
    still runs and IPIV is set (its pivots), ITER = 0, A untouched."""
    L = oba.lib()
    rng = np.random.default_rng(87)
    n = 64
    A0 = np.asfortranarray(rng.standard_normal((n, n)))
    A = A0.copy(order="F")
    piv = np.zeros(n, np.int32)
    it, info = c_int(), c_int()
    L.dsgesv_(oba.ip(n), oba.ip(0), A.ctypes.data, oba.ip(n), piv.ctypes.data, np.zeros(n).ctypes.data, oba.ip(n),
              np.zeros(n).ctypes.data, oba.ip(n), None, None, ctypes.byref(it), ctypes.byref(info))
    assert it.value == 0 and info.value == 0
    _, spiv, _ = oracle.getf2(A0.astype(np.float32))
    assert np.array_equal(piv, spiv)
    assert np.array_equal(A, A0)


def _lu_whole_rows(A):
    """Partial pivoting with whole-row interchanges and OpenBLAS's scaling
    rule (multipliers divided by the pivot only when |pivot| >= DBL_MIN,
    getf2_k.c:103-116): the factorization this library computes."""
    A = np.array(A, dtype=np.float64)
    n = A.shape[0]
    piv = np.zeros(n, np.int32)
    for j in range(n):
        p = j + int(np.argmax(np.abs(A[j:, j])))
        piv[j] = p + 1
        A[[j, p], :] = A[[p, j], :]
        if A[j, j] != 0 and abs(A[j, j]) >= np.finfo(np.float64).tiny:
            A[j + 1:, j] /= A[j, j]
        A[j + 1:, j + 1:] -= np.outer(A[j + 1:, j], A[j, j + 1:])
    return A, piv


@pytest.mark.parametrize("n", [8, 40])
def test_getrf_subnormal_pivot_deviation(oba, extra, n):
    """A nonzero pivot below DBL_MIN (column 3 here).  The reference records
    the interchange in IPIV but, inside its unblocked getf2 leaf, applies it
    only to the columns right of the pivot and does not scale
    (getf2_k.c:103-116).  Which columns count as "inside the leaf" depends on
    its blocking (GEMM_UNROLL_N, GEMM_Q of the CPU), and later columns then
    combine the multipliers of one row with the entries of the other, so the
    reference's factors (fixture) are not a row permutation of L*U from that
    column on.  This library exchanges whole rows and keeps the no-scaling
    rule.  Pinned: IPIV up to and including the subnormal column and INFO
    equal the reference's; IPIV and the factors equal the whole-row
    factorization above; the reference's factors differ from it (the
    deviation is real and starts at the subnormal column)."""
    L = oba.lib()
    A0, LUr, pr = extra[f"getrf_tiny_{n}_A"], extra[f"getrf_tiny_{n}_LU"], extra[f"getrf_tiny_{n}_ipiv"]
    A = np.asfortranarray(A0.copy())
    piv = np.zeros(n, np.int32)
    info = c_int()
    L.dgetrf_(oba.ip(n), oba.ip(n), A.ctypes.data, oba.ip(n), piv.ctypes.data, ctypes.byref(info))
    j = 3
    assert info.value == int(extra[f"getrf_tiny_{n}_info"][0]) == 0
    assert np.array_equal(piv[:j + 1], pr[:j + 1])
    jp = int(pr[j]) - 1
    assert jp == 6   # the subnormal entry is the reference's pivot of column 3 as well
    model, pm = _lu_whole_rows(A0)
    assert np.array_equal(piv, pm)
    tol = 64 * n * EPS64 * max(1.0, np.abs(model).max())
    assert np.abs(A - model).max() <= tol
    assert np.abs(LUr - model).max() > 1e3 * tol
    assert np.abs(LUr - model)[:j].max() <= tol   # rows above the subnormal column agree


@pytest.mark.parametrize("t", [0, 1])
def test_getrf_nan_pivot_column(oba, extra, t):
    """NaN in a pivot column follows the C IAMAX semantic of
    kernel/arm/iamax.c:51-74 (a NaN wins only at the first position), as the
    oracle restates it.  The installed reference's x86 IAMAX kernel
    (kernel/x86_64/iamax_sse2.S, maxpd-based) picks differently when the NaN
    is not at the first position (fixture 1): that choice depends on its SSE
    unrolling, so the oracle's rule is the one checked; fixture 0 (NaN at the
    diagonal position) agrees with the reference library as well."""
    L = oba.lib()
    A0 = extra[f"getrf_nan_{t}_A"]
    n = A0.shape[0]
    A = np.asfortranarray(A0.copy())
    piv = np.zeros(n, np.int32)
    info = c_int()
    L.dgetrf_(oba.ip(n), oba.ip(n), A.ctypes.data, oba.ip(n), piv.ctypes.data, ctypes.byref(info))
    _, po, io = oracle.getf2(A0)
    assert np.array_equal(piv, po) and info.value == io
    if t == 0:
        assert np.array_equal(piv, extra["getrf_nan_0_ipiv"])


def _ozaki(L, A, B, C, alpha=1.0, beta=0.0, slices=0):
    m, k = A.shape
    n = B.shape[1]
    return L.openblas_amd_dgemm_ozaki(102, 111, 111, m, n, k, alpha, A.ctypes.data, m, B.ctypes.data, k, beta,
                                      C.ctypes.data, m, slices)


def test_ozaki_wide_spread_falls_back_to_native(oba):
    """Operands whose exponent spread needs more than 12 slices: the result
    keeps FP64 accuracy (native DGEMM) and the return value says so (-s)."""
    L = oba.lib()
    rng = np.random.default_rng(89)
    n = 256
    A = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, n)) * np.exp(8.0 * rng.standard_normal((n, n))))
    B = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, n)) * np.exp(8.0 * rng.standard_normal((n, n))))
    C = np.zeros((n, n), order="F")
    s = _ozaki(L, A, B, C)
    assert s <= -13
    Cn = np.zeros((n, n), order="F")
    L.cblas_dgemm(102, 111, 111, n, n, n, 1.0, A.ctypes.data, n, B.ctypes.data, n, 0.0, Cn.ctypes.data, n)
    assert np.array_equal(C, Cn)
    ref = A @ B
    gauge = np.abs(A) @ np.abs(B)
    assert np.max(np.abs(C - ref) / gauge) < 1e-13


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_ozaki_nonfinite_operands(oba, bad):
    """NaN / Inf in A or B: return -2 and the native DGEMM's result, so they
    propagate exactly as in cblas_dgemm."""
    L = oba.lib()
    rng = np.random.default_rng(91)
    m, n, k = 70, 50, 90
    for where in ("A", "B"):
        A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
        B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
        (A if where == "A" else B)[5, 7] = bad
        C0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        C, Cn = C0.copy(order="F"), C0.copy(order="F")
        assert _ozaki(L, A, B, C, 0.5, 0.25) == -2
        L.cblas_dgemm(102, 111, 111, m, n, k, 0.5, A.ctypes.data, m, B.ctypes.data, k, 0.25, Cn.ctypes.data, m)
        assert np.array_equal(np.isnan(C), np.isnan(Cn)) and np.array_equal(np.isinf(C), np.isinf(Cn))
        fin = np.isfinite(Cn)
        assert np.array_equal(C[fin], Cn[fin])
        assert not np.all(fin)
