"""CPU tests that pin the C oracle (oracle/oracle.c) before it is trusted:
against the reference library's own outputs (golden fixtures written by
scipy's bundled OpenBLAS 0.3.28, tests/golden/make_golden.py), against the
reference library called live, against closed forms and hand-derived values,
and by planting mistakes that the checks must catch."""
import numpy as np
import pytest

import oracle
from oracle import refblas

EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


def _cases(golden, prefix):
    return sorted(k[:-len("_meta")] for k in golden.files if k.startswith(prefix) and k.endswith("_meta"))


@pytest.mark.parametrize("tag", ["d", "s"])
def test_gemm_matches_reference_fixtures(golden, tag):
    n_cases = 0
    for key in _cases(golden, f"gemm{tag}_"):
        ta, tb, m, n, k, lda, ldb, ldc = (int(x) for x in golden[key + "_meta"])
        alpha, beta = (float(x) for x in golden[key + "_ab"])
        A, B, Cin, Cout = golden[key + "_A"], golden[key + "_B"], golden[key + "_Cin"], golden[key + "_Cout"]
        Ar = k if ta else m
        Br = n if tb else k
        got = oracle.gemm(ta, tb, alpha, A[:Ar], B[:Br], beta, Cin[:m])
        assert not np.isnan(got).any(), "beta == 0 must overwrite (kernel/generic/gemm_beta.c:52)"
        eps = EPS64 if tag == "d" else EPS32
        scale = abs(alpha) * np.abs(A[:Ar].T if ta else A[:Ar]).astype(np.float64) @ np.abs(
            B[:Br].T if tb else B[:Br]).astype(np.float64) + abs(beta) * np.nan_to_num(np.abs(Cin[:m]))
        err = np.abs(got.astype(np.float64) - Cout.astype(np.float64))
        assert np.all(err <= 4 * (k + 2) * eps * scale + 1e-300), key
        n_cases += 1
    assert n_cases == 48


def test_gemm_closed_forms():
    rng = np.random.default_rng(1)
    B = rng.uniform(-1, 1, (9, 7)); C = rng.uniform(-1, 1, (9, 7))
    # A = I: C <- alpha*B + beta*C exactly
    got = oracle.gemm(0, 0, 0.5, np.eye(9), B, 2.0, C)
    assert np.array_equal(got, 0.5 * B + 2.0 * C)
    # integer data: exact
    Ai = rng.integers(-50, 50, (13, 11)).astype(np.float64); Bi = rng.integers(-50, 50, (11, 6)).astype(np.float64)
    assert np.array_equal(oracle.gemm(0, 0, 1.0, Ai, Bi, 0.0, np.zeros((13, 6))), Ai @ Bi)
    # transposes are operands transposed, not swapped
    assert np.array_equal(oracle.gemm(1, 1, 1.0, Bi, Ai, 0.0, np.zeros((6, 13))), (Ai @ Bi).T)
    # k == 0 and alpha == 0 only scale; beta == 1 leaves C (NaN included) untouched
    Cn = C.copy(); Cn[2, 3] = np.nan
    assert np.array_equal(oracle.gemm(0, 0, 0.0, np.zeros((9, 4)), np.zeros((4, 7)), 3.0, C), 3.0 * C)
    out = oracle.gemm(0, 0, 1.0, np.zeros((9, 0)), np.zeros((0, 7)), 1.0, Cn)
    assert np.isnan(out[2, 3]) and np.array_equal(np.nan_to_num(out), np.nan_to_num(Cn))
    assert not np.isnan(oracle.gemm(0, 0, 1.0, np.ones((9, 2)), np.ones((2, 7)), 0.0, Cn)).any()


def test_gemm_live_reference_and_planted_mistake():
    rng = np.random.default_rng(2)
    for ta, tb in ((0, 0), (0, 1), (1, 0), (1, 1)):
        A = np.asfortranarray(rng.uniform(-1, 1, (23, 31) if not ta else (31, 23)))
        B = np.asfortranarray(rng.uniform(-1, 1, (31, 17) if not tb else (17, 31)))
        C = np.asfortranarray(rng.uniform(-1, 1, (23, 17)))
        ref = refblas.cblas_gemm(ta, tb, 0.7, A, B, 1.3, C.copy(order="F"))
        got = oracle.gemm(ta, tb, 0.7, A, B, 1.3, C)
        assert oracle.dmmch_ratio(ta, tb, 0.7, A, B, 1.3, C, got) < 16
        assert oracle.dmmch_ratio(ta, tb, 0.7, A, B, 1.3, C, ref) < 16
        # a transposed operand or a dropped beta term must fail the checker
        wrong = oracle.gemm(ta, tb, 0.7, A, B, 0.0, C)
        assert oracle.dmmch_ratio(ta, tb, 0.7, A, B, 1.3, C, wrong) > 1e6


def test_dbeg_generator_hand_values():
    # I <- I*891 mod 1000 from 7, one value in six skipped: 237,167,797,127,(157 skipped),887
    v = [oracle.dbeg(i == 0) for i in range(5)]
    assert np.allclose(v, [(237 - 500) / 1001, (167 - 500) / 1001, (797 - 500) / 1001, (127 - 500) / 1001,
                           (887 - 500) / 1001], rtol=0, atol=1e-15)


@pytest.mark.parametrize("tag", ["d", "s"])
def test_trsm_matches_reference_fixtures(golden, tag):
    i = 0
    for key in _cases(golden, f"trsm{tag}_"):
        side, uplo, tr, diag, m, n = (int(x) for x in golden[key + "_meta"])
        got = oracle.trsm(side, uplo, tr, diag, 0.7, golden[key + "_A"], golden[key + "_B"])
        tol = 50 * (EPS64 if tag == "d" else EPS32)
        X = golden[key + "_X"]
        assert np.max(np.abs(got - X)) <= tol * max(1.0, np.max(np.abs(X))), key
        i += 1
    assert i == 32


def test_trsm_multiply_back_and_planted_mistake():
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, (12, 12)) + 12 * np.eye(12)
    B = rng.uniform(-1, 1, (12, 5))
    X = oracle.trsm(0, 1, 0, 1, 2.0, A, B)
    assert np.allclose(np.tril(A) @ X, 2.0 * B, rtol=0, atol=1e-13)
    # unit diagonal: the stored diagonal must be ignored
    Xu = oracle.trsm(0, 1, 0, 0, 1.0, A, B)
    Lu = np.tril(A, -1) + np.eye(12)
    assert np.allclose(Lu @ Xu, B, atol=1e-13)
    # the upper triangle is not read (NaN-poisoned)
    Ap = A.copy(); Ap[np.triu_indices(12, 1)] = np.nan
    assert np.array_equal(oracle.trsm(0, 1, 0, 1, 2.0, Ap, B), X)
    # alpha == 0 zeroes B
    assert np.array_equal(oracle.trsm(0, 1, 0, 1, 0.0, A, B), np.zeros_like(B))


@pytest.mark.parametrize("tag", ["d", "s"])
def test_getf2_matches_reference_fixtures(golden, tag):
    j = 0
    while f"getrf{tag}_{j}_A" in golden.files:
        key = f"getrf{tag}_{j}"
        LU, ipiv, info = oracle.getf2(golden[key + "_A"])
        assert np.array_equal(ipiv, golden[key + "_ipiv"]), key
        assert info == int(golden[key + "_info"][0]), key
        tol = (200 if tag == "d" else 50) * (EPS64 if tag == "d" else EPS32) * LU.shape[0]
        assert np.max(np.abs(LU - golden[key + "_LU"])) <= tol * max(1, np.max(np.abs(LU))), key
        for tr in (0, 1):
            if key + f"_rs{tr}_B" in golden.files:
                X = oracle.getrs(tr, LU, ipiv, golden[key + f"_rs{tr}_B"])
                Xr = golden[key + f"_rs{tr}_X"]
                assert np.max(np.abs(X - Xr)) <= 1e3 * tol * max(1, np.max(np.abs(Xr))), key
        j += 1
    assert j == 5


def test_getf2_properties():
    rng = np.random.default_rng(4)
    # diagonally dominant -> no interchanges
    A = rng.uniform(-1, 1, (15, 15)) + 20 * np.eye(15)
    _, ipiv, info = oracle.getf2(A)
    assert info == 0 and np.array_equal(ipiv, np.arange(1, 16))
    # reconstruction P*L*U = A (dget01-style ratio)
    A = rng.standard_normal((30, 30))
    LU, ipiv, info = oracle.getf2(A)
    Lm = np.tril(LU, -1) + np.eye(30); Um = np.triu(LU)
    PA = oracle.laswp(A, 1, 30, ipiv, 1)
    assert np.linalg.norm(Lm @ Um - PA, 1) / (30 * np.linalg.norm(A, 1) * EPS64) < 30
    # ties: the first maximal |x| wins (kernel/arm/iamax.c:64-74)
    T = np.array([[1.0, 2.0], [-1.0, 5.0]])
    assert list(oracle.getf2(T)[1]) == [1, 2]
    # exact zero column: info = first zero pivot, factorization continues
    Z = rng.standard_normal((8, 8)); Z[:, 3] = 0
    assert oracle.getf2(Z)[2] == 4
    # planted mistake: a factorization without interchanges fails the ratio
    L0 = np.tril(LU, -1) + np.eye(30)
    assert np.linalg.norm(L0 @ Um - A, 1) / (30 * np.linalg.norm(A, 1) * EPS64) > 1e3


def test_laswp_live_reference():
    import scipy.linalg.lapack as la
    rng = np.random.default_rng(5)
    A = np.asfortranarray(rng.standard_normal((10, 4)))
    piv = np.array([3, 3, 7, 4, 10, 6], np.int32)
    for incx in (1, -1):
        ref = la.dlaswp(A.copy(order="F"), piv - 1, k1=0, k2=5, off=0, inc=incx)
        got = oracle.laswp(A, 1, 6, piv, incx)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("tag", ["d", "s"])
def test_potf2_matches_reference_fixtures(golden, tag):
    for j in range(4):
        for uplo in (0, 1):
            key = f"potrf{tag}_{j}_{uplo}"
            F, info = oracle.potf2(uplo, golden[key + "_A"])
            assert info == int(golden[key + "_info"][0])
            tri = np.tril if uplo else np.triu
            tol = 50 * (EPS64 if tag == "d" else EPS32) * F.shape[0]
            assert np.max(np.abs(tri(F) - tri(golden[key + "_F"]))) <= tol * np.max(np.abs(F)), key
    for uplo in (0, 1):
        F, info = oracle.potf2(uplo, golden[f"potrfd_np_{uplo}_A"])
        assert info == int(golden[f"potrfd_np_{uplo}_info"][0]) == 8


def test_potf2_kat_and_planted_mistake():
    # utest/test_potrs.c:403: A = [[2,.3],[.3,3]]
    F, info = oracle.potf2(1, np.array([[2.0, 0.3], [0.3, 3.0]]))
    assert info == 0
    assert np.allclose(np.tril(F), [[np.sqrt(2), 0], [0.3 / np.sqrt(2), np.sqrt(3 - 0.045)]], atol=1e-15)
    rng = np.random.default_rng(6)
    Bm = rng.standard_normal((20, 20)); A = Bm @ Bm.T + 20 * np.eye(20)
    F, _ = oracle.potf2(1, A)
    Lm = np.tril(F)
    assert np.linalg.norm(Lm @ Lm.T - A) / (20 * np.linalg.norm(A) * EPS64) < 10
    # upper triangle untouched by uplo = L
    assert np.array_equal(np.triu(F, 1), np.triu(A, 1))
    # planted mistake: dropping the off-diagonal coupling fails reconstruction
    D = np.diag(np.sqrt(np.diag(A)))
    assert np.linalg.norm(D @ D.T - A) / (20 * np.linalg.norm(A) * EPS64) > 1e3


def test_dsgesv_matches_reference_fixtures(golden):
    for j in range(3):
        for nrhs in (1, 2):
            key = f"dsgesv_{j}_{nrhs}"
            _, _, X, it, info = oracle.dsgesv(golden[key + "_A"], golden[key + "_B"])
            assert info == int(golden[key + "_info"][0])
            assert it == int(golden[key + "_iter"][0]), key
            A, B = golden[key + "_A"], golden[key + "_B"]
            bwd = np.max(np.abs(B - A @ X)) / (np.max(np.abs(A).sum(1)) * np.max(np.abs(X)))
            assert bwd < 1e-15
    # kappa = 1e10 cannot converge in FP32 and falls back (ddrvab.f:354-364 rule inverted)
    assert int(golden["dsgesv_2_1_iter"][0]) == -31


def test_dsgesv_live_reference_kappa2():
    rng = np.random.default_rng(7)
    A = np.asfortranarray(rng.standard_normal((50, 50)) + 10 * np.eye(50))
    B = np.asfortranarray(rng.standard_normal((50, 3)))
    Xr, _, itr, infor = refblas.dsgesv(A.copy(order="F"), B.copy(order="F"))
    _, _, X, it, info = oracle.dsgesv(A, B)
    assert it == itr >= 0 and info == infor == 0
    assert np.max(np.abs(X - Xr)) < 1e-13


@pytest.mark.parametrize("name", ["B_overflow", "A_overflow", "A_single_singular"])
def test_dsgesv_fallback_codes_match_reference(dsgesv_fallback_cases, name):
    A, B, want = dsgesv_fallback_cases[name]
    Xr, pivr, itr, infor = refblas.dsgesv(np.asfortranarray(A), np.asfortranarray(B))
    Ao, pivo, X, it, info = oracle.dsgesv(A, B)
    assert it == itr == want and info == infor == 0
    # fallback = DGETRF + DGETRS on A: same pivots, factors in A, and a
    # solution within a few ulps of the reference's
    assert np.array_equal(pivo, pivr)
    scale = np.abs(Xr).max(axis=0)
    assert np.all(np.abs(X - Xr).max(axis=0) <= 1e-12 * scale)
    bwd = np.abs(B - A @ X).max() / (np.abs(A).sum(1).max() * np.abs(X).max())
    assert bwd < 1e-15


def test_ctest_grid_harness_on_reference_library():
    """The din3 grid harness used by the GPU test, run against the reference
    library itself: pins the row-/column-major argument mapping of the
    harness and the dmmch checker (ctest/c_dblat3c.c:2956) on real data."""
    import ctypes
    L = refblas.lib()
    f = L.scipy_cblas_dgemm if hasattr(L, "scipy_cblas_dgemm") else None
    if f is None:
        pytest.skip("installed reference library is the ILP64 build")
    I, D, P = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    TR = {0: 111, 1: 112}
    sizes = [1, 2, 3, 7, 35]
    worst = 0.0
    oracle.dbeg(True)
    for (m, n, k) in [(m, n, k) for m in sizes for n in sizes for k in sizes if (m + n + k) % 3 == 0]:
        for ta in (0, 1):
            for tb in (0, 1):
                Ar, Ac = (k, m) if ta else (m, k)
                Br, Bc = (n, k) if tb else (k, n)
                A = np.array([[oracle.dbeg() for _ in range(Ac)] for _ in range(Ar + 1)], order="F")
                B = np.array([[oracle.dbeg() for _ in range(Bc)] for _ in range(Br + 1)], order="F")
                C0 = np.array([[oracle.dbeg() for _ in range(n)] for _ in range(m + 1)], order="F")
                for alpha, beta in ((0.0, 1.3), (1.0, 0.0), (0.7, 1.3)):
                    C = C0.copy(order="F")
                    f(I(102), I(TR[ta]), I(TR[tb]), I(m), I(n), I(k), D(alpha), P(A.ctypes.data), I(Ar + 1),
                      P(B.ctypes.data), I(Br + 1), D(beta), P(C.ctypes.data), I(m + 1))
                    worst = max(worst, oracle.dmmch_ratio(ta, tb, alpha, A[:Ar], B[:Br], beta, C0[:m], C[:m]))
                    Cr = np.array(C0[:m], order="C", copy=True)
                    Ac_ = np.array(A[:Ar], order="C", copy=True)
                    Bc_ = np.array(B[:Br], order="C", copy=True)
                    f(I(101), I(TR[ta]), I(TR[tb]), I(m), I(n), I(k), D(alpha), P(Ac_.ctypes.data), I(Ac),
                      P(Bc_.ctypes.data), I(Bc), D(beta), P(Cr.ctypes.data), I(n))
                    worst = max(worst, oracle.dmmch_ratio(ta, tb, alpha, A[:Ar], B[:Br], beta, C0[:m], Cr))
    assert worst < 16


def test_ozaki2_oracle_pinned():
    """oracle.ozaki2 (scheme II by its definition) against exact rational
    products: with enough integer bits the truncation is exact and the
    result is the correctly rounded exact product; with t bits the error
    obeys the truncation bound K 2^(1-t) max|a_i| max|b_j| (per operand,
    both operands); integer inputs are exact at modest t; a planted mistake
    (floor instead of truncation toward zero) breaks the bound's symmetry
    test on negative entries."""
    from fractions import Fraction
    import math
    rng = np.random.default_rng(5)
    A = rng.uniform(-1, 1, (6, 9)) * np.exp2(rng.integers(-8, 1, (6, 9)))
    B = rng.uniform(-1, 1, (9, 5))
    exact = [[sum(Fraction(A[i, k]) * Fraction(B[k, j]) for k in range(9)) for j in range(5)] for i in range(6)]
    ex = np.array([[float(v) for v in row] for row in exact])
    assert np.array_equal(oracle.ozaki2(A, B, 1200, 1200), ex)
    for t in (20, 30, 40):
        C = oracle.ozaki2(A, B, t, t)
        for i in range(6):
            for j in range(5):
                bound = 9 * (2.0 ** (1 - t)) * 2 * np.abs(A[i]).max() * np.abs(B[:, j]).max() * 2
                assert abs(float(Fraction(C[i, j]) - exact[i][j])) <= bound + abs(float(exact[i][j])) * 2 ** -52
    Ai = rng.integers(-100, 101, (7, 8)).astype(np.float64)
    Bi = rng.integers(-100, 101, (8, 3)).astype(np.float64)
    assert np.array_equal(oracle.ozaki2(Ai, Bi, 12, 12), Ai @ Bi)
    # the modulus bits the library splits between the operands
    tA, tB = oracle.crt_bits(16, 500)
    assert tA + tB == math.floor(oracle.crt_log2P(16) - 10 - 1e-9) and tA == (tA + tB) // 2
    # planted mistake: floor() for the truncation (rounds negative entries away
    # from zero): integer-exactness still holds, but at t = 20 the result is
    # biased and differs from the oracle
    def floored(A, B, t):
        ea = [math.frexp(float(np.abs(r).max()))[1] for r in A]
        fb = [math.frexp(float(np.abs(c).max()))[1] for c in B.T]
        Aa = [[math.floor(Fraction(x) * Fraction(2) ** (t - ea[i])) for x in A[i]] for i in range(A.shape[0])]
        Bb = [[math.floor(Fraction(x) * Fraction(2) ** (t - fb[j])) for x in B[:, j]] for j in range(B.shape[1])]
        return np.array([[float(Fraction(sum(x * y for x, y in zip(Aa[i], Bb[j]))) * Fraction(2) ** (ea[i] + fb[j] - 2 * t))
                          for j in range(B.shape[1])] for i in range(A.shape[0])])
    assert not np.array_equal(floored(A, B, 20), oracle.ozaki2(A, B, 20, 20))
