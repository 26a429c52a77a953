"""GPU tests of Ozaki scheme II (openblas_amd_dgemm_ozaki2): FP64 GEMM from
INT8 GEMMs modulo pairwise coprime moduli and the Chinese remainder theorem.

No reference counterpart; its contract is cblas_dgemm's
(interface/gemm.c:294).  The oracle restates the scheme by its definition
(oracle.ozaki2: operands truncated to tA / tB-bit integers per row of op(A)
and column of op(B), exact integer product, one rounding), pinned on the
CPU against exact rational products (tests/test_oracle.py):
* with B = A^T both operands have the same spread, so the library's split
  of the modulus product's bits is tA = floor(total/2) (oracle.crt_bits);
  for 8, 12 and 17 moduli the device result equals the oracle's bit for bit
  (both round the exact integer product once, to nearest, ties to even: the
  device through a normalized 64-bit head with a sticky bit);
* integer operands are exact (bitwise) in all four transposes and in the
  row-major layout, also when K needs two GEMM launches (K > 131008);
* with the automatic modulus count, uniform data meet the FP64 dmmch bar;
* beta = 0 over NaN, alpha = 0, NaN operands (native path, -2), a spread
  beyond the table (native path, -n with n > 30), argument errors.
"""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def run2(oba, A, B, C, alpha, beta, moduli, ta=0, tb=0, order=102):
    L = oba.lib()
    m, n = C.shape
    k = A.shape[0] if ta else A.shape[1]
    Cout = np.array(C, order="F", copy=True)
    nm = L.openblas_amd_dgemm_ozaki2(order, 112 if ta else 111, 112 if tb else 111, m, n, k, alpha, A.ctypes.data,
                                     A.shape[0], B.ctypes.data, B.shape[0], beta, Cout.ctypes.data, m, moduli)
    return Cout, nm


def ulps(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    sp = np.spacing(np.maximum(np.abs(a), np.abs(b)))
    return float(np.max(np.abs(a - b) / sp))


@pytest.mark.parametrize("nm", [8, 12, 17])
def test_ozaki2_matches_oracle(oba, nm):
    rng = np.random.default_rng(900 + nm)
    m, k = 40, 70
    # entries spread over a few binades, so the truncation matters
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)) * np.exp2(rng.integers(-6, 1, (m, k))))
    C = np.zeros((m, m), order="F")
    # op(B) = A^T: B is A itself, transposed
    got, used = run2(oba, A, A, C, 1.0, 0.0, nm, 0, 1)
    assert used == nm
    tA, tB = oracle.crt_bits(nm, k)
    want = oracle.ozaki2(A, A.T, tA, tB)
    assert np.array_equal(got, want), (nm, tA, tB, ulps(got, want))


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_ozaki2_exact_on_integers(oba, ta, tb):
    rng = np.random.default_rng(910 + 2 * ta + tb)
    m, n, k = 97, 130, 300
    A = np.asfortranarray(rng.integers(-1000, 1001, (k, m) if ta else (m, k)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-1000, 1001, (n, k) if tb else (k, n)).astype(np.float64))
    C = np.full((m, n), np.nan, order="F")
    got, nm = run2(oba, A, B, C, 1.0, 0.0, 0, ta, tb)
    assert 1 <= nm <= 30
    ref = (A.T if ta else A) @ (B.T if tb else B)
    assert np.array_equal(got, ref)


def test_ozaki2_row_major_and_two_k_launches(oba):
    L = oba.lib()
    rng = np.random.default_rng(920)
    # row-major: C (m x n) = A (m x k) B (k x n), all stored row by row
    m, n, k = 33, 45, 61
    A = np.ascontiguousarray(rng.integers(-50, 51, (m, k)).astype(np.float64))
    B = np.ascontiguousarray(rng.integers(-50, 51, (k, n)).astype(np.float64))
    C = np.ascontiguousarray(rng.integers(-9, 10, (m, n)).astype(np.float64))
    C0 = C.copy()
    nm = L.openblas_amd_dgemm_ozaki2(101, 111, 111, m, n, k, 2.0, A.ctypes.data, k, B.ctypes.data, n, -1.0,
                                     C.ctypes.data, n, 0)
    assert nm > 0 and np.array_equal(C, 2.0 * (A @ B) - C0)
    # K = 140000 > 131008: the residues of the second k chunk are added to
    # those of the first before the reconstruction
    m, n, k = 64, 40, 140000
    A = np.asfortranarray(rng.integers(-8, 9, (m, k)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-8, 9, (k, n)).astype(np.float64))
    got, nm = run2(oba, A, B, np.zeros((m, n), order="F"), 1.0, 0.0, 0)
    assert nm > 0
    ref = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    assert np.array_equal(got, ref)


def test_ozaki2_large_tiles_exact(oba):
    """4096 x 4096 (at least one 256 x 256 tile per CU per modulus): the
    large-tile GEMM kernel, integer data, bitwise exact; and with B = A^T
    equal to the oracle on sampled rows (oracle restricted to those rows)."""
    rng = np.random.default_rng(950)
    m = n = 4096
    k = 96
    A = np.asfortranarray(rng.integers(-100, 101, (m, k)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-100, 101, (k, n)).astype(np.float64))
    got, nm = run2(oba, A, B, np.zeros((m, n), order="F"), 1.0, 0.0, 0)
    assert nm > 0 and np.array_equal(got, A @ B)
    Af = np.asfortranarray(rng.uniform(-1, 1, (m, 40)))
    got, nm = run2(oba, Af, Af, np.zeros((m, m), order="F"), 1.0, 0.0, 9, 0, 1)
    assert nm == 9
    tA, tB = oracle.crt_bits(9, 40)
    rows = rng.choice(m, 3, replace=False)
    cols = rng.choice(m, 5, replace=False)
    want = oracle.ozaki2(Af[rows], Af[cols].T, tA, tB)
    assert np.array_equal(got[np.ix_(rows, cols)], want)


def test_ozaki2_fp64_level_on_uniform_data(oba):
    rng = np.random.default_rng(930)
    m, n, k = 300, 200, 500
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    got, nm = run2(oba, A, B, C, 0.7, 1.3, 0)
    assert 14 <= nm <= 18, nm
    assert oracle.dmmch_ratio(0, 0, 0.7, A, B, 1.3, C, got) < 16
    # fewer moduli: fewer integer bits, a larger error, monotone
    errs = []
    for q in (6, 9, 12):
        g, used = run2(oba, A, B, C, 0.7, 1.3, q)
        assert used == q
        errs.append(oracle.dmmch_ratio(0, 0, 0.7, A, B, 1.3, C, g))
    assert errs[0] > errs[1] > errs[2] > 16


def test_ozaki2_edges_and_fallbacks(oba):
    L = oba.lib()
    rng = np.random.default_rng(940)
    m, n, k = 50, 60, 70
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    # alpha = 0: only beta * C, returns 0
    got, nm = run2(oba, A, B, C, 0.0, 0.5, 0)
    assert nm == 0 and np.array_equal(got, 0.5 * C)
    # NaN in A: the native FP64 GEMM runs (NaN propagates as in cblas_dgemm)
    An = A.copy(order="F")
    An[3, 5] = np.nan
    got, nm = run2(oba, An, B, C, 1.0, 0.0, 0)
    assert nm == -2 and np.isnan(got[3]).all() and np.isfinite(np.delete(got, 3, axis=0)).all()
    # a spread beyond 30 moduli: native FP64 GEMM, same bits as cblas_dgemm
    Aw = A.copy(order="F")
    Aw[:, ::2] *= 1e-150
    got, nm = run2(oba, Aw, B, C, 1.0, 0.0, 0)
    assert nm <= -31, nm
    ref = np.array(C, order="F", copy=True)
    L.cblas_dgemm(102, 111, 111, m, n, k, 1.0, Aw.ctypes.data, m, B.ctypes.data, k, 0.0, ref.ctypes.data, m)
    assert np.array_equal(got, ref)
    # argument errors: xerbla numbers of cblas_dgemm, 15 for the modulus count
    oba.last_xerbla()
    assert L.openblas_amd_dgemm_ozaki2(102, 111, 111, m, n, k, 1.0, A.ctypes.data, m - 1, B.ctypes.data, k, 0.0,
                                       C.ctypes.data, m, 0) == -1
    assert oba.last_xerbla() == ("DGEMM", 8)
    assert L.openblas_amd_dgemm_ozaki2(102, 111, 111, m, n, k, 1.0, A.ctypes.data, m, B.ctypes.data, k, 0.0,
                                       C.ctypes.data, m, 31) == -1
    assert oba.last_xerbla() == ("DGEMM", 15)
