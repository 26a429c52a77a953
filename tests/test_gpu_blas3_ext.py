"""GPU parity of SYRK, GEMMT, SYMM and TRMM through the C ABI.

* Golden vectors (tests/golden/level3_ext_golden.npz, written by the
  reference library, tests/golden/make_golden.py --ext3): every uplo /
  trans / side / diag form, FP64 and FP32, both CBLAS layouts, beta = 0 with
  NaN in C (overwritten, not propagated: driver/level3/level3_syrk.c:150-156,
  interface/gemmt.c SCAL_K), NaN in the triangle of C that must stay
  untouched, NaN in the unreferenced triangle (and, for unit TRMM, on the
  diagonal) of A.  The inputs are multiples of 0.5 and alpha, beta dyadic,
  so every result is exact: the comparison is bitwise.  (The installed
  reference build's row-major cblas_?gemmt misses the result that its own
  utest/test_extensions/test_dgemmt.c requires; the fixture is the
  column-major call of the same logical problem, the definition that test
  checks, and the row-major call here must match it.)
* Larger random problems against an FP64 numpy product, normwise error
  within c*k*eps of |A||B| (FP64 and FP32), device-resident operands.
* TRMM never forms 0 * Inf with the unstored triangle: an Inf in B reaches
  only the rows the triangle couples it to.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_golden as G  # noqa: E402  (fixture layout and poisoning rules)

F64, F32 = ctypes.c_double, ctypes.c_float


@pytest.fixture(scope="module")
def x3():
    return np.load(os.path.join(ROOT, "tests", "golden", "level3_ext_golden.npz"))


def _ld(M, layout):
    return max(M.shape[0] if layout == 102 else M.shape[1], 1)


def _call(L, routine, params, layout, alpha, beta, A, B, C, tag):
    """libopenblas_amd on logical matrices in one layout; the logical result."""
    conv = np.asfortranarray if layout == 102 else np.ascontiguousarray
    A = conv(A)
    B = None if B is None else conv(B)
    C = None if C is None else conv(C)
    F = F64 if tag == "d" else F32
    if routine == "syrk":
        uplo, tr, _ = params
        n = C.shape[0]
        k = A.shape[0] if tr == 112 else A.shape[1]
        getattr(L, f"cblas_{tag}syrk")(layout, uplo, tr, n, k, alpha, A.ctypes.data, _ld(A, layout), beta,
                                         C.ctypes.data, _ld(C, layout))
        return C
    if routine == "gemmt":
        uplo, ta, tb, _ = params
        m = C.shape[0]
        k = A.shape[0] if ta == 112 else A.shape[1]
        getattr(L, f"cblas_{tag}gemmt")(layout, uplo, ta, tb, m, k, alpha, A.ctypes.data, _ld(A, layout),
                                          B.ctypes.data, _ld(B, layout), beta, C.ctypes.data, _ld(C, layout))
        return C
    if routine == "symm":
        side, uplo, _ = params
        m, n = C.shape
        getattr(L, f"cblas_{tag}symm")(layout, side, uplo, m, n, alpha, A.ctypes.data, _ld(A, layout),
                                         B.ctypes.data, _ld(B, layout), beta, C.ctypes.data, _ld(C, layout))
        return C
    side, uplo, tr, diag, _ = params
    m, n = B.shape
    getattr(L, f"cblas_{tag}trmm")(layout, side, uplo, tr, diag, m, n, alpha, A.ctypes.data, _ld(A, layout),
                                     B.ctypes.data, _ld(B, layout))
    return B


@pytest.mark.parametrize("tag", ["d", "s"])
@pytest.mark.parametrize("routine", ["syrk", "gemmt", "symm", "trmm"])
def test_ext3_golden_vectors(oba, x3, routine, tag):
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    A0, B0, C0 = x3["x3_A"], x3["x3_B"], x3["x3_C"]
    n_cases = 0
    for rt, params in G.x3_cases():
        if rt != routine:
            continue
        A, B, C = G.x3_operands(routine, params, A0, B0, C0, dt)
        alpha, beta = G.x3_alpha_beta(routine, params)
        want = x3[G.x3_key(routine, params, tag)].astype(dt)
        for layout in (102, 101):
            got = _call(L, routine, params, layout, alpha, beta, A, None if B is None else B.copy(),
                        None if C is None else C.copy(), tag)
            assert np.array_equal(np.array(got), want, equal_nan=True), (routine, params, layout, tag)
        n_cases += 1
    assert n_cases >= 8


def _rand(rng, shape, dt):
    return rng.standard_normal(shape).astype(dt)


@pytest.mark.parametrize("tag", ["d", "s"])
def test_ext3_large_against_numpy(oba, tag):
    """Multi-tile sizes on device-resident operands."""
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    eps = np.finfo(dt).eps
    rng = np.random.default_rng(33)
    n, k, m2 = 333, 257, 190

    def dev(X):
        return oba.DeviceBuffer.from_array(np.asfortranarray(X))

    def back(d, shape):
        return d.to_array(np.empty(shape, dtype=dt, order="F"))

    # SYRK, lower, trans = T: C = alpha A^T A + beta C
    A = _rand(rng, (k, n), dt)
    C = _rand(rng, (n, n), dt)
    dA, dC = dev(A), dev(C)
    getattr(L, f"cblas_{tag}syrk")(102, 122, 112, n, k, 0.5, dA.ptr, k, -1.5, dC.ptr, n)
    got = back(dC, (n, n))
    ref = 0.5 * A.T.astype(np.float64) @ A.astype(np.float64) - 1.5 * C
    gauge = 0.5 * np.abs(A.T).astype(np.float64) @ np.abs(A) + 1.5 * np.abs(C)
    low = np.tril(np.ones((n, n), bool))
    assert np.max(np.abs(got - ref)[low] / gauge[low]) < 4 * k * eps
    assert np.array_equal(got[~low], C[~low])
    # GEMMT, upper, A^T B^T
    A2 = _rand(rng, (k, n), dt)
    B2 = _rand(rng, (n, k), dt)
    dA2, dB2 = dev(A2), dev(B2)
    dC2 = dev(C)
    getattr(L, f"cblas_{tag}gemmt")(102, 121, 112, 112, n, k, 1.25, dA2.ptr, k, dB2.ptr, n, 0.0, dC2.ptr, n)
    got = back(dC2, (n, n))
    ref = 1.25 * A2.T.astype(np.float64) @ B2.T.astype(np.float64)
    gauge = 1.25 * np.abs(A2.T).astype(np.float64) @ np.abs(B2.T)
    up = np.triu(np.ones((n, n), bool))
    assert np.max(np.abs(got - ref)[up] / gauge[up]) < 4 * k * eps
    assert np.array_equal(got[~up], C[~up])
    # SYMM, right side, upper: C = alpha B A + beta C, A m2 x m2 symmetric
    S = _rand(rng, (m2, m2), dt)
    Sfull = np.triu(S) + np.triu(S, 1).T
    Bs = _rand(rng, (n, m2), dt)
    Cs = _rand(rng, (n, m2), dt)
    dS, dBs, dCs = dev(S), dev(Bs), dev(Cs)
    getattr(L, f"cblas_{tag}symm")(102, 142, 121, n, m2, 0.75, dS.ptr, m2, dBs.ptr, n, 0.5, dCs.ptr, n)
    got = back(dCs, (n, m2))
    ref = 0.75 * Bs.astype(np.float64) @ Sfull.astype(np.float64) + 0.5 * Cs
    gauge = 0.75 * np.abs(Bs).astype(np.float64) @ np.abs(Sfull) + 0.5 * np.abs(Cs)
    assert np.max(np.abs(got - ref) / gauge) < 4 * m2 * eps
    # TRMM, every side/uplo/trans with non-unit diagonal, m x n2 B
    for side in (141, 142):
        for uplo in (121, 122):
            for tr in (111, 112):
                na = n if side == 141 else m2
                T = _rand(rng, (na, na), dt)
                Bt = _rand(rng, (n, m2), dt)
                dT, dBt = dev(T), dev(Bt)
                getattr(L, f"cblas_{tag}trmm")(102, side, uplo, tr, 131, n, m2, -0.5, dT.ptr, na, dBt.ptr, n)
                got = back(dBt, (n, m2))
                Tt = (np.tril(T) if uplo == 122 else np.triu(T)).astype(np.float64)
                op = Tt.T if tr == 112 else Tt
                ref = -0.5 * (op @ Bt if side == 141 else Bt @ op)
                gauge = 0.5 * (np.abs(op) @ np.abs(Bt) if side == 141 else np.abs(Bt) @ np.abs(op))
                assert np.max(np.abs(got - ref) / gauge) < 4 * na * eps, (side, uplo, tr)
                for d in (dT, dBt):
                    d.free()
    for d in (dA, dC, dA2, dB2, dC2, dS, dBs, dCs):
        d.free()


def test_trmm_inf_meets_no_unstored_zero(oba):
    """B := L B (left, lower, no transpose): row i of the result combines
    rows k <= i of B.  An Inf in B(60, :) must not reach rows above 60 as a
    NaN (that would be 0 * Inf with the unstored upper triangle)."""
    L = oba.lib()
    n = 150
    rng = np.random.default_rng(4)
    T = np.asfortranarray(np.tril(rng.standard_normal((n, n))) + 5 * np.eye(n))
    B = np.asfortranarray(rng.standard_normal((n, 3)))
    B[60, 1] = np.inf
    L.cblas_dtrmm(102, 141, 122, 111, 131, n, 3, 1.0, T.ctypes.data, n, B.ctypes.data, n)
    assert np.all(np.isfinite(B[:60])), "rows above the Inf must stay finite"
    assert np.all(np.isfinite(B[:, [0, 2]]))
    assert np.all(~np.isfinite(B[60:, 1]))
