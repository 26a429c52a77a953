"""GPU tests of the INT8-slice emulated DGEMM (openblas_amd_dgemm_ozaki).

It has no reference counterpart; its contract is cblas_dgemm's
(interface/gemm.c:294), so it is checked against the oracle GEMM (long
double accumulation) and against the native FP64 path:
* integer operands small enough to be represented by two slices give the
  exact product (the scheme is exact when no digit is truncated);
* for uniform data, s = 8 slices reach the FP64 dmmch bar (ratio < 16);
* the error shrinks with the slice count and the auto choice reports it;
* the configs' phi-scaled inputs u*exp(phi*g) at M=N=K=8192 are compared
  with native FP64 DGEMM on sampled rows, error reported per phi.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def run(oba, A, B, C, alpha, beta, slices, ta=0, tb=0):
    L = oba.lib()
    m, n = C.shape
    k = A.shape[0] if ta else A.shape[1]
    Cout = np.array(C, order="F", copy=True)
    ns = L.openblas_amd_dgemm_ozaki(102, 112 if ta else 111, 112 if tb else 111, m, n, k, alpha, A.ctypes.data,
                                    A.shape[0], B.ctypes.data, B.shape[0], beta, Cout.ctypes.data, m, slices)
    return Cout, ns


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_ozaki_exact_on_small_integers(oba, ta, tb):
    rng = np.random.default_rng(70 + 2 * ta + tb)
    m, n, k = 97, 70, 300
    A = np.asfortranarray(rng.integers(-1000, 1001, (k, m) if ta else (m, k)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-1000, 1001, (n, k) if tb else (k, n)).astype(np.float64))
    C = np.zeros((m, n), order="F")
    got, ns = run(oba, A, B, C, 1.0, 0.0, 3, ta, tb)
    assert ns == 3
    ref = (A.T if ta else A) @ (B.T if tb else B)
    assert np.array_equal(got, ref)


def test_ozaki_fp64_level_on_uniform_data(oba):
    rng = np.random.default_rng(71)
    m, n, k = 130, 129, 257
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    errs = []
    for s in (4, 6, 8):
        got, ns = run(oba, A, B, C, 0.7, 1.3, s)
        assert ns == s
        errs.append(oracle.dmmch_ratio(0, 0, 0.7, A, B, 1.3, C, got))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 16
    # auto slice count reaches the same bar
    got, ns = run(oba, A, B, C, 0.7, 1.3, 0)
    assert 7 <= ns <= 12 and oracle.dmmch_ratio(0, 0, 0.7, A, B, 1.3, C, got) < 16


def test_ozaki_edges_and_errors(oba):
    L = oba.lib()
    A = np.asfortranarray(np.ones((5, 3)))
    C = np.full((5, 4), np.nan, order="F")
    # beta == 0 overwrites NaN; k == 0 scales
    got, _ = run(oba, A, np.ones((3, 4), order="F"), C, 2.0, 0.0, 2)
    assert np.array_equal(got, np.full((5, 4), 6.0))
    got, _ = run(oba, np.zeros((5, 0), order="F"), np.zeros((0, 4), order="F"), np.ones((5, 4), order="F"), 1.0, 3.0, 2)
    assert np.array_equal(got, np.full((5, 4), 3.0))
    # zero rows/columns (exponent 0) are exact zeros
    A2 = np.asfortranarray(np.vstack([np.zeros((1, 3)), np.ones((4, 3))]))
    got, _ = run(oba, A2, np.ones((3, 4), order="F"), np.zeros((5, 4), order="F"), 1.0, 0.0, 2)
    assert np.array_equal(got[0], np.zeros(4)) and np.array_equal(got[1:], np.full((4, 4), 3.0))
    oba.last_xerbla()
    assert L.openblas_amd_dgemm_ozaki(102, 111, 111, 2, 2, 2, 1.0, A.ctypes.data, 5, A.ctypes.data, 3, 0.0,
                                      C.ctypes.data, 5, 13) == -1
    assert oba.last_xerbla() == ("DGEMM", 15)


@pytest.mark.parametrize("phi", [0.5, 2.0, 8.0])
def test_ozaki_8192_phi_inputs(oba, phi):
    """configs: M=N=K=8192, entries u*exp(phi*g); error vs native FP64 DGEMM
    on 32 sampled rows, normwise relative to |A||B|: FP64 level for every
    phi (phi = 8 through the native-GEMM fallback)."""
    L = oba.lib()
    N = 8192
    rng = np.random.default_rng(int(phi * 10))
    A = np.asfortranarray(rng.uniform(-0.5, 0.5, (N, N)) * np.exp(phi * rng.standard_normal((N, N))))
    B = np.asfortranarray(rng.uniform(-0.5, 0.5, (N, N)) * np.exp(phi * rng.standard_normal((N, N))))
    dA, dB, dC = oba.DeviceBuffer.from_array(A), oba.DeviceBuffer.from_array(B), oba.DeviceBuffer(8 * N * N)
    ns = L.openblas_amd_dgemm_ozaki(102, 111, 111, N, N, N, 1.0, dA.ptr, N, dB.ptr, N, 0.0, dC.ptr, N, 0)
    out = dC.to_array(np.empty((N, N), order="F"))
    rows = rng.choice(N, 32, replace=False)
    ref = A[rows] @ B
    gauge = np.abs(A[rows]) @ np.abs(B)
    rel = float(np.max(np.abs(out[rows] - ref) / gauge))
    print(f"phi={phi} slices={ns} max |C-Cref|/(|A||B|) = {rel:.3e}")
    if phi <= 2.0:
        assert 2 <= ns <= 12
    else:
        # the spread needs more than 12 slices: the native FP64 GEMM runs and
        # the return value reports the required count
        assert ns <= -13
    assert rel < 1e-13
    for d in (dA, dB, dC):
        d.free()
