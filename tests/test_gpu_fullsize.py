"""GPU parity tests at the benchmark configurations' full sizes, through the C
ABI, with size-independent properties.

* Mixed-precision dense solve, N = 16384: A = Q1 diag(s) Q2 (Q1, Q2 products
  of 8 Householder reflectors with Gaussian vectors, s geometric from 1 to
  1/kappa), kappa in {1e2, 1e6, 1e10}, NRHS in {1, 16}, b = A x_true with
  Gaussian x_true.  dsgesv_ and dgesv_ both reach normwise backward error
  <= 1e-14; the forward error against x_true is within c*kappa*eps64
  (LAPACK dget04's ratio, threshold 30); ITER follows the ddrvab.f:354-364
  rule: ITER >= 0 requires ||b - A x|| / (||A|| ||x|| eps) < sqrt(N), ITER < 0
  (the FP64 fallback of dsgesv.f:405-423) requires it < 30.  For kappa <= 1e6
  the refinement must converge (ITER >= 0); for kappa = 1e10 (kappa * eps32
  ~ 600) it may converge slowly or give up (ITER = -31), depending on the
  FP32 factor's rounding.
* Batched small problems, 10,000 problems with n ~ U{8..256}, FP64 and FP32:
  every GEMM passes the ctest dmmch ratio, every getrf the dget01
  reconstruction ratio, every getrs the dget02 residual ratio, every potrf
  the dpot01 ratio; 200 random problems equal the oracle (FP64 ipiv
  bit-exact).
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


@pytest.fixture(scope="module", params=[1e2, 1e6, 1e10], ids=["k1e2", "k1e6", "k1e10"])
def system16k(request, oba):
    """The N = 16384 system of one kappa on the device, its host copy and
    ||A||_inf."""
    import bench
    L = oba.lib()
    n = 16384
    print(f"\n[c3] building A n={n} kappa={request.param:g}", flush=True)
    dA, _ = bench.make_system(L, n, request.param, 1, seed=2024)
    A = np.empty((n, n), order="F")
    dA.to_array(A)
    anorm = float(np.abs(A).sum(axis=1).max())
    yield n, request.param, dA, A, anorm
    dA.free()


def _solve_and_check(oba, L, n, kappa, dA, A, anorm, nrhs, mixed):
    rng = np.random.default_rng(int(np.log10(kappa)) * 100 + nrhs)
    xt = np.asfortranarray(rng.standard_normal((n, nrhs)))
    dxt = oba.DeviceBuffer.from_array(xt)
    db = oba.DeviceBuffer(8 * n * nrhs)
    # b = A x_true on the device (FP64 GEMM)
    L.cblas_dgemm(102, 111, 111, n, nrhs, n, 1.0, dA.ptr, n, dxt.ptr, n, 0.0, db.ptr, n)
    b = db.to_array(np.empty((n, nrhs), order="F"))
    dW = oba.DeviceBuffer(8 * n * n)   # the solvers may overwrite A
    L.openblas_amd_memcpy(dW.ptr, dA.ptr, 8 * n * n, 2)
    dpiv = oba.DeviceBuffer(4 * n)
    info, it = ctypes.c_int(-99), ctypes.c_int(-99)
    if mixed:
        dX = oba.DeviceBuffer(8 * n * nrhs)
        L.dsgesv_(oba.ip(n), oba.ip(nrhs), dW.ptr, oba.ip(n), dpiv.ptr, db.ptr, oba.ip(n), dX.ptr, oba.ip(n), None,
                  None, ctypes.byref(it), ctypes.byref(info))
    else:
        dX = db
        L.dgesv_(oba.ip(n), oba.ip(nrhs), dW.ptr, oba.ip(n), dpiv.ptr, dX.ptr, oba.ip(n), ctypes.byref(info))
    x = dX.to_array(np.empty((n, nrhs), order="F"))
    # residual r = b - A x on the device
    dR = oba.DeviceBuffer.from_array(b)
    L.cblas_dgemm(102, 111, 111, n, nrhs, n, -1.0, dA.ptr, n, dX.ptr, n, 1.0, dR.ptr, n)
    r = dR.to_array(np.empty((n, nrhs), order="F"))
    for d in (dxt, db, dW, dpiv, dR) + ((dX,) if mixed else ()):
        d.free()
    return x, xt, b, r, info.value, it.value


@pytest.mark.parametrize("nrhs", [1, 16])
@pytest.mark.parametrize("mixed", [True, False], ids=["dsgesv", "dgesv"])
def test_c3_solve_16384(oba, system16k, nrhs, mixed):
    L = oba.lib()
    n, kappa, dA, A, anorm = system16k
    print(f"[c3] {'dsgesv' if mixed else 'dgesv'} n={n} kappa={kappa:g} nrhs={nrhs}", flush=True)
    x, xt, b, r, info, it = _solve_and_check(oba, L, n, kappa, dA, A, anorm, nrhs, mixed)
    assert info == 0
    xn = np.abs(x).max(axis=0)
    rn = np.abs(r).max(axis=0)
    bwd = rn / (anorm * xn + np.abs(b).max(axis=0))
    # FP64-level backward error; refinement that converges slowly (kappa =
    # 1e10) stops at dsgesv.f's own test, ||r|| <= ||x|| ||A|| eps sqrt(N)
    slow = mixed and kappa > 1e6
    assert np.all(bwd <= (np.sqrt(n) * EPS64 if slow else 1e-14)), bwd.max()
    fwd = np.abs(x - xt).max(axis=0) / np.abs(xt).max(axis=0)
    assert np.all(fwd <= 30 * kappa * EPS64), (fwd.max(), kappa)
    if mixed:
        ratio = float((rn / (anorm * xn * EPS64)).max())   # dget08's test ratio
        # kappa <= 1e6 is well inside what an FP32 LU can refine.  At kappa =
        # 1e10 (kappa * eps32 ~ 600) refinement may still converge or may
        # give up after ITERMAX sweeps (ITER = -31, FP64 fallback): which one
        # depends on the FP32 factor's rounding, so the outcome is not fixed,
        # and lapack-netlib/TESTING/LIN/ddrvab.f:354-364 checks the solution
        # by the branch taken, as below.
        if kappa <= 1e6:
            assert it >= 0
        assert it >= 0 or it == -31
        if it >= 0:
            assert ratio < np.sqrt(n)
        else:
            assert ratio < 30


# ---- batched, 10,000 problems -----------------------------------------------------
def _ptrs(base, off, es):
    return (ctypes.c_void_p * len(off))(*[base + int(o) * es for o in off])


def _i32(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32))


def _pi(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


@pytest.fixture(scope="module")
def batch10k():
    rng = np.random.default_rng(5)
    ns = rng.integers(8, 257, 10000).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(ns * ns)[:-1]])
    voff = np.concatenate([[0], np.cumsum(ns)[:-1]])
    return ns, off, voff


def _mat(flat, o, n):
    return flat[o:o + n * n].reshape(n, n, order="F")


@pytest.mark.parametrize("tag", ["d", "s"])
def test_c5_batched_10000(oba, batch10k, tag):
    L = oba.lib()
    ns, off, voff = batch10k
    count = len(ns)
    dt = np.float64 if tag == "d" else np.float32
    eps = EPS64 if tag == "d" else EPS32
    es = np.dtype(dt).itemsize
    total, vtotal = int((ns * ns).sum()), int(ns.sum())
    rng = np.random.default_rng(55 if tag == "d" else 56)
    A = rng.standard_normal(total).astype(dt)
    B = rng.standard_normal(total).astype(dt)
    n32 = _i32(ns)
    print(f"\n[c5] {tag} count={count} elements={total}", flush=True)
    dA, dB, dC = oba.DeviceBuffer.from_array(A), oba.DeviceBuffer.from_array(B), oba.DeviceBuffer(es * total)
    pA, pB, pC = (_ptrs(d.ptr, off, es) for d in (dA, dB, dC))
    ct = ctypes.c_double if tag == "d" else ctypes.c_float
    one, zero = (ct * count)(*([1.0] * count)), (ct * count)(*([0.0] * count))
    tn, tt, gs = _i32(np.full(count, 111)), _i32(np.full(count, 112)), _i32(np.ones(count))
    gemm_b = L.cblas_dgemm_batch if tag == "d" else L.cblas_sgemm_batch
    # (i) GEMM: C = A B, every problem passes the dmmch ratio
    gemm_b(102, _pi(tn), _pi(tn), _pi(n32), _pi(n32), _pi(n32), one, pA, _pi(n32), pB, _pi(n32), zero, pC,
           _pi(n32), count, _pi(gs))
    C = dC.to_array(np.empty(total, dt))
    worst = 0.0
    for i in range(count):
        n, o = int(ns[i]), int(off[i])
        Ai, Bi = _mat(A, o, n).astype(np.float64), _mat(B, o, n).astype(np.float64)
        err = np.abs(_mat(C, o, n) - Ai @ Bi)
        gauge = np.abs(Ai) @ np.abs(Bi)
        worst = max(worst, float((err / (eps * gauge)).max()))
    assert worst < 16, worst
    # (ii) getrf: dget01 ratio ||P A - L U|| / (n ||A|| eps) for every problem
    dW = oba.DeviceBuffer.from_array(A)
    pW = _ptrs(dW.ptr, off, es)
    dpiv = oba.DeviceBuffer(4 * vtotal)
    ppiv = _ptrs(dpiv.ptr, voff, 4)
    info = np.full(count, -99, np.int32)
    fr = L.openblas_amd_dgetrf_batch if tag == "d" else L.openblas_amd_sgetrf_batch
    assert fr(_pi(n32), _pi(n32), pW, _pi(n32), ppiv, _pi(info), count) == 0
    assert np.all(info == 0)
    LU = dW.to_array(np.empty(total, dt))
    piv = dpiv.to_array(np.empty(vtotal, np.int32))
    worst = 0.0
    for i in range(count):
        n, o, v = int(ns[i]), int(off[i]), int(voff[i])
        F = _mat(LU, o, n).astype(np.float64)
        Lf = np.tril(F, -1) + np.eye(n)
        Uf = np.triu(F)
        perm = list(range(n))
        for k, p in enumerate(piv[v:v + n].tolist()):
            assert k < p <= n or p == k + 1
            perm[k], perm[p - 1] = perm[p - 1], perm[k]
        PA = _mat(A, o, n).astype(np.float64)[perm]
        worst = max(worst, float(np.abs(PA - Lf @ Uf).max() / (n * np.abs(PA).sum(0).max() * eps)))
    assert worst < 30, worst
    # 200 random problems equal the oracle's getf2
    for i in np.random.default_rng(57).choice(count, 200, replace=False):
        n, o, v = int(ns[i]), int(off[i]), int(voff[i])
        Fo, po, io = oracle.getf2(_mat(A, o, n))
        if tag == "d":
            assert np.array_equal(piv[v:v + n], po), i
            assert np.abs(_mat(LU, o, n) - Fo).max() <= 50 * n * eps * max(1.0, float(np.abs(Fo).max())), i
        else:   # FP32 near-ties may order differently: bound |l_ij| <= 1 instead
            assert np.abs(np.tril(_mat(LU, o, n), -1)).max() <= 1.0 + 1e-6
    # (iii) getrs: dget02 ratio ||b - A x|| / (||A|| ||x|| eps) for every problem
    bv = rng.standard_normal(vtotal).astype(dt)
    dx = oba.DeviceBuffer.from_array(bv)
    pX = _ptrs(dx.ptr, voff, es)
    fs = L.openblas_amd_dgetrs_batch if tag == "d" else L.openblas_amd_sgetrs_batch
    one_rhs = _i32(np.ones(count))
    assert fs(b"N", _pi(n32), _pi(one_rhs), pW, _pi(n32), ppiv, pX, _pi(n32), _pi(info), count) == 0
    X = dx.to_array(np.empty(vtotal, dt))
    worst = 0.0
    for i in range(count):
        n, o, v = int(ns[i]), int(off[i]), int(voff[i])
        Ai = _mat(A, o, n).astype(np.float64)
        xi, bi = X[v:v + n].astype(np.float64), bv[v:v + n].astype(np.float64)
        worst = max(worst, float(np.abs(bi - Ai @ xi).max() / (np.abs(Ai).sum(1).max() * np.abs(xi).max() * eps)))
    assert worst < 30, worst   # dget02 ratio, LAPACK threshold 30
    # (iv) potrf on S = B B^T + n I: dpot01 ratio ||S - L L^T|| / (n ||S|| eps)
    Sh = np.zeros(total, dt)
    for i in range(count):
        n, o = int(ns[i]), int(off[i])
        Sh[o:o + n * n:n + 1] = n
    dS = oba.DeviceBuffer.from_array(Sh)
    pS = _ptrs(dS.ptr, off, es)
    gemm_b(102, _pi(tn), _pi(tt), _pi(n32), _pi(n32), _pi(n32), one, pB, _pi(n32), pB, _pi(n32), one, pS, _pi(n32),
           count, _pi(gs))
    S = dS.to_array(np.empty(total, dt))
    fp = L.openblas_amd_dpotrf_batch if tag == "d" else L.openblas_amd_spotrf_batch
    assert fp(b"L", _pi(n32), pS, _pi(n32), _pi(info), count) == 0
    assert np.all(info == 0)
    Fh = dS.to_array(np.empty(total, dt))
    worst = 0.0
    for i in range(count):
        n, o = int(ns[i]), int(off[i])
        Si = _mat(S, o, n).astype(np.float64)
        Lf = np.tril(_mat(Fh, o, n).astype(np.float64))
        worst = max(worst, float(np.abs(np.tril(Si - Lf @ Lf.T)).max() / (n * np.abs(Si).sum(0).max() * eps)))
        assert np.array_equal(np.triu(_mat(Fh, o, n), 1), np.triu(_mat(S, o, n), 1)), i   # upper untouched
    assert worst < 30, worst
    for i in np.random.default_rng(58).choice(count, 200, replace=False):
        n, o = int(ns[i]), int(off[i])
        Fo, io = oracle.potf2(1, _mat(S, o, n))
        assert io == 0
        assert np.abs(np.tril(_mat(Fh, o, n)) - np.tril(Fo)).max() <= 50 * n * eps * float(np.abs(Fo).max()), i
    for d in (dA, dB, dC, dW, dpiv, dx, dS):
        d.free()
