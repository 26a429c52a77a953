"""GPU parity tests for the level-3 BLAS entry points (cblas_?gemm, ?gemm_,
cblas_?trsm, cblas_?gemm_batch), called through the C ABI of
libopenblas_amd.so and checked against the C oracle, the reference
library's golden vectors and, at full size, exact-arithmetic properties.

Tolerances: FP64/FP32 GEMM results must pass the ctest dmmch test ratio
(ctest/c_dblat3c.c:2956, threshold 16 from ctest/din3); TRSM results must
match the oracle to 50*eps*max|X| and multiply back (c_dblat3c.c:1860-1895).
"""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps}
TR = {0: 111, 1: 112}


def gemm_ratio(ta, tb, alpha, A, B, beta, Cin, Cout, dt):
    """dmmch ratio computed with numpy (any size)."""
    A64, B64, C64 = A.astype(np.float64), B.astype(np.float64), np.nan_to_num(Cin.astype(np.float64))
    oA = A64.T if ta else A64
    oB = B64.T if tb else B64
    ref = alpha * (oA @ oB) + beta * C64
    g = abs(alpha) * (np.abs(oA) @ np.abs(oB)) + abs(beta) * np.abs(C64)
    err = np.abs(ref - Cout.astype(np.float64)) / EPS[dt]
    err = np.where(g != 0, err / np.where(g != 0, g, 1), err)
    return float(err.max()) if err.size else 0.0


@pytest.mark.parametrize("tag", ["d", "s"])
def test_gemm_golden_vectors(oba, golden, tag):
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    keys = sorted(k[:-5] for k in golden.files if k.startswith(f"gemm{tag}_") and k.endswith("_meta"))
    assert len(keys) == 48
    for key in keys:
        ta, tb, m, n, k, lda, ldb, ldc = (int(x) for x in golden[key + "_meta"])
        alpha, beta = (float(x) for x in golden[key + "_ab"])
        A, B = golden[key + "_A"], golden[key + "_B"]
        C = np.array(golden[key + "_Cin"], order="F")
        f = L.cblas_dgemm if tag == "d" else L.cblas_sgemm
        f(102, TR[ta], TR[tb], m, n, k, alpha, A.ctypes.data, lda, B.ctypes.data, ldb, beta, C.ctypes.data, ldc)
        got = C[:m]
        assert not np.isnan(got).any(), key
        Cin = golden[key + "_Cin"][:m]
        Ar, Br = (k if ta else m), (n if tb else k)
        assert gemm_ratio(ta, tb, alpha, A[:Ar], B[:Br], beta, Cin, got, dt) < 16, key
        # rows beyond m (the lda padding of C) are untouched
        assert np.array_equal(np.nan_to_num(C[m:]), np.nan_to_num(golden[key + "_Cin"][m:])), key


def test_dgemm_ctest_grid(oba):
    """ctest/din3 grid: N in {1,2,3,5,7,9,35}, alpha in {0,1,.7}, beta in
    {0,1,1.3}, 4 transposes, both layouts, data from the dbeg generator."""
    L = oba.lib()
    sizes = [1, 2, 3, 5, 7, 9, 35]
    shapes = [(m, n, k) for m in sizes for n in sizes for k in sizes if (m + 2 * n + 3 * k) % 5 == 0 or m == n == k]
    worst = 0.0
    bad = []
    oracle.dbeg(True)
    for (m, n, k) in shapes:
        for ta in (0, 1):
            for tb in (0, 1):
                Ar, Ac = (k, m) if ta else (m, k)
                Br, Bc = (n, k) if tb else (k, n)
                A = np.array([[oracle.dbeg() for _ in range(Ac)] for _ in range(Ar + 1)], order="F")
                B = np.array([[oracle.dbeg() for _ in range(Bc)] for _ in range(Br + 1)], order="F")
                C0 = np.array([[oracle.dbeg() for _ in range(n)] for _ in range(m + 1)], order="F")
                for alpha in (0.0, 1.0, 0.7):
                    for beta in (0.0, 1.0, 1.3):
                        C = C0.copy(order="F")
                        L.cblas_dgemm(102, TR[ta], TR[tb], m, n, k, alpha, A.ctypes.data, Ar + 1, B.ctypes.data,
                                      Br + 1, beta, C.ctypes.data, m + 1)
                        r = oracle.dmmch_ratio(ta, tb, alpha, A[:Ar], B[:Br], beta, C0[:m], C[:m])
                        worst = max(worst, r)
                        if r >= 16:
                            bad.append(("col", m, n, k, ta, tb, alpha, beta, r))
                        # row-major: C^T = op(B)^T op(A)^T on the same memory
                        # fresh C-ordered copies (ascontiguousarray may return a view of C0)
                        Cr = np.array(C0[:m], order="C", copy=True)
                        Ac_ = np.array(A[:Ar], order="C", copy=True)
                        Bc_ = np.array(B[:Br], order="C", copy=True)
                        L.cblas_dgemm(101, TR[ta], TR[tb], m, n, k, alpha, Ac_.ctypes.data, Ac, Bc_.ctypes.data, Bc,
                                      beta, Cr.ctypes.data, n)
                        r2 = oracle.dmmch_ratio(ta, tb, alpha, A[:Ar], B[:Br], beta, C0[:m], Cr)
                        worst = max(worst, r2)
                        if r2 >= 16:
                            bad.append(("row", m, n, k, ta, tb, alpha, beta, r2))
    assert worst < 16, (worst, len(bad), bad[:12])


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_gemm_odd_ld_and_edges(oba, dt):
    L = oba.lib()
    f = L.cblas_dgemm if dt == np.float64 else L.cblas_sgemm
    rng = np.random.default_rng(11)
    # odd leading dimensions force the scalar-load kernels
    for (m, n, k, pad) in ((129, 131, 133, 1), (200, 64, 300, 3), (1, 300, 17, 5)):
        A = np.asfortranarray(rng.uniform(-1, 1, (m + pad, k)).astype(dt))
        B = np.asfortranarray(rng.uniform(-1, 1, (k + pad, n)).astype(dt))
        C = np.asfortranarray(rng.uniform(-1, 1, (m + pad, n)).astype(dt))
        C0 = C.copy(order="F")
        f(102, 111, 111, m, n, k, 0.7, A.ctypes.data, m + pad, B.ctypes.data, k + pad, 1.3, C.ctypes.data, m + pad)
        assert gemm_ratio(0, 0, 0.7, A[:m], B[:k], 1.3, C0[:m], C[:m], dt) < 16
    # alpha == 0: C <- beta*C exactly; beta == 0 with NaN in C: overwritten
    C = np.asfortranarray(rng.uniform(-1, 1, (40, 30)).astype(dt)); C0 = C.copy(order="F")
    A = np.asfortranarray(rng.uniform(-1, 1, (40, 20)).astype(dt)); B = np.asfortranarray(rng.uniform(-1, 1, (20, 30)).astype(dt))
    f(102, 111, 111, 40, 30, 20, 0.0, A.ctypes.data, 40, B.ctypes.data, 20, 2.0, C.ctypes.data, 40)
    assert np.array_equal(C, (dt(2.0) * C0).astype(dt))
    C[:] = np.nan
    f(102, 111, 111, 40, 30, 20, 1.0, A.ctypes.data, 40, B.ctypes.data, 20, 0.0, C.ctypes.data, 40)
    assert not np.isnan(C).any()
    # k == 0 with beta == 0 writes exact zeros
    C[:] = np.nan
    f(102, 111, 111, 40, 30, 0, 1.0, A.ctypes.data, 40, B.ctypes.data, 1, 0.0, C.ctypes.data, 40)
    assert np.array_equal(C, np.zeros_like(C))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_gemm_few_columns(oba, dt):
    """The few-column path (N <= 4, no transposes, large M and K: split-k
    matrix-vector products with an ordered combine): dmmch ratio for every
    N in 1..4, odd leading dimensions, beta == 0 over NaN."""
    L = oba.lib()
    f = L.cblas_dgemm if dt == np.float64 else L.cblas_sgemm
    rng = np.random.default_rng(17)
    m, k, pad = 3001, 2600, 3
    A = np.asfortranarray(rng.uniform(-1, 1, (m + pad, k)).astype(dt))
    for n in (1, 2, 3, 4):
        B = np.asfortranarray(rng.uniform(-1, 1, (k + 1, n)).astype(dt))
        C = np.asfortranarray(rng.uniform(-1, 1, (m + pad, n)).astype(dt))
        C0 = C.copy(order="F")
        f(102, 111, 111, m, n, k, -0.8, A.ctypes.data, m + pad, B.ctypes.data, k + 1, 1.1, C.ctypes.data, m + pad)
        assert gemm_ratio(0, 0, -0.8, A[:m], B[:k], 1.1, C0[:m], C[:m], dt) < 16, n
        C[:] = np.nan
        f(102, 111, 111, m, n, k, 1.0, A.ctypes.data, m + pad, B.ctypes.data, k + 1, 0.0, C.ctypes.data, m + pad)
        assert not np.isnan(C[:m]).any()
        assert gemm_ratio(0, 0, 1.0, A[:m], B[:k], 0.0, np.zeros((m, n), dt), C[:m], dt) < 16, n


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_gemm_split_k(oba, dt):
    """The deep-K paths for few tiles: 300 x 250 x 3001 (20 tiles: K split
    across workgroups, partials summed by a second kernel), 1000 x 1000 x
    1100 (256 tiles: K sliced over the 4 wave groups of one workgroup, K not
    a multiple of 4 slabs, so every group's range is guarded, M/N edges
    inside a tile) and 1024^3 (sliced, unguarded); 65 x 1030 x 259 and
    1 x 7 x 333 stay on the plain tile kernel.  dmmch ratio for all four
    transpose forms, odd leading dimensions, rows of C beyond M untouched,
    beta == 0 over NaN."""
    L = oba.lib()
    f = L.cblas_dgemm if dt == np.float64 else L.cblas_sgemm
    rng = np.random.default_rng(23)
    for (m, n, k) in ((300, 250, 3001), (1000, 1000, 1100), (1024, 1024, 1024), (65, 1030, 259), (1, 7, 333)):
      for ta in (0, 1):
        for tb in (0, 1):
            A = np.asfortranarray(rng.uniform(-1, 1, (k + 1, m + 1) if ta else (m + 1, k)).astype(dt))
            B = np.asfortranarray(rng.uniform(-1, 1, (n + 1, k) if tb else (k + 1, n)).astype(dt))
            C = np.asfortranarray(rng.uniform(-1, 1, (m + 3, n)).astype(dt))
            C0 = C.copy(order="F")
            f(102, TR[ta], TR[tb], m, n, k, 0.9, A.ctypes.data, A.shape[0], B.ctypes.data, B.shape[0], -0.4,
              C.ctypes.data, m + 3)
            Av = A[:k, :m] if ta else A[:m, :k]
            Bv = B[:n, :k] if tb else B[:k, :n]
            assert gemm_ratio(ta, tb, 0.9, Av, Bv, -0.4, C0[:m], C[:m], dt) < 16, (m, n, k, ta, tb)
            assert np.array_equal(C[m:], C0[m:])   # rows beyond m untouched
    A = np.asfortranarray(rng.uniform(-1, 1, (1024, 1024)).astype(dt))
    C = np.full((1024, 1024), np.nan, dt, order="F")
    f(102, 111, 111, 1024, 1024, 1024, 1.0, A.ctypes.data, 1024, A.ctypes.data, 1024, 0.0, C.ctypes.data, 1024)
    assert not np.isnan(C).any()
    assert gemm_ratio(0, 0, 1.0, A, A, 0.0, np.zeros_like(C), C, dt) < 16
    # device-resident operands with odd leading dimensions (scalar loads on
    # the sliced path), twice: the fixed summation order gives equal bits
    m, n, k = 1000, 999, 1500
    A = np.asfortranarray(rng.uniform(-1, 1, (m + 1, k)).astype(dt))
    B = np.asfortranarray(rng.uniform(-1, 1, (k + 3, n)).astype(dt))
    C0 = np.asfortranarray(rng.uniform(-1, 1, (m + 1, n)).astype(dt))
    dA, dB = oba.DeviceBuffer.from_array(A), oba.DeviceBuffer.from_array(B)
    outs = []
    for _ in range(2):
        dC = oba.DeviceBuffer.from_array(C0)
        f(102, 111, 111, m, n, k, -1.5, dA.ptr, m + 1, dB.ptr, k + 3, 0.25, dC.ptr, m + 1)
        outs.append(dC.to_array(np.empty_like(C0)))
        dC.free()
    dA.free(); dB.free()
    assert np.array_equal(outs[0], outs[1])
    assert gemm_ratio(0, 0, -1.5, A[:m], B[:k], 0.25, C0[:m], outs[0][:m], dt) < 16
    assert np.array_equal(outs[0][m:], C0[m:])


def test_fortran_gemm_matches_cblas(oba):
    L = oba.lib()
    rng = np.random.default_rng(12)
    A = np.asfortranarray(rng.uniform(-1, 1, (50, 40))); B = np.asfortranarray(rng.uniform(-1, 1, (30, 40)))
    C1 = np.asfortranarray(rng.uniform(-1, 1, (50, 30))); C2 = C1.copy(order="F")
    al, be = ctypes.c_double(0.3), ctypes.c_double(-1.0)
    L.dgemm_(b"n", b"t", oba.ip(50), oba.ip(30), oba.ip(40), ctypes.byref(al), A.ctypes.data, oba.ip(50),
             B.ctypes.data, oba.ip(30), ctypes.byref(be), C1.ctypes.data, oba.ip(50))
    L.cblas_dgemm(102, 111, 112, 50, 30, 40, 0.3, A.ctypes.data, 50, B.ctypes.data, 30, -1.0, C2.ctypes.data, 50)
    assert np.array_equal(C1, C2)


@pytest.mark.parametrize("dt,lim", [(np.float64, 8), (np.float32, 2)])
def test_gemm_8192_exact_integers(oba, dt, lim):
    """configs: M=N=K=8192, all four transposes, device-resident operands.
    Integer entries in [-lim, lim] make every partial sum exact in the
    working precision, so the result must equal the exact product bitwise."""
    L = oba.lib()
    f = L.cblas_dgemm if dt == np.float64 else L.cblas_sgemm
    N = 8192
    rng = np.random.default_rng(13)
    A = np.asfortranarray(rng.integers(-lim, lim + 1, (N, N)).astype(dt))
    B = np.asfortranarray(rng.integers(-lim, lim + 1, (N, N)).astype(dt))
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    dA, dB = oba.DeviceBuffer.from_array(A), oba.DeviceBuffer.from_array(B)
    dC = oba.DeviceBuffer(A.nbytes)
    out = np.empty((N, N), dtype=dt, order="F")
    rows = rng.choice(N, 64, replace=False)
    for ta in (0, 1):
        for tb in (0, 1):
            f(102, TR[ta], TR[tb], N, N, N, 1.0, dA.ptr, N, dB.ptr, N, 0.0, dC.ptr, N)
            dC.to_array(out)
            oA = A64.T if ta else A64
            oB = B64.T if tb else B64
            ref = oA[rows] @ oB
            assert np.array_equal(out[rows].astype(np.float64), ref), (ta, tb)
    for d in (dA, dB, dC):
        d.free()


@pytest.mark.parametrize("tag", ["d", "s"])
def test_trsm_golden_vectors(oba, golden, tag):
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    keys = sorted(k[:-5] for k in golden.files if k.startswith(f"trsm{tag}_") and k.endswith("_meta"))
    assert len(keys) == 32
    for key in keys:
        side, uplo, tr, diag, m, n = (int(x) for x in golden[key + "_meta"])
        A = golden[key + "_A"]
        X = np.array(golden[key + "_B"], order="F")
        f = L.cblas_dtrsm if tag == "d" else L.cblas_strsm
        f(102, 142 if side else 141, 122 if uplo else 121, TR[tr], 131 if diag else 132, m, n, 0.7,
          A.ctypes.data, A.shape[0], X.ctypes.data, m)
        ref = golden[key + "_X"]
        orc = oracle.trsm(side, uplo, tr, diag, 0.7, A, golden[key + "_B"])
        tol = 50 * EPS[dt] * max(1.0, float(np.abs(ref).max()))
        assert np.abs(X - ref).max() <= tol, key
        assert np.abs(X - orc).max() <= tol, key


@pytest.mark.parametrize("tag", ["d", "s"])
@pytest.mark.parametrize("m", [200, 256, 300])
def test_trsm_left_multiblock_leaves(oba, tag, m):
    """Left-side solves whose recursion leaves hold 2-4 blocks of 64 rows
    (the multi-block leaf kernel): every uplo / trans / diag, the unstored
    triangle NaN-poisoned, against an FP64 solve of the stored triangle."""
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    rng = np.random.default_rng(m)
    nr = 37
    f = L.cblas_dtrsm if tag == "d" else L.cblas_strsm
    for uplo in (121, 122):
        for tr in (111, 112):
            for diag in (131, 132):
                T = rng.uniform(-1, 1, (m, m)) / np.sqrt(m)   # well conditioned with a unit diagonal too
                T[np.diag_indices(m)] = 4 + rng.uniform(0, 1, m)
                keep = np.tril(np.ones((m, m), bool)) if uplo == 122 else np.triu(np.ones((m, m), bool))
                Tw = np.where(keep, T, 0.0)
                if diag == 132:
                    Tw[np.diag_indices(m)] = 1.0
                A = np.where(keep, T, np.nan).astype(dt, order="F")
                if diag == 132:
                    A[np.diag_indices(m)] = np.nan   # unit: the diagonal is not read either
                B = np.asfortranarray(rng.uniform(-1, 1, (m, nr)).astype(dt))
                X = B.copy(order="F")
                f(102, 141, uplo, tr, diag, m, nr, 1.5, A.ctypes.data, m, X.ctypes.data, m)
                op = Tw.T if tr == 112 else Tw
                ref = np.linalg.solve(op, 1.5 * B.astype(np.float64))
                assert not np.isnan(X).any(), (uplo, tr, diag)
                err = np.abs(X - ref).max() / np.abs(ref).max()
                assert err < 64 * m * EPS[dt], (uplo, tr, diag, err)


def test_trsm_8192x1024_multiply_back(oba):
    """configs: dtrsm side=L uplo=L trans=N diag=N, M=8192, NRHS=1024, the
    triangle U[-1,1] off the diagonal and n+|U(0,1)| on it; the strictly
    upper part is NaN-poisoned to prove it is never read."""
    L = oba.lib()
    M, NR = 8192, 1024
    rng = np.random.default_rng(14)
    A = rng.uniform(-1, 1, (M, M))
    A[np.diag_indices(M)] = M + rng.uniform(0, 1, M)
    Lw = np.tril(A)
    A[np.triu_indices(M, 1)] = np.nan
    A = np.asfortranarray(A)
    B = np.asfortranarray(rng.uniform(-1, 1, (M, NR)))
    dA, dB = oba.DeviceBuffer.from_array(A), oba.DeviceBuffer.from_array(B)
    L.cblas_dtrsm(102, 141, 122, 111, 131, M, NR, 1.0, dA.ptr, M, dB.ptr, M)
    X = dB.to_array(np.empty((M, NR), order="F"))
    assert not np.isnan(X).any()
    res = np.abs(Lw @ X - B).max() / (np.abs(Lw).sum(1).max() * np.abs(X).max() * EPS[np.float64] * M)
    assert res < 1.0, res
    # FP32 variant on the same data
    Af = np.asfortranarray(A.astype(np.float32)); Bf = np.asfortranarray(B.astype(np.float32))
    dAf, dBf = oba.DeviceBuffer.from_array(Af), oba.DeviceBuffer.from_array(Bf)
    L.cblas_strsm(102, 141, 122, 111, 131, M, NR, 1.0, dAf.ptr, M, dBf.ptr, M)
    Xf = dBf.to_array(np.empty((M, NR), dtype=np.float32, order="F")).astype(np.float64)
    Lf = np.tril(Af.astype(np.float64))
    Lf[np.isnan(Lf)] = 0
    res = np.abs(Lf @ Xf - Bf.astype(np.float64)).max() / (np.abs(Lf).sum(1).max() * np.abs(Xf).max()
                                                             * EPS[np.float32] * M)
    assert res < 1.0, res
    for d in (dA, dB, dAf, dBf):
        d.free()


def test_dgemm_batch_groups(oba):
    L = oba.lib()
    rng = np.random.default_rng(15)
    groups = [(111, 111, 5, 7, 3, 2), (112, 111, 33, 17, 9, 3), (111, 112, 64, 65, 66, 1)]
    tas, tbs, ms, ns, ks, gs, als, bes = [], [], [], [], [], [], [], []
    As, Bs, Cs, Cref = [], [], [], []
    for (ta, tb, m, n, k, g) in groups:
        tas.append(ta); tbs.append(tb); ms.append(m); ns.append(n); ks.append(k); gs.append(g)
        als.append(0.5); bes.append(2.0)
        for _ in range(g):
            A = np.asfortranarray(rng.uniform(-1, 1, (k, m) if ta == 112 else (m, k)))
            B = np.asfortranarray(rng.uniform(-1, 1, (n, k) if tb == 112 else (k, n)))
            C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
            As.append(A); Bs.append(B); Cs.append(C)
            Cref.append(oracle.gemm(ta == 112, tb == 112, 0.5, A, B, 2.0, C))
    I = lambda v: (ctypes.c_int * len(v))(*v)
    D = lambda v: (ctypes.c_double * len(v))(*v)
    P = lambda v: (ctypes.c_void_p * len(v))(*[x.ctypes.data for x in v])
    lda = [m if ta == 111 else k for (ta, tb, m, n, k, g) in groups]
    ldb = [k if tb == 111 else n for (ta, tb, m, n, k, g) in groups]
    L.cblas_dgemm_batch(102, I(tas), I(tbs), I(ms), I(ns), I(ks), D(als), P(As), I(lda), P(Bs), I(ldb), D(bes),
                        P(Cs), I(ms), len(groups), I(gs))
    for C, R in zip(Cs, Cref):
        assert np.abs(C - R).max() < 1e-13
