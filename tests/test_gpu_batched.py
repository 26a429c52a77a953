"""GPU parity tests for the batched engine, through the C ABI.

* cblas_?gemm_batch (reference interface/gemm_batch.c:117): every problem
  equals the oracle GEMM (dmmch ratio), all four transpose pairs, uneven
  sizes, alpha/beta edge values, NaN in C with beta == 0, host and device
  operands, row-major groups; the strided-batch extension bit-for-bit
  against the group API on the same problems;
* openblas_amd_?getrf_batch / ?getrs_batch / ?potrf_batch (no reference
  batched routine): every problem gives the single-problem oracle's result
  (getf2 ipiv and info bit-exact on tie-free data, factors within c*n*eps,
  dget02-style solve residuals, potf2 info on non-PD problems); problems
  above 512 rows take the single-problem path inside the same call;
* graph replay: the same pointer table twice (the second call replays the
  captured graph) with new matrix contents in between, every replay checked.
"""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps}


def iarr(v):
    return np.ascontiguousarray(np.array(v, dtype=np.int32))


def parr(ptrs):
    return (ctypes.c_void_p * len(ptrs))(*ptrs)


def pi(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class DevPool:
    """One device allocation carved into 256-byte aligned pieces."""

    def __init__(self, oba, arrays):
        self.oba = oba
        self.offs, off = [], 0
        for a in arrays:
            self.offs.append(off)
            off += (a.nbytes + 255) // 256 * 256
        self.buf = oba.DeviceBuffer(max(off, 256))
        for a, o in zip(arrays, self.offs):
            oba.lib().openblas_amd_memcpy(self.buf.ptr + o, a.ctypes.data, a.nbytes, 0)

    def ptr(self, i):
        return self.buf.ptr + self.offs[i]

    def get(self, i, out):
        self.oba.lib().openblas_amd_memcpy(out.ctypes.data, self.ptr(i), out.nbytes, 1)
        return out


def sizes(rng, count, lo, hi):
    return [int(x) for x in rng.integers(lo, hi + 1, count)]


# ---- GEMM batch ------------------------------------------------------------------
@pytest.mark.parametrize("tag", ["d", "s"])
@pytest.mark.parametrize("where", ["host", "device"])
def test_gemm_batch_groups(oba, tag, where):
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    f = L.cblas_dgemm_batch if tag == "d" else L.cblas_sgemm_batch
    ct = ctypes.c_double if tag == "d" else ctypes.c_float
    rng = np.random.default_rng(5 if tag == "d" else 6)
    # groups: (ta, tb, m, n, k, alpha, beta, size)
    groups = [(111, 111, 37, 50, 70, 1.0, 0.0, 5), (112, 111, 64, 64, 64, 0.7, 1.3, 3),
              (111, 112, 1, 130, 9, -1.0, 1.0, 4), (112, 112, 200, 33, 128, 2.0, 0.5, 2),
              (111, 111, 8, 8, 0, 1.0, 3.0, 2), (111, 111, 16, 17, 18, 0.0, 2.0, 2),
              (111, 111, 600, 520, 530, 1.0, 0.0, 1)]
    As, Bs, Cs, Cin, meta = [], [], [], [], []
    for (ta, tb, m, n, k, al, be, gs) in groups:
        for _ in range(gs):
            A = np.asfortranarray(rng.uniform(-1, 1, (k, m) if ta == 112 else (m, k)).astype(dt))
            B = np.asfortranarray(rng.uniform(-1, 1, (n, k) if tb == 112 else (k, n)).astype(dt))
            C = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(dt))
            if be == 0.0:
                C[:] = np.nan
            As.append(A); Bs.append(B); Cs.append(C.copy(order="F")); Cin.append(C); meta.append((ta, tb, al, be))
    if where == "device":
        pool = DevPool(oba, As + Bs + Cs)
        nA = len(As)
        pa = [pool.ptr(i) for i in range(nA)]
        pb = [pool.ptr(nA + i) for i in range(nA)]
        pc = [pool.ptr(2 * nA + i) for i in range(nA)]
    else:
        pa = [a.ctypes.data for a in As]; pb = [b.ctypes.data for b in Bs]; pc = [c.ctypes.data for c in Cs]
    G = len(groups)
    ta = iarr([g[0] for g in groups]); tb = iarr([g[1] for g in groups])
    m = iarr([g[2] for g in groups]); n = iarr([g[3] for g in groups]); k = iarr([g[4] for g in groups])
    lda = iarr([max(1, g[4] if g[0] == 112 else g[2]) for g in groups])
    ldb = iarr([max(1, g[3] if g[1] == 112 else g[4]) for g in groups])
    ldc = iarr([max(1, g[2]) for g in groups])
    al = (ct * G)(*[g[5] for g in groups]); be = (ct * G)(*[g[6] for g in groups])
    gsz = iarr([g[7] for g in groups])
    f(102, pi(ta), pi(tb), pi(m), pi(n), pi(k), al, parr(pa), pi(lda), parr(pb), pi(ldb), be, parr(pc), pi(ldc), G,
      pi(gsz))
    for i, (A, B, C0, (t1, t2, a_, b_)) in enumerate(zip(As, Bs, Cin, meta)):
        out = pool.get(2 * len(As) + i, np.empty_like(Cs[i])) if where == "device" else Cs[i]
        if b_ == 0.0:
            assert not np.isnan(out).any(), i
        ref = oracle.gemm(t1 == 112, t2 == 112, a_, A, B, b_, np.nan_to_num(C0) if b_ == 0.0 else C0)
        tol = 1e-12 if dt == np.float64 else 2e-5
        assert np.abs(out - ref).max() <= tol * max(1.0, A.shape[0] * 1.0, A.shape[1] * 1.0) * 4, i
        if a_ != 0.0 and A.size:
            assert oracle.dmmch_ratio(t1 == 112, t2 == 112, a_, A, B, b_, np.nan_to_num(C0) if b_ == 0.0 else C0,
                                      out, eps=EPS[dt]) < 16, i


def test_gemm_batch_row_major_and_errors(oba):
    L = oba.lib()
    rng = np.random.default_rng(8)
    m, n, k = 23, 41, 17
    A = np.ascontiguousarray(rng.standard_normal((m, k)))
    B = np.ascontiguousarray(rng.standard_normal((k, n)))
    C = np.zeros((m, n))
    one = (ctypes.c_double * 1)(1.0); zero = (ctypes.c_double * 1)(0.0)
    L.cblas_dgemm_batch(101, pi(iarr([111])), pi(iarr([111])), pi(iarr([m])), pi(iarr([n])), pi(iarr([k])), one,
                        parr([A.ctypes.data]), pi(iarr([k])), parr([B.ctypes.data]), pi(iarr([n])), zero,
                        parr([C.ctypes.data]), pi(iarr([n])), 1, pi(iarr([1])))
    assert np.abs(C - A @ B).max() < 1e-12
    # a bad group: xerbla with the group's info, nothing computed
    oba.last_xerbla()
    C2 = np.full((m, n), 7.0)
    L.cblas_dgemm_batch(101, pi(iarr([111, 111])), pi(iarr([111, 111])), pi(iarr([m, m])), pi(iarr([n, n])),
                        pi(iarr([k, k])), (ctypes.c_double * 2)(1.0, 1.0), parr([A.ctypes.data] * 2),
                        pi(iarr([k, k])), parr([B.ctypes.data] * 2), pi(iarr([n, n - 1])),
                        (ctypes.c_double * 2)(0.0, 0.0), parr([C2.ctypes.data] * 2), pi(iarr([n, n])), 2,
                        pi(iarr([1, 1])))
    assert oba.last_xerbla() == ("DGEMM_BATCH", 8)   # row-major: ldb of the row-major call checks as lda
    assert (C2 == 7.0).all()


@pytest.mark.parametrize("tag", ["d", "s"])
@pytest.mark.parametrize("where", ["host", "device"])
def test_gemm_batch_strided(oba, tag, where):
    """cblas_?gemm_batch_strided (extension): for every transpose pair and
    both layouts, 40 problems at strides with gaps between the matrices
    (and a B stride of 0, one operand shared by the batch) give exactly the
    bits of the group API called with the same pointers (both run the same
    descriptors through the same engine), and each problem passes the dmmch
    ratio against the FP64 product; beta == 0 over NaN; the gaps between the
    C matrices are not written."""
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    fs = L.cblas_dgemm_batch_strided if tag == "d" else L.cblas_sgemm_batch_strided
    fg = L.cblas_dgemm_batch if tag == "d" else L.cblas_sgemm_batch
    cty = ctypes.c_double if tag == "d" else ctypes.c_float
    rng = np.random.default_rng(71)
    cnt, m, n, k = 40, 37, 29, 45
    for order in (102, 101):
        for ta in (111, 112):
            for tb in (111, 112):
                # logical op(A) m x k, op(B) k x n; stored per the layout
                ra, ca = (k, m) if ta == 112 else (m, k)
                rb, cb = (n, k) if tb == 112 else (k, n)
                if order == 101:
                    ra, ca, rb, cb, rc, cc = ca, ra, cb, rb, n, m   # column-major view of the row-major data
                else:
                    rc, cc = m, n
                lda, ldb, ldc = ra + 3, rb + 1, rc + 2
                sa, sb, sc = lda * ca + 5, 0 if ta == tb else ldb * cb + 7, ldc * cc + 11
                beta = 0.0 if ta == 111 else -0.5
                A = rng.uniform(-1, 1, sa * (cnt - 1) + lda * ca).astype(dt)
                B = rng.uniform(-1, 1, max(sb * (cnt - 1), 0) + ldb * cb).astype(dt)
                C0 = rng.uniform(-1, 1, sc * (cnt - 1) + ldc * cc).astype(dt)
                if beta == 0.0:
                    C0[::3] = np.nan
                outs = []
                for api in ("strided", "group"):
                    C = C0.copy()
                    if where == "device":
                        pool = DevPool(oba, [A, B, C])
                        pa, pb, pc = pool.ptr(0), pool.ptr(1), pool.ptr(2)
                    else:
                        pa, pb, pc = A.ctypes.data, B.ctypes.data, C.ctypes.data
                    es = A.itemsize
                    if api == "strided":
                        fs(order, ta, tb, m, n, k, 0.75, pa, lda, sa, pb, ldb, sb, beta, pc, ldc, sc, cnt)
                    else:
                        fg(order, pi(iarr([ta])), pi(iarr([tb])), pi(iarr([m])), pi(iarr([n])), pi(iarr([k])),
                           (cty * 1)(0.75), parr([pa + i * sa * es for i in range(cnt)]), pi(iarr([lda])),
                           parr([pb + i * sb * es for i in range(cnt)]), pi(iarr([ldb])), (cty * 1)(beta),
                           parr([pc + i * sc * es for i in range(cnt)]), pi(iarr([ldc])), 1, pi(iarr([cnt])))
                    if where == "device":
                        pool.get(2, C)
                        pool.buf.free()
                    outs.append(C)
                assert np.array_equal(outs[0], outs[1], equal_nan=True), (order, ta, tb)
                C = outs[0]
                for i in range(cnt):
                    def mat(X, off, ld, r, c):
                        return X[off:off + ld * c].reshape(c, ld).T[:r]
                    Ai = mat(A, i * sa, lda, ra, ca).astype(np.float64)
                    Bi = mat(B, i * sb, ldb, rb, cb).astype(np.float64)
                    Ci0 = np.nan_to_num(mat(C0, i * sc, ldc, rc, cc).astype(np.float64)) if beta != 0.0 else 0.0
                    Ci = mat(C, i * sc, ldc, rc, cc).astype(np.float64)
                    if order == 101:   # the column-major views hold C^T = op(B)^T op(A)^T, A^T, B^T
                        oA = Bi if tb == 111 else Bi.T
                        oB = Ai if ta == 111 else Ai.T
                    else:
                        oA = Ai.T if ta == 112 else Ai
                        oB = Bi.T if tb == 112 else Bi
                    ref = 0.75 * oA @ oB + beta * Ci0
                    g = 0.75 * np.abs(oA) @ np.abs(oB) + abs(beta) * np.abs(Ci0)
                    assert np.max(np.abs(Ci - ref) / g) / EPS[dt] < 16, (order, ta, tb, i)
                    # the gap after C_i (rows ldc-rc of each column, and the stride slack) keeps C0's bits
                    end = i * sc + ldc * cc
                    full = C[i * sc:end].reshape(cc, ldc)[:, rc:]
                    assert np.array_equal(full, C0[i * sc:end].reshape(cc, ldc)[:, rc:], equal_nan=True)
                    if i < cnt - 1:
                        assert np.array_equal(C[end:(i + 1) * sc], C0[end:(i + 1) * sc], equal_nan=True)


# ---- batched LU ----------------------------------------------------------------
def make_lu_batch(rng, dt, shapes):
    mats = [np.asfortranarray(rng.standard_normal(s).astype(dt)) for s in shapes]
    return mats


@pytest.mark.parametrize("tag", ["d", "s"])
def test_getrf_batch_vs_oracle(oba, tag):
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    f = L.openblas_amd_dgetrf_batch if tag == "d" else L.openblas_amd_sgetrf_batch
    rng = np.random.default_rng(30 if tag == "d" else 31)
    ns = sizes(rng, 60, 1, 256)
    shapes = [(n, n) for n in ns] + [(300, 120), (90, 200), (512, 512), (257, 257), (1, 40), (40, 1), (600, 600)]
    mats = make_lu_batch(rng, dt, shapes)
    mats[3][:, min(5, mats[3].shape[1] - 1)] = 0.0   # exact zero column -> info
    cnt = len(mats)
    pool = DevPool(oba, mats)
    pivs = [np.zeros(min(s), np.int32) for s in shapes]
    ppool = DevPool(oba, pivs)
    info = np.full(cnt, -99, np.int32)
    m = iarr([s[0] for s in shapes]); n = iarr([s[1] for s in shapes])
    assert f(pi(m), pi(n), parr([pool.ptr(i) for i in range(cnt)]), pi(m), parr([ppool.ptr(i) for i in range(cnt)]),
             pi(info), cnt) == 0
    for i, A0 in enumerate(mats):
        LU = pool.get(i, np.empty_like(A0))
        piv = ppool.get(i, np.empty_like(pivs[i]))
        LUo, pivo, infoo = oracle.getf2(A0)
        assert info[i] == infoo, (i, shapes[i])
        mi, ni = shapes[i]
        if dt == np.float64:
            # FP64: pivots bit-exact, factors within c*n*eps of the oracle's
            assert np.array_equal(piv, pivo), (i, shapes[i])
            tol = 200 * max(shapes[i]) * EPS[dt] * max(1, float(np.abs(LUo).max()))
            assert np.abs(LU - LUo).max() <= tol, (i, shapes[i])
        # dget01 on the returned factors: ||P L U - A||_1 / (n ||A||_1 eps) < 30,
        # partial pivoting bound |l_ij| <= 1 (FP32 near-ties may pick another
        # pivot than the oracle's different summation order; the factorization
        # must still be a valid partial-pivoting LU)
        k = min(mi, ni)
        Lm = np.tril(LU, -1)[:, :k] + np.eye(mi, k)
        Um = np.triu(LU)[:k, :]
        assert np.abs(np.tril(LU, -1)).max(initial=0) <= 1.0, (i, shapes[i])
        if infoo == 0:
            PA = oracle.laswp(A0.astype(np.float64), 1, k, piv, 1)
            r = np.linalg.norm(Lm.astype(np.float64) @ Um.astype(np.float64) - PA, 1) / (
                max(mi, ni) * np.linalg.norm(A0.astype(np.float64), 1) * EPS[dt])
            assert r < 30, (i, shapes[i], r)


def test_getrf_batch_host_operands_and_errors(oba):
    L = oba.lib()
    rng = np.random.default_rng(33)
    shapes = [(70, 70), (5, 9), (130, 130)]
    mats = make_lu_batch(rng, np.float64, shapes)
    work = [a.copy(order="F") for a in mats]
    pivs = [np.zeros(min(s), np.int32) for s in shapes] + [np.zeros(4, np.int32)]
    info = np.full(4, -99, np.int32)
    m = iarr([70, 5, 130, 4]); n = iarr([70, 9, 130, 4]); lda = iarr([70, 5, 130, 3])   # last: lda < m
    bad = np.zeros((4, 4), order="F")
    oba.last_xerbla()
    assert L.openblas_amd_dgetrf_batch(pi(m), pi(n), parr([w.ctypes.data for w in work] + [bad.ctypes.data]), pi(lda),
                                       parr([p.ctypes.data for p in pivs]), pi(info), 4) == 0
    assert oba.last_xerbla() == ("DGETRF", 4)
    assert info[3] == -4
    for i, A0 in enumerate(mats):
        LUo, pivo, infoo = oracle.getf2(A0)
        assert info[i] == infoo and np.array_equal(pivs[i], pivo)
        assert np.abs(work[i] - LUo).max() <= 1e-11 * max(1, np.abs(LUo).max())


@pytest.mark.parametrize("tag", ["d", "s"])
@pytest.mark.parametrize("trans", [b"N", b"T"])
def test_getrs_batch(oba, tag, trans):
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    fr = L.openblas_amd_dgetrf_batch if tag == "d" else L.openblas_amd_sgetrf_batch
    fs = L.openblas_amd_dgetrs_batch if tag == "d" else L.openblas_amd_sgetrs_batch
    rng = np.random.default_rng(40 + (trans == b"T"))
    ns = sizes(rng, 40, 1, 300) + [64, 65, 128, 129, 512, 700]
    nr = [1] * 30 + [3] * (len(ns) - 30)
    mats = [np.asfortranarray(rng.standard_normal((n, n)).astype(dt)) for n in ns]
    rhs = [np.asfortranarray(rng.standard_normal((n, r)).astype(dt)) for n, r in zip(ns, nr)]
    cnt = len(ns)
    pool = DevPool(oba, mats + rhs)
    pivs = [np.zeros(n, np.int32) for n in ns]
    ppool = DevPool(oba, pivs)
    info = np.full(cnt, -99, np.int32)
    nn = iarr(ns)
    pa = parr([pool.ptr(i) for i in range(cnt)]); pp = parr([ppool.ptr(i) for i in range(cnt)])
    assert fr(pi(nn), pi(nn), pa, pi(nn), pp, pi(info), cnt) == 0
    assert (info == 0).all()
    assert fs(trans, pi(nn), pi(iarr(nr)), pa, pi(nn), pp, parr([pool.ptr(cnt + i) for i in range(cnt)]), pi(nn),
              pi(info), cnt) == 0
    assert (info == 0).all()
    for i in range(cnt):
        A0, B0 = mats[i].astype(np.float64), rhs[i].astype(np.float64)
        X = pool.get(cnt + i, np.empty_like(rhs[i])).astype(np.float64)
        Aop = A0.T if trans == b"T" else A0
        # dget02: ||B - op(A) X||_1 / (||op(A)||_1 ||X||_1 n eps) < 30
        r = np.linalg.norm(B0 - Aop @ X, 1) / (np.linalg.norm(Aop, 1) * np.linalg.norm(X, 1) * ns[i] * EPS[dt])
        assert r < 30, (i, ns[i], r)


@pytest.mark.parametrize("tag", ["d", "s"])
@pytest.mark.parametrize("uplo", [b"L", b"U"])
def test_potrf_batch(oba, tag, uplo):
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    f = L.openblas_amd_dpotrf_batch if tag == "d" else L.openblas_amd_spotrf_batch
    rng = np.random.default_rng(50 + (uplo == b"U"))
    ns = sizes(rng, 40, 1, 256) + [300, 512, 600]
    mats = []
    for n in ns:
        Bm = rng.standard_normal((n, n))
        mats.append(np.asfortranarray((Bm @ Bm.T + n * np.eye(n)).astype(dt)))
    # a non-PD problem: a negative diagonal entry at j -> info = j+1 (lapack/potf2/potf2_L.c:71-75)
    dg = np.ones(ns[2]); dg[min(7, ns[2] - 1)] = -1.0
    mats[2] = np.asfortranarray(np.diag(dg).astype(dt))
    cnt = len(ns)
    pool = DevPool(oba, mats)
    info = np.full(cnt, -99, np.int32)
    nn = iarr(ns)
    assert f(uplo, pi(nn), parr([pool.ptr(i) for i in range(cnt)]), pi(nn), pi(info), cnt) == 0
    for i, A0 in enumerate(mats):
        got = pool.get(i, np.empty_like(A0))
        Fo, infoo = oracle.potf2(1 if uplo == b"L" else 0, A0)
        assert info[i] == infoo, (i, ns[i])
        if infoo == 0:
            tri = np.tril if uplo == b"L" else np.triu
            tol = 100 * ns[i] * EPS[dt] * max(1, float(np.abs(Fo).max()))
            assert np.abs(tri(got) - tri(Fo)).max() <= tol, (i, ns[i])
            # the other triangle is untouched
            other = np.triu if uplo == b"L" else np.tril
            assert np.array_equal(other(got, 1 if uplo == b"L" else -1), other(A0, 1 if uplo == b"L" else -1))


@pytest.mark.parametrize("tag", ["d", "s"])
def test_batched_small_sizes_mfma_updates(oba, tag):
    """Batches whose largest problem has at most 256 rows run the one-row-
    per-thread kernels, whose trailing updates are MFMA tiles (getrf: GEMM
    tiles over the deferred-pivot physical rows, rectangular shapes
    included; potrf: lower-triangle SYRK tiles, the other triangle
    untouched): pivots and info as the oracle, factors to 100*n*eps."""
    L = oba.lib()
    dt = np.float64 if tag == "d" else np.float32
    rng = np.random.default_rng(70 if tag == "d" else 71)
    shapes = [(n, n) for n in sizes(rng, 30, 1, 256)] + [(256, 256), (200, 37), (37, 200), (256, 100), (17, 256)]
    mats = make_lu_batch(rng, dt, shapes)
    cnt = len(mats)
    pool = DevPool(oba, mats)
    pivs = [np.zeros(min(sh), np.int32) for sh in shapes]
    ppool = DevPool(oba, pivs)
    info = np.full(cnt, -99, np.int32)
    m = iarr([sh[0] for sh in shapes]); n = iarr([sh[1] for sh in shapes])
    f = L.openblas_amd_dgetrf_batch if tag == "d" else L.openblas_amd_sgetrf_batch
    assert f(pi(m), pi(n), parr([pool.ptr(i) for i in range(cnt)]), pi(m), parr([ppool.ptr(i) for i in range(cnt)]),
             pi(info), cnt) == 0
    for i, A0 in enumerate(mats):
        LU = pool.get(i, np.empty_like(A0))
        piv = ppool.get(i, np.empty_like(pivs[i]))
        mi, ni = shapes[i]
        k = min(mi, ni)
        LUo, pivo, infoo = oracle.getf2(A0)
        assert info[i] == infoo == 0, (i, shapes[i])
        if tag == "d":
            assert np.array_equal(piv, pivo), (i, shapes[i])
            assert np.abs(LU - LUo).max() <= 100 * max(mi, ni) * EPS[dt] * max(1, float(np.abs(LUo).max())), i
        Lm = (np.tril(LU, -1)[:, :k] + np.eye(mi, k)).astype(np.float64)
        Um = np.triu(LU)[:k, :].astype(np.float64)
        PA = oracle.laswp(A0.astype(np.float64), 1, k, piv, 1)
        assert np.linalg.norm(Lm @ Um - PA, 1) / (max(mi, ni) * np.linalg.norm(A0, 1) * EPS[dt]) < 30, i
    for uplo in (b"L", b"U"):
        ns = sizes(rng, 30, 1, 256) + [256]
        spd = []
        for nn in ns:
            Bm = rng.standard_normal((nn, nn))
            spd.append(np.asfortranarray((Bm @ Bm.T + nn * np.eye(nn)).astype(dt)))
        pool = DevPool(oba, spd)
        info = np.full(len(ns), -99, np.int32)
        na = iarr(ns)
        fp = L.openblas_amd_dpotrf_batch if tag == "d" else L.openblas_amd_spotrf_batch
        assert fp(uplo, pi(na), parr([pool.ptr(i) for i in range(len(ns))]), pi(na), pi(info), len(ns)) == 0
        for i, A0 in enumerate(spd):
            got = pool.get(i, np.empty_like(A0))
            Fo, infoo = oracle.potf2(1 if uplo == b"L" else 0, A0)
            assert info[i] == infoo == 0, i
            tri = np.tril if uplo == b"L" else np.triu
            tol = 100 * ns[i] * EPS[dt] * max(1, float(np.abs(Fo).max()))
            assert np.abs(tri(got) - tri(Fo)).max() <= tol, (uplo, i, ns[i])
            other = np.triu if uplo == b"L" else np.tril
            assert np.array_equal(other(got, 1 if uplo == b"L" else -1), other(A0, 1 if uplo == b"L" else -1))


def test_batched_graph_replay(oba):
    """Same pointer table three times (graph replays), new contents each time,
    with other library work in between; every replay checked."""
    L = oba.lib()
    rng = np.random.default_rng(60)
    ns = sizes(rng, 25, 2, 150)
    cnt = len(ns)
    pivs = [np.zeros(n, np.int32) for n in ns]
    ppool = DevPool(oba, pivs)
    mats0 = [np.zeros((n, n), order="F") for n in ns]
    pool = DevPool(oba, mats0)
    nn = iarr(ns)
    pa = parr([pool.ptr(i) for i in range(cnt)]); pp = parr([ppool.ptr(i) for i in range(cnt)])
    for rep in range(3):
        mats = [np.asfortranarray(rng.standard_normal((n, n))) for n in ns]
        for i, a in enumerate(mats):
            L.openblas_amd_memcpy(pool.ptr(i), a.ctypes.data, a.nbytes, 0)
        info = np.full(cnt, -99, np.int32)
        assert L.openblas_amd_dgetrf_batch(pi(nn), pi(nn), pa, pi(nn), pp, pi(info), cnt) == 0
        # unrelated library work between replays
        X = np.asfortranarray(rng.standard_normal((64, 64)))
        Y = np.zeros((64, 64), order="F")
        L.cblas_dgemm(102, 111, 111, 64, 64, 64, 1.0, X.ctypes.data, 64, X.ctypes.data, 64, 0.0, Y.ctypes.data, 64)
        assert np.abs(Y - X @ X).max() < 1e-12
        for i, A0 in enumerate(mats):
            LUo, pivo, infoo = oracle.getf2(A0)
            assert info[i] == infoo == 0
            assert np.array_equal(ppool.get(i, np.empty_like(pivs[i])), pivo), (rep, i)
            assert np.abs(pool.get(i, np.empty_like(A0)) - LUo).max() <= 1e-10 * max(1, np.abs(LUo).max()), (rep, i)
