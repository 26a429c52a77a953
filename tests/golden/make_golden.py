"""Writes tests/golden/level3_golden.npz: input/output vectors produced by
the reference library itself (scipy's bundled OpenBLAS 0.3.28, called through
its exported entry points by oracle/refblas.py) for the functions on the
level-3 hot path.  The oracle tests check the C oracle against these, the
GPU tests check libopenblas_amd against them.

    python tests/golden/make_golden.py           # level3_golden.npz
    python tests/golden/make_golden.py --extra   # level3_golden_extra.npz
    python tests/golden/make_golden.py --ext3    # level3_ext_golden.npz
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import refblas as R  # noqa: E402


def householder_matrix(n, kappa, rng, nref=8):
    """Q1 diag(s) Q2, Qi = products of nref Householder reflectors with
    Gaussian vectors, s geometric from 1 to 1/kappa (dlatms MODE 3)."""
    s = kappa ** (-np.arange(n) / max(n - 1, 1))
    A = np.diag(s)
    for _ in range(nref):
        v = rng.standard_normal(n)
        A = A - np.outer(v, (2.0 / (v @ v)) * (v @ A))
    for _ in range(nref):
        v = rng.standard_normal(n)
        A = A - np.outer((2.0 / (v @ v)) * (A @ v), v)
    return np.asfortranarray(A)


def main():
    rng = np.random.default_rng(20241202)
    out = {}
    # --- gemm: 4 trans x shapes (odd, lda padding), alpha/beta incl. 0, NaN in C with beta 0
    cases = [(1, 1, 1), (7, 5, 3), (35, 33, 37), (67, 17, 29)]
    i = 0
    for dt, tag in ((np.float64, "d"), (np.float32, "s")):
        for (m, n, k) in cases:
            for ta in (0, 1):
                for tb in (0, 1):
                    for (alpha, beta) in ((1.0, 0.0), (0.7, 1.3), (0.0, 1.3)):
                        pad = 3
                        Ar, Ac = (k, m) if ta else (m, k)
                        Br, Bc = (n, k) if tb else (k, n)
                        A = np.asfortranarray(rng.uniform(-1, 1, (Ar + pad, Ac)).astype(dt))
                        B = np.asfortranarray(rng.uniform(-1, 1, (Br + pad, Bc)).astype(dt))
                        C = np.asfortranarray(rng.uniform(-1, 1, (m + pad, n)).astype(dt))
                        if beta == 0.0:
                            C[0, 0] = np.nan   # beta == 0 must not propagate NaN
                        Cin = C.copy(order="F")
                        Cv = np.asfortranarray(C[:m])
                        R.cblas_gemm(ta, tb, alpha, np.asfortranarray(A[:Ar]), np.asfortranarray(B[:Br]), beta, Cv)
                        key = f"gemm{tag}_{i}"
                        out[key + "_meta"] = np.array([ta, tb, m, n, k, Ar + pad, Br + pad, m + pad], np.int64)
                        out[key + "_ab"] = np.array([alpha, beta])
                        out[key + "_A"], out[key + "_B"], out[key + "_Cin"], out[key + "_Cout"] = A, B, Cin, Cv
                        i += 1
    # --- trsm: all 16 side/uplo/trans/diag, two sizes
    i = 0
    for dt, tag in ((np.float64, "d"), (np.float32, "s")):
        for (m, n) in ((9, 5), (70, 33)):
            for side in (0, 1):
                for uplo in (0, 1):
                    for tr in (0, 1):
                        for diag in (0, 1):
                            na = n if side else m
                            A = rng.uniform(-1, 1, (na, na))
                            A[np.diag_indices(na)] = na + rng.uniform(0, 1, na)
                            A = np.asfortranarray(A.astype(dt))
                            B = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(dt))
                            X = B.copy(order="F")
                            R.cblas_trsm(side, uplo, tr, diag, 0.7, A, X)
                            key = f"trsm{tag}_{i}"
                            out[key + "_meta"] = np.array([side, uplo, tr, diag, m, n], np.int64)
                            out[key + "_A"], out[key + "_B"], out[key + "_X"] = A, B, X
                            i += 1
    # --- getrf: random square/wide/tall, exact zero column (info), reversed magnitudes
    mats = []
    mats.append(rng.standard_normal((50, 50)))
    mats.append(rng.standard_normal((37, 61)))
    mats.append(rng.standard_normal((61, 37)))
    Z = rng.standard_normal((40, 40)); Z[:, 10] = 0.0; mats.append(Z)
    rev = np.diag(np.arange(1, 21, dtype=np.float64)[::-1].copy()) + 0.01 * rng.standard_normal((20, 20))
    mats.append(rev[::-1].copy())
    for j, M in enumerate(mats):
        for dt, tag in ((np.float64, "d"), (np.float32, "s")):
            A = np.asfortranarray(M.astype(dt))
            LU = A.copy(order="F")
            ipiv, info = R.getrf(LU)
            key = f"getrf{tag}_{j}"
            out[key + "_A"], out[key + "_LU"], out[key + "_ipiv"] = A, LU, ipiv
            out[key + "_info"] = np.array([info])
            if A.shape[0] == A.shape[1] and info == 0:
                for tr in (0, 1):
                    B = np.asfortranarray(rng.standard_normal((A.shape[0], 3)).astype(dt))
                    X = B.copy(order="F")
                    R.getrs(tr, LU.copy(order="F"), ipiv, X)
                    out[key + f"_rs{tr}_B"], out[key + f"_rs{tr}_X"] = B, X
    # --- potrf: SPD (B B^T + n I), both triangles; a non-PD matrix (info)
    for j, n in enumerate((1, 2, 40, 70)):
        Bm = rng.standard_normal((n, n))
        A = Bm @ Bm.T + n * np.eye(n)
        if n == 2:
            A = np.array([[2.0, 0.3], [0.3, 3.0]])   # utest/test_potrs.c:403 smoketest_trivial
        for dt, tag in ((np.float64, "d"), (np.float32, "s")):
            for uplo in (0, 1):
                F = np.asfortranarray(A.astype(dt))
                info = R.potrf(uplo, F)
                key = f"potrf{tag}_{j}_{uplo}"
                out[key + "_A"], out[key + "_F"], out[key + "_info"] = np.asfortranarray(A.astype(dt)), F, np.array([info])
    NP = np.eye(12); NP[7, 7] = -1.0; NP[3, 5] = NP[5, 3] = 0.2
    for uplo in (0, 1):
        F = np.asfortranarray(NP.copy())
        info = R.potrf(uplo, F)
        out[f"potrfd_np_{uplo}_A"], out[f"potrfd_np_{uplo}_F"], out[f"potrfd_np_{uplo}_info"] = NP, F, np.array([info])
    # --- dsgesv: kappa in {1e2, 1e6, 1e10}, nrhs in {1, 2}
    for j, kappa in enumerate((1e2, 1e6, 1e10)):
        A = householder_matrix(64, kappa, rng)
        for nrhs in (1, 2):
            B = np.asfortranarray(rng.standard_normal((64, nrhs)))
            X, ipiv, it, info = R.dsgesv(A.copy(order="F"), B.copy(order="F"))
            key = f"dsgesv_{j}_{nrhs}"
            out[key + "_A"], out[key + "_B"], out[key + "_X"] = A, B, X
            out[key + "_iter"], out[key + "_info"] = np.array([it]), np.array([info])
    out["reference_config"] = np.array(R.config())
    path = os.path.join(HERE, "level3_golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes,", len(out), "arrays")


def spd_matrix(n, kappa, rng, nref=8):
    """Q diag(s) Q^T with Q a product of nref Householder reflectors and s
    geometric from 1 to 1/kappa: symmetric positive definite, cond = kappa."""
    s = kappa ** (-np.arange(n) / max(n - 1, 1))
    A = np.diag(s)
    for _ in range(nref):
        v = rng.standard_normal(n)
        w = (2.0 / (v @ v)) * v
        A = A - np.outer(v, w @ A)          # H A
        A = A - np.outer(A @ v, w)          # (H A) H
    A = 0.5 * (A + A.T)
    return np.asfortranarray(A)


def main_extra():
    """tests/golden/level3_golden_extra.npz: dsposv (both triangles, kappa in
    {1e2, 1e6, 1e10}), getrf with a nonzero pivot below the smallest normal
    number (getf2_k.c:103-116) and with NaN in a pivot column (IAMAX rule),
    LAPACKE_dpotrf's NaN pre-check on the referenced triangle."""
    rng = np.random.default_rng(20261203)
    out = {}
    for j, kappa in enumerate((1e2, 1e6, 1e10)):
        A = spd_matrix(64, kappa, rng)
        for uplo in (0, 1):
            for nrhs in (1, 2):
                B = np.asfortranarray(rng.standard_normal((64, nrhs)))
                Aw = A.copy(order="F")
                X, it, info = R.dsposv(uplo, Aw, B.copy(order="F"))
                key = f"dsposv_{j}_{uplo}_{nrhs}"
                out[key + "_A"], out[key + "_B"], out[key + "_X"], out[key + "_Aout"] = A, B, X, Aw
                out[key + "_iter"], out[key + "_info"] = np.array([it]), np.array([info])
    # getrf: column 3's largest remaining entry is subnormal (nonzero, < DBL_MIN)
    for n in (8, 40):
        M = rng.standard_normal((n, n))
        M[:, 3] = 0.0
        M[6, 3] = 1e-310
        M[5, 3] = -3e-311
        LU = np.asfortranarray(M.copy())
        ipiv, info = R.getrf(LU)
        out[f"getrf_tiny_{n}_A"], out[f"getrf_tiny_{n}_LU"] = M, LU
        out[f"getrf_tiny_{n}_ipiv"], out[f"getrf_tiny_{n}_info"] = ipiv, np.array([info])
    # getrf: NaN entries in pivot columns, one at the diagonal position, one below it
    for t, (r, c) in enumerate(((2, 2), (7, 4))):
        M = rng.standard_normal((12, 12))
        M[r, c] = np.nan
        LU = np.asfortranarray(M.copy())
        ipiv, info = R.getrf(LU)
        out[f"getrf_nan_{t}_A"], out[f"getrf_nan_{t}_LU"] = M, LU
        out[f"getrf_nan_{t}_ipiv"], out[f"getrf_nan_{t}_info"] = ipiv, np.array([info])
    # LAPACKE_dpotrf NaN pre-check: NaN in the referenced / the other triangle
    n = 6
    Bm = rng.standard_normal((n, n))
    S = Bm @ Bm.T + n * np.eye(n)
    for layout in (101, 102):
        for uplo in (b"L", b"U"):
            for where in ("ref", "other"):
                A = S.copy()
                lower_ref = uplo == b"L"   # the referenced triangle of the logical matrix, in either layout
                i, j = (4, 1) if (where == "ref") == lower_ref else (1, 4)
                A[i, j] = np.nan
                buf = np.array(A if layout == 102 else A.T, order="F")   # memory: column-major view
                info = R.lapacke_potrf(layout, uplo, buf)
                key = f"lapacke_potrf_{layout}_{uplo.decode()}_{where}"
                out[key + "_A"], out[key + "_info"] = A, np.array([info])
    out["reference_config"] = np.array(R.config())
    path = os.path.join(HERE, "level3_golden_extra.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes,", len(out), "arrays")


# ---- SYRK / GEMMT / SYMM / TRMM ---------------------------------------------------
# Inputs are multiples of 0.5 in [-2, 2] and alpha, beta dyadic, so every
# product and sum is exact in FP32 and FP64: the expected outputs are exact
# and a correct implementation must reproduce them bit for bit.  The
# reference is called in both CBLAS layouts on the same logical matrices
# (row-major: C-ordered arrays) and must give the same logical result.
X3_N, X3_K, X3_SN, X3_TN = 66, 20, 45, 70
X3_AB = ((0.75, 1.25), (1.5, 0.0))   # the second: beta = 0 with NaN in C


def x3_inputs(rng):
    base = lambda: (rng.integers(-4, 5, (X3_TN, X3_TN)) / 2.0).astype(np.float32)
    return base(), base(), base()


def x3_poison_tri(C, lower, also_ref):
    """NaN in the triangle of C that the update must not touch (strictly
    above for lower, below for upper); with also_ref, in the updated
    triangle as well (beta = 0 must overwrite it)."""
    C = C.copy()
    n = C.shape[0]
    other = np.triu(np.ones((n, n), bool), 1) if lower else np.tril(np.ones((n, n), bool), -1)
    C[other] = np.nan
    if also_ref:
        C[~other] = np.nan
    return C


def x3_poison_a(A, lower, unit):
    """NaN in the unreferenced strict triangle of A (and its diagonal when
    unit)."""
    A = A.copy()
    n = A.shape[0]
    other = np.triu(np.ones((n, n), bool), 1) if lower else np.tril(np.ones((n, n), bool), -1)
    A[other] = np.nan
    if unit:
        A[np.eye(n, dtype=bool)] = np.nan
    return A


def x3_cases():
    """(routine, params) of every fixture; params are the CBLAS enums."""
    cases = []
    for uplo in (121, 122):
        for tr in (111, 112):
            for ab in range(2):
                cases.append(("syrk", (uplo, tr, ab)))
    for uplo in (121, 122):
        for ta in (111, 112):
            for tb in (111, 112):
                for ab in range(2):
                    cases.append(("gemmt", (uplo, ta, tb, ab)))
    for side in (141, 142):
        for uplo in (121, 122):
            for ab in range(2):
                cases.append(("symm", (side, uplo, ab)))
    for side in (141, 142):
        for uplo in (121, 122):
            for tr in (111, 112):
                for diag in (131, 132):
                    cases.append(("trmm", (side, uplo, tr, diag, 0)))
    cases.append(("trmm", (141, 122, 111, 131, 1)))   # alpha = 0: B <- 0
    return cases


def x3_operands(routine, params, A0, B0, C0, dt):
    """Logical input matrices of one case (float arrays of dtype dt)."""
    n, k = X3_N, X3_K
    if routine == "syrk":
        uplo, tr, ab = params
        A = A0[:k, :n] if tr == 112 else A0[:n, :k]
        C = x3_poison_tri(C0[:n, :n], uplo == 122, ab == 1)
        return A.astype(dt), None, C.astype(dt)
    if routine == "gemmt":
        uplo, ta, tb, ab = params
        A = A0[:k, :n] if ta == 112 else A0[:n, :k]
        B = B0[:n, :k] if tb == 112 else B0[:k, :n]
        C = x3_poison_tri(C0[:n, :n], uplo == 122, ab == 1)
        return A.astype(dt), B.astype(dt), C.astype(dt)
    if routine == "symm":
        side, uplo, ab = params
        m, nn = n, X3_SN
        na = nn if side == 142 else m
        A = x3_poison_a(A0[:na, :na], uplo == 122, False)
        C = C0[:m, :nn].copy()
        if ab == 1:
            C[:] = np.nan
        return A.astype(dt), B0[:m, :nn].astype(dt), C.astype(dt)
    side, uplo, tr, diag, az = params
    m, nn = n, X3_TN
    na = nn if side == 142 else m
    A = x3_poison_a(A0[:na, :na], uplo == 122, diag == 132)
    return A.astype(dt), B0[:m, :nn].astype(dt), None


def x3_alpha_beta(routine, params):
    if routine == "trmm":
        return (0.0 if params[-1] else 0.75), None
    return X3_AB[params[-1]]


def x3_key(routine, params, tag):
    return f"{routine}_{tag}_" + "_".join(str(p) for p in params)


def x3_call(routine, params, layout, alpha, beta, A, B, C):
    """The reference on logical matrices in one layout; returns the logical
    result."""
    conv = (lambda X: np.asfortranarray(X)) if layout == 102 else (lambda X: np.ascontiguousarray(X))
    A = conv(A)
    if B is not None:
        B = conv(B)
    if C is not None:
        C = conv(C)
    if routine == "syrk":
        return R.cblas_syrk(layout, params[0], params[1], alpha, A, beta, C)
    if routine == "gemmt":
        return R.cblas_gemmt(layout, params[0], params[1], params[2], alpha, A, B, beta, C)
    if routine == "symm":
        return R.cblas_symm(layout, params[0], params[1], alpha, A, B, beta, C)
    return R.cblas_trmm(layout, params[0], params[1], params[2], params[3], alpha, A, B)


def main_ext3():
    """tests/golden/level3_ext_golden.npz: SYRK, GEMMT, SYMM, TRMM (every
    uplo/trans/side/diag, beta = 0 with NaN in C, NaN in the parts of A and C
    that must not be read or written), FP64 and FP32, exact outputs."""
    rng = np.random.default_rng(20261204)
    A0, B0, C0 = x3_inputs(rng)
    out = {"x3_A": A0, "x3_B": B0, "x3_C": C0}
    for dt, tag in ((np.float64, "d"), (np.float32, "s")):
        for routine, params in x3_cases():
            A, B, C = x3_operands(routine, params, A0, B0, C0, dt)
            alpha, beta = x3_alpha_beta(routine, params)
            res = []
            for layout in (102, 101):
                Cin = None if C is None else C.copy()
                Bin = None if B is None else B.copy()
                res.append(np.array(x3_call(routine, params, layout, alpha, beta, A, Bin, Cin)))
            if routine == "gemmt":
                # The installed 0.3.28 build's row-major cblas_?gemmt does not
                # give the triangle of op(A) op(B) that the reference's own
                # utest/test_extensions/test_dgemmt.c requires (it compares with
                # cblas_dgemm); the column-major call of the same logical
                # problem does, and is the fixture.  The mismatch is recorded.
                out.setdefault("gemmt_rowmajor_mismatch", [])
                if not np.array_equal(res[0], res[1], equal_nan=True):
                    out["gemmt_rowmajor_mismatch"].append(x3_key(routine, params, tag))
            else:
                assert np.array_equal(res[0], res[1], equal_nan=True), (routine, params, tag)
            out[x3_key(routine, params, tag)] = res[0].astype(np.float32)   # exact: small dyadic values
    out["gemmt_rowmajor_mismatch"] = np.array(out.get("gemmt_rowmajor_mismatch", []), dtype=str)
    out["reference_config"] = np.array(R.config())
    path = os.path.join(HERE, "level3_ext_golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    if "--extra" in sys.argv:
        main_extra()
    elif "--ext3" in sys.argv:
        main_ext3()
    else:
        main()
