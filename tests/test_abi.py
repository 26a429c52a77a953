"""CPU tests of the drop-in boundary: the C-ABI library loads, exports every
symbol include/openblas_amd.h declares (and nothing the header lacks among
the reference names), and its host-side argument checking reports the
reference's xerbla info numbers and quick returns without touching a GPU
(interface/gemm.c:374-476, interface/trsm.c:258-354, interface/lapack/*.c,
ctest/c_d3chke.c + c_xerbla.c for the expected numbering)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import openblas_amd as oa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "openblas_amd.h")


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(oa.LIB_PATH):
        subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "openblas_amd", "csrc")], check=True)
    return oa.lib()


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"^#.*$", "", txt, flags=re.M)   # preprocessor lines (macros are not symbols)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt)
    skip = {"if", "defined", "sizeof"}
    return sorted({n for n in names if n not in skip and not n.startswith("CBLAS_")})


def test_header_symbols_exported(L):
    out = subprocess.run(["nm", "-D", "--defined-only", oa.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    declared = header_functions()
    assert len(declared) >= 55
    missing = [n for n in declared if n not in exported]
    assert not missing, missing
    # every declared symbol is also typed in the Python view and vice versa
    assert sorted(oa.exported_symbols()) == declared


def test_reference_symbol_set_covered():
    # the hot-path entry points of exports/gensymbol that the scope names
    ref = ["cblas_dgemm", "cblas_sgemm", "cblas_dtrsm", "cblas_strsm", "cblas_dgemm_batch", "cblas_sgemm_batch",
           "dgemm_", "sgemm_", "dtrsm_", "strsm_", "dgetrf_", "sgetrf_", "dgetrs_", "sgetrs_", "dgesv_", "sgesv_",
           "dpotrf_", "spotrf_", "dlaswp_", "slaswp_", "dgetf2_", "sgetf2_", "dpotf2_", "spotf2_", "dsgesv_",
           "LAPACKE_dgesv", "LAPACKE_dsgesv", "LAPACKE_dgetrf", "LAPACKE_dgetrs", "LAPACKE_dpotrf", "xerbla_",
           "openblas_get_config", "openblas_set_num_threads",
           "cblas_dsyrk", "cblas_ssyrk", "cblas_dsymm", "cblas_ssymm", "cblas_dtrmm", "cblas_strmm",
           "cblas_dgemmt", "cblas_sgemmt", "dsyrk_", "ssyrk_", "dsymm_", "ssymm_", "dtrmm_", "strmm_",
           "dgemmt_", "sgemmt_"]
    declared = set(header_functions())
    assert not [r for r in ref if r not in declared]


def _gemm_info(L, order, ta, tb, m, n, k, lda, ldb, ldc):
    A = np.zeros(4); B = np.zeros(4); C = np.zeros(4)
    oa.last_xerbla()
    L.cblas_dgemm(order, ta, tb, m, n, k, 1.0, A.ctypes.data, lda, B.ctypes.data, ldb, 0.0, C.ctypes.data, ldc)
    return oa.last_xerbla()


def _ctest_gemm_number(order, info):
    """What ctest/c_xerbla.c makes of the Fortran info OpenBLAS reports."""
    info += 1
    if order == oa.RowMajor:
        info = {5: 4, 4: 5, 11: 9, 9: 11}.get(info, info)
    return info


@pytest.mark.parametrize("args,expected", [
    ((0, 111, 111, 0, 0, 0, 1, 1, 1), 1),          # invalid order
    ((102, 0, 111, 0, 0, 0, 1, 1, 1), 2),          # TransA
    ((102, 111, 0, 0, 0, 0, 1, 1, 1), 3),          # TransB
    ((102, 111, 111, -1, 0, 0, 1, 1, 1), 4),       # M
    ((102, 111, 111, 0, -1, 0, 1, 1, 1), 5),       # N
    ((102, 111, 111, 0, 0, -1, 1, 1, 1), 6),       # K
    ((102, 111, 111, 2, 0, 0, 1, 1, 2), 9),        # lda < M
    ((102, 112, 111, 0, 0, 2, 1, 1, 1), 9),        # lda < K (A transposed)
    ((102, 111, 111, 0, 0, 2, 1, 1, 1), 11),       # ldb < K
    ((102, 111, 111, 2, 0, 0, 2, 1, 1), 14),       # ldc < M
    ((101, 111, 111, -1, 0, 0, 1, 1, 1), 4),       # row-major M
    ((101, 111, 111, 0, -1, 0, 1, 1, 1), 5),       # row-major N
    ((101, 111, 111, 0, 2, 0, 1, 1, 2), 11),       # row-major ldb < N
    ((101, 111, 111, 2, 0, 0, 1, 1, 1), 9),        # row-major lda < K... (K=0 -> lda>=1 ok) ldc < N? see below
])
def test_cblas_dgemm_error_exits(L, args, expected):
    order = args[0]
    name, info = _gemm_info(L, *args)
    if expected == 9 and order == 101:
        # row-major M=2,K=0: lda >= max(K,1) holds, ldc=1 >= N=0 holds: no error
        assert name == "" and info == 0
        return
    assert name == "DGEMM"
    assert _ctest_gemm_number(order, info) == expected


def test_fortran_dgemm_error_exits(L):
    ip = oa.ip
    A = np.zeros(4)
    a = ctypes.c_double(1.0)
    for (ta, tb, m, n, k, lda, ldb, ldc), exp in [((b"X", b"N", 0, 0, 0, 1, 1, 1), 1), ((b"n", b"Q", 0, 0, 0, 1, 1, 1), 2),
                                                  ((b"N", b"N", -1, 0, 0, 1, 1, 1), 3), ((b"N", b"T", 0, -1, 0, 1, 1, 1), 4),
                                                  ((b"t", b"N", 0, 0, -1, 1, 1, 1), 5), ((b"N", b"N", 3, 0, 0, 2, 1, 3), 8),
                                                  ((b"N", b"N", 0, 0, 3, 1, 2, 1), 10), ((b"N", b"N", 3, 0, 0, 3, 1, 2), 13)]:
        oa.last_xerbla()
        L.dgemm_(ta, tb, ip(m), ip(n), ip(k), ctypes.byref(a), A.ctypes.data, ip(lda), A.ctypes.data, ip(ldb),
                 ctypes.byref(a), A.ctypes.data, ip(ldc))
        assert oa.last_xerbla() == ("DGEMM", exp)


def test_cblas_dtrsm_error_exits(L):
    A = np.zeros(4)
    cases = [((102, 0, 121, 111, 131, 0, 0, 1, 1), 1), ((102, 141, 0, 111, 131, 0, 0, 1, 1), 2),
             ((102, 141, 121, 0, 131, 0, 0, 1, 1), 3), ((102, 141, 121, 111, 0, 0, 0, 1, 1), 4),
             ((102, 141, 121, 111, 131, -1, 0, 1, 1), 5), ((102, 141, 121, 111, 131, 0, -1, 1, 1), 6),
             ((102, 141, 121, 111, 131, 2, 0, 1, 2), 9), ((102, 142, 121, 111, 131, 0, 2, 1, 1), 9),
             ((102, 141, 121, 111, 131, 2, 0, 2, 1), 11), ((0, 141, 121, 111, 131, 0, 0, 1, 1), 0)]
    for (order, side, uplo, tr, diag, m, n, lda, ldb), exp in cases:
        oa.last_xerbla()
        L.cblas_dtrsm(order, side, uplo, tr, diag, m, n, 1.0, A.ctypes.data, lda, A.ctypes.data, ldb)
        assert oa.last_xerbla() == ("DTRSM", exp), (order, side, uplo, tr, diag, m, n, lda, ldb)


def test_lapack_error_exits_and_quick_returns(L):
    ip = oa.ip
    A = np.zeros(4)
    piv = np.zeros(4, np.int32)
    info = ctypes.c_int(99)
    for (m, n, lda), exp in [((-1, 0, 1), 1), ((0, -1, 1), 2), ((3, 3, 2), 4)]:
        L.dgetrf_(ip(m), ip(n), A.ctypes.data, ip(lda), piv.ctypes.data, ctypes.byref(info))
        assert info.value == -exp and oa.last_xerbla() == ("DGETRF", exp)
    info.value = 99
    L.dgetrf_(ip(0), ip(5), A.ctypes.data, ip(1), piv.ctypes.data, ctypes.byref(info))
    assert info.value == 0 and oa.last_xerbla() == ("", 0)
    for (n, nrhs, lda, ldb), exp in [((-1, 0, 1, 1), 1), ((0, -1, 1, 1), 2), ((3, 1, 2, 3), 4), ((3, 1, 3, 2), 7)]:
        L.dgesv_(ip(n), ip(nrhs), A.ctypes.data, ip(lda), piv.ctypes.data, A.ctypes.data, ip(ldb), ctypes.byref(info))
        assert info.value == -exp and oa.last_xerbla() == ("DGESV", exp)
    # getrs reports through xerbla and leaves INFO untouched (interface/lapack/getrs.c:95-100)
    info.value = 77
    L.dgetrs_(b"X", ip(1), ip(1), A.ctypes.data, ip(1), piv.ctypes.data, A.ctypes.data, ip(1), ctypes.byref(info))
    assert info.value == 77 and oa.last_xerbla() == ("DGETRS", 1)
    for (uplo, n, lda), exp in [((b"X", 1, 1), 1), ((b"L", -1, 1), 2), ((b"U", 3, 2), 4)]:
        L.dpotrf_(uplo, ip(n), A.ctypes.data, ip(lda), ctypes.byref(info))
        assert info.value == -exp and oa.last_xerbla() == ("DPOTRF", exp)
    it = ctypes.c_int(5)
    for (n, nrhs, lda, ldb, ldx), exp in [((-1, 1, 1, 1, 1), 1), ((2, -1, 2, 2, 2), 2), ((2, 1, 1, 2, 2), 4),
                                          ((2, 1, 2, 1, 2), 7), ((2, 1, 2, 2, 1), 9)]:
        L.dsgesv_(ip(n), ip(nrhs), A.ctypes.data, ip(lda), piv.ctypes.data, A.ctypes.data, ip(ldb), A.ctypes.data,
                  ip(ldx), None, None, ctypes.byref(it), ctypes.byref(info))
        assert info.value == -exp and oa.last_xerbla() == ("DSGESV", exp)
    # LAPACKE: bad layout -> -1; row-major lda < n -> -5
    assert L.LAPACKE_dgesv(7, 2, 1, A.ctypes.data, 2, piv.ctypes.data, A.ctypes.data, 1) == -1
    assert L.LAPACKE_dgesv(101, 2, 1, A.ctypes.data, 1, piv.ctypes.data, A.ctypes.data, 1) == -5


def test_gemm_quick_returns_touch_nothing(L):
    C = np.full((3, 3), np.nan, order="F")
    # m == 0 / n == 0: return before any device work (interface/gemm.c:494)
    L.cblas_dgemm(102, 111, 111, 0, 3, 3, 1.0, C.ctypes.data, 1, C.ctypes.data, 3, 0.0, C.ctypes.data, 1)
    # alpha == 0 and beta == 1: nothing to do at all
    L.cblas_dgemm(102, 111, 111, 3, 3, 3, 0.0, C.ctypes.data, 3, C.ctypes.data, 3, 1.0, C.ctypes.data, 3)
    assert np.isnan(C).all()


def test_python_layer_refuses_missing_library(tmp_path, monkeypatch):
    monkeypatch.setattr(oa, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(oa, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        oa.lib()


def test_batched_argument_errors_without_device(L):
    """Argument checks of the batched LAPACK entry points happen before any
    device work: each problem gets the single routine's info numbering."""
    import ctypes
    i32 = lambda v: np.ascontiguousarray(np.array(v, np.int32))
    pi = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    nul = (ctypes.c_void_p * 3)(0, 0, 0)
    info = i32([9, 9, 9])
    oa.last_xerbla()
    # getrf: m < 0 -> -1, n < 0 -> -2, lda < max(1, m) -> -4
    assert L.openblas_amd_dgetrf_batch(pi(i32([-1, 3, 5])), pi(i32([2, -1, 5])), nul, pi(i32([1, 3, 4])), nul,
                                       pi(info), 3) == 0
    assert list(info) == [-1, -2, -4] and oa.last_xerbla() == ("DGETRF", 1)
    assert L.openblas_amd_dgetrf_batch(pi(info), pi(info), nul, pi(info), nul, pi(info), -1) == -7
    # getrs: bad trans is argument 1 of the batched call
    assert L.openblas_amd_dgetrs_batch(b"X", pi(info), pi(info), nul, pi(info), nul, nul, pi(info), pi(info), 1) == -1
    info = i32([9, 9])
    assert L.openblas_amd_sgetrs_batch(b"n", pi(i32([4, 4])), pi(i32([-1, 1])), nul, pi(i32([4, 4])), nul, nul,
                                       pi(i32([4, 3])), pi(info), 2) == 0
    assert list(info) == [-3, -8]
    # potrf: lda < n -> -4
    info = i32([9])
    assert L.openblas_amd_dpotrf_batch(b"L", pi(i32([6])), nul, pi(i32([5])), pi(info), 1) == 0
    assert list(info) == [-4]
    assert L.openblas_amd_spotrf_batch(b"Q", pi(info), nul, pi(info), pi(info), 1) == -1


def test_ext3_argument_errors(L):
    """xerbla numbers of interface/syrk.c, gemmt.c, symm.c, trsm.c (TRMM)."""
    z = np.zeros(4)
    oa.last_xerbla()
    L.cblas_dsyrk(102, 121, 111, 4, 2, 1.0, z.ctypes.data, 3, 0.0, z.ctypes.data, 4)
    assert oa.last_xerbla() == ("DSYRK", 7)
    L.cblas_dsyrk(102, 121, 111, 4, 2, 1.0, z.ctypes.data, 4, 0.0, z.ctypes.data, 3)
    assert oa.last_xerbla() == ("DSYRK", 10)
    L.cblas_dsyrk(102, 99, 111, 4, 2, 1.0, z.ctypes.data, 4, 0.0, z.ctypes.data, 4)
    assert oa.last_xerbla() == ("DSYRK", 1)
    L.cblas_dgemmt(102, 121, 111, 111, 4, 2, 1.0, z.ctypes.data, 4, z.ctypes.data, 1, 0.0, z.ctypes.data, 4)
    assert oa.last_xerbla() == ("DGEMMT", 10)
    L.cblas_dgemmt(101, 121, 111, 111, 4, 2, 1.0, z.ctypes.data, 1, z.ctypes.data, 4, 0.0, z.ctypes.data, 4)
    assert oa.last_xerbla() == ("DGEMMT", 8)
    L.cblas_dsymm(102, 142, 121, 4, 5, 1.0, z.ctypes.data, 4, z.ctypes.data, 4, 0.0, z.ctypes.data, 4)
    assert oa.last_xerbla() == ("DSYMM", 7)
    L.cblas_dtrmm(102, 141, 121, 111, 131, 4, 2, 1.0, z.ctypes.data, 3, z.ctypes.data, 4)
    assert oa.last_xerbla() == ("DTRMM", 9)
    L.dtrmm_(b"L", b"U", b"N", b"X", oa.ip(4), oa.ip(2), ctypes.byref(ctypes.c_double(1.0)), z.ctypes.data, oa.ip(4),
             z.ctypes.data, oa.ip(4))
    assert oa.last_xerbla() == ("DTRMM", 4)
    # quick returns: n == 0 touches nothing
    C = np.full((2, 2), 7.0)
    L.cblas_dsyrk(102, 121, 111, 0, 2, 1.0, z.ctypes.data, 1, 0.0, C.ctypes.data, 1)
    assert np.all(C == 7.0) and oa.last_xerbla() == ("", 0)


def test_gemm_batch_strided_argument_errors(L):
    """xerbla numbers of the strided-batch extension (checks of interface/gemm.c
    numbered after Order, plus the strides and the batch size); a zero batch,
    M or N touches nothing.  Nothing here reaches the device."""
    z = np.zeros(64)
    C = np.full(8, 7.0)
    f = L.cblas_dgemm_batch_strided
    cases = [
        ((102, 99, 111, 2, 2, 2, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 2, 4, 0.0, C.ctypes.data, 2, 4, 2), 1),
        ((102, 111, 99, 2, 2, 2, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 2, 4, 0.0, C.ctypes.data, 2, 4, 2), 2),
        ((102, 111, 111, -1, 2, 2, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 2, 4, 0.0, C.ctypes.data, 2, 4, 2), 3),
        ((102, 111, 111, 2, -1, 2, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 2, 4, 0.0, C.ctypes.data, 2, 4, 2), 4),
        ((102, 111, 111, 2, 2, -1, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 2, 4, 0.0, C.ctypes.data, 2, 4, 2), 5),
        ((102, 111, 111, 2, 2, 2, 1.0, z.ctypes.data, 1, 4, z.ctypes.data, 2, 4, 0.0, C.ctypes.data, 2, 4, 2), 8),
        ((102, 111, 111, 2, 2, 2, 1.0, z.ctypes.data, 2, -4, z.ctypes.data, 2, 4, 0.0, C.ctypes.data, 2, 4, 2), 9),
        ((102, 111, 111, 2, 2, 2, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 1, 4, 0.0, C.ctypes.data, 2, 4, 2), 11),
        ((102, 111, 111, 2, 2, 2, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 2, -1, 0.0, C.ctypes.data, 2, 4, 2), 12),
        ((102, 111, 111, 2, 2, 2, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 2, 4, 0.0, C.ctypes.data, 1, 4, 2), 15),
        ((102, 111, 111, 2, 2, 2, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 2, 4, 0.0, C.ctypes.data, 2, 3, 2), 16),
        ((102, 111, 111, 2, 2, 2, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 2, 4, 0.0, C.ctypes.data, 2, 4, -1), 17),
        # row-major: A and B swap before the checks, so the row-major lda is checked as ldb (11)
        ((101, 111, 111, 2, 3, 2, 1.0, z.ctypes.data, 1, 4, z.ctypes.data, 3, 6, 0.0, C.ctypes.data, 3, 6, 2), 11),
        ((101, 111, 111, 2, 3, 2, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 2, 6, 0.0, C.ctypes.data, 3, 6, 2), 8),
    ]
    for args, exp in cases:
        oa.last_xerbla()
        f(*args)
        assert oa.last_xerbla() == ("DGEMM_STRIDED", exp), (args, exp)
        assert (C == 7.0).all()
    oa.last_xerbla()
    L.cblas_sgemm_batch_strided(102, 111, 111, 2, 2, 2, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 2, 4, 0.0,
                                C.ctypes.data, 2, 3, 2)
    assert oa.last_xerbla() == ("SGEMM_STRIDED", 16)
    # quick returns: batch 0, M 0, N 0 (a single problem may use stride 0 for C)
    for args in ((102, 111, 111, 2, 2, 2, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 2, 4, 0.0, C.ctypes.data, 2, 4, 0),
                 (102, 111, 111, 0, 2, 2, 1.0, z.ctypes.data, 1, 4, z.ctypes.data, 2, 4, 0.0, C.ctypes.data, 1, 4, 3),
                 (102, 111, 111, 2, 0, 2, 1.0, z.ctypes.data, 2, 4, z.ctypes.data, 2, 4, 0.0, C.ctypes.data, 2, 0, 3)):
        oa.last_xerbla()
        f(*args)
        assert oa.last_xerbla() == ("", 0) and (C == 7.0).all()
