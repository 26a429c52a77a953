"""LU of tall matrices, through the C ABI.

The panel kernels exchange pivot candidates between all of their
workgroups every column, so a launch needs all of them resident at once;
the library computes the device's capacity for each panel form before
launching (one 512-row workgroup per 512 rows up to 65,536 rows, then the
second form with 512 rows per 256-thread workgroup) and reports a panel
it cannot hold as a device error instead of launching it.

* m = 24,576 (FP64) and 32,768 (FP32) square: the LU satisfies the
  backward-error bound |P A - L U| <= gamma_n |L| |U| (the property that
  lapack-netlib/TESTING/LIN/dget01.f measures normwise), checked on random
  probe vectors x: ||P A x - L (U x)|| / (||L| (|U| |x|)|| eps) < n.  The
  pivots are checked to be a permutation with |l_ij| <= 1.
* m = 131,072 x n = 64 (FP64): 256 workgroups of the second panel form;
  ipiv bit-exact and the factors within c*n*eps of the oracle's getf2
  (lapack/getf2/getf2_k.c restated in oracle/oracle.c).
"""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _perm_from_ipiv(ipiv):
    perm = np.arange(len(ipiv))
    for i, p in enumerate(ipiv - 1):
        if p != i:
            perm[[i, p]] = perm[[p, i]]
    return perm


def _tri_apply(LU, x, upper, absval, chunk=2048):
    """U x (upper) or L x with unit diagonal (lower) from the packed factors,
    a block of rows at a time; |U| |x| or |L| |x| with absval."""
    n = LU.shape[0]
    y = np.empty_like(x, dtype=np.float64)
    for r0 in range(0, n, chunk):
        r1 = min(n, r0 + chunk)
        blk = LU[r0:r1].astype(np.float64)
        blk = np.triu(blk, r0) if upper else np.tril(blk, r0 - 1)
        if absval:
            blk = np.abs(blk)
        y[r0:r1] = blk @ x
        if not upper:
            y[r0:r1] += x[r0:r1]
    return y


@pytest.mark.parametrize("n,dtype", [(24576, np.float64), (32768, np.float32)], ids=["d24576", "s32768"])
def test_getrf_tall_square(oba, n, dtype):
    L = oba.lib()
    tag = "d" if dtype == np.float64 else "s"
    print(f"\n[tall] {tag}getrf n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = np.asfortranarray(rng.standard_normal((n, n), dtype=np.float32).astype(dtype, copy=False))
    dA = oba.DeviceBuffer.from_array(A)
    piv = np.zeros(n, np.int32)
    info = ctypes.c_int(-99)
    fn = L.dgetrf_ if tag == "d" else L.sgetrf_
    fn(oba.ip(n), oba.ip(n), dA.ptr, oba.ip(n), piv.ctypes.data, ctypes.byref(info))
    assert info.value == 0
    LU = dA.to_array(np.empty((n, n), dtype=dtype, order="F"))
    dA.free()
    assert piv.min() >= 1 and piv.max() <= n and np.all(piv >= np.arange(1, n + 1))
    perm = _perm_from_ipiv(piv)
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert np.abs(np.tril(LU[:4096], -1)).max() <= 1.0   # partial pivoting: |l_ij| <= 1
    x = rng.standard_normal((n, 2))
    lhs = A[perm].astype(np.float64) @ x
    rhs = _tri_apply(LU, _tri_apply(LU, x, True, False), False, False)
    gauge = _tri_apply(LU, _tri_apply(LU, np.abs(x), True, True), False, True)
    eps = np.finfo(dtype).eps
    ratio = float((np.abs(lhs - rhs) / (gauge * eps)).max())
    print(f"[tall] {tag}getrf n={n}: |PAx - LUx| / (|L||U||x| eps) = {ratio:.2f}", flush=True)
    assert ratio < n


def test_getrf_tall_skinny_second_form(oba):
    L = oba.lib()
    m, n = 131072, 64
    print(f"\n[tall] dgetrf {m} x {n}", flush=True)
    rng = np.random.default_rng(7)
    A = np.asfortranarray(rng.standard_normal((m, n)))
    dA = oba.DeviceBuffer.from_array(A)
    piv = np.zeros(n, np.int32)
    info = ctypes.c_int(-99)
    L.dgetrf_(oba.ip(m), oba.ip(n), dA.ptr, oba.ip(m), piv.ctypes.data, ctypes.byref(info))
    LU = dA.to_array(np.empty((m, n), order="F"))
    dA.free()
    LUo, po, io = oracle.getf2(A)
    assert info.value == io == 0
    assert np.array_equal(piv, po)
    eps = np.finfo(np.float64).eps
    assert np.abs(LU - LUo).max() <= 64 * n * eps * max(1.0, np.abs(LUo).max())
