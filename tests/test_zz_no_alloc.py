"""openblas_amd_dgetrs_noalloc (dgetrs_ whose kernels' workspace is a
static device array instead of the memory pool) against dgetrs_ itself,
through the C ABI: two calls in a row, the second smaller, for both
transposes and host and device IPIV.  Up to 16 right-hand sides dgetrs_
runs the same kernels (the look-back triangular solves), so the results
are bitwise equal; with more, dgetrs_ takes the TRSM recursion while the
no-allocation form keeps solving 8 columns per launch, so the two agree to
rounding (both backward stable: FP64 backward error, and the difference
within c*kappa*eps of the solution).  The size limit and the
device-memory requirement are reported as argument errors."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _factor(oba, L, n, rng):
    A = np.asfortranarray(rng.standard_normal((n, n)))
    dA = oba.DeviceBuffer.from_array(A)
    piv = np.zeros(n, np.int32)
    info = ctypes.c_int(-99)
    L.dgetrf_(oba.ip(n), oba.ip(n), dA.ptr, oba.ip(n), piv.ctypes.data, ctypes.byref(info))
    assert info.value == 0
    return A, dA, piv


@pytest.mark.parametrize("trans", [b"N", b"T"])
def test_noalloc_matches_dgetrs(oba, trans):
    L = oba.lib()
    rng = np.random.default_rng(11)
    # second call smaller than the first; 19 columns = 3 solve chunks
    for n, nrhs, dev_ipiv in ((3000, 19, False), (1100, 5, True)):
        print(f"[noalloc] trans={trans.decode()} n={n} nrhs={nrhs} device ipiv={dev_ipiv}", flush=True)
        A, dA, piv = _factor(oba, L, n, rng)
        B = np.asfortranarray(rng.standard_normal((n, nrhs)))
        dB1 = oba.DeviceBuffer.from_array(B)
        dB2 = oba.DeviceBuffer.from_array(B)
        info1, info2 = ctypes.c_int(-99), ctypes.c_int(-99)
        L.dgetrs_(trans, oba.ip(n), oba.ip(nrhs), dA.ptr, oba.ip(n), piv.ctypes.data, dB1.ptr, oba.ip(n),
                  ctypes.byref(info1))
        dpiv = oba.DeviceBuffer.from_array(piv) if dev_ipiv else None
        L.openblas_amd_dgetrs_noalloc(trans, oba.ip(n), oba.ip(nrhs), dA.ptr, oba.ip(n),
                                      dpiv.ptr if dev_ipiv else piv.ctypes.data, dB2.ptr, oba.ip(n),
                                      ctypes.byref(info2))
        assert info1.value == 0 and info2.value == 0
        X1 = dB1.to_array(np.empty((n, nrhs), order="F"))
        X2 = dB2.to_array(np.empty((n, nrhs), order="F"))
        op = A if trans == b"N" else A.T
        bwd = np.abs(B - op @ X2).max() / (np.abs(op).sum(1).max() * np.abs(X2).max())
        assert bwd < 1e-14
        if nrhs <= 16:
            assert np.array_equal(X1, X2)
        else:
            kappa = np.linalg.cond(A, 1)
            assert np.abs(X1 - X2).max() <= 16 * kappa * np.finfo(np.float64).eps * np.abs(X1).max()
        for d in (dA, dB1, dB2) + ((dpiv,) if dev_ipiv else ()):
            d.free()


def test_noalloc_limits(oba):
    L = oba.lib()
    info = ctypes.c_int(-99)
    oba.last_xerbla()
    dummy = oba.DeviceBuffer(64)
    L.openblas_amd_dgetrs_noalloc(b"N", oba.ip(40000), oba.ip(1), dummy.ptr, oba.ip(40000), None, dummy.ptr,
                                  oba.ip(40000), ctypes.byref(info))
    assert info.value == -2 and oba.last_xerbla() == ("DGETRS_NOALLOC", 2)
    A = np.eye(4, order="F")
    B = np.ones((4, 1), order="F")
    piv = np.arange(1, 5, dtype=np.int32)
    L.openblas_amd_dgetrs_noalloc(b"N", oba.ip(4), oba.ip(1), A.ctypes.data, oba.ip(4), piv.ctypes.data,
                                  dummy.ptr, oba.ip(4), ctypes.byref(info))
    assert info.value == -4 and oba.last_xerbla() == ("DGETRS_NOALLOC", 4)
    L.openblas_amd_dgetrs_noalloc(b"N", oba.ip(4), oba.ip(1), dummy.ptr, oba.ip(4), piv.ctypes.data,
                                  B.ctypes.data, oba.ip(4), ctypes.byref(info))
    assert info.value == -7 and oba.last_xerbla() == ("DGETRS_NOALLOC", 7)
    # argument errors of dgetrs_ itself come first, as there, with INFO untouched
    info.value = 5
    L.openblas_amd_dgetrs_noalloc(b"X", oba.ip(4), oba.ip(1), dummy.ptr, oba.ip(4), piv.ctypes.data,
                                  dummy.ptr, oba.ip(4), ctypes.byref(info))
    assert info.value == 5 and oba.last_xerbla() == ("DGETRS", 1)
    dummy.free()
