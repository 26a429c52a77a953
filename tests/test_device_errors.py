"""Failure handling without a device (CPU only): a call that cannot run on
the GPU reports an error instead of aborting the caller's process.  LAPACK
routines set INFO = OPENBLAS_AMD_DEVICE_ERROR (-1000), LAPACKE routines and
the extension API return it, BLAS routines return with their outputs
untouched, and openblas_amd_last_error() holds the message.  (The reference
only aborts on allocation failure, driver/others/memory.c:3134.)  Argument
errors are still reported first, through xerbla, exactly as before."""
import ctypes
import subprocess
import sys
import os

import numpy as np
import pytest

import openblas_amd as oa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEVICE_ERROR = -1000


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(oa.LIB_PATH):
        subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "openblas_amd", "csrc")], check=True)
    L = oa.lib()
    if L.openblas_amd_device_count() > 0:
        pytest.skip("a device is visible: the no-device path cannot be exercised here")
    return L


def last_error(L):
    buf = ctypes.create_string_buffer(256)
    had = L.openblas_amd_last_error(buf, 256)
    return had, buf.value.decode()


def test_lapack_reports_device_error(L):
    n = 8
    A = np.asfortranarray(np.eye(n) * 2.0)
    B = np.ones((n, 1), order="F")
    piv = np.zeros(n, np.int32)
    info = ctypes.c_int(123)
    last_error(L)
    L.dgesv_(oa.ip(n), oa.ip(1), A.ctypes.data, oa.ip(n), piv.ctypes.data, B.ctypes.data, oa.ip(n),
             ctypes.byref(info))
    assert info.value == DEVICE_ERROR
    had, msg = last_error(L)
    assert had == 1 and "device" in msg.lower()
    assert np.array_equal(B, np.ones((n, 1)))   # outputs untouched
    L.dpotrf_(b"L", oa.ip(n), A.ctypes.data, oa.ip(n), ctypes.byref(info))
    assert info.value == DEVICE_ERROR
    it = ctypes.c_int(0)
    X = np.zeros((n, 1), order="F")
    L.dsgesv_(oa.ip(n), oa.ip(1), A.ctypes.data, oa.ip(n), piv.ctypes.data, B.ctypes.data, oa.ip(n),
              X.ctypes.data, oa.ip(n), None, None, ctypes.byref(it), ctypes.byref(info))
    assert info.value == DEVICE_ERROR


def test_argument_errors_still_first(L):
    info = ctypes.c_int(0)
    oa.last_xerbla()
    L.dgesv_(oa.ip(-1), oa.ip(1), None, oa.ip(1), None, None, oa.ip(1), ctypes.byref(info))
    assert info.value == -1 and oa.last_xerbla() == ("DGESV", 1)
    # quick returns touch no device and succeed
    L.dgesv_(oa.ip(0), oa.ip(1), None, oa.ip(1), None, None, oa.ip(1), ctypes.byref(info))
    assert info.value == 0


def test_blas_and_extensions_return(L):
    A = np.eye(4, order="F")
    C = np.full((4, 4), 3.0, order="F")
    L.cblas_dgemm(102, 111, 111, 4, 4, 4, 1.0, A.ctypes.data, 4, A.ctypes.data, 4, 0.0, C.ctypes.data, 4)
    assert np.all(C == 3.0)
    assert last_error(L)[0] == 1
    assert L.LAPACKE_dgetrf(102, 4, 4, A.ctypes.data, 4, np.zeros(4, np.int32).ctypes.data) == DEVICE_ERROR
    assert L.openblas_amd_synchronize() == DEVICE_ERROR
    assert L.openblas_amd_malloc(64) is None
    info = np.zeros(1, np.int32)
    n = np.array([4], np.int32)
    ptrs = (ctypes.c_void_p * 1)(A.ctypes.data)
    r = L.openblas_amd_dpotrf_batch(b"L", n.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ptrs,
                                    n.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                    info.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 1)
    assert r == DEVICE_ERROR


def test_process_survives(L):
    """The same sequence in a fresh interpreter exits normally (no abort)."""
    code = ("import sys, ctypes, numpy as np; sys.path.insert(0, %r); import openblas_amd as oa; L = oa.lib(); "
            "A = np.asfortranarray(np.eye(8)); info = ctypes.c_int(0); "
            "L.dgetrf_(oa.ip(8), oa.ip(8), A.ctypes.data, oa.ip(8), np.zeros(8, np.int32).ctypes.data, "
            "ctypes.byref(info)); print(info.value)") % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == str(DEVICE_ERROR)
