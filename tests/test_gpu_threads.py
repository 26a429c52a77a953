"""Concurrent callers, modelled on the reference's
cpp_thread_test/dgemm_thread_safety.cpp:64-92 (many threads call cblas_dgemm
at once on identical inputs; every thread's C must agree with the first
thread's within 1e-13).

Here 8 host threads call the library at the same time (ctypes releases the
interpreter lock during each call), mixing cblas_dgemm on host and device
operands (1024^3, the thread-safety test's size, and a 300 x 250 x 3001
shape that takes the pooled split-K workspace), dgesv_ (N = 1500: the
blocked LU with its second stream and pooled buffers) and the strided-batch
GEMM.  The library serializes calls, so the bar is stricter than the
reference's: every result must equal the single-threaded result bit for bit,
in every round.
"""
import ctypes
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = 8
ROUNDS = 3


def _gemm_case(oba, rng, m, n, k):
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    return A, B, C


def test_concurrent_callers_bitwise(oba):
    L = oba.lib()
    rng = np.random.default_rng(2024)
    big = _gemm_case(oba, rng, 1024, 1024, 1024)
    skinny = _gemm_case(oba, rng, 300, 250, 3001)
    n = 1500
    As = np.asfortranarray(rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n))
    bs = np.asfortranarray(rng.standard_normal((n, 2)))
    cnt, m_, n_, k_ = 30, 33, 21, 40
    Ab = rng.uniform(-1, 1, m_ * k_ * cnt)
    Bb = rng.uniform(-1, 1, k_ * n_ * cnt)
    dA, dB = oba.DeviceBuffer.from_array(big[0]), oba.DeviceBuffer.from_array(big[1])

    def gemm_host(case):
        A, B, C0 = case
        C = C0.copy(order="F")
        m, k = A.shape
        L.cblas_dgemm(102, 111, 111, m, C.shape[1], k, 1.0, A.ctypes.data, m, B.ctypes.data, k, 0.5,
                      C.ctypes.data, m)
        return C

    def gemm_dev():
        C0 = big[2]
        dC = oba.DeviceBuffer.from_array(C0)
        L.cblas_dgemm(102, 111, 111, 1024, 1024, 1024, 1.0, dA.ptr, 1024, dB.ptr, 1024, 0.5, dC.ptr, 1024)
        out = dC.to_array(np.empty_like(C0))
        dC.free()
        return out

    def gesv():
        A = As.copy(order="F")
        b = bs.copy(order="F")
        piv = np.zeros(n, dtype=np.int32)
        info = ctypes.c_int(-7)
        L.dgesv_(oba.ip(n), oba.ip(2), A.ctypes.data, oba.ip(n), piv.ctypes.data, b.ctypes.data, oba.ip(n),
                 ctypes.byref(info))
        assert info.value == 0
        return np.concatenate([b.ravel(order="F"), A.ravel(order="F"), piv.astype(np.float64)])

    def strided():
        C = np.zeros(m_ * n_ * cnt)
        L.cblas_dgemm_batch_strided(102, 111, 111, m_, n_, k_, 1.0, Ab.ctypes.data, m_, m_ * k_, Bb.ctypes.data,
                                    k_, k_ * n_, 0.0, C.ctypes.data, m_, m_ * n_, cnt)
        return C

    jobs = [lambda: gemm_host(big), lambda: gemm_host(skinny), gemm_dev, gesv, strided]
    want = [j() for j in jobs]   # one thread first
    # every job is checked against its own single-threaded result
    assert np.isfinite(want[0]).all() and np.isfinite(want[3]).all()
    for r in range(ROUNDS):
        tasks = [(t % len(jobs)) for t in range(THREADS * 2)]
        with ThreadPoolExecutor(max_workers=THREADS) as ex:
            outs = list(ex.map(lambda j: (j, jobs[j]()), tasks))
        for j, got in outs:
            assert np.array_equal(got, want[j]), (r, j)
    dA.free()
    dB.free()
