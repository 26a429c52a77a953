#!/usr/bin/env python3
"""Benchmark of the level-3 hot path on one MI355X.

Headline workload (the metric "dgesv N=16384 time to FP64 backward error:
FP32-LU+IR vs FP64 LU; DGEMM TF/s"): A = Q1*diag(s)*Q2 of order 16384 in
FP64, Q1 and Q2 products of 8 Householder reflectors with Gaussian vectors,
s geometric from 1 to 1/kappa (kappa = 1e6), b Gaussian, NRHS = 1.  One
step = one mixed-precision solve (dsgesv_: FP32 LU on MFMA + FP64 iterative
refinement) on device-resident A and b.  The plain FP64 solve (dgesv_) is
timed in the same run.

Every configuration of the benchmark set is measured beside it:
  C1  cblas_dgemm NN 1024^3, alpha 1, beta 0.5
  C2  d/sgemm 8192^3 in all four transposes; d/strsm L,L,N,N 8192 x 1024
  C3  dsgesv_ vs dgesv_ at N = 16384 for kappa in {1e2, 1e6, 1e10} x NRHS in
      {1, 16}, with backward and forward error (b = A x_true)
  C4  INT8-slice emulated DGEMM 8192^3, phi in {0.5, 2, 8}
  C5  10,000 uneven small problems (gemm, getrf, getrs, potrf; FP64, FP32)
  f1  dsposv_ (FP32 Cholesky + FP64 refinement) at N = 16384 on an SPD
      system of condition kappa, beside dpotrf_ and spotrf_ alone
Each reports TFLOP/s and a roofline record: the algorithmic flops (or HBM
bytes) of its dominant kernel over that kernel's time, from HIP events the
library records around its launches on the launching stream
(openblas_amd_ktimer_*), against the published MI355X peaks.  The
reference CPU path (the OpenBLAS library installed on the host, called
through its exported entry points) is timed beside each on a bounded sample
(1 warm-up, median of >= 3), with the core count stated.

Prints ONE JSON line.  Data are synthetic (seeded).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import openblas_amd as oa  # noqa: E402

PEAK_FP64 = 78.6e12   # MI355X dense FP64 MFMA, published
PEAK_FP32 = 157.3e12
PEAK_INT8 = 5.0e15
PEAK_HBM = 8.0e12
EPS64 = float(np.finfo(np.float64).eps)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, calibrated with a copy of known size)."""
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


# ---- kernel timer ----------------------------------------------------------------
KERNEL_NAMES = ["lu_panel", "lu_trailing_gemm", "lu_next_gemm", "lu_node_gemm", "trsv_lookback", "dgemm",
                "sgemm", "dtrsm", "strsm", "ozaki_gemm", "batched_gemm", "batched_getrf", "batched_getrs",
                "batched_potrf"]


def ktimer_profile(L, fn, peak_of):
    """Runs fn once with the library's kernel timer on; returns per named
    kernel: launches, total/average time, algorithmic rate and roofline
    fraction (peak_of(name) -> (bound, peak))."""
    L.openblas_amd_synchronize()
    L.openblas_amd_ktimer_reset()
    L.openblas_amd_ktimer_enable(1)
    fn()
    L.openblas_amd_synchronize()
    L.openblas_amd_ktimer_enable(0)
    out = {}
    for name in KERNEL_NAMES:
        cnt, ms, fl, by = ctypes.c_long(0), ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
        if L.openblas_amd_ktimer_read(name.encode(), ctypes.byref(cnt), ctypes.byref(ms), ctypes.byref(fl),
                                      ctypes.byref(by)) != 0 or cnt.value == 0:
            continue
        bound, peak = peak_of(name)
        t = ms.value * 1e-3
        rate = (fl.value if bound == "mfma" else by.value) / t
        out[name] = {"launches": cnt.value, "total_ms": ms.value, "avg_us": ms.value * 1e3 / cnt.value,
                     "flops_per_launch": fl.value / cnt.value, "bytes_per_launch": by.value / cnt.value,
                     "bound": bound, "achieved": rate / 1e12 if bound == "mfma" else rate / 1e9,
                     "unit": ("TOP/s" if name == "ozaki_gemm" else "TFLOP/s") if bound == "mfma" else "GB/s",
                     "peak": peak / 1e12 if bound == "mfma" else peak / 1e9, "frac": rate / peak}
    L.openblas_amd_ktimer_reset()
    return out


def roofline_of(kstats, name, traffic_file=None):
    k = kstats[name]
    return {"kernel": name, "bound": k["bound"], "achieved": k["achieved"], "peak": k["peak"], "unit": k["unit"],
            "frac": k["frac"], "traffic": profiled_traffic(traffic_file) if traffic_file else None,
            "launches": k["launches"], "avg_us": k["avg_us"], "flops_per_launch": k["flops_per_launch"],
            "bytes_per_launch": k["bytes_per_launch"]}


def dominant(kstats):
    return max(kstats, key=lambda n: kstats[n]["total_ms"]) if kstats else None


# ---- CPU reference timing ----------------------------------------------------------
def refblas_or_none():
    try:
        from oracle import refblas
        cfg = refblas.config()
        return refblas, cfg
    except Exception as e:   # no installed OpenBLAS: the CPU columns are reported absent
        log(f"cpu reference absent: {e}")
        return None, None


def cpu_median(fn, reps=3, warmup=1):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


# ---- C3: the N = 16384 system ---------------------------------------------------------
def make_system(L, n, kappa, nrhs, seed):
    """A = Q1 diag(s) Q2 built on the device with rank-1 updates (16 x O(n^2)),
    b Gaussian.  Returns the device buffer of A and b."""
    rng = np.random.default_rng(seed)
    s = kappa ** (-np.arange(n) / (n - 1))
    A0 = np.zeros((n, n), order="F")
    A0[np.diag_indices(n)] = s
    dA = oa.DeviceBuffer.from_array(A0)
    del A0
    dv = oa.DeviceBuffer(8 * n)
    dw = oa.DeviceBuffer(8 * n)
    for side in ("left", "right"):
        for _ in range(8):
            v = rng.standard_normal(n)
            L.openblas_amd_memcpy(dv.ptr, v.ctypes.data, 8 * n, 0)
            c = -2.0 / float(v @ v)
            if side == "left":   # A <- A - c' v (A^T v)^T
                L.cblas_dgemm(102, 112, 111, n, 1, n, 1.0, dA.ptr, n, dv.ptr, n, 0.0, dw.ptr, n)
                L.cblas_dgemm(102, 111, 112, n, n, 1, c, dv.ptr, n, dw.ptr, n, 1.0, dA.ptr, n)
            else:                # A <- A - c' (A v) v^T
                L.cblas_dgemm(102, 111, 111, n, 1, n, 1.0, dA.ptr, n, dv.ptr, n, 0.0, dw.ptr, n)
                L.cblas_dgemm(102, 111, 112, n, n, 1, c, dw.ptr, n, dv.ptr, n, 1.0, dA.ptr, n)
    b = np.asfortranarray(rng.standard_normal((n, nrhs)))
    dv.free()
    dw.free()
    return dA, b


def make_spd(L, n, kappa, seed):
    """A = Q diag(s) Q^T (Q the product of 8 Householder reflectors, each
    applied on both sides), s geometric from 1 to 1/kappa: symmetric
    positive definite with 2-norm condition kappa; built on the device."""
    rng = np.random.default_rng(seed)
    s = kappa ** (-np.arange(n) / (n - 1))
    A0 = np.zeros((n, n), order="F")
    A0[np.diag_indices(n)] = s
    dA = oa.DeviceBuffer.from_array(A0)
    del A0
    dv, dw = oa.DeviceBuffer(8 * n), oa.DeviceBuffer(8 * n)
    for _ in range(8):
        v = rng.standard_normal(n)
        L.openblas_amd_memcpy(dv.ptr, v.ctypes.data, 8 * n, 0)
        c = -2.0 / float(v @ v)
        L.cblas_dgemm(102, 112, 111, n, 1, n, 1.0, dA.ptr, n, dv.ptr, n, 0.0, dw.ptr, n)
        L.cblas_dgemm(102, 111, 112, n, n, 1, c, dv.ptr, n, dw.ptr, n, 1.0, dA.ptr, n)
        L.cblas_dgemm(102, 111, 111, n, 1, n, 1.0, dA.ptr, n, dv.ptr, n, 0.0, dw.ptr, n)
        L.cblas_dgemm(102, 111, 112, n, n, 1, c, dw.ptr, n, dv.ptr, n, 1.0, dA.ptr, n)
    dv.free()
    dw.free()
    return dA


def bench_cholesky(L, n, kappa, reps, refblas=None, cores=1, cfg=None):
    """Row f1: dsposv_ (FP32 Cholesky + FP64 refinement, lower triangle) on an
    SPD system, beside the FP64 and FP32 factorizations alone (dpotrf_,
    spotrf_); the reference CPU dsposv_ on the same system."""
    log(f"f1 building SPD A (n={n}, kappa={kappa:g})")
    dA = make_spd(L, n, kappa, seed=77)
    Ah = np.empty((n, n), order="F")
    dA.to_array(Ah)
    anorm = float(np.abs(Ah).sum(axis=1).max())
    rng = np.random.default_rng(78)
    xt = np.asfortranarray(rng.standard_normal((n, 1)))
    dxt = oa.DeviceBuffer.from_array(xt)
    db = oa.DeviceBuffer(8 * n)
    L.cblas_dgemm(102, 111, 111, n, 1, n, 1.0, dA.ptr, n, dxt.ptr, n, 0.0, db.ptr, n)
    b = db.to_array(np.empty((n, 1), order="F"))
    dW, dX = oa.DeviceBuffer(8 * n * n), oa.DeviceBuffer(8 * n)
    it, info = ctypes.c_int(0), ctypes.c_int(0)
    copy_ws = lambda: L.openblas_amd_memcpy(dW.ptr, dA.ptr, 8 * n * n, 2)
    ds = lambda: L.dsposv_(b"L", oa.ip(n), oa.ip(1), dW.ptr, oa.ip(n), db.ptr, oa.ip(n), dX.ptr, oa.ip(n), None,
                           None, ctypes.byref(it), ctypes.byref(info))
    ms_ds = float(np.median([timed_ms(L, ds, copy_ws) for _ in range(reps + 1)][1:]))
    it_ds, info_ds = it.value, info.value
    bwd, fwd = errors(L, n, dA, anorm, dX, b, xt, 1)
    dpo = lambda: L.dpotrf_(b"L", oa.ip(n), dW.ptr, oa.ip(n), ctypes.byref(info))
    ms_d = float(np.median([timed_ms(L, dpo, copy_ws) for _ in range(reps + 1)][1:]))
    info_d = info.value
    dW.free()
    Af = np.asfortranarray(Ah.astype(np.float32))
    dAf, dWf = oa.DeviceBuffer.from_array(Af), oa.DeviceBuffer(4 * n * n)
    del Af
    spo = lambda: L.spotrf_(b"L", oa.ip(n), dWf.ptr, oa.ip(n), ctypes.byref(info))
    ms_s = float(np.median([timed_ms(L, spo, lambda: L.openblas_amd_memcpy(dWf.ptr, dAf.ptr, 4 * n * n, 2))
                            for _ in range(reps + 1)][1:]))
    info_s = info.value
    for d in (dAf, dWf, dxt, db, dX, dA):
        d.free()
    fl = n ** 3 / 3.0
    rec = {"n": n, "kappa": kappa, "dsposv_ms": ms_ds, "iter": it_ds, "info": info_ds, "bwd_dsposv": bwd,
           "fwd_dsposv": fwd, "dpotrf_ms": ms_d, "dpotrf_tflops": fl / (ms_d * 1e-3) / 1e12, "dpotrf_info": info_d,
           "spotrf_ms": ms_s, "spotrf_tflops": fl / (ms_s * 1e-3) / 1e12, "spotrf_info": info_s,
           "dsposv_vs_dpotrf_factor_only": ms_d / ms_ds}
    log(f"f1 dsposv {ms_ds:.1f} ms (ITER {it_ds}, bwd {bwd:.1e}, fwd {fwd:.1e}); dpotrf {ms_d:.1f} ms, "
        f"spotrf {ms_s:.1f} ms")
    if refblas:
        t = cpu_median(lambda: refblas.dsposv(True, np.array(Ah, order="F"), np.array(b, order="F")), reps=3)
        rec["cpu"] = {"dsposv_s": t, "cores": cores, "kind": "reference",
                      "sample": f"dsposv_ N={n} NRHS=1, the full solve, 1 warm-up + median of 3, via {cfg}"}
        log(f"f1 cpu dsposv {t:.3f} s")
    del Ah
    return rec


class Solver:
    """Working storage for dsgesv_ / dgesv_ on a device-resident system: A is
    kept pristine in dA and every solve runs on a fresh copy dW (dgesv_
    overwrites it; dsgesv_ does when it falls back to FP64), made before the
    timer starts."""

    def __init__(self, L, n, dA):
        self.L, self.n, self.dA = L, n, dA
        self.dW = oa.DeviceBuffer(8 * n * n)
        self.dpiv = oa.DeviceBuffer(4 * n)
        self.it, self.info = ctypes.c_int(0), ctypes.c_int(0)

    def free(self):
        self.dW.free()
        self.dpiv.free()


def timed_ms(L, fn, prep=None):
    if prep:
        prep()
    L.openblas_amd_synchronize()
    L.openblas_amd_timer_start()
    fn()
    return L.openblas_amd_timer_stop()


def errors(L, n, dA, anorm, dX, b, xt, nrhs):
    """normwise backward error ||b - A x|| / (||A|| ||x|| + ||b||) and forward
    error ||x - x_true|| / ||x_true|| (infinity norms, worst column)."""
    x = dX.to_array(np.empty((n, nrhs), order="F"))
    dR = oa.DeviceBuffer.from_array(b)
    L.cblas_dgemm(102, 111, 111, n, nrhs, n, -1.0, dA.ptr, n, dX.ptr, n, 1.0, dR.ptr, n)
    r = dR.to_array(np.empty((n, nrhs), order="F"))
    dR.free()
    bwd = float((np.abs(r).max(0) / (anorm * np.abs(x).max(0) + np.abs(b).max(0))).max())
    fwd = float((np.abs(x - xt).max(0) / np.abs(xt).max(0)).max()) if xt is not None else None
    return bwd, fwd


def c3_case(L, n, kappa, nrhs, dA, anorm, reps, seed):
    """One (kappa, NRHS) case: b = A x_true, dsgesv_ and dgesv_ timed (median
    of reps after a warm-up), errors of both."""
    rng = np.random.default_rng(seed)
    xt = np.asfortranarray(rng.standard_normal((n, nrhs)))
    dxt = oa.DeviceBuffer.from_array(xt)
    db = oa.DeviceBuffer(8 * n * nrhs)
    L.cblas_dgemm(102, 111, 111, n, nrhs, n, 1.0, dA.ptr, n, dxt.ptr, n, 0.0, db.ptr, n)
    b = db.to_array(np.empty((n, nrhs), order="F"))
    S = Solver(L, n, dA)
    dX = oa.DeviceBuffer(8 * n * nrhs)
    dB2 = oa.DeviceBuffer(8 * n * nrhs)
    copy_ws = lambda: L.openblas_amd_memcpy(S.dW.ptr, dA.ptr, 8 * n * n, 2)
    # timing excludes the working copy of A (done before the timer starts)
    def mixed_only():
        L.dsgesv_(oa.ip(n), oa.ip(nrhs), S.dW.ptr, oa.ip(n), S.dpiv.ptr, db.ptr, oa.ip(n), dX.ptr, oa.ip(n), None,
                  None, ctypes.byref(S.it), ctypes.byref(S.info))

    def fp64_only():
        L.dgesv_(oa.ip(n), oa.ip(nrhs), S.dW.ptr, oa.ip(n), S.dpiv.ptr, dB2.ptr, oa.ip(n), ctypes.byref(S.info))
    reset_b = lambda: (copy_ws(), L.openblas_amd_memcpy(dB2.ptr, db.ptr, 8 * n * nrhs, 2))
    ms_mixed = [timed_ms(L, mixed_only, copy_ws) for _ in range(reps + 1)][1:]
    it, info_m = S.it.value, S.info.value
    bwd_m, fwd_m = errors(L, n, dA, anorm, dX, b, xt, nrhs)
    ms_fp64 = [timed_ms(L, fp64_only, reset_b) for _ in range(reps + 1)][1:]
    info_d = S.info.value
    bwd_d, fwd_d = errors(L, n, dA, anorm, dB2, b, xt, nrhs)
    m_ms, d_ms = float(np.median(ms_mixed)), float(np.median(ms_fp64))
    res = {"kappa": kappa, "nrhs": nrhs, "dsgesv_ms": m_ms, "dgesv_ms": d_ms, "speedup": d_ms / m_ms,
           "iter": it, "info": [info_m, info_d], "bwd_dsgesv": bwd_m, "bwd_dgesv": bwd_d, "fwd_dsgesv": fwd_m,
           "fwd_dgesv": fwd_d, "fwd_bound_30_kappa_eps": 30 * kappa * EPS64}
    for d in (dxt, db, dX, dB2):
        d.free()
    S.free()
    return res, b


# ---- C1 / C2 dense level 3 --------------------------------------------------------------
def gpu_gemm(L, n, dt, reps, ta, tb, beta, dA, dB, dC):
    f = L.cblas_dgemm if dt == np.float64 else L.cblas_sgemm
    call = lambda: f(102, ta, tb, n, n, n, 1.0, dA.ptr, n, dB.ptr, n, beta, dC.ptr, n)
    call()
    ts = [timed_ms(L, call) for _ in range(reps)]
    ms = float(np.median(ts))
    return ms, 2.0 * n ** 3 / (ms * 1e-3), call


def c1c2(L, reps, refblas, cores):
    out = {}
    peak = {np.float64: PEAK_FP64, np.float32: PEAK_FP32}
    rng = np.random.default_rng(7)
    # C1: 1024^3 NN, beta 0.5
    n = 1024
    A = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
    B = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
    C = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
    dA, dB, dC = (oa.DeviceBuffer.from_array(x) for x in (A, B, C))
    ms, fl, call = gpu_gemm(L, n, np.float64, 4 * reps, 111, 111, 0.5, dA, dB, dC)
    ks = ktimer_profile(L, call, lambda name: ("mfma", PEAK_FP64))
    # the reference harness's method (benchmark/gemm.c: wall clock around
    # OPENBLAS_LOOPS back-to-back calls, averaged): 10 synchronous calls per
    # loop, median over 5 loops; the HIP-event time of a single call is kept
    loops = []
    for _ in range(5):
        w0 = time.perf_counter()
        for _ in range(10):
            call()
        loops.append((time.perf_counter() - w0) / 10 * 1e3)
    wall = float(np.median(loops))
    c1 = {"ms": wall, "tflops": 2.0 * n ** 3 / (wall * 1e-3) / 1e12, "frac_of_fp64_peak": 2.0 * n ** 3 / (wall * 1e-3) / PEAK_FP64,
          "method": "wall clock per call, 10-call loop (benchmark/gemm.c), median of 5 loops",
          "event_ms_single_call": ms, "event_tflops_single_call": fl / 1e12,
          "roofline": roofline_of(ks, "dgemm", "dgemm_1024_nn_fp64.md") if "dgemm" in ks else None}
    if refblas:
        Cw = C.copy(order="F")
        t = cpu_median(lambda: refblas.cblas_gemm(False, False, 1.0, A, B, 0.5, Cw), reps=10)
        c1["cpu"] = {"ms": t * 1e3, "gflops": 2.0 * n ** 3 / t / 1e9, "cores": cores, "kind": "reference",
                     "sample": "the full call, 1 warm-up + median of 10"}
    out["c1_dgemm_1024"] = c1
    for d in (dA, dB, dC):
        d.free()
    log(f"C1 dgemm 1024: {wall:.4f} ms per call (10-call loop) {c1['tflops']:.1f} TF/s; single call, HIP events: "
        f"{ms:.4f} ms {fl / 1e12:.1f} TF/s")
    # C2: 8192^3, four transposes, FP64 and FP32; TRSM 8192 x 1024
    n = 8192
    c2 = {}
    A64 = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
    B64 = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
    for dt, tag in ((np.float64, "d"), (np.float32, "s")):
        A, B = A64.astype(dt, order="F"), B64.astype(dt, order="F")
        dA, dB, dC = oa.DeviceBuffer.from_array(A), oa.DeviceBuffer.from_array(B), oa.DeviceBuffer(A.nbytes)
        for ta, tb, name in ((111, 111, "NN"), (111, 112, "NT"), (112, 111, "TN"), (112, 112, "TT")):
            ms, fl, call = gpu_gemm(L, n, dt, reps, ta, tb, 0.0, dA, dB, dC)
            rec = {"ms": ms, "tflops": fl / 1e12, "frac": fl / peak[dt]}
            if name == "NN":
                ks = ktimer_profile(L, call, lambda nm: ("mfma", peak[dt]))
                kn = "dgemm" if tag == "d" else "sgemm"
                if kn in ks:
                    rec["roofline"] = roofline_of(ks, kn, "dgemm_8192_nn_fp64.md" if tag == "d" else None)
            if refblas:
                # bounded sample: the first 1024 columns of C (1/8 of the product)
                Cs = np.zeros((n, 1024), dtype=dt, order="F")
                Bs = np.asfortranarray(B[:, :1024]) if tb == 111 else np.asfortranarray(B[:1024, :])
                t = cpu_median(lambda: refblas.cblas_gemm(ta == 112, tb == 112, 1.0, A, Bs, 0.0, Cs), reps=3)
                rec["cpu"] = {"gflops": 2.0 * n * n * 1024 / t / 1e9, "cores": cores, "kind": "reference",
                              "sample": f"{n}x1024x{n} slab (1/8 of C), 1 warm-up + median of 3"}
            c2[f"{tag}gemm_{name}"] = rec
            log(f"C2 {tag}gemm {name}: {ms:.2f} ms {fl / 1e12:.1f} TF/s")
        for d in (dA, dB, dC):
            d.free()
        # TRSM L,L,N,N M=8192 NRHS=1024: unit-lower garbage poisoned above the diagonal
        nr = 1024
        T = np.tril(A64, -1) + np.diag(n + rng.uniform(0, 1, n))
        T = T + np.triu(np.full((n, n), np.nan), 1)
        T = T.astype(dt, order="F")
        Bt = np.asfortranarray(rng.uniform(-1, 1, (n, nr)).astype(dt))
        dT, dBt, dW = oa.DeviceBuffer.from_array(T), oa.DeviceBuffer.from_array(Bt), oa.DeviceBuffer(Bt.nbytes)
        f = L.cblas_dtrsm if tag == "d" else L.cblas_strsm
        call = lambda: f(102, 141, 122, 111, 131, n, nr, 1.0, dT.ptr, n, dW.ptr, n)
        reset = lambda: L.openblas_amd_memcpy(dW.ptr, dBt.ptr, Bt.nbytes, 2)
        ts = [timed_ms(L, call, reset) for _ in range(reps + 1)][1:]
        ms = float(np.median(ts))
        fl = float(n) * n * nr
        reset()
        ks = ktimer_profile(L, call, lambda nm: ("mfma", peak[dt]))
        rec = {"ms": ms, "tflops": fl / (ms * 1e-3) / 1e12, "frac": fl / (ms * 1e-3) / peak[dt]}
        kn = "dtrsm" if tag == "d" else "strsm"
        if kn in ks:
            rec["roofline"] = roofline_of(ks, kn)
        X = dW.to_array(np.empty_like(Bt))
        Tl = np.tril(T.astype(np.float64))
        rows = rng.choice(n, 16, replace=False)
        rec["multiply_back_rel"] = float(np.abs(Tl[rows] @ X.astype(np.float64) - Bt[rows]).max()
                                         / (np.abs(Tl[rows]).sum(1).max() * np.abs(X).max()))
        if refblas:
            Tc = np.asfortranarray(np.tril(T))   # the CPU path gets the same triangle
            Bc = Bt.copy(order="F")
            t = cpu_median(lambda: refblas.cblas_trsm(0, 1, 0, 1, 1.0, Tc, Bc.copy(order="F")), reps=3)
            rec["cpu"] = {"gflops": fl / t / 1e9, "cores": cores, "kind": "reference",
                          "sample": "the full call, 1 warm-up + median of 3"}
        c2[f"{tag}trsm_LLNN_8192x1024"] = rec
        log(f"C2 {tag}trsm: {ms:.2f} ms {rec['tflops']:.1f} TF/s")
        for d in (dT, dBt, dW):
            d.free()
    out["c2"] = c2
    return out


# ---- C4: INT8-slice emulated DGEMM ------------------------------------------------------
def bench_ozaki(L, n=8192, phis=(0.5, 2.0, 8.0), reps=3, refblas=None, cores=0):
    res = []
    for phi in phis:
        rng = np.random.default_rng(int(phi * 100) + 11)
        A = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, n)) * np.exp(phi * rng.standard_normal((n, n))))
        B = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, n)) * np.exp(phi * rng.standard_normal((n, n))))
        dA, dB = oa.DeviceBuffer.from_array(A), oa.DeviceBuffer.from_array(B)
        dC, dC64 = oa.DeviceBuffer(8 * n * n), oa.DeviceBuffer(8 * n * n)
        nat = lambda: L.cblas_dgemm(102, 111, 111, n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.0, dC64.ptr, n)
        nat()
        t64 = [timed_ms(L, nat) for _ in range(reps)]
        oz = lambda: L.openblas_amd_dgemm_ozaki(102, 111, 111, n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.0, dC.ptr, n, 0)
        s = oz()
        tz = [timed_ms(L, oz) for _ in range(reps)]
        ks = ktimer_profile(L, oz, lambda nm: ("mfma", PEAK_INT8 if nm == "ozaki_gemm" else PEAK_FP64))
        C = np.empty((n, n), order="F"); C64 = np.empty((n, n), order="F")
        dC.to_array(C); dC64.to_array(C64)
        rows = rng.choice(n, 32, replace=False)
        ref = A[rows] @ B
        gauge = np.abs(A[rows]) @ np.abs(B)
        err_oz = float(np.max(np.abs(C[rows] - ref) / gauge))
        err_64 = float(np.max(np.abs(C64[rows] - ref) / gauge))
        ms_z, ms_64 = float(np.median(tz)), float(np.median(t64))
        fl = 2.0 * n ** 3
        r = {"phi": phi, "slices": s, "native_fallback": s < 0, "ms": ms_z, "fp64_equiv_tflops": fl / ms_z / 1e9,
             "frac_of_fp64_peak": fl / (ms_z * 1e-3) / PEAK_FP64, "native_dgemm_ms": ms_64,
             "speedup_vs_native": ms_64 / ms_z, "err_vs_ref_normwise": err_oz, "native_err_vs_ref_normwise": err_64}
        if s > 0:
            pairs = s * (s + 1) // 2
            r["int8_tops"] = pairs * fl / ms_z / 1e9
            r["frac_of_int8_peak"] = pairs * fl / (ms_z * 1e-3) / PEAK_INT8
            if "ozaki_gemm" in ks:
                r["roofline"] = roofline_of(ks, "ozaki_gemm")
        if refblas:
            Cs = np.zeros((1024, n), order="F")
            As = np.asfortranarray(A[:1024])
            t = cpu_median(lambda: refblas.cblas_gemm(False, False, 1.0, As, B, 0.0, Cs), reps=3)
            r["cpu"] = {"dgemm_tflops": 2.0 * 1024 * n * n / t / 1e12, "cores": cores, "kind": "reference",
                        "sample": "native dgemm of the first 1024 rows of A, 1 warm-up + median of 3"}
        # scheme II (integer moduli + CRT) on the same operands
        oz2 = lambda: L.openblas_amd_dgemm_ozaki2(102, 111, 111, n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.0, dC.ptr, n, 0)
        nm = oz2()
        t2 = [timed_ms(L, oz2) for _ in range(reps)]
        ks2 = ktimer_profile(L, oz2, lambda nm_: ("mfma", PEAK_INT8 if nm_ == "ozaki_gemm" else PEAK_FP64))
        dC.to_array(C)
        ms_2 = float(np.median(t2))
        r2 = {"moduli": nm, "native_fallback": nm < 0, "ms": ms_2, "fp64_equiv_tflops": fl / ms_2 / 1e9,
              "speedup_vs_native": ms_64 / ms_2, "err_vs_ref_normwise": float(np.max(np.abs(C[rows] - ref) / gauge))}
        if nm > 0:
            r2["int8_tops"] = nm * fl / ms_2 / 1e9
            r2["frac_of_int8_peak"] = nm * fl / (ms_2 * 1e-3) / PEAK_INT8
            if "ozaki_gemm" in ks2:
                r2["roofline"] = roofline_of(ks2, "ozaki_gemm")
        r["scheme2"] = r2
        log(f"C4 ozaki phi={phi}: s={s} {ms_z:.1f} ms ({r['fp64_equiv_tflops']:.1f} TF/s eq, native {ms_64:.1f} ms) "
            f"err={err_oz:.2e} native_err={err_64:.2e}; scheme II: {nm} moduli {ms_2:.1f} ms "
            f"(x{ms_64 / ms_2:.2f} vs native) err={r2['err_vs_ref_normwise']:.2e}")
        res.append(r)
        for d in (dA, dB, dC, dC64):
            d.free()
        del A, B, C, C64
    return res


# ---- C5: batched small problems (10,000 problems, n ~ U{8..256}) -----------------------
def _ptrs(base, off, es):
    return (ctypes.c_void_p * len(off))(*[base + int(o) * es for o in off])


def _i32(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32))


def _pi(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def bench_batched(L, count=10000, seed=5, reps=3, refblas=None, cores=0, loop_sample=300, cpu_sample=1000):
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    ns = rng.integers(8, 257, count).astype(np.int64)
    sq = ns * ns
    off = np.concatenate([[0], np.cumsum(sq)[:-1]])
    voff = np.concatenate([[0], np.cumsum(ns)[:-1]])
    total, vtotal = int(sq.sum()), int(ns.sum())
    n32 = _i32(ns)
    gemm_flops = float((2 * ns ** 3).sum())
    out = {"count": count, "n_dist": "U{8..256}", "seed": seed}
    for tag, dt in (("d", np.float64), ("s", np.float32)):
        es = np.dtype(dt).itemsize
        peak = PEAK_FP64 if tag == "d" else PEAK_FP32
        peak_of = lambda nm: ("hbm", PEAK_HBM) if nm in ("batched_getrs",) else ("mfma", peak)
        r = {}
        A = rng.standard_normal(total).astype(dt)
        Bm = rng.standard_normal(total).astype(dt)
        dA, dB, dC, dW = (oa.DeviceBuffer.from_array(A), oa.DeviceBuffer.from_array(Bm), oa.DeviceBuffer(es * total),
                          oa.DeviceBuffer(es * total))
        pA, pB, pC, pW = (_ptrs(d.ptr, off, es) for d in (dA, dB, dC, dW))
        ct = ctypes.c_double if dt == np.float64 else ctypes.c_float
        gemm_b = L.cblas_dgemm_batch if tag == "d" else L.cblas_sgemm_batch
        one = (ct * count)(*([1.0] * count)); zero = (ct * count)(*([0.0] * count))
        tn = _i32(np.full(count, 111)); tt = _i32(np.full(count, 112)); gs = _i32(np.ones(count))

        def timed(fn, prep=None):
            if prep:
                prep()
            fn()   # first call: tuning + graph capture
            ts, ws = [], []
            for _ in range(reps):
                if prep:
                    prep()
                L.openblas_amd_synchronize()
                w0 = time.perf_counter()
                L.openblas_amd_timer_start()
                fn()
                ts.append(L.openblas_amd_timer_stop())
                ws.append((time.perf_counter() - w0) * 1e3)
            if prep:
                prep()
            ks = ktimer_profile(L, fn, peak_of)
            return float(np.median(ts)), float(np.median(ws)), ks

        def rec_of(ms, wall, ks, kname, flops):
            d = {"ms": ms, "wall_ms": wall, "tflops": flops / ms / 1e9, "frac_of_peak": flops / (ms * 1e-3) / peak}
            if kname in ks:
                d["roofline"] = roofline_of(ks, kname)
            return d

        g = lambda: gemm_b(102, _pi(tn), _pi(tn), _pi(n32), _pi(n32), _pi(n32), one, pA, _pi(n32), pB, _pi(n32), zero,
                           pC, _pi(n32), count, _pi(gs))
        r["gemm"] = rec_of(*timed(g), "batched_gemm", gemm_flops)
        f1 = L.cblas_dgemm if tag == "d" else L.cblas_sgemm
        L.openblas_amd_synchronize()
        w0 = time.perf_counter()
        for i in range(loop_sample):
            n = int(ns[i])
            f1(102, 111, 111, n, n, n, 1.0, pA[i], n, pB[i], n, 0.0, pC[i], n)
        L.openblas_amd_synchronize()
        r["gemm"]["launch_loop_ms_est"] = (time.perf_counter() - w0) * 1e3 * count / loop_sample

        dpiv = oa.DeviceBuffer(4 * vtotal)
        ppiv = _ptrs(dpiv.ptr, voff, 4)
        dbv, dx = oa.DeviceBuffer.from_array(rng.standard_normal(vtotal).astype(dt)), oa.DeviceBuffer(es * vtotal)
        pX = _ptrs(dx.ptr, voff, es)
        info = np.zeros(count, np.int32)
        fr = L.openblas_amd_dgetrf_batch if tag == "d" else L.openblas_amd_sgetrf_batch
        fs = L.openblas_amd_dgetrs_batch if tag == "d" else L.openblas_amd_sgetrs_batch
        reset_a = lambda: L.openblas_amd_memcpy(dW.ptr, dA.ptr, es * total, 2)
        ms, wall, ks = timed(lambda: fr(_pi(n32), _pi(n32), pW, _pi(n32), ppiv, _pi(info), count), reset_a)
        lu_flops = float((ns ** 3 - ns ** 3 / 3.0).sum())   # m n^2 - n^3/3 at m = n: 2/3 n^3
        r["getrf"] = rec_of(ms, wall, ks, "batched_getrf", lu_flops)
        r["getrf"]["info_nonzero"] = int((info != 0).sum())
        reset_b = lambda: L.openblas_amd_memcpy(dx.ptr, dbv.ptr, es * vtotal, 2)
        one_rhs = _i32(np.ones(count))
        ms, wall, ks = timed(lambda: fs(b"N", _pi(n32), _pi(one_rhs), pW, _pi(n32), ppiv, pX, _pi(n32), _pi(info),
                                        count), reset_b)
        byt = float((sq + 2 * ns).sum()) * es
        r["getrs"] = {"ms": ms, "wall_ms": wall, "gflops": float((2 * sq).sum()) / ms / 1e6,
                      "hbm_frac": byt / (ms * 1e-3) / PEAK_HBM}
        if "batched_getrs" in ks:
            r["getrs"]["roofline"] = roofline_of(ks, "batched_getrs")
        Xh = np.empty(vtotal, dt); dx.to_array(Xh)
        bh = np.empty(vtotal, dt); dbv.to_array(bh)
        worst = 0.0
        for i in rng.choice(count, 16, replace=False):
            n, o, v = int(ns[i]), int(off[i]), int(voff[i])
            Ai = A[o:o + n * n].reshape(n, n, order="F").astype(np.float64)
            xi, bi = Xh[v:v + n].astype(np.float64), bh[v:v + n].astype(np.float64)
            worst = max(worst, float(np.abs(bi - Ai @ xi).max() / (np.abs(Ai).sum(1).max() * np.abs(xi).max())))
        r["getrs"]["worst_backward_error_16"] = worst
        f1 = L.dgetrf_ if tag == "d" else L.sgetrf_
        reset_a()
        L.openblas_amd_synchronize()
        ii = ctypes.c_int(0)
        w0 = time.perf_counter()
        for i in range(loop_sample):
            n = int(ns[i])
            f1(oa.ip(n), oa.ip(n), pW[i], oa.ip(n), ppiv[i], ctypes.byref(ii))
        r["getrf"]["launch_loop_ms_est"] = (time.perf_counter() - w0) * 1e3 * count / loop_sample

        Ch = np.zeros(total, dt)
        for i in range(count):
            n, o = int(ns[i]), int(off[i])
            Ch[o:o + n * n:n + 1] = n
        dS = oa.DeviceBuffer.from_array(Ch)
        del Ch
        pS = _ptrs(dS.ptr, off, es)
        gemm_b(102, _pi(tn), _pi(tt), _pi(n32), _pi(n32), _pi(n32), one, pB, _pi(n32), pB, _pi(n32), one, pS,
               _pi(n32), count, _pi(gs))
        fp = L.openblas_amd_dpotrf_batch if tag == "d" else L.openblas_amd_spotrf_batch
        reset_s = lambda: L.openblas_amd_memcpy(dW.ptr, dS.ptr, es * total, 2)
        ms, wall, ks = timed(lambda: fp(b"L", _pi(n32), pW, _pi(n32), _pi(info), count), reset_s)
        r["potrf"] = rec_of(ms, wall, ks, "batched_potrf", float((ns ** 3 / 3.0).sum()))
        r["potrf"]["info_nonzero"] = int((info != 0).sum())

        if refblas and tag == "d":
            sel = range(cpu_sample)
            As = [A[off[i]:off[i] + sq[i]].reshape(ns[i], ns[i], order="F").copy(order="F") for i in sel]
            Bs = [Bm[off[i]:off[i] + sq[i]].reshape(ns[i], ns[i], order="F").copy(order="F") for i in sel]
            Cs = [np.zeros((ns[i], ns[i]), order="F") for i in sel]
            Ss = [np.asfortranarray(b @ b.T + b.shape[0] * np.eye(b.shape[0])) for b in Bs]
            scale = count / cpu_sample

            def loops(parallel):
                """gemm / getrf+getrs / potrf loops; parallel: one problem per
                host thread with the library single-threaded (an OpenMP
                parallel-for over problems)"""
                def gemm_i(i):
                    refblas.cblas_gemm(False, False, 1.0, As[i], Bs[i], 0.0, Cs[i])

                def lu_i(i):
                    a = As[i].copy(order="F")
                    pv, _ = refblas.getrf(a)
                    refblas.getrs(False, a, pv, np.ones((a.shape[0], 1), order="F"))

                def po_i(i):
                    refblas.potrf(True, Ss[i].copy(order="F"))
                res = {}
                for key, fn in (("gemm", gemm_i), ("getrf_getrs", lu_i), ("potrf", po_i)):
                    if parallel:
                        with ThreadPoolExecutor(max_workers=cores) as ex:
                            fn(0)
                            t0 = time.perf_counter()
                            list(ex.map(fn, range(cpu_sample)))
                            res[key + "_ms_est"] = (time.perf_counter() - t0) * 1e3 * scale
                    else:
                        res[key + "_ms_est"] = cpu_median(lambda: [fn(i) for i in range(cpu_sample)], reps=3) * 1e3 * scale
                return res
            # the installed library's cblas_dgemm_batch crashes on problems with
            # m*n*k <= 100^3 (its small-matrix path), so its CPU timing is a loop
            refblas.set_num_threads(cores)
            lib_thr = loops(False)
            refblas.set_num_threads(1)
            per_prob = loops(True)
            refblas.set_num_threads(cores)
            r["cpu"] = {"cores": cores, "kind": "reference",
                        "sample": f"first {cpu_sample} problems, scaled to {count}",
                        "library_threads": lib_thr, "thread_per_problem": per_prob}
        log(f"C5 batched {tag}: " + json.dumps({k: {kk: (round(vv, 4) if isinstance(vv, float) else vv)
                                                   for kk, vv in v.items() if not isinstance(vv, dict)}
                                               for k, v in r.items() if isinstance(v, dict)}))
        out[tag] = r
        for d in (dA, dB, dC, dW, dpiv, dbv, dx, dS):
            d.free()
        del A, Bm
    return out


def step_peak(name):
    if name in ("lu_panel", "trsv_lookback"):
        return "hbm", PEAK_HBM
    return "mfma", PEAK_FP32 if name in ("lu_trailing_gemm", "lu_next_gemm", "lu_node_gemm") else PEAK_FP64


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--kappa", type=float, default=1e6)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--section", default="all",
                    choices=["all", "solve", "sweep", "dense", "ozaki", "batched", "cholesky"],
                    help="run one workload only (for profiling); prints a partial JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    L = oa.lib()
    arch = ctypes.create_string_buffer(64)
    ncu, clk = ctypes.c_int(0), ctypes.c_int(0)
    L.openblas_amd_device_info(arch, 64, ctypes.byref(ncu), ctypes.byref(clk))
    log(f"device {arch.value.decode()} CUs={ncu.value} clock={clk.value // 1000} MHz")
    refblas, cfg = (None, None) if args.no_cpu or args.section != "all" else refblas_or_none()
    cores = os.cpu_count() or 1
    if refblas:
        refblas.set_num_threads(cores)
    sec = args.section
    out_extra = {}
    if sec in ("dense",):
        print(json.dumps({"section": sec, **c1c2(L, 5, None, cores)}), flush=True)
        return
    if sec == "ozaki":
        print(json.dumps({"section": sec, "ozaki": bench_ozaki(L, reps=2)}), flush=True)
        return
    if sec == "batched":
        print(json.dumps({"section": sec, "batched": bench_batched(L, reps=2, loop_sample=50)}), flush=True)
        return
    if sec == "cholesky":
        print(json.dumps({"section": sec, "dsposv_16384": bench_cholesky(L, args.n, args.kappa, reps=2)}), flush=True)
        return

    # ---- headline: the N=16384, kappa=1e6, NRHS=1 system ----
    n, nrhs = args.n, args.nrhs
    log(f"building A (n={n}, kappa={args.kappa:g})")
    dA, b = make_system(L, n, args.kappa, nrhs, seed=2024)
    A_host = np.empty((n, n), order="F")
    dA.to_array(A_host)
    anorm = float(np.abs(A_host).sum(axis=1).max())
    db = oa.DeviceBuffer.from_array(b)
    dX = oa.DeviceBuffer(8 * n * nrhs)
    S = Solver(L, n, dA)
    solve = lambda: L.dsgesv_(oa.ip(n), oa.ip(nrhs), S.dW.ptr, oa.ip(n), S.dpiv.ptr, db.ptr, oa.ip(n), dX.ptr,
                              oa.ip(n), None, None, ctypes.byref(S.it), ctypes.byref(S.info))
    copy_ws = lambda: L.openblas_amd_memcpy(S.dW.ptr, dA.ptr, 8 * n * n, 2)
    for _ in range(args.warmup):
        copy_ws()
        solve()
    ir_ms = []
    L.openblas_amd_synchronize()
    t_wall0 = time.perf_counter()
    for _ in range(args.steps):
        copy_ws()   # dsgesv_ keeps A unless it falls back; the copy keeps every step identical (not timed)
        L.openblas_amd_timer_start()
        solve()
        ir_ms.append(L.openblas_amd_timer_stop())
    L.openblas_amd_synchronize()
    t_wall = time.perf_counter() - t_wall0
    ir_iter, ir_info = S.it.value, S.info.value
    bwd_ir, _ = errors(L, n, dA, anorm, dX, b, None, nrhs)
    ir_med = float(np.median(ir_ms))
    log(f"dsgesv: {ir_med:.1f} ms/solve ITER={ir_iter} info={ir_info} bwd={bwd_ir:.2e}")
    copy_ws()
    step_k = ktimer_profile(L, solve, step_peak)
    # plain FP64 solve on the same system
    dB2 = oa.DeviceBuffer(8 * n * nrhs)
    fp64 = lambda: L.dgesv_(oa.ip(n), oa.ip(nrhs), S.dW.ptr, oa.ip(n), S.dpiv.ptr, dB2.ptr, oa.ip(n),
                            ctypes.byref(S.info))
    reset64 = lambda: (copy_ws(), L.openblas_amd_memcpy(dB2.ptr, db.ptr, 8 * n * nrhs, 2))
    f64_ms = [timed_ms(L, fp64, reset64) for _ in range(args.warmup + args.steps)][args.warmup:]
    bwd_64, _ = errors(L, n, dA, anorm, dB2, b, None, nrhs)
    f64_med = float(np.median(f64_ms))
    log(f"dgesv: {f64_med:.1f} ms/solve info={S.info.value} bwd={bwd_64:.2e}")
    reset64()
    step64_k = ktimer_profile(L, fp64, lambda nm: ("hbm", PEAK_HBM) if nm in ("lu_panel", "trsv_lookback")
                              else ("mfma", PEAK_FP64))
    dB2.free()
    S.free()
    for d in (db, dX):
        d.free()
    # the CPU reference on the headline system
    cpu = None
    if refblas:
        log("timing the reference CPU path (dsgesv_, dgesv_) ...")
        A1 = np.array(A_host, order="F")
        Xc = [None]

        def cpu_ds():
            Xc[0], _, it_, _ = refblas.dsgesv(np.array(A_host, order="F"), np.array(b, order="F"))

        def cpu_dg():
            refblas.gesv(np.array(A_host, order="F"), np.array(b, order="F"))
        t_ds = cpu_median(cpu_ds, reps=3)
        t_dg = cpu_median(cpu_dg, reps=3)
        cpu = {"value": 1.0 / t_ds, "unit": "solves/s", "cores": cores, "kind": "reference",
               "sample": f"dsgesv_ N={n} NRHS={nrhs}, the full solve, 1 warm-up + median of 3, via {cfg}",
               "dsgesv_s": t_ds, "dgesv_s": t_dg,
               "bwd_err": float(np.abs(b - A_host @ Xc[0]).max() / (anorm * np.abs(Xc[0]).max() + np.abs(b).max()))}
        del A1
        log(f"cpu: dsgesv {t_ds:.3f} s dgesv {t_dg:.3f} s on {cores} cores")
    del A_host
    dA.free()

    # ---- C3 sweep ----
    sweep = []
    if sec in ("all", "sweep"):
        for kappa in (1e2, 1e6, 1e10):
            log(f"C3 building A kappa={kappa:g}")
            dAk, _ = make_system(L, n, kappa, 1, seed=2024)
            Ah = np.empty((n, n), order="F")
            dAk.to_array(Ah)
            an = float(np.abs(Ah).sum(axis=1).max())
            del Ah
            for nr in (1, 16):
                res, _ = c3_case(L, n, kappa, nr, dAk, an, reps=2, seed=int(np.log10(kappa)) * 10 + nr)
                log(f"C3 kappa={kappa:g} nrhs={nr}: dsgesv {res['dsgesv_ms']:.1f} ms (ITER {res['iter']}) "
                    f"dgesv {res['dgesv_ms']:.1f} ms x{res['speedup']:.2f} bwd {res['bwd_dsgesv']:.1e}/"
                    f"{res['bwd_dgesv']:.1e} fwd {res['fwd_dsgesv']:.1e}/{res['fwd_dgesv']:.1e}")
                sweep.append(res)
            dAk.free()
    if sec in ("all",):
        out_extra.update(c1c2(L, 5, refblas, cores))
        out_extra["ozaki_dgemm_8192"] = bench_ozaki(L, refblas=refblas, cores=cores)
        out_extra["batched_10000"] = bench_batched(L, refblas=refblas, cores=cores)
        out_extra["dsposv_16384"] = bench_cholesky(L, n, args.kappa, reps=2, refblas=refblas, cores=cores, cfg=cfg)

    flops_solve = 2.0 / 3.0 * n ** 3 + 2.0 * n * n * nrhs
    dom = dominant(step_k)
    roof = roofline_of(step_k, dom, "lu_panel_16384_fp32.md" if dom == "lu_panel" else None) if dom else None
    out = {
        "metric": "dgesv N=16384 time to FP64 backward error: FP32-LU+IR vs FP64 LU; DGEMM TF/s",
        "value": 1000.0 / ir_med,
        "unit": "solves/s (FP32-LU+FP64-IR dgesv, N=16384, NRHS=1, to FP64 backward error)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ir_med,
        "wall_s_timed_region": t_wall,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 LU + fp64 refinement",
        "data": "synthetic: A=Q1*diag(s)*Q2 (8+8 Householder reflectors, s geometric 1..1/kappa), b~N(0,1), seeded",
        "config": {"workload": f"dsgesv_ N={n} NRHS={nrhs} kappa={args.kappa:g} (device-resident)", "n": n,
                   "nrhs": nrhs, "kappa": args.kappa},
        "ir": {"ms": ir_med, "iter": ir_iter, "info": ir_info, "backward_error": bwd_ir,
               "effective_tflops": flops_solve / (ir_med * 1e-3) / 1e12, "step_kernels": step_k},
        "fp64_lu": {"ms": f64_med, "backward_error": bwd_64, "tflops": flops_solve / (f64_med * 1e-3) / 1e12,
                    "step_kernels": step64_k},
        "speedup_ir_vs_fp64": f64_med / ir_med,
        "roofline": roof,
        "cpu_baseline": cpu,
        "c3_sweep": sweep,
        **out_extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
