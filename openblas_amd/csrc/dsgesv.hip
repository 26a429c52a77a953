// Mixed-precision solves: an FP32 factorization on MFMA plus FP64 iterative
// refinement, with the contracts of netlib
//   dsgesv (LU,       lapack-netlib/SRC/dsgesv.f:193-430) and
//   dsposv (Cholesky, lapack-netlib/SRC/dsposv.f:197-440):
//   ANRM = ||A||_inf (DLANGE / DLANSY 'I'), CTE = ANRM*eps*sqrt(N)*BWDMAX
//   (BWDMAX = 1, eps = 2^-53); demote B and A (DLAG2S / DLAT2S overflow ->
//   ITER = -2); SGETRF / SPOTRF (info != 0 -> ITER = -3); solve; promote;
//   R = B - A*X (DGEMM / DSYMM); converged when for every RHS column
//   max|R| <= max|X|*CTE (ITER = number of refinement steps), otherwise up to
//   ITERMAX = 30 corrections; no convergence -> ITER = -31; every negative
//   ITER falls back to DGETRF+DGETRS / DPOTRF+DPOTRS on the original A.
//
// Everything stays on the GPU.  The LU works on the demoted transpose (the
// layout the LU kernels want, plus the permutation-tracking row), Cholesky on
// the demoted triangle in place.  The host only learns, once per step, the
// overflow flags and the per-column (max|R|, max|X|) pairs: a demotion
// overflow is detected at the next synchronization rather than right away
// (the work in between is discarded and the FP64 fallback runs, as in the
// reference), so the FP32 factorization starts without a host round trip.
#include "kernels.h"
#include "lu_internal.h"
#include "dsgesv.h"
#include <cfloat>
#include <cmath>
#include <vector>

namespace oba {

static inline unsigned nb256(long total) { return (unsigned)((total + 255) / 256); }

// ---- ||A||_inf ----------------------------------------------------------------
// row sums of |A| over the columns in chunk blockIdx.y (thread per row,
// coalesced down the columns); tri: 0 every column, 1 only j <= i (lower
// triangle), 2 only j >= i (upper triangle)
constexpr int RS_CH = 32;
__global__ void row_abs_sum_part_kernel(long n, const double* __restrict__ a, long lda, int tri,
                                        double* __restrict__ part) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  long c0 = n * blockIdx.y / RS_CH, c1 = n * (blockIdx.y + 1) / RS_CH;
  if (tri == 1) c1 = min(c1, i + 1);
  if (tri == 2) c0 = max(c0, i);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  long j = c0;
  for (; j + 4 <= c1; j += 4) {
    s0 += fabs(a[i + j * lda]);
    s1 += fabs(a[i + (j + 1) * lda]);
    s2 += fabs(a[i + (j + 2) * lda]);
    s3 += fabs(a[i + (j + 3) * lda]);
  }
  for (; j < c1; j++) s0 += fabs(a[i + j * lda]);
  part[blockIdx.y * n + i] = (s0 + s1) + (s2 + s3);
}

// symmetric storage: the mirrored part of row i is the stored column i
// strictly below (tri 1) or above (tri 2) the diagonal; one wavefront per
// column, 4 columns per workgroup
__global__ __launch_bounds__(256) void sym_col_abs_sum_kernel(long n, const double* __restrict__ a, long lda, int tri,
                                                              double* __restrict__ colsum) {
  const long c = blockIdx.x * 4L + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= n) return;
  const long r0 = tri == 1 ? c + 1 : 0, r1 = tri == 1 ? n : c;
  double s = 0.0;
  for (long r = r0 + lane; r < r1; r += 64) s += fabs(a[r + c * lda]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) colsum[c] = s;
}

// rs[i] = sum of the chunk partials (+ the mirrored column sum)
__global__ void row_abs_sum_fin_kernel(long n, const double* __restrict__ part, const double* __restrict__ colsum,
                                       double* __restrict__ rs) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int c = 0; c < RS_CH; c++) s += part[c * n + i];
  if (colsum) s += colsum[i];
  rs[i] = s;
}

// max with NaN propagation (DLANGE / DLANSY 'I' return NaN if any entry is NaN)
__global__ void max_nan_kernel(long n, const double* __restrict__ v, double* __restrict__ out) {
  __shared__ double sm[256];
  double m = 0.0;
  for (long i = threadIdx.x; i < n; i += 256) {
    double x = v[i];
    if (isnan(x) || x > m) m = isnan(m) ? m : x;
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      double a = sm[threadIdx.x], b = sm[threadIdx.x + s];
      sm[threadIdx.x] = (isnan(a) || isnan(b)) ? NAN : (a > b ? a : b);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sm[0];
}

// ||A||_inf of a general (tri 0) or symmetric (tri 1 lower / 2 upper
// stored) matrix into *out (device)
static void anorm_inf(long n, const double* a, long lda, int tri, double* out, hipStream_t s) {
  DevBuf rs(sizeof(double) * n), part(sizeof(double) * n * RS_CH), cs(tri ? sizeof(double) * n : 0);
  row_abs_sum_part_kernel<<<dim3(nb256(n), RS_CH), 256, 0, s>>>(n, a, lda, tri, part.as<double>());
  OBA_HIP(hipGetLastError());
  if (tri) {
    sym_col_abs_sum_kernel<<<(unsigned)((n + 3) / 4), 256, 0, s>>>(n, a, lda, tri, cs.as<double>());
    OBA_HIP(hipGetLastError());
  }
  row_abs_sum_fin_kernel<<<nb256(n), 256, 0, s>>>(n, part.as<double>(), tri ? cs.as<double>() : nullptr,
                                                  rs.as<double>());
  OBA_HIP(hipGetLastError());
  max_nan_kernel<<<1, 256, 0, s>>>(n, rs.as<double>(), out);
  OBA_HIP(hipGetLastError());
}

// ---- per-column norms of the stopping test -----------------------------------
// per-column max|.| of two n x nrhs matrices (IDAMAX-based norms of
// dsgesv.f:335-337); one workgroup per column, NaN entries are skipped.
__global__ void col_absmax2_kernel(long n, const double* __restrict__ x, long ldx, const double* __restrict__ r,
                                   long ldr, double* __restrict__ out) {
  __shared__ double sx[256], sr[256];
  const long c = blockIdx.x;
  double mx = 0.0, mr = 0.0;
  for (long i = threadIdx.x; i < n; i += 256) {
    double a = fabs(x[i + c * ldx]), b = fabs(r[i + c * ldr]);
    if (a > mx) mx = a;
    if (b > mr) mr = b;
  }
  sx[threadIdx.x] = mx;
  sr[threadIdx.x] = mr;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sx[threadIdx.x] = fmax(sx[threadIdx.x], sx[threadIdx.x + s]);
      sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[2 * c] = sx[0]; out[2 * c + 1] = sr[0]; }
}

// ---- demotion / promotion ----------------------------------------------------------
// demote an n x nrhs block with the DLAG2S overflow test
__global__ void demote_kernel(long n, long nrhs, const double* __restrict__ a, long lda, float* __restrict__ s,
                              long lds, int* overflow) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= n * nrhs) return;
  long i = idx % n, c = idx / n;
  double v = a[i + c * lda];
  if (v < -(double)FLT_MAX || v > (double)FLT_MAX) atomicOr(overflow, 1);
  s[i + c * lds] = (float)v;
}

// DLAT2S: demote the stored triangle (tri 1 lower / 2 upper) with the
// overflow test; the other triangle of the FP32 copy is zeroed
__global__ void demote_tri_kernel(long n, const double* __restrict__ a, long lda, int tri, float* __restrict__ s,
                                  long lds, int* overflow) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool ovf = false;
  for (long j = blockIdx.y; j < n; j += gridDim.y) {
    const bool stored = tri == 1 ? i >= j : i <= j;
    float out = 0.f;
    if (stored) {
      const double v = a[i + j * lda];
      ovf |= v < -(double)FLT_MAX || v > (double)FLT_MAX;
      out = (float)v;
    }
    s[i + j * lds] = out;
  }
  if (ovf) atomicOr(overflow, 1);
}

// x (+)= (double) s   (SLAG2D, then the DAXPY of dsgesv.f:376-378)
__global__ void promote_kernel(long n, long nrhs, const float* __restrict__ s, long lds, double* __restrict__ x,
                               long ldx, bool add) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= n * nrhs) return;
  long i = idx % n, c = idx / n;
  double v = (double)s[i + c * lds];
  x[i + c * ldx] = add ? x[i + c * ldx] + v : v;
}

// ---- symmetric residual R = B - A X (DSYMM of dsposv.f:326) --------------------------
// A symmetric with only the tri triangle stored: A X = S X + E^T X, S = the
// stored triangle with its diagonal, E = the stored triangle strictly off the
// diagonal.  Part 1 (S X): thread per row, columns in chunks over blockIdx.y
// (coalesced down the columns); part 2 (E^T X): wavefront per row of the
// result = a dot product down the stored column.  Up to SR_NR right-hand
// sides per pass; the partial sums are added in a fixed order.
constexpr int SR_NR = 4;
constexpr int SR_CH = 32;

template <int NR>
__global__ __launch_bounds__(256) void sym_resid_rows_kernel(long n, const double* __restrict__ a, long lda, int tri,
                                                             const double* __restrict__ x, long ldx,
                                                             double* __restrict__ part) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= n) return;
  long c0 = n * blockIdx.y / SR_CH, c1 = n * (blockIdx.y + 1) / SR_CH;
  if (tri == 1) c1 = min(c1, i + 1);
  else c0 = max(c0, i);
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) acc[r] = 0.0;
#pragma unroll 8
  for (long j = c0; j < c1; j++) {
    const double v = a[i + j * lda];
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r] = fma(v, x[j + r * ldx], acc[r]);
  }
#pragma unroll
  for (int r = 0; r < NR; r++) part[((long)blockIdx.y * NR + r) * n + i] = acc[r];
}

template <int NR>
__global__ __launch_bounds__(256) void sym_resid_cols_kernel(long n, const double* __restrict__ a, long lda, int tri,
                                                             const double* __restrict__ x, long ldx,
                                                             double* __restrict__ colpart) {
  const long c = blockIdx.x * 4L + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= n) return;
  const long r0 = tri == 1 ? c + 1 : 0, r1 = tri == 1 ? n : c;
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) acc[r] = 0.0;
  for (long k = r0 + lane; k < r1; k += 64) {
    const double v = a[k + c * lda];
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r] = fma(v, x[k + r * ldx], acc[r]);
  }
#pragma unroll
  for (int r = 0; r < NR; r++) {
    double s = acc[r];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) colpart[(long)r * n + c] = s;
  }
}

template <int NR>
__global__ void sym_resid_fin_kernel(long n, const double* __restrict__ part, const double* __restrict__ colpart,
                                     const double* __restrict__ b, long ldb, double* __restrict__ rr, long ldr) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    double s = 0.0;
    for (int c = 0; c < SR_CH; c++) s += part[((long)c * NR + r) * n + i];
    s += colpart[(long)r * n + i];
    rr[i + r * ldr] = b[i + r * ldb] - s;
  }
}

template <int NR>
static void sym_resid_pass(long n, const double* a, long lda, int tri, const double* x, long ldx, const double* b,
                           long ldb, double* rr, long ldr, double* part, double* colpart, hipStream_t s) {
  sym_resid_rows_kernel<NR><<<dim3(nb256(n), SR_CH), 256, 0, s>>>(n, a, lda, tri, x, ldx, part);
  OBA_HIP(hipGetLastError());
  sym_resid_cols_kernel<NR><<<(unsigned)((n + 3) / 4), 256, 0, s>>>(n, a, lda, tri, x, ldx, colpart);
  OBA_HIP(hipGetLastError());
  sym_resid_fin_kernel<NR><<<nb256(n), 256, 0, s>>>(n, part, colpart, b, ldb, rr, ldr);
  OBA_HIP(hipGetLastError());
}

// R (ld n) = B - A_sym X
static void sym_residual(long n, long nrhs, const double* a, long lda, int tri, const double* x, long ldx,
                         const double* b, long ldb, double* rr, hipStream_t s) {
  DevBuf part(sizeof(double) * (size_t)SR_CH * SR_NR * n), colpart(sizeof(double) * (size_t)SR_NR * n);
  for (long c = 0; c < nrhs; c += SR_NR) {
    const int k = (int)std::min<long>(SR_NR, nrhs - c);
    const double* xc = x + c * ldx;
    const double* bc = b + c * ldb;
    double* rc = rr + c * n;
    switch (k) {
      case 1: sym_resid_pass<1>(n, a, lda, tri, xc, ldx, bc, ldb, rc, n, part.as<double>(), colpart.as<double>(), s); break;
      case 2: sym_resid_pass<2>(n, a, lda, tri, xc, ldx, bc, ldb, rc, n, part.as<double>(), colpart.as<double>(), s); break;
      case 3: sym_resid_pass<3>(n, a, lda, tri, xc, ldx, bc, ldb, rc, n, part.as<double>(), colpart.as<double>(), s); break;
      default: sym_resid_pass<4>(n, a, lda, tri, xc, ldx, bc, ldb, rc, n, part.as<double>(), colpart.as<double>(), s); break;
    }
  }
}

// ---- Cholesky solves ---------------------------------------------------------------------
// Solve A X = B with the Cholesky factor in the tri triangle of `f` (ld ldf),
// in place on B (ld ldb): lower L L^T, upper U^T U (potrs semantics).
template <typename T>
void potrs_device(bool lower, long n, long nrhs, const T* f, long ldf, T* b, long ldb, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return;
  if (nrhs <= 16) {   // one trsv_lookback launch per triangle (a chain per column)
    if (lower) {   // L y = b: L(i,k) = f[i + k*ldf]; L^T x = y: L^T(i,k) = f[k + i*ldf]
      trsv_lookback<T>(true, false, n, f, 1, ldf, b, ldb, nrhs, s);
      trsv_lookback<T>(false, false, n, f, ldf, 1, b, ldb, nrhs, s);
    } else {       // U^T y = b: U^T(i,k) = f[k + i*ldf]; U x = y: U(i,k) = f[i + k*ldf]
      trsv_lookback<T>(true, false, n, f, ldf, 1, b, ldb, nrhs, s);
      trsv_lookback<T>(false, false, n, f, 1, ldf, b, ldb, nrhs, s);
    }
    return;
  }
  if (lower) {
    trsm<T>(false, true, false, false, n, nrhs, T(1), f, ldf, b, ldb, s);
    trsm<T>(false, true, true, false, n, nrhs, T(1), f, ldf, b, ldb, s);
  } else {
    trsm<T>(false, false, true, false, n, nrhs, T(1), f, ldf, b, ldb, s);
    trsm<T>(false, false, false, false, n, nrhs, T(1), f, ldf, b, ldb, s);
  }
}
template void potrs_device<double>(bool, long, long, const double*, long, double*, long, hipStream_t);
template void potrs_device<float>(bool, long, long, const float*, long, float*, long, hipStream_t);

// ---- the refinement loop shared by both solvers ---------------------------------------
// State of one solve: device scratch and the host-side loop test.
struct IrState {
  long n, nrhs;
  const double* b;
  long ldb;
  double* x;
  long ldx;
  hipStream_t s;
  DevBuf small, nrm_big, sxb, rbuf;
  double* d_anrm = nullptr;
  int* d_ovf = nullptr;       // demotion overflow flag (DLAG2S / DLAT2S INFO)
  double* d_norms = nullptr;  // per column (max|X|, max|R|)
  float* SX = nullptr;
  double* R = nullptr;
  double cte = 0.0;
  bool have_cte = false;

  IrState(long n_, long nrhs_, const double* b_, long ldb_, double* x_, long ldx_, hipStream_t s_)
      : n(n_), nrhs(nrhs_), b(b_), ldb(ldb_), x(x_), ldx(ldx_), s(s_), small(4096) {
    d_anrm = small.as<double>();
    d_ovf = (int*)(small.as<char>() + 64);
    d_norms = (double*)(small.as<char>() + 128);   // up to (4096-128)/16 columns
    if (nrhs > (4096 - 128) / 16) { nrm_big = DevBuf(sizeof(double) * 2 * nrhs); d_norms = nrm_big.as<double>(); }
    if (nrhs > 0) {
      sxb = DevBuf(sizeof(float) * (size_t)n * (size_t)nrhs);
      rbuf = DevBuf(sizeof(double) * (size_t)n * (size_t)nrhs);
      SX = sxb.as<float>();
      R = rbuf.as<double>();
    }
    OBA_HIP(hipMemsetAsync(d_ovf, 0, sizeof(int), s));
  }
  // SX <- (float) src (n x nrhs), overflow into d_ovf
  void demote(const double* src, long lds) {
    if (nrhs <= 0) return;
    demote_kernel<<<nb256(n * nrhs), 256, 0, s>>>(n, nrhs, src, lds, SX, n, d_ovf);
    OBA_HIP(hipGetLastError());
  }
  void promote(bool add) {
    if (nrhs <= 0) return;
    promote_kernel<<<nb256(n * nrhs), 256, 0, s>>>(n, nrhs, SX, n, x, ldx, add);
    OBA_HIP(hipGetLastError());
  }
  // one synchronization: reads the overflow flag, ANRM (first time) and the
  // per-column norms (when `norms`); returns the overflow flag
  bool sync(bool norms, bool* conv) {
    if (norms && nrhs > 0) {
      col_absmax2_kernel<<<(unsigned)nrhs, 256, 0, s>>>(n, x, ldx, R, n, d_norms);
      OBA_HIP(hipGetLastError());
    }
    int ovf = 0;
    double anrm = 0.0;
    std::vector<double> h(norms ? 2 * std::max(nrhs, 1L) : 0);
    OBA_HIP(hipMemcpyAsync(&ovf, d_ovf, sizeof(int), hipMemcpyDeviceToHost, s));
    if (!have_cte) OBA_HIP(hipMemcpyAsync(&anrm, d_anrm, sizeof(double), hipMemcpyDeviceToHost, s));
    if (norms && nrhs > 0) OBA_HIP(hipMemcpyAsync(h.data(), d_norms, sizeof(double) * 2 * nrhs, hipMemcpyDeviceToHost, s));
    OBA_HIP(hipStreamSynchronize(s));
    check_device_error();
    if (!have_cte) {
      cte = anrm * std::ldexp(1.0, -53) * std::sqrt((double)n) * 1.0;   // BWDMAX = 1
      have_cte = true;
    }
    if (conv) {
      *conv = true;
      for (long c = 0; c < nrhs; c++)
        if (h[2 * c + 1] > h[2 * c] * cte) *conv = false;
    }
    return ovf != 0;
  }
};

// The loop of dsgesv.f:318-395 / dsposv.f:318-395 after the FP32 factorization:
// solve32(SX) solves with the FP32 factors in place, residual() computes
// R = B - A X.  Returns ITER >= 0 on convergence, or -2 / -31.
template <typename Solve, typename Resid>
static int refine(IrState& st, Solve&& solve32, Resid&& residual) {
  const int ITERMAX = 30;
  solve32();
  st.promote(false);
  residual();
  bool conv = false;
  if (st.sync(true, &conv)) return -2;   // overflow in the demotion of B or A
  if (conv) return 0;
  for (int iiter = 1; iiter <= ITERMAX; iiter++) {
    st.demote(st.R, st.n);
    solve32();
    st.promote(true);
    residual();
    if (st.sync(true, &conv)) return -2;   // overflow in the demotion of R
    if (conv) return iiter;
  }
  return -ITERMAX - 1;
}

void dsgesv_device(long n, long nrhs, double* a, long lda, int* ipiv_dev, const double* b, long ldb, double* x,
                   long ldx, int* iter, int* info, DsgesvStats* stats, hipStream_t s) {
  *iter = 0;
  *info = 0;
  if (n == 0) return;
  IrState st(n, nrhs, b, ldb, x, ldx, s);
  anorm_inf(n, a, lda, 0, st.d_anrm, s);
  const long ldt = (n + 16) & ~15L;
  DevBuf sat(sizeof(float) * (size_t)ldt * (size_t)n);
  DevBuf permb(sizeof(int) * n);
  float* SAt = sat.as<float>();
  int* perm = permb.as<int>();
  int it = 0;
  // demote B and A (fused with the transpose the LU wants); the overflow
  // flag is read together with SGETRF's info
  st.demote(b, ldb);
  transpose<double, float>(n, n, a, lda, SAt, ldt, st.d_ovf, s);
  const int sinfo = getrf_t<float>(n, n, SAt, ldt, ipiv_dev, true, s);   // synchronizes
  if (stats) stats->sgetrf_done = true;
  if (st.sync(false, nullptr)) it = -2;
  else if (sinfo != 0) it = -3;
  if (it == 0) {
    perm_from_tracked<float>(SAt, ldt, n, n, perm, s);
    it = refine(
        st, [&] { getrs_t<float>(n, nrhs, SAt, ldt, perm, st.SX, n, s); },
        [&] {
          if (nrhs <= 0) return;
          OBA_HIP(hipMemcpy2DAsync(st.R, sizeof(double) * n, b, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                                   hipMemcpyDeviceToDevice, s));
          gemm<double>(false, false, n, nrhs, n, -1.0, a, lda, x, ldx, 1.0, st.R, n, s);
        });
    if (it >= 0) { *iter = it; return; }
  }
  // FP64 fallback on the original A (A is overwritten by its factors)
  *iter = it;
  if (stats) stats->fallback = true;
  int dinfo = getrf_perm<double>(n, n, a, lda, ipiv_dev, perm, s);
  if (dinfo != 0) { *info = dinfo; return; }
  if (nrhs <= 0) return;
  OBA_HIP(hipMemcpy2DAsync(x, sizeof(double) * ldx, b, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                           hipMemcpyDeviceToDevice, s));
  getrs_perm<double>(false, n, nrhs, a, lda, perm, x, ldx, s);
}

void dsposv_device(bool lower, long n, long nrhs, double* a, long lda, const double* b, long ldb, double* x, long ldx,
                   int* iter, int* info, DsgesvStats* stats, hipStream_t s) {
  *iter = 0;
  *info = 0;
  if (n == 0) return;
  const int tri = lower ? 1 : 2;
  IrState st(n, nrhs, b, ldb, x, ldx, s);
  anorm_inf(n, a, lda, tri, st.d_anrm, s);
  const long lds = (n + 15) & ~15L;
  DevBuf sab(sizeof(float) * (size_t)lds * (size_t)n);
  float* SA = sab.as<float>();
  int it = 0;
  st.demote(b, ldb);
  demote_tri_kernel<<<dim3(nb256(n), (unsigned)std::min(n, 4096L)), 256, 0, s>>>(n, a, lda, tri, SA, lds, st.d_ovf);
  OBA_HIP(hipGetLastError());
  const int sinfo = potrf<float>(lower, n, SA, lds, s);   // synchronizes
  if (stats) stats->sgetrf_done = true;
  if (st.sync(false, nullptr)) it = -2;
  else if (sinfo != 0) it = -3;
  if (it == 0) {
    it = refine(
        st, [&] { potrs_device<float>(lower, n, nrhs, SA, lds, st.SX, n, s); },
        [&] {
          if (nrhs <= 0) return;
          sym_residual(n, nrhs, a, lda, tri, x, ldx, b, ldb, st.R, s);
        });
    if (it >= 0) { *iter = it; return; }
  }
  *iter = it;
  if (stats) stats->fallback = true;
  const int dinfo = potrf<double>(lower, n, a, lda, s);
  if (dinfo != 0) { *info = dinfo; return; }
  if (nrhs <= 0) return;
  OBA_HIP(hipMemcpy2DAsync(x, sizeof(double) * ldx, b, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                           hipMemcpyDeviceToDevice, s));
  potrs_device<double>(lower, n, nrhs, a, lda, x, ldx, s);
}

}  // namespace oba
