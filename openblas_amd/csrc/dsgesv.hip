// Mixed-precision solve: FP32 LU on MFMA + FP64 iterative refinement, with
// the contract of netlib dsgesv (lapack-netlib/SRC/dsgesv.f:193-430):
//   ANRM = ||A||_inf, CTE = ANRM*eps*sqrt(N)*BWDMAX (BWDMAX = 1, eps = 2^-53)
//   demote B and A (overflow -> ITER = -2), SGETRF (info != 0 -> ITER = -3),
//   SGETRS, promote, R = B - A*X, converged when for every RHS column
//   max|R| <= max|X|*CTE (ITER = number of refinement steps), otherwise up
//   to ITERMAX = 30 corrections; no convergence -> ITER = -31; every
//   negative ITER falls back to DGETRF + DGETRS on the original A.
//
// Everything stays on the GPU: the demotion is fused with the transpose the
// LU kernels want (At layout, plus the permutation-tracking row), the
// residual is an FP64 MFMA GEMM, the norms are device reductions; only the
// per-column (max|R|, max|X|) pairs cross to the host for the loop test.
#include "kernels.h"
#include "lu_internal.h"
#include "dsgesv.h"
#include <cfloat>
#include <cmath>
#include <vector>

namespace oba {

// row sums of |A|: thread per row (coalesced down the columns), the columns
// split into RS_CH chunks over blockIdx.y for parallelism; the partial sums
// are added in a fixed order (deterministic)
constexpr int RS_CH = 32;
__global__ void row_abs_sum_part_kernel(long n, const double* __restrict__ a, long lda, double* __restrict__ part) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long c0 = n * blockIdx.y / RS_CH, c1 = n * (blockIdx.y + 1) / RS_CH;
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
__global__ void row_abs_sum_fin_kernel(long n, const double* __restrict__ part, double* __restrict__ rs) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int c = 0; c < RS_CH; c++) s += part[c * n + i];
  rs[i] = s;
}

// max with NaN propagation (DLANGE 'I' returns NaN if any entry is NaN)
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

// x (+)= (double) s   (SLAG2D, then the DAXPY of dsgesv.f:376-378)
__global__ void promote_kernel(long n, long nrhs, const float* __restrict__ s, long lds, double* __restrict__ x,
                               long ldx, bool add) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= n * nrhs) return;
  long i = idx % n, c = idx / n;
  double v = (double)s[i + c * lds];
  x[i + c * ldx] = add ? x[i + c * ldx] + v : v;
}

static inline unsigned nb256(long total) { return (unsigned)((total + 255) / 256); }

void dsgesv_device(long n, long nrhs, double* a, long lda, int* ipiv_dev, const double* b, long ldb, double* x,
                   long ldx, int* iter, int* info, DsgesvStats* stats, hipStream_t s) {
  *iter = 0;
  *info = 0;
  if (n == 0) return;
  const int ITERMAX = 30;
  const double BWDMAX = 1.0;
  const double eps = std::ldexp(1.0, -53);

  DevBuf small(4096);
  double* d_anrm = small.as<double>();
  int* d_ovf = (int*)(small.as<char>() + 64);
  double* d_norms = (double*)(small.as<char>() + 128);   // up to (4096-128)/16 columns
  DevBuf nrm_big;
  if (nrhs > (4096 - 128) / 16) { nrm_big = DevBuf(sizeof(double) * 2 * nrhs); d_norms = nrm_big.as<double>(); }

  // ANRM, CTE
  {
    DevBuf rs(sizeof(double) * n), part(sizeof(double) * n * RS_CH);
    row_abs_sum_part_kernel<<<dim3(nb256(n), RS_CH), 256, 0, s>>>(n, a, lda, part.as<double>());
    row_abs_sum_fin_kernel<<<nb256(n), 256, 0, s>>>(n, part.as<double>(), rs.as<double>());
    max_nan_kernel<<<1, 256, 0, s>>>(n, rs.as<double>(), d_anrm);
    OBA_HIP(hipGetLastError());
  }
  double anrm = 0.0;
  OBA_HIP(hipMemcpyAsync(&anrm, d_anrm, sizeof(double), hipMemcpyDeviceToHost, s));
  OBA_HIP(hipStreamSynchronize(s));
  const double cte = anrm * eps * std::sqrt((double)n) * BWDMAX;

  const long ldt = (n + 16) & ~15L;
  DevBuf sat(sizeof(float) * (size_t)ldt * (size_t)n);
  DevBuf sx(sizeof(float) * (size_t)n * (size_t)nrhs);
  DevBuf rbuf(sizeof(double) * (size_t)n * (size_t)nrhs);
  DevBuf permb(sizeof(int) * n);
  float* SAt = sat.as<float>();
  float* SX = sx.as<float>();
  double* R = rbuf.as<double>();
  int* perm = permb.as<int>();
  int h_ovf = 0;
  int it = 0;
  bool fallback = false;

  auto read_ovf = [&]() {
    OBA_HIP(hipMemcpyAsync(&h_ovf, d_ovf, sizeof(int), hipMemcpyDeviceToHost, s));
    OBA_HIP(hipStreamSynchronize(s));
    return h_ovf != 0;
  };
  auto residual = [&]() {   // R = B - A*X
    OBA_HIP(hipMemcpy2DAsync(R, sizeof(double) * n, b, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                             hipMemcpyDeviceToDevice, s));
    gemm<double>(false, false, n, nrhs, n, -1.0, a, lda, x, ldx, 1.0, R, n, s);
  };
  auto converged = [&]() {
    col_absmax2_kernel<<<(unsigned)nrhs, 256, 0, s>>>(n, x, ldx, R, n, d_norms);
    OBA_HIP(hipGetLastError());
    std::vector<double> h(2 * nrhs);
    OBA_HIP(hipMemcpyAsync(h.data(), d_norms, sizeof(double) * 2 * nrhs, hipMemcpyDeviceToHost, s));
    OBA_HIP(hipStreamSynchronize(s));
    for (long c = 0; c < nrhs; c++)
      if (h[2 * c + 1] > h[2 * c] * cte) return false;
    return true;
  };

  OBA_HIP(hipMemsetAsync(d_ovf, 0, sizeof(int), s));
  demote_kernel<<<nb256(n * nrhs), 256, 0, s>>>(n, nrhs, b, ldb, SX, n, d_ovf);
  OBA_HIP(hipGetLastError());
  if (read_ovf()) { it = -2; fallback = true; }
  if (!fallback) {
    transpose<double, float>(n, n, a, lda, SAt, ldt, d_ovf, s);
    if (read_ovf()) { it = -2; fallback = true; }
  }
  if (!fallback) {
    int sinfo = getrf_t<float>(n, n, SAt, ldt, ipiv_dev, true, s);
    if (stats) stats->sgetrf_done = true;
    if (sinfo != 0) { it = -3; fallback = true; }
  }
  if (!fallback) {
    perm_from_tracked<float>(SAt, ldt, n, n, perm, s);
    getrs_t<float>(n, nrhs, SAt, ldt, perm, SX, n, s);
    promote_kernel<<<nb256(n * nrhs), 256, 0, s>>>(n, nrhs, SX, n, x, ldx, false);
    OBA_HIP(hipGetLastError());
    residual();
    if (converged()) { *iter = 0; return; }
    bool ok = false;
    for (int iiter = 1; iiter <= ITERMAX; iiter++) {
      OBA_HIP(hipMemsetAsync(d_ovf, 0, sizeof(int), s));
      demote_kernel<<<nb256(n * nrhs), 256, 0, s>>>(n, nrhs, R, n, SX, n, d_ovf);
      OBA_HIP(hipGetLastError());
      if (read_ovf()) { it = -2; fallback = true; break; }
      getrs_t<float>(n, nrhs, SAt, ldt, perm, SX, n, s);
      promote_kernel<<<nb256(n * nrhs), 256, 0, s>>>(n, nrhs, SX, n, x, ldx, true);
      OBA_HIP(hipGetLastError());
      residual();
      if (converged()) { *iter = iiter; ok = true; break; }
    }
    if (ok) return;
    if (!fallback) { it = -ITERMAX - 1; fallback = true; }
  }
  // FP64 fallback on the original A (A is overwritten by its factors)
  *iter = it;
  if (stats) stats->fallback = true;
  int dinfo = getrf_perm<double>(n, n, a, lda, ipiv_dev, perm, s);
  if (dinfo != 0) { *info = dinfo; return; }
  OBA_HIP(hipMemcpy2DAsync(x, sizeof(double) * ldx, b, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                           hipMemcpyDeviceToDevice, s));
  getrs_perm<double>(false, n, nrhs, a, lda, perm, x, ldx, s);
}

}  // namespace oba
