// LU factorization with partial pivoting (getrf), row interchanges (laswp),
// and the solve with the factors (getrs), device pointers, column-major API.
//
// Reference semantics (lapack/getf2/getf2_k.c:46-125, getrf_single.c:47-180,
// getrs_single.c:42-68, laswp/generic/laswp_k_1.c:48):
//  * pivot = first row of maximal |a| in the column (IAMAX), ipiv 1-based
//    and absolute;
//  * rows are exchanged and the column scaled by 1/pivot only when
//    pivot != 0 and |pivot| >= the smallest normal number; an exact zero
//    pivot sets info = j+1 (first one wins) and the factorization goes on;
//  * the interchanges are applied to all columns left and right of a panel.
//
// MI355X design (not the reference's packed GotoBLAS loops):
//  * the factorization runs on the TRANSPOSE At = A^T kept column-major in a
//    pool buffer, so that an interchange of two rows of A is an exchange of
//    two contiguous columns of At: every laswp is a coalesced HBM stream
//    instead of 8-byte gathers with stride lda;
//  * Toledo-style recursion on column halves puts the flops into large MFMA
//    GEMMs (At22 -= At21*At12) and one right-side unit-upper TRSM;
//  * leaves of 32 columns are factored by one cooperative kernel: each
//    workgroup keeps 512 rows x 32 columns of the panel in registers, the
//    per-column pivot search is a wavefront max-reduction (ties -> lowest
//    row), then one grid barrier per column exchanges the per-workgroup
//    candidates (value, row index and the whole 32-wide candidate row), so
//    every workgroup applies the interchange, scaling and rank-1 update to
//    its own rows without a second barrier;
//  * an extra row of At carries the original row numbers through every
//    interchange, which yields the composed permutation for the solves.
#include "kernels.h"
#include "lu_internal.h"
#include <cfloat>
#include <climits>
#include <cmath>
#include <utility>
#include <vector>

namespace oba {

constexpr int LU_W = 32;        // leaf panel width
constexpr int LU_RPT = 2;       // rows per thread in the panel kernel
constexpr int LU_PMAX = 256;    // max workgroups of the panel kernel

template <typename F, int... J>
__device__ __attribute__((always_inline)) inline void static_for_impl(F&& f, std::integer_sequence<int, J...>) {
  (f(std::integral_constant<int, J>{}), ...);
}
template <int N, typename F>
__device__ __attribute__((always_inline)) inline void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <typename T> struct MinNormal;
template <> struct MinNormal<double> { static constexpr double v = DBL_MIN; };
template <> struct MinNormal<float> { static constexpr float v = FLT_MIN; };

// ---- transposes (optionally demoting, with an overflow flag) --------------
template <typename Ti, typename To>
__global__ __launch_bounds__(256) void transpose_kernel(long rows, long cols, const Ti* __restrict__ in, long ldi,
                                                        To* __restrict__ out, long ldo, int* overflow) {
  __shared__ To tile[32][33];
  const long r0 = (long)blockIdx.x * 32, c0 = (long)blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  bool ovf = false;
  for (int k = ty; k < 32; k += 8) {
    long r = r0 + tx, c = c0 + k;
    if (r < rows && c < cols) {
      Ti v = in[r + c * ldi];
      if (sizeof(To) < sizeof(Ti)) ovf |= !(fabs((double)v) <= (double)FLT_MAX) && !isnan((double)v);
      tile[k][tx] = (To)v;
    }
  }
  if (overflow && ovf) atomicOr(overflow, 1);
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    long c = c0 + tx, r = r0 + k;   // out(c, r) = in(r, c)
    if (r < rows && c < cols) out[c + r * ldo] = tile[tx][k];
  }
}

template <typename Ti, typename To>
void transpose(long rows, long cols, const Ti* in, long ldi, To* out, long ldo, int* overflow, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return;
  dim3 grid((unsigned)((rows + 31) / 32), (unsigned)((cols + 31) / 32));
  transpose_kernel<Ti, To><<<grid, 256, 0, s>>>(rows, cols, in, ldi, out, ldo, overflow);
  OBA_HIP(hipGetLastError());
}

template <typename T>
__global__ void iota_row_kernel(T* row, long n, long stride) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) row[i * stride] = (T)i;
}

// ---- cooperative panel kernel ----------------------------------------------
// Factors A(j0:m, j0:j0+w) held as At(j0:j0+w, j0:m).  Workgroup g owns
// panel rows [g*R, (g+1)*R), thread t rows g*R + q*256 + t.
template <typename T, int W, int RPT>
__global__ __launch_bounds__(256) void lu_panel_kernel(T* __restrict__ At, long ldt, long m, long j0, int w,
                                                       int* __restrict__ ipiv, int* __restrict__ info,
                                                       unsigned* counter, unsigned base, unsigned* err,
                                                       double* cval, int* cidx, T* crow, T* rowj) {
  const int P = gridDim.x, g = blockIdx.x, t = threadIdx.x;
  const long mp = m - j0;
  constexpr long R = 256L * RPT;
  __shared__ double s_val[4];
  __shared__ int s_idx[4];
  __shared__ int s_best, s_win, s_dead;
  __shared__ T u[W], rj[W];

  T x[RPT][W];
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const long lr = g * R + q * 256 + t;
    const T* src = At + j0 + (j0 + lr) * ldt;
#pragma unroll
    for (int c = 0; c < W; c++) x[q][c] = (lr < mp && c < w) ? src[c] : T(0);
  }
  if (t == 0) s_dead = 0;
  __syncthreads();

  bool dead = false;
  // the column loop is expanded at compile time so that x[][] stays in VGPRs
  static_for<W>([&](auto jc) __attribute__((always_inline)) {
    constexpr int j = decltype(jc)::value;
    if (j < w && !dead) {
    const int buf = j & 1;
    // (a) local candidate: max |x|, lowest row on ties, NaN never wins
    double bv = -1.0;
    int bi = INT_MAX;
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const long lr = g * R + q * 256 + t;
      if (lr >= j && lr < mp) {
        double v = fabs((double)x[q][j]);
        if (v > bv) { bv = v; bi = (int)lr; }
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      double ov = __shfl_xor(bv, off);
      int oi = __shfl_xor(bi, off);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if ((t & 63) == 0) { s_val[t >> 6] = bv; s_idx[t >> 6] = bi; }
    __syncthreads();
    if (t == 0) {
      double v = s_val[0];
      int id = s_idx[0];
      for (int k = 1; k < 4; k++)
        if (s_val[k] > v || (s_val[k] == v && s_idx[k] < id)) { v = s_val[k]; id = s_idx[k]; }
      s_best = id;
      cval[buf * LU_PMAX + g] = v;
      cidx[buf * LU_PMAX + g] = id;
    }
    __syncthreads();
    // (b) publish the candidate row and the current row j
    {
      const int best = s_best;
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        const long lr = g * R + q * 256 + t;
        if (lr == best) {
#pragma unroll
          for (int c = 0; c < W; c++)
            if (c < w) crow[((long)buf * LU_PMAX + g) * W + c] = x[q][c];
        }
        if (lr == j) {
#pragma unroll
          for (int c = 0; c < W; c++)
            if (c < w) rowj[buf * W + c] = x[q][c];
        }
      }
    }
    __syncthreads();
    // (c) grid barrier (bounded spin; a timeout is reported, never hangs)
    if (t == 0 && !s_dead) {
      __threadfence();
      atomicAdd(counter, 1u);
      const unsigned target = base + (unsigned)P * (unsigned)(j + 1);
      long spins = 0;
      while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++spins > (1L << 22)) { atomicExch(err, 1u); s_dead = 1; break; }
      }
      __threadfence();
    }
    __syncthreads();
    dead = s_dead != 0;
    if (!dead) {
    // (d) global pivot: reduce the P candidates
    {
      double gv = -1.0;
      int gi = INT_MAX;
      if (t < P) { gv = cval[buf * LU_PMAX + t]; gi = cidx[buf * LU_PMAX + t]; }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        double ov = __shfl_xor(gv, off);
        int oi = __shfl_xor(gi, off);
        if (ov > gv || (ov == gv && oi < gi)) { gv = ov; gi = oi; }
      }
      if ((t & 63) == 0) { s_val[t >> 6] = gv; s_idx[t >> 6] = gi; }
      __syncthreads();
      if (t == 0) {
        double v = s_val[0];
        int id = s_idx[0];
        for (int k = 1; k < 4; k++)
          if (s_val[k] > v || (s_val[k] == v && s_idx[k] < id)) { v = s_val[k]; id = s_idx[k]; }
        s_win = id;
      }
      __syncthreads();
    }
    const int win = s_win;
    const int p = (win == INT_MAX) ? j : win;   // all-NaN column: keep row j
    if (t < w) {
      rj[t] = rowj[buf * W + t];
      u[t] = (win == INT_MAX) ? rj[t] : crow[((long)buf * LU_PMAX + p / R) * W + t];
    }
    __syncthreads();
    const T piv = u[j];
    const bool nz = piv != T(0);
    const bool doscale = nz && fabs(piv) >= MinNormal<T>::v;
    if (g == 0 && t == 0) {
      ipiv[j0 + j] = (int)(j0 + p + 1);
      if (!nz && *info == 0) *info = (int)(j0 + j + 1);
    }
    const T rinv = doscale ? T(1) / piv : T(1);
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const long lr = g * R + q * 256 + t;
      if (lr < mp) {
        if (p != j) {
          if (lr == j) {
#pragma unroll
            for (int c = 0; c < W; c++)
              if (c < w && (c > j || doscale)) x[q][c] = u[c];
          } else if (lr == p) {
#pragma unroll
            for (int c = 0; c < W; c++)
              if (c < w && (c > j || doscale)) x[q][c] = rj[c];
          }
        }
        if (lr > j) {
          T l = x[q][j];
          if (doscale) { l *= rinv; x[q][j] = l; }
#pragma unroll
          for (int c = j + 1; c < W; c++)
            if (c < w) x[q][c] -= l * u[c];
        }
      }
    }
    __syncthreads();
    }  // !dead
    }  // j < w
  });

#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const long lr = g * R + q * 256 + t;
    if (lr < mp) {
      T* dst = At + j0 + (j0 + lr) * ldt;
#pragma unroll
      for (int c = 0; c < W; c++)
        if (c < w) dst[c] = x[q][c];
    }
  }
}

// Apply the interchanges of one panel (rows j0..j0+w-1 of A) to every other
// column of A, i.e. to rows [0, nrows_t) \ [j0, j0+w) of At.
template <typename T>
__global__ void lu_swap_kernel(T* __restrict__ At, long ldt, long nrows_t, long j0, int w, const int* __restrict__ ipiv) {
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= nrows_t || (r >= j0 && r < j0 + w)) return;
  for (int k = 0; k < w; k++) {
    const long a = j0 + k, p = ipiv[j0 + k] - 1;
    if (p != a) {
      T* x = At + r + a * ldt;
      T* y = At + r + p * ldt;
      T tmp = *x;
      *x = *y;
      *y = tmp;
    }
  }
}

template <typename T>
struct LuCtx {
  T* At;
  long ldt, m, nrows_t;
  int* ipiv;
  int* info;
  unsigned* counter;
  unsigned* err;
  unsigned base = 0;
  double* cval;
  int* cidx;
  T* crow;
  T* rowj;
  hipStream_t s;
  int max_coop_blocks;
};

template <typename T>
static void lu_panel(LuCtx<T>& c, long j0, int w) {
  const long mp = c.m - j0;
  const long R = 256L * LU_RPT;
  long P = (mp + R - 1) / R;
  if (P > LU_PMAX || P > c.max_coop_blocks) fatal("getrf: too many rows for the panel kernel", nullptr);
  void* args[] = {&c.At, &c.ldt, &c.m, &j0, &w, &c.ipiv, &c.info, &c.counter, &c.base, &c.err,
                  &c.cval, &c.cidx, &c.crow, &c.rowj};
  OBA_HIP(hipLaunchCooperativeKernel((const void*)lu_panel_kernel<T, LU_W, LU_RPT>, dim3((unsigned)P), dim3(256),
                                     args, 0, c.s));
  c.base += (unsigned)P * (unsigned)w;
  long others = c.nrows_t - w;
  if (others > 0) {
    lu_swap_kernel<T><<<(unsigned)((c.nrows_t + 255) / 256), 256, 0, c.s>>>(c.At, c.ldt, c.nrows_t, j0, w, c.ipiv);
    OBA_HIP(hipGetLastError());
  }
}

template <typename T>
static void lu_rec(LuCtx<T>& c, long j0, long nc) {
  if (nc <= LU_W) { lu_panel<T>(c, j0, (int)nc); return; }
  long n1 = ((nc / 2) + LU_W - 1) / LU_W * LU_W;
  if (n1 >= nc) n1 = nc - LU_W;
  const long n2 = nc - n1;
  lu_rec<T>(c, j0, n1);
  T* At = c.At;
  const long ldt = c.ldt;
  // A12 <- L11^{-1} A12   <=>   X = At(j0+n1:j0+nc, j0:j0+n1) <- X * U^{-1},
  // U = unit upper part of At(j0:j0+n1, j0:j0+n1)
  trsm<T>(true, false, false, true, n2, n1, T(1), At + j0 + j0 * ldt, ldt, At + (j0 + n1) + j0 * ldt, ldt, c.s);
  // A22 -= A21*A12   <=>   At(j0+n1:j0+nc, j0+n1:m) -= X * At(j0:j0+n1, j0+n1:m)
  const long below = c.m - (j0 + n1);
  if (below > 0)
    gemm<T>(false, false, n2, below, n1, T(-1), At + (j0 + n1) + j0 * ldt, ldt, At + j0 + (j0 + n1) * ldt, ldt,
            T(1), At + (j0 + n1) + (j0 + n1) * ldt, ldt, c.s);
  lu_rec<T>(c, j0 + n1, n2);
}

template <typename T>
int getrf_t(long m, long n, T* At, long ldt, int* ipiv_dev, bool track_perm, hipStream_t s) {
  const long mn = std::min(m, n);
  if (mn <= 0) return 0;
  DevBuf ws(4096 + sizeof(double) * 2 * LU_PMAX + sizeof(int) * 2 * LU_PMAX + sizeof(T) * (2 * LU_PMAX * LU_W + 2 * LU_W));
  char* p = ws.as<char>();
  LuCtx<T> c;
  c.At = At; c.ldt = ldt; c.m = m; c.nrows_t = n + (track_perm ? 1 : 0); c.ipiv = ipiv_dev; c.s = s;
  c.info = (int*)p;
  c.counter = (unsigned*)(p + 64);
  c.err = (unsigned*)(p + 128);
  c.cval = (double*)(p + 4096);
  c.cidx = (int*)(c.cval + 2 * LU_PMAX);
  c.crow = (T*)(c.cidx + 2 * LU_PMAX);
  c.rowj = c.crow + 2 * LU_PMAX * LU_W;
  OBA_HIP(hipMemsetAsync(p, 0, 4096, s));
  int per_cu = 0;
  OBA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)lu_panel_kernel<T, LU_W, LU_RPT>, 256, 0));
  c.max_coop_blocks = per_cu * rt().num_cu;
  if (track_perm) {
    iota_row_kernel<T><<<(unsigned)((m + 255) / 256), 256, 0, s>>>(At + n, m, ldt);
    OBA_HIP(hipGetLastError());
  }
  lu_rec<T>(c, 0, mn);
  if (n > mn) trsm<T>(true, false, false, true, n - mn, mn, T(1), At, ldt, At + mn, ldt, s);
  int h[3] = {0, 0, 0};
  OBA_HIP(hipMemcpyAsync(&h[0], c.info, sizeof(int), hipMemcpyDeviceToHost, s));
  OBA_HIP(hipMemcpyAsync(&h[1], c.err, sizeof(int), hipMemcpyDeviceToHost, s));
  OBA_HIP(hipStreamSynchronize(s));
  if (h[1]) fatal("getrf: panel grid barrier timed out", nullptr);
  return h[0];
}

template <typename T>
void perm_from_tracked(const T* At, long ldt, long m, long n, int* perm_dev, hipStream_t s);

template <typename T>
int getrf_perm(long m, long n, T* a, long lda, int* ipiv_dev, int* perm_dev, hipStream_t s) {
  if (m <= 0 || n <= 0) return 0;
  const long ldt = (n + 16) & ~15L;   // room for the permutation row
  DevBuf at(sizeof(T) * (size_t)ldt * (size_t)m);
  transpose<T, T>(m, n, a, lda, at.as<T>(), ldt, nullptr, s);
  int info = getrf_t<T>(m, n, at.as<T>(), ldt, ipiv_dev, perm_dev != nullptr, s);
  if (perm_dev) perm_from_tracked<T>(at.as<T>(), ldt, m, n, perm_dev, s);
  transpose<T, T>(n, m, at.as<T>(), ldt, a, lda, nullptr, s);
  return info;
}

template <typename T>
int getrf(long m, long n, T* a, long lda, int* ipiv_dev, hipStream_t s) {
  return getrf_perm<T>(m, n, a, lda, ipiv_dev, nullptr, s);
}

// ---- permutations and solves --------------------------------------------------
// b'(i, :) = b(perm[i], :) (inverse=false) or b'(perm[i], :) = b(i, :).
template <typename T>
__global__ void permute_rows_kernel(long n, long nrhs, const int* __restrict__ perm, const T* __restrict__ src,
                                    long lds, T* __restrict__ dst, long ldd, bool inverse) {
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= n * nrhs) return;
  const long i = idx % n, c = idx / n;
  const long p = perm[i];
  if (!inverse) dst[i + c * ldd] = src[p + c * lds];
  else dst[p + c * ldd] = src[i + c * lds];
}

template <typename T>
void permute_rows(bool inverse, long n, long nrhs, const int* perm_dev, T* b, long ldb, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return;
  DevBuf tmp(sizeof(T) * (size_t)n * (size_t)nrhs);
  const long total = n * nrhs;
  permute_rows_kernel<T><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(n, nrhs, perm_dev, b, ldb, tmp.as<T>(), n,
                                                                         inverse);
  OBA_HIP(hipGetLastError());
  OBA_HIP(hipMemcpy2DAsync(b, sizeof(T) * ldb, tmp.p, sizeof(T) * n, sizeof(T) * n, nrhs, hipMemcpyDeviceToDevice, s));
}

template <typename T>
__global__ void perm_from_row_kernel(const T* row, long stride, long n, int* perm) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) perm[i] = (int)row[i * stride];
}

template <typename T>
void perm_from_tracked(const T* At, long ldt, long m, long n, int* perm_dev, hipStream_t s) {
  perm_from_row_kernel<T><<<(unsigned)((m + 255) / 256), 256, 0, s>>>(At + n, ldt, m, perm_dev);
  OBA_HIP(hipGetLastError());
}

// LAPACK ipiv (1-based, applied in order k1..k2) -> composed row permutation:
// applying the interchanges to x gives x'(i) = x(perm[i]).
void ipiv_to_perm(long n, const int* ipiv_host, std::vector<int>& perm) {
  perm.resize(n);
  for (long i = 0; i < n; i++) perm[i] = (int)i;
  for (long i = 0; i < n; i++) {
    long p = ipiv_host[i] - 1;
    if (p != i && p >= 0 && p < n) std::swap(perm[i], perm[p]);
  }
}

// Solve op(A) X = B with the column-major LU of A and its row permutation.
template <typename T>
void getrs_perm(bool trans, long n, long nrhs, const T* a, long lda, const int* perm_dev, T* b, long ldb,
                hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return;
  if (!trans) {
    permute_rows<T>(false, n, nrhs, perm_dev, b, ldb, s);
    trsm<T>(false, true, false, true, n, nrhs, T(1), a, lda, b, ldb, s);
    trsm<T>(false, false, false, false, n, nrhs, T(1), a, lda, b, ldb, s);
  } else {
    trsm<T>(false, false, true, false, n, nrhs, T(1), a, lda, b, ldb, s);
    trsm<T>(false, true, true, true, n, nrhs, T(1), a, lda, b, ldb, s);
    permute_rows<T>(true, n, nrhs, perm_dev, b, ldb, s);
  }
}

// Same with the factors stored transposed (At = LU^T as left by getrf_t).
template <typename T>
void getrs_t(long n, long nrhs, const T* At, long ldt, const int* perm_dev, T* b, long ldb, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return;
  permute_rows<T>(false, n, nrhs, perm_dev, b, ldb, s);
  // L y = b: L = (unit upper part of At)^T ; U x = y: U = (lower part of At)^T
  trsm<T>(false, false, true, true, n, nrhs, T(1), At, ldt, b, ldb, s);
  trsm<T>(false, true, true, false, n, nrhs, T(1), At, ldt, b, ldb, s);
}

// Row interchanges on a column-major matrix, one thread per column
// (lapack/laswp/generic/laswp_k_1.c semantics, incx = +1 or -1).
template <typename T>
__global__ void laswp_kernel(long n, T* a, long lda, long k1, long k2, const int* ipiv, int incx) {
  const long col = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (col >= n) return;
  T* c = a + col * lda;
  if (incx > 0) {
    for (long i = k1; i <= k2; i++) {
      long p = ipiv[(k1 - 1) + (i - k1) * incx] - 1;  // IPIV(K1+(I-K1)*|INCX|)
      if (p != i - 1) { T t = c[i - 1]; c[i - 1] = c[p]; c[p] = t; }
    }
  } else {
    for (long i = k2; i >= k1; i--) {
      long p = ipiv[(k1 - 1) + (i - k1) * (-incx)] - 1;
      if (p != i - 1) { T t = c[i - 1]; c[i - 1] = c[p]; c[p] = t; }
    }
  }
}

template <typename T>
void laswp(long n, T* a, long lda, long k1, long k2, const int* ipiv_dev, int incx, hipStream_t s) {
  if (n <= 0 || k2 < k1 || incx == 0) return;
  laswp_kernel<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, a, lda, k1, k2, ipiv_dev, incx);
  OBA_HIP(hipGetLastError());
}

template void transpose<double, double>(long, long, const double*, long, double*, long, int*, hipStream_t);
template void transpose<float, float>(long, long, const float*, long, float*, long, int*, hipStream_t);
template void transpose<double, float>(long, long, const double*, long, float*, long, int*, hipStream_t);
template void transpose<float, double>(long, long, const float*, long, double*, long, int*, hipStream_t);
template int getrf_t<double>(long, long, double*, long, int*, bool, hipStream_t);
template int getrf_t<float>(long, long, float*, long, int*, bool, hipStream_t);
template int getrf<double>(long, long, double*, long, int*, hipStream_t);
template int getrf_perm<double>(long, long, double*, long, int*, int*, hipStream_t);
template int getrf_perm<float>(long, long, float*, long, int*, int*, hipStream_t);
template int getrf<float>(long, long, float*, long, int*, hipStream_t);
template void permute_rows<double>(bool, long, long, const int*, double*, long, hipStream_t);
template void permute_rows<float>(bool, long, long, const int*, float*, long, hipStream_t);
template void perm_from_tracked<double>(const double*, long, long, long, int*, hipStream_t);
template void perm_from_tracked<float>(const float*, long, long, long, int*, hipStream_t);
template void getrs_perm<double>(bool, long, long, const double*, long, const int*, double*, long, hipStream_t);
template void getrs_perm<float>(bool, long, long, const float*, long, const int*, float*, long, hipStream_t);
template void getrs_t<double>(long, long, const double*, long, const int*, double*, long, hipStream_t);
template void getrs_t<float>(long, long, const float*, long, const int*, float*, long, hipStream_t);
template void laswp<double>(long, double*, long, long, long, const int*, int, hipStream_t);
template void laswp<float>(long, float*, long, long, long, const int*, int, hipStream_t);

}  // namespace oba
