// Batched small problems (GEMM, LU, LU solve, Cholesky) in one persistent
// launch per kind.  The reference has a batched GEMM only
// (interface/gemm_batch.c:117, one problem per thread in
// driver/level3/gemm_batch_thread.c:79); batched getrf/getrs/potrf have no
// reference entry point and follow the single-problem semantics of
// lapack/getf2/getf2_k.c, lapack/getrs/getrs_single.c and
// lapack/potf2/potf2_L.c problem by problem.
//
// MI355X design:
//  * a problem table (BDesc) sorted largest-first lives in HBM; workgroups
//    (GEMM tiles, LU, Cholesky) or wavefronts (solves) pull the next item
//    from an atomic queue, so 10^4 uneven problems keep all 256 CUs busy
//    without one launch per problem;
//  * GEMM: BT x BT output tiles (BT = 32 or 64) on the FP64/FP32 16x16x4
//    MFMA, k-slabs staged through LDS;
//  * LU (m <= 512): one workgroup per problem; each thread owns one or two
//    rows of the current W-column panel in registers, the pivot search is
//    a wavefront max-reduction plus one barrier per column, interchanges
//    are deferred (a position -> row map in LDS; rows are permuted once at
//    the end), the U12 block is a column-per-thread substitution and the
//    trailing update a rank-W FMA sweep with U12 broadcast from LDS;
//  * Cholesky (n <= 512): same row-owner panel, one barrier per column,
//    SYRK-shaped trailing update of the lower triangle;
//  * LU solve (n <= 512): one wavefront per problem, the right-hand side in
//    registers (lane l owns rows l, l+64, ...), row interchanges by lane
//    reads/writes, triangular sweeps with 8 columns of the factor loaded
//    per step;
//  * the launch sequence (queue reset by a fill kernel, the kernel) is
//    captured once in a hipGraph per problem table and replayed while the
//    table is unchanged; tile width, panel width and persistent grid size
//    are chosen per batch shape by timing the candidates on scratch copies.
#include "batched.h"
#include "kernels.h"
#include "lu_device.h"
#include "lu_internal.h"
#include <algorithm>
#include <climits>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

namespace oba {

// The problems' pointers come from the descriptor table, so the compiler
// sees generic pointers and emits flat loads and stores, which also count
// against the LDS wait counter: every LDS wait then waits for outstanding
// memory operations too.  They point to device memory: use them as global.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

// ---- small device helpers ----------------------------------------------------
__device__ inline double bcast_lane(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
__device__ inline float bcast_lane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline int bcast_lane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__global__ void bzero_kernel(unsigned* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// copies operand `which` (0: a, 1: b, 2: c) of every problem from src to dst
// descriptors (rows x cols, each with its own leading dimension); used to
// give the tuner scratch copies of in/out operands
template <typename T>
__global__ void bcopy_kernel(const BDesc* __restrict__ src, const BDesc* __restrict__ dst, int which, int rows_field,
                             int cols_field) {
  const BDesc& s = src[blockIdx.y];
  const BDesc& d = dst[blockIdx.y];
  const int* fs = &s.m;
  const long rows = fs[rows_field], cols = fs[cols_field];
  const long lds = which == 2 ? s.ldc : which ? s.ldb : s.lda, ldd = which == 2 ? d.ldc : which ? d.ldb : d.lda;
  gptr<const T> a = (gptr<const T>)(which == 2 ? s.c : which ? s.b : s.a);
  gptr<T> b = (gptr<T>)(which == 2 ? d.c : which ? d.b : d.a);
  const long total = rows * cols;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long i = e % rows, j = e / rows;
    b[i + j * ldd] = a[i + j * lds];
  }
}

// ---- batched GEMM --------------------------------------------------------------
template <typename T> struct BMF;
template <> struct BMF<double> {
  typedef double acc __attribute__((ext_vector_type(4)));
  __device__ static inline acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int orow(int l, int r) { return 4 * r + (l >> 4); }
};
template <> struct BMF<float> {
  typedef float acc __attribute__((ext_vector_type(4)));
  __device__ static inline acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int orow(int l, int r) { return 4 * (l >> 4) + r; }
};

// work[i] = (problem, 64x64 tile of C) in launch order; per queue item one
// workgroup computes that tile of C = alpha*op(A)*op(B) + beta*C, as one
// 64-wide tile (BT = 64) or as its up to four 32-wide quarters that
// intersect C (BT = 32, less padding for small problems)
template <typename T, int BT, bool TA, bool TB>
__global__ __launch_bounds__(256) void bgemm_kernel(const BDesc* __restrict__ desc, const int2* __restrict__ work,
                                                    int nwork, unsigned* __restrict__ counter) {
  constexpr int BK = 16, PAD = 16, LDS_ = BT + PAD;
  constexpr int WT = BT / 2, F = WT / 16, NE = BT * BK / 256;
  typedef BMF<T> mf;
  typedef typename mf::acc acc_t;
  __shared__ T sa[BK * LDS_];
  __shared__ T sb[BK * LDS_];
  __shared__ int s_item[2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave & 1) * WT, wn0 = (wave >> 1) * WT;

  if (tid == 0) s_item[0] = (int)atomicAdd(counter, 1u);
  __syncthreads();
  int cur = 0;
  int it = s_item[0];
  while (it < nwork) {
    if (tid == 0) s_item[cur ^ 1] = (int)atomicAdd(counter, 1u);   // next item, read after the last barrier
    const int2 wk = work[it];
    const BDesc& d = desc[wk.x];
    const long M = d.m, N = d.n, K = d.k, lda = d.lda, ldb = d.ldb, ldc = d.ldc;
    gptr<const T> A = (gptr<const T>)d.a;
    gptr<const T> B = (gptr<const T>)d.b;
    gptr<T> C = (gptr<T>)d.c;
    const T alpha = (T)d.alpha, beta = (T)d.beta;
    const int st_m = (int)((M + 63) / 64);
    const long sm0 = (long)(wk.y % st_m) * 64, sn0 = (long)(wk.y / st_m) * 64;
    for (int sub = 0; sub < (BT == 64 ? 1 : 4); sub++) {
      const long m0 = sm0 + (BT == 64 ? 0 : (sub & 1) * 32), n0 = sn0 + (BT == 64 ? 0 : (sub >> 1) * 32);
      if (m0 >= M || n0 >= N) continue;   // uniform

      acc_t acc[F][F];
#pragma unroll
      for (int i = 0; i < F; i++)
#pragma unroll
        for (int j = 0; j < F; j++) acc[i][j] = acc_t{};

      if (alpha != T(0)) {
        T ra[NE], rb[NE];
        auto load = [&](long k0) __attribute__((always_inline)) {
#pragma unroll
          for (int e = 0; e < NE; e++) {
            const int v = tid + e * 256;
            int i, kk;
            if (!TA) { i = v % BT; kk = v / BT; } else { kk = v % BK; i = v / BK; }
            const long gi = m0 + i, gk = k0 + kk;
            ra[e] = (gi < M && gk < K) ? (TA ? A[gk + gi * lda] : A[gi + gk * lda]) : T(0);
            int j, kb;
            if (!TB) { kb = v % BK; j = v / BK; } else { j = v % BT; kb = v / BT; }
            const long gj = n0 + j, gkb = k0 + kb;
            rb[e] = (gj < N && gkb < K) ? (TB ? B[gj + gkb * ldb] : B[gkb + gj * ldb]) : T(0);
          }
        };
        auto store = [&]() __attribute__((always_inline)) {
#pragma unroll
          for (int e = 0; e < NE; e++) {
            const int v = tid + e * 256;
            int i, kk;
            if (!TA) { i = v % BT; kk = v / BT; } else { kk = v % BK; i = v / BK; }
            sa[kk * LDS_ + i] = ra[e];
            int j, kb;
            if (!TB) { kb = v % BK; j = v / BK; } else { j = v % BT; kb = v / BT; }
            sb[kb * LDS_ + j] = rb[e];
          }
        };
        load(0);
        for (long k0 = 0; k0 < K; k0 += BK) {
          store();
          __syncthreads();
          if (k0 + BK < K) load(k0 + BK);   // next slab in flight during the MFMAs
#pragma unroll
          for (int kk = 0; kk < BK / 4; kk++) {
            const int kr = kk * 4 + (lane >> 4);
            T fa[F], fb[F];
#pragma unroll
            for (int i = 0; i < F; i++) fa[i] = sa[kr * LDS_ + wm0 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < F; j++) fb[j] = sb[kr * LDS_ + wn0 + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < F; i++)
#pragma unroll
              for (int j = 0; j < F; j++) acc[i][j] = mf::mma(fb[j], fa[i], acc[i][j]);
          }
          __syncthreads();
        }
      }
      const bool bz = beta == T(0);
#pragma unroll
      for (int i = 0; i < F; i++) {
        const long m = m0 + wm0 + i * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < F; j++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const long n = n0 + wn0 + j * 16 + mf::orow(lane, r);
            if (m < M && n < N) {
              auto p = C + m + n * ldc;
              T v = alpha * acc[i][j][r];
              if (!bz) v += beta * *p;
              *p = v;
            }
          }
      }
    }
    __syncthreads();
    cur ^= 1;
    it = s_item[cur];
  }
}

// ---- batched LU (m <= 256*RPT) -------------------------------------------------
template <typename T, int RPT, int W>
__global__ __launch_bounds__(256) void bgetrf_kernel(const BDesc* __restrict__ desc, int count,
                                                     unsigned* __restrict__ counter, int* __restrict__ info_out) {
  constexpr int NT = 256, MAXM = NT * RPT, UB = NT, LDU = W + 1, CH = 8, BU = RPT == 1 ? 32 : 16;
  typedef typename RowVec<T, W>::type RV;
  __shared__ int s_perm[MAXM];            // position -> physical row
  __shared__ T s_row[2][4][W];            // per-wave pivot candidates (by column parity)
  __shared__ double s_key[2][4];
  __shared__ int s_pos[2][4];
  __shared__ T s_l11[W * W];              // strictly lower part of the diagonal block
  __shared__ T s_u[UB * LDU];             // U12 chunk, one column per thread
  __shared__ T s_l21[RPT == 1 ? MAXM * LDU : 1];   // multipliers of the trailing rows (MFMA update)
  __shared__ int s_prow[RPT == 1 ? MAXM : 1];      // their physical rows
  __shared__ int s_item[2];
  __shared__ int s_moved;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;

  if (t == 0) s_item[0] = (int)atomicAdd(counter, 1u);
  __syncthreads();
  int cur = 0;
  int it = s_item[0];
  while (it < count) {
    if (t == 0) { s_item[cur ^ 1] = (int)atomicAdd(counter, 1u); s_moved = 0; }
    const BDesc& d = desc[it];
    const int m = d.m, n = d.n, kmax = min(d.m, d.n);
    const long lda = d.lda;
    gptr<T> A = (gptr<T>)d.a;
    gptr<int> ipiv = (gptr<int>)d.ipiv;
    int info = 0;
    for (int i = t; i < m; i += NT) s_perm[i] = i;
    __syncthreads();

    #pragma unroll 1
    for (int j0 = 0; j0 < kmax; j0 += W) {
      const int w = min(W, kmax - j0);
      RV x[RPT];
      int pos[RPT], phys[RPT];
      bool live[RPT];
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        pos[q] = j0 + t + NT * q;
        live[q] = pos[q] < m;
        phys[q] = live[q] ? s_perm[pos[q]] : 0;
#pragma unroll
        for (int c = 0; c < W; c++) x[q][c] = (live[q] && c < w) ? A[phys[q] + (j0 + c) * lda] : T(0);
      }
      #pragma unroll 1
      for (int k = 0; k < w; k++) {
        const int j = j0 + k, buf = k & 1;
        double bv = -2.0;
        int bp = INT_MAX, bq = -1;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
          if (live[q] && pos[q] >= j) {
            const double v = pivot_key((T)pick<RV, W>(x[q], k), pos[q], j);
            if (better(v, pos[q], bv, bp)) { bv = v; bp = pos[q]; bq = q; }
          }
        }
        double wv = bv;
        int wp = bp;
        lu_wave_best(wv, wp);   // DPP row moves + readlanes (no LDS shuffles)
        if (lane == 0) { s_key[buf][wave] = wv; s_pos[buf][wave] = wp; }
        if (bq >= 0 && bp == wp) {
#pragma unroll
          for (int q = 0; q < RPT; q++)
            if (q == bq) {
#pragma unroll
              for (int c = 0; c < W; c++) s_row[buf][wave][c] = x[q][c];
            }
        }
        __syncthreads();
        double gv = s_key[buf][0];
        int gp = s_pos[buf][0], gw = 0;
#pragma unroll
        for (int v = 1; v < 4; v++)
          if (better(s_key[buf][v], s_pos[buf][v], gv, gp)) { gv = s_key[buf][v]; gp = s_pos[buf][v]; gw = v; }
        const T* prow = s_row[buf][gw];
        const T piv = prow[k];
        const bool nz = piv != T(0);
        const bool doscale = nz && fabs(piv) >= MinNormal<T>::v;
        if (t == 0) {
          ipiv[j] = gp + 1;
          if (!nz && info == 0) info = j + 1;
          if (gp != j) { const int tmp = s_perm[j]; s_perm[j] = s_perm[gp]; s_perm[gp] = tmp; s_moved = 1; }
        }
        const T rinv = doscale ? T(1) / piv : T(1);
#pragma unroll
        for (int q = 0; q < RPT; q++) {
          if (!live[q]) continue;
          if (pos[q] == gp) pos[q] = j;
          else if (pos[q] == j) pos[q] = gp;
          if (pos[q] > j) {
            T l = pick<RV, W>(x[q], k);
            if (doscale) l *= rinv;
#pragma unroll
            for (int c = 0; c < W; c++) {
              if (c == k) x[q][c] = l;
              else if (c > k && c < w) x[q][c] -= l * prow[c];
            }
          }
        }
      }
      // panel back to its physical rows; L11 to LDS
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        if (!live[q]) continue;
#pragma unroll
        for (int c = 0; c < W; c++)
          if (c < w) A[phys[q] + (j0 + c) * lda] = x[q][c];
        const int rr = pos[q] - j0;
        if (rr < w) {
#pragma unroll
          for (int c = 0; c < W; c++)
            if (c < rr) s_l11[rr * W + c] = x[q][c];
        }
      }
      __syncthreads();
      #pragma unroll 1
      for (int c0 = j0 + w; c0 < n; c0 += UB) {
        const int cw = min(UB, n - c0);
        if (t < cw) {   // U12 column c0+t: forward substitution with unit L11
          auto colp = A + (long)(c0 + t) * lda;
          T u[W];
#pragma unroll
          for (int k = 0; k < W; k++) u[k] = (k < w) ? colp[s_perm[j0 + k]] : T(0);
#pragma unroll
          for (int kp = 0; kp < W; kp++) {
            // keep each step's L11 column reads in that step (bounded registers)
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = kp + 1; k < W; k++)
              if (k < w) u[k] -= s_l11[k * W + kp] * u[kp];
          }
#pragma unroll
          for (int k = 0; k < W; k++)
            if (k < w) { colp[s_perm[j0 + k]] = u[k]; s_u[t * LDU + k] = u[k]; }
        }
        __syncthreads();
        if constexpr (RPT == 1) {
          // A22 -= L21 U12 on MFMA 16x16 tiles (rows: positions >= j0+w at
          // their physical rows, staged with their multipliers in LDS;
          // 4 wavefronts, GT tiles per step with their C elements loaded
          // before the products)
          typedef BMF<T> mf;
          const int r0 = j0 + w, nr = m - r0;
          if (c0 == j0 + w && live[0] && pos[0] >= r0) {   // once per panel
            s_prow[pos[0] - r0] = phys[0];
#pragma unroll
            for (int k = 0; k < W; k++) s_l21[(pos[0] - r0) * LDU + k] = k < w ? x[0][k] : T(0);
          }
          __syncthreads();
          const int ntr = (nr + 15) / 16, ntc = (cw + 15) / 16, ntiles = ntr * ntc;
          constexpr int GT = 4;
          #pragma unroll 1
          for (int tb = wave * GT; tb < ntiles; tb += 4 * GT) {
            int ti[GT], tj[GT];
            T cv[GT][4];
            typename mf::acc acc[GT];
#pragma unroll
            for (int g = 0; g < GT; g++) {
              const int tile = min(tb + g, ntiles - 1);
              ti[g] = tile % ntr;
              tj[g] = tile / ntr;
              const int ri = ti[g] * 16 + (lane & 15);
              const int pr = ri < nr ? s_prow[ri] : 0;
#pragma unroll
              for (int r = 0; r < 4; r++) {
                const int cc = tj[g] * 16 + mf::orow(lane, r);
                cv[g][r] = (tb + g < ntiles && ri < nr && cc < cw) ? A[pr + (long)(c0 + cc) * lda] : T(0);
              }
            }
#pragma unroll
            for (int g = 0; g < GT; g++) {
              acc[g] = typename mf::acc{T(0), T(0), T(0), T(0)};
              const int ra = ti[g] * 16 + (lane & 15), cb = tj[g] * 16 + (lane & 15);
#pragma unroll
              for (int kk = 0; kk < W; kk += 4) {
                const int k = kk + (lane >> 4);
                const T fa = ra < nr ? s_l21[ra * LDU + k] : T(0);
                const T fb = (cb < cw && k < w) ? s_u[cb * LDU + k] : T(0);
                acc[g] = mf::mma(fb, fa, acc[g]);
              }
            }
#pragma unroll
            for (int g = 0; g < GT; g++) {
              const int ri = ti[g] * 16 + (lane & 15);
              const int pr = ri < nr ? s_prow[ri] : 0;
#pragma unroll
              for (int r = 0; r < 4; r++) {
                const int cc = tj[g] * 16 + mf::orow(lane, r);
                if (tb + g < ntiles && ri < nr && cc < cw) A[pr + (long)(c0 + cc) * lda] = cv[g][r] - acc[g][r];
              }
            }
          }
        } else {
#pragma unroll
        for (int q = 0; q < RPT; q++) {
          if (!live[q] || pos[q] < j0 + w) continue;
          auto rowp = A + phys[q];
          // BU columns per step, their loads issued together (one exposed
          // memory latency per step instead of per column)
          #pragma unroll 1
          for (int cb = 0; cb < cw; cb += BU) {
            T a[BU];
#pragma unroll
            for (int u = 0; u < BU; u++) a[u] = cb + u < cw ? rowp[(long)(c0 + cb + u) * lda] : T(0);
#pragma unroll
            for (int u = 0; u < BU; u++) {
              if (cb + u < cw) {
                const T* uc = s_u + (cb + u) * LDU;
                T v = a[u];
#pragma unroll
                for (int k = 0; k < W; k++)
                  if (k < w) v -= x[q][k] * uc[k];
                rowp[(long)(c0 + cb + u) * lda] = v;
              }
            }
          }
        }
        }
        __syncthreads();
      }
    }
    // rows to their final positions (position i <- physical row s_perm[i])
    if (s_moved) {
      #pragma unroll 1
      for (int c0 = 0; c0 < n; c0 += CH) {
        T v[RPT][CH];
#pragma unroll
        for (int q = 0; q < RPT; q++) {
          const int i = t + NT * q;
          const int src = i < m ? s_perm[i] : i;
          if (i < m && src != i) {
#pragma unroll
            for (int cc = 0; cc < CH; cc++)
              if (c0 + cc < n) v[q][cc] = A[src + (c0 + cc) * lda];
          }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RPT; q++) {
          const int i = t + NT * q;
          const int src = i < m ? s_perm[i] : i;
          if (i < m && src != i) {
#pragma unroll
            for (int cc = 0; cc < CH; cc++)
              if (c0 + cc < n) A[i + (c0 + cc) * lda] = v[q][cc];
          }
        }
        __syncthreads();
      }
    }
    if (t == 0) info_out[it] = info;
    __syncthreads();
    cur ^= 1;
    it = s_item[cur];
  }
}

// ---- batched Cholesky (n <= 256*RPT) -------------------------------------------
// LOWER: A = L*L^T with L(i,c) = a[i + c*lda]; upper: A = U^T*U, U = L^T,
// L(i,c) = a[c + i*lda].  Failure rule of lapack/potf2/potf2_L.c:71-75: an
// updated diagonal <= 0 is stored and info = j+1; NaN passes through.
template <typename T, int RPT, int W, bool LOWER>
__global__ __launch_bounds__(256) void bpotrf_kernel(const BDesc* __restrict__ desc, int count,
                                                     unsigned* __restrict__ counter, int* __restrict__ info_out) {
  constexpr int NT = 256, UB = NT, LDL = W + 1, BU = RPT == 1 ? 32 : 16;
  typedef typename RowVec<T, W>::type RV;
  __shared__ T s_col[2][W];
  __shared__ T s_l[UB * LDL];
  __shared__ int s_item[2];
  const int t = threadIdx.x;

  if (t == 0) s_item[0] = (int)atomicAdd(counter, 1u);
  __syncthreads();
  int cur = 0;
  int it = s_item[0];
  while (it < count) {
    if (t == 0) s_item[cur ^ 1] = (int)atomicAdd(counter, 1u);
    const BDesc& d = desc[it];
    const int n = d.n;
    const long lda = d.lda;
    gptr<T> A = (gptr<T>)d.a;
    auto at = [&](long i, long c) -> decltype(A[0])& { return LOWER ? A[i + c * lda] : A[c + i * lda]; };
    int info = 0;
    #pragma unroll 1
    for (int j0 = 0; j0 < n; j0 += W) {
      const int w = min(W, n - j0);
      RV x[RPT];
      int row[RPT];
      bool live[RPT];
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        row[q] = j0 + t + NT * q;
        live[q] = row[q] < n;
#pragma unroll
        for (int c = 0; c < W; c++) x[q][c] = (live[q] && c < w && j0 + c <= row[q]) ? at(row[q], j0 + c) : T(0);
      }
      int fail = 0;
      #pragma unroll 1
      for (int k = 0; k < w; k++) {
        const int j = j0 + k, buf = k & 1;
#pragma unroll
        for (int q = 0; q < RPT; q++)
          if (live[q] && row[q] >= j && row[q] < j0 + w) s_col[buf][row[q] - j0] = pick<RV, W>(x[q], k);
        __syncthreads();
        const T dj = s_col[buf][k];
        if (dj <= T(0)) {   // uniform: every thread read the same value
#pragma unroll
          for (int q = 0; q < RPT; q++)
            if (live[q] && row[q] == j) at(j, j) = dj;
          fail = j + 1;
          break;
        }
        const T sj = sqrt(dj);
        const T rinv = T(1) / sj;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
          if (!live[q] || row[q] < j) continue;
          if (row[q] == j) {
#pragma unroll
            for (int c = 0; c < W; c++)
              if (c == k) x[q][c] = sj;
          } else {
            const T l = pick<RV, W>(x[q], k) * rinv;
#pragma unroll
            for (int c = 0; c < W; c++) {
              if (c == k) x[q][c] = l;
              else if (c > k && c < w && j0 + c <= row[q]) x[q][c] -= l * (s_col[buf][c] * rinv);
            }
          }
        }
      }
      const int kdone = fail ? fail - 1 - j0 : w;
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        if (!live[q]) continue;
#pragma unroll
        for (int c = 0; c < W; c++)
          if (c < kdone && j0 + c <= row[q]) at(row[q], j0 + c) = x[q][c];
      }
      if (fail) { info = fail; break; }
      if constexpr (RPT == 1) {
        // trailing SYRK A22 -= L21 L21^T on MFMA 16x16 tiles of the lower
        // triangle (4 wavefronts, GT tiles each per step, their C elements
        // loaded before the products so one memory latency covers GT tiles)
        typedef BMF<T> mf;
        const int c0 = j0 + w, nr = n - c0;
        if (nr > 0) {
          if (live[0] && row[0] >= c0) {
#pragma unroll
            for (int c = 0; c < W; c++) s_l[(row[0] - c0) * LDL + c] = c < w ? x[0][c] : T(0);
          }
          __syncthreads();
          const int lane = t & 63, wave = t >> 6;
          const int nt = (nr + 15) / 16, ntiles = nt * (nt + 1) / 2;
          constexpr int GT = 4;
          #pragma unroll 1
          for (int tb = wave * GT; tb < ntiles; tb += 4 * GT) {
            int ti[GT], tj[GT];
            T cv[GT][4];
            typename mf::acc acc[GT];
#pragma unroll
            for (int g = 0; g < GT; g++) {
              const int tile = min(tb + g, ntiles - 1);
              int a0 = (int)((sqrtf(8.0f * (float)tile + 1.0f) - 1.0f) * 0.5f);
              while (a0 * (a0 + 1) / 2 > tile) a0--;
              while ((a0 + 1) * (a0 + 2) / 2 <= tile) a0++;
              ti[g] = a0;
              tj[g] = tile - a0 * (a0 + 1) / 2;
#pragma unroll
              for (int r = 0; r < 4; r++) {
                const int ri = c0 + ti[g] * 16 + (lane & 15), cj = c0 + tj[g] * 16 + mf::orow(lane, r);
                cv[g][r] = (tb + g < ntiles && ri < n && ri >= cj) ? at(ri, cj) : T(0);
              }
            }
#pragma unroll
            for (int g = 0; g < GT; g++) {
              acc[g] = typename mf::acc{T(0), T(0), T(0), T(0)};
              const int ra = ti[g] * 16 + (lane & 15), rb = tj[g] * 16 + (lane & 15);
#pragma unroll
              for (int kk = 0; kk < W; kk += 4) {
                const int k = kk + (lane >> 4);
                const T fa = ra < nr ? s_l[ra * LDL + k] : T(0);
                const T fb = rb < nr ? s_l[rb * LDL + k] : T(0);
                acc[g] = mf::mma(fb, fa, acc[g]);
              }
            }
#pragma unroll
            for (int g = 0; g < GT; g++)
#pragma unroll
              for (int r = 0; r < 4; r++) {
                const int ri = c0 + ti[g] * 16 + (lane & 15), cj = c0 + tj[g] * 16 + mf::orow(lane, r);
                if (tb + g < ntiles && ri < n && ri >= cj) at(ri, cj) = cv[g][r] - acc[g][r];
              }
          }
        }
      } else {
      #pragma unroll 1
      for (int c0 = j0 + w; c0 < n; c0 += UB) {
#pragma unroll
        for (int q = 0; q < RPT; q++)
          if (live[q] && row[q] >= c0 && row[q] < c0 + UB) {
#pragma unroll
            for (int c = 0; c < W; c++)
              if (c < w) s_l[(row[q] - c0) * LDL + c] = x[q][c];
          }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RPT; q++) {
          if (!live[q] || row[q] < c0) continue;
          const int cend = min(row[q], c0 + UB - 1);
          // BU columns per step, their loads issued together
          #pragma unroll 1
          for (int cb = c0; cb <= cend; cb += BU) {
            T a[BU];
#pragma unroll
            for (int u = 0; u < BU; u++) a[u] = cb + u <= cend ? at(row[q], cb + u) : T(0);
#pragma unroll
            for (int u = 0; u < BU; u++) {
              if (cb + u <= cend) {
                const T* lr = s_l + (cb + u - c0) * LDL;
                T v = a[u];
#pragma unroll
                for (int k = 0; k < W; k++)
                  if (k < w) v -= x[q][k] * lr[k];
                at(row[q], cb + u) = v;
              }
            }
          }
        }
        __syncthreads();
      }
      }
      __syncthreads();
    }
    if (t == 0) info_out[it] = info;
    __syncthreads();
    cur ^= 1;
    it = s_item[cur];
  }
}

// ---- batched LU solve (n <= 64*S), one wavefront per problem -------------------
// x <- A^-1 b (TRANS: A^-T b) with the factors and pivots of getrf, as
// lapack/getrs/getrs_single.c: interchanges, unit-L sweep, U sweep (or the
// transposed sweeps and the interchanges in reverse).  Lane l owns rows
// l + 64s of the right-hand side; slot indices s are compile-time constants
// (static_for) so x stays in registers.
template <int I, int N, typename F>
__device__ __attribute__((always_inline)) inline void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

template <typename T, int S>
__device__ __attribute__((always_inline)) inline void lane_swap(T (&x)[S], int s, int l, int p) {
  const int ps = p >> 6, pl = p & 63, lane = threadIdx.x & 63;
  T a = x[0], src = x[0];
#pragma unroll
  for (int s2 = 1; s2 < S; s2++) {
    a = (s == s2) ? x[s2] : a;
    src = (ps == s2) ? x[s2] : src;
  }
  const T va = bcast_lane(a, l);
  const T vb = bcast_lane(src, pl);
#pragma unroll
  for (int s2 = 0; s2 < S; s2++) {
    if (s2 == s) x[s2] = (lane == l) ? vb : x[s2];
    if (s2 == ps) x[s2] = (lane == pl) ? va : x[s2];
  }
}

template <typename T, int S, bool TRANS>
__global__ __launch_bounds__(256) void bgetrs_kernel(const BDesc* __restrict__ desc, int count,
                                                     unsigned* __restrict__ counter) {
  constexpr int U = 8;   // factor columns loaded per step
  const int lane = threadIdx.x & 63;
  int it = 0;
  if (lane == 0) it = (int)atomicAdd(counter, 1u);
  it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
  while (it < count) {
    int nx = 0;
    if (lane == 0) nx = (int)atomicAdd(counter, 1u);
    const BDesc& d = desc[it];
    const int n = d.n, nrhs = d.k;
    const long lda = d.lda, ldb = d.ldb;
    gptr<const T> A = (gptr<const T>)d.a;
    gptr<T> B = (gptr<T>)d.b;
    int ip[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
      const int i = 64 * s + lane;
      ip[s] = i < n ? ((gptr<const int>)d.ipiv)[i] - 1 : i;
    }
#pragma unroll 1
    for (int r = 0; r < nrhs; r++) {
      auto b = B + r * ldb;
      T x[S];
#pragma unroll
      for (int s = 0; s < S; s++) {
        const int i = 64 * s + lane;
        x[s] = i < n ? b[i] : T(0);
      }
      // interchanges k = 0 .. n-1 (forward) or n-1 .. 0 (TRANS, after the sweeps)
      auto swaps = [&](bool forward) __attribute__((always_inline)) {
#pragma unroll 1
        for (int kk = 0; kk < n; kk++) {
          const int k = forward ? kk : n - 1 - kk;
          const int s = k >> 6, l = k & 63;
          int pv = ip[0];
#pragma unroll
          for (int s2 = 1; s2 < S; s2++) pv = (s == s2) ? ip[s2] : pv;
          const int p = bcast_lane(pv, l);
          if (p != k) lane_swap<T, S>(x, s, l, p);
        }
      };
      if (!TRANS) {
        swaps(true);
        // unit lower sweep: x_i -= L(i,k) x_k for i > k
        static_for<0, S>([&](auto sc) __attribute__((always_inline)) {
          constexpr int s = decltype(sc)::value;
#pragma unroll 1
          for (int l0 = 0; l0 < 64 && 64 * s + l0 < n; l0 += U) {
            T cv[U][S];
#pragma unroll
            for (int e = 0; e < U; e++) {
              const int k = 64 * s + l0 + e;
#pragma unroll
              for (int s2 = s; s2 < S; s2++) {
                const int i = 64 * s2 + lane;
                cv[e][s2] = (k < n && i > k && i < n) ? A[i + k * lda] : T(0);
              }
            }
#pragma unroll
            for (int e = 0; e < U; e++) {
              const int k = 64 * s + l0 + e;
              if (k < n) {
                const T xk = bcast_lane(x[s], l0 + e);
#pragma unroll
                for (int s2 = s; s2 < S; s2++)
                  if (64 * s2 + lane > k) x[s2] -= cv[e][s2] * xk;
              }
            }
          }
        });
        // upper sweep, k = n-1 .. 0: x_k /= U(k,k); x_i -= U(i,k) x_k for i < k
        static_for<0, S>([&](auto sc) __attribute__((always_inline)) {
          constexpr int s = S - 1 - decltype(sc)::value;
#pragma unroll 1
          for (int l0 = 64 - U; l0 >= 0; l0 -= U) {
            if (64 * s + l0 >= n) continue;
            T cv[U][S];
#pragma unroll
            for (int e = 0; e < U; e++) {
              const int k = 64 * s + l0 + e;
#pragma unroll
              for (int s2 = 0; s2 <= s; s2++) {
                const int i = 64 * s2 + lane;
                cv[e][s2] = (k < n && i <= k) ? A[i + k * lda] : T(0);
              }
            }
#pragma unroll
            for (int e = U - 1; e >= 0; e--) {
              const int k = 64 * s + l0 + e;
              if (k < n) {
                x[s] = (lane == l0 + e) ? x[s] / cv[e][s] : x[s];
                const T xk = bcast_lane(x[s], l0 + e);
#pragma unroll
                for (int s2 = 0; s2 <= s; s2++)
                  if (64 * s2 + lane < k) x[s2] -= cv[e][s2] * xk;
              }
            }
          }
        });
      } else {
        // U^T sweep, i = 0 .. n-1: x_i = (x_i - sum_{k<i} U(k,i) x_k) / U(i,i)
        static_for<0, S>([&](auto sc) __attribute__((always_inline)) {
          constexpr int s = decltype(sc)::value;
#pragma unroll 1
          for (int l0 = 0; l0 < 64 && 64 * s + l0 < n; l0 += U) {
            T cv[U][S];
#pragma unroll
            for (int e = 0; e < U; e++) {
              const int i = 64 * s + l0 + e;
#pragma unroll
              for (int s2 = 0; s2 <= s; s2++) {
                const int k = 64 * s2 + lane;
                cv[e][s2] = (i < n && k <= i) ? A[k + i * lda] : T(0);
              }
            }
#pragma unroll
            for (int e = 0; e < U; e++) {
              const int i = 64 * s + l0 + e;
              if (i < n) {
                T part = T(0);
#pragma unroll
                for (int s2 = 0; s2 <= s; s2++)
                  if (64 * s2 + lane < i) part += cv[e][s2] * x[s2];
                const T sum = wave_sum(part);
                const T dia = bcast_lane(cv[e][s], l0 + e);
                x[s] = (lane == l0 + e) ? (x[s] - sum) / dia : x[s];
              }
            }
          }
        });
        // L^T sweep (unit), i = n-1 .. 0: x_i -= sum_{k>i} L(k,i) x_k
        static_for<0, S>([&](auto sc) __attribute__((always_inline)) {
          constexpr int s = S - 1 - decltype(sc)::value;
#pragma unroll 1
          for (int l0 = 64 - U; l0 >= 0; l0 -= U) {
            if (64 * s + l0 >= n) continue;
            T cv[U][S];
#pragma unroll
            for (int e = 0; e < U; e++) {
              const int i = 64 * s + l0 + e;
#pragma unroll
              for (int s2 = s; s2 < S; s2++) {
                const int k = 64 * s2 + lane;
                cv[e][s2] = (i < n && k > i && k < n) ? A[k + i * lda] : T(0);
              }
            }
#pragma unroll
            for (int e = U - 1; e >= 0; e--) {
              const int i = 64 * s + l0 + e;
              if (i < n) {
                T part = T(0);
#pragma unroll
                for (int s2 = s; s2 < S; s2++)
                  if (64 * s2 + lane > i && 64 * s2 + lane < n) part += cv[e][s2] * x[s2];
                const T sum = wave_sum(part);
                x[s] = (lane == l0 + e) ? x[s] - sum : x[s];
              }
            }
          }
        });
        swaps(false);
      }
#pragma unroll
      for (int s = 0; s < S; s++) {
        const int i = 64 * s + lane;
        if (i < n) b[i] = x[s];
      }
    }
    it = __builtin_amdgcn_readfirstlane(__shfl(nx, 0));
  }
}

// ---- host side: plans, graphs, tuning -------------------------------------------
namespace {

struct Cfg {
  int var = 0;    // kernel variant
  int grid = 2;   // persistent workgroups per CU
};

struct Plan {
  std::vector<char> key;
  DevBuf tab;       // BDesc table (+ GEMM work list)
  DevBuf ctr;       // queue counter
  DevBuf info;      // per-problem info
  hipGraphExec_t exec = nullptr;
};

std::map<int, Plan>& plans() {
  static std::map<int, Plan> p;
  return p;
}
std::map<std::string, Cfg>& tuned() {
  static std::map<std::string, Cfg> t;
  return t;
}

bool env_flag(const char* name, bool dflt) {
  const char* s = getenv(name);
  return s ? atoi(s) != 0 : dflt;
}

// a launcher enqueues the persistent kernel for config c on stream s; the
// kernels write the info of table entry p to info[p]
typedef void (*Launcher)(const Cfg& c, const void* ctx, const BDesc* tab, const void* extra, int nitems, int count,
                         unsigned* ctr, int* info, hipStream_t s);

void enqueue(Launcher L, const Cfg& c, const void* ctx, const BDesc* tab, const void* extra, int nitems, int count,
             unsigned* ctr, int* info, hipStream_t s) {
  bzero_kernel<<<1, 64, 0, s>>>(ctr, 1);
  L(c, ctx, tab, extra, nitems, count, ctr, info, s);
}

// Runs the batch described by host bytes `tab` (BDesc[count], then `extra`
// bytes) through the cached graph of plan slot `slot`, re-capturing only
// when the table or config changed.  Returns the device info array.
int* run_plan(int slot, Launcher L, const Cfg& c, const void* ctx, size_t ctx_bytes, const std::vector<BDesc>& probs,
              const std::vector<char>& extra, int nitems) {
  Runtime& r = rt();
  Plan& P = plans()[slot];
  const int count = (int)probs.size();
  const size_t tb = sizeof(BDesc) * probs.size();
  std::vector<char> key(sizeof(Cfg) + ctx_bytes + tb + extra.size());
  memcpy(key.data(), &c, sizeof(Cfg));
  if (ctx_bytes) memcpy(key.data() + sizeof(Cfg), ctx, ctx_bytes);
  memcpy(key.data() + sizeof(Cfg) + ctx_bytes, probs.data(), tb);
  if (!extra.empty()) memcpy(key.data() + sizeof(Cfg) + ctx_bytes + tb, extra.data(), extra.size());
  const bool use_graph = env_flag("OPENBLAS_AMD_BATCH_GRAPH", true);
  if (P.exec && use_graph && key == P.key) {
    OBA_HIP(hipGraphLaunch(P.exec, r.stream));
    return P.info.as<int>();
  }
  // the previous graph may still run from the last call: let it finish
  // before its table goes back to the pool
  OBA_HIP(hipStreamSynchronize(r.stream));
  if (P.exec) { OBA_HIP(hipGraphExecDestroy(P.exec)); P.exec = nullptr; }
  P.key.clear();
  P.tab = DevBuf(tb + extra.size() + 16);
  P.ctr = DevBuf(256);
  P.info = DevBuf(sizeof(int) * (size_t)std::max(count, 1));
  OBA_HIP(hipMemcpyAsync(P.tab.p, probs.data(), tb, hipMemcpyHostToDevice, r.stream));
  if (!extra.empty())
    OBA_HIP(hipMemcpyAsync(P.tab.as<char>() + tb, extra.data(), extra.size(), hipMemcpyHostToDevice, r.stream));
  const BDesc* dt = P.tab.as<BDesc>();
  const void* dx = extra.empty() ? nullptr : (const void*)(P.tab.as<char>() + tb);
  if (!use_graph) {
    enqueue(L, c, ctx, dt, dx, nitems, count, P.ctr.as<unsigned>(), P.info.as<int>(), r.stream);
    return P.info.as<int>();
  }
  hipGraph_t g = nullptr;
  OBA_HIP(hipStreamBeginCapture(r.stream, hipStreamCaptureModeThreadLocal));
  enqueue(L, c, ctx, dt, dx, nitems, count, P.ctr.as<unsigned>(), P.info.as<int>(), r.stream);
  OBA_HIP(hipStreamEndCapture(r.stream, &g));
  OBA_HIP(hipGraphInstantiate(&P.exec, g, nullptr, nullptr, 0));
  OBA_HIP(hipGraphDestroy(g));
  P.key.swap(key);
  OBA_HIP(hipGraphLaunch(P.exec, r.stream));
  return P.info.as<int>();
}

// Times every candidate config on scratch copies (operand `which` of each
// problem, rows x cols given by BDesc field indices) and keeps the fastest
// for this batch signature.
Cfg tune(const std::string& sig, const std::vector<Cfg>& cands, Launcher L, const void* ctx,
         const std::vector<BDesc>& probs, const std::vector<char>& extra, int nitems, int which, int rows_field,
         int cols_field, size_t elem) {
  auto& tt = tuned();
  auto f = tt.find(sig);
  if (f != tt.end()) return f->second;
  if (cands.size() == 1 || !env_flag("OPENBLAS_AMD_BATCH_TUNE", true)) { tt[sig] = cands[0]; return cands[0]; }
  Runtime& r = rt();
  const int count = (int)probs.size();
  // scratch copies of the in/out operand, packed
  std::vector<BDesc> scr = probs;
  size_t total = 0;
  for (auto& d : scr) {
    const int* fs = &d.m;
    total += elem * (size_t)std::max(fs[rows_field], 1) * (size_t)std::max(fs[cols_field], 1);
  }
  DevBuf sbuf(total + 256), tsrc(sizeof(BDesc) * count), tdst(sizeof(BDesc) * count + extra.size() + 16),
      ctr(256), info(sizeof(int) * (size_t)std::max(count, 1));
  size_t off = 0;
  for (auto& d : scr) {
    const int* fs = &d.m;
    const int rows = std::max(fs[rows_field], 1);
    void* p = sbuf.as<char>() + off;
    off += elem * (size_t)rows * (size_t)std::max(fs[cols_field], 1);
    if (which == 0) { d.a = p; d.lda = rows; }
    else if (which == 1) { d.b = p; d.ldb = rows; }
    else { d.c = p; d.ldc = rows; }
  }
  OBA_HIP(hipMemcpyAsync(tsrc.p, probs.data(), sizeof(BDesc) * count, hipMemcpyHostToDevice, r.stream));
  OBA_HIP(hipMemcpyAsync(tdst.p, scr.data(), sizeof(BDesc) * count, hipMemcpyHostToDevice, r.stream));
  if (!extra.empty())
    OBA_HIP(hipMemcpyAsync(tdst.as<char>() + sizeof(BDesc) * count, extra.data(), extra.size(), hipMemcpyHostToDevice,
                           r.stream));
  const void* dx = extra.empty() ? nullptr : (const void*)(tdst.as<char>() + sizeof(BDesc) * count);
  auto refresh = [&]() {
    if (count == 0) return;
    dim3 g(64, (unsigned)count);
    if (elem == 8) bcopy_kernel<double><<<g, 256, 0, r.stream>>>(tsrc.as<BDesc>(), tdst.as<BDesc>(), which, rows_field, cols_field);
    else bcopy_kernel<float><<<g, 256, 0, r.stream>>>(tsrc.as<BDesc>(), tdst.as<BDesc>(), which, rows_field, cols_field);
  };
  hipEvent_t e0, e1;
  OBA_HIP(hipEventCreate(&e0));
  OBA_HIP(hipEventCreate(&e1));
  refresh();
  enqueue(L, cands[0], ctx, tdst.as<BDesc>(), dx, nitems, count, ctr.as<unsigned>(), info.as<int>(), r.stream);
  float best = 1e30f;
  Cfg bc = cands[0];
  for (const Cfg& c : cands) {
    refresh();
    OBA_HIP(hipEventRecord(e0, r.stream));
    enqueue(L, c, ctx, tdst.as<BDesc>(), dx, nitems, count, ctr.as<unsigned>(), info.as<int>(), r.stream);
    OBA_HIP(hipEventRecord(e1, r.stream));
    OBA_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    OBA_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) { best = ms; bc = c; }
  }
  OBA_HIP(hipEventDestroy(e0));
  OBA_HIP(hipEventDestroy(e1));
  tt[sig] = bc;
  return bc;
}

std::string signature(int kind, size_t elem, int flags, const std::vector<BDesc>& probs) {
  long mx = 0, cnt = (long)probs.size();
  double work = 0;
  for (auto& d : probs) {
    mx = std::max<long>(mx, std::max(d.m, std::max(d.n, d.k)));
    work += (double)std::max(d.m, 1) * std::max(d.n, 1) * std::max(d.k, std::max(std::min(d.m, d.n), 1));
  }
  int lc = 0, lw = 0;
  while ((1L << lc) < cnt) lc++;
  while (work > 1.0 && lw < 80) { work /= 2; lw++; }
  char b[128];
  snprintf(b, sizeof b, "%d/%zu/%d/%d/%ld/%d", kind, elem, flags, lc, (mx + 31) / 32, lw);
  return b;
}

int grid_of(const Cfg& c) { return std::max(1, rt().num_cu * c.grid); }

// ---- launchers ---------------------------------------------------------------
template <typename T>
void launch_gemm(const Cfg& c, const void* ctx, const BDesc* tab, const void* extra, int nitems, int, unsigned* ctr,
                 int*, hipStream_t s) {
  const int ta = ((const int*)ctx)[0], tb = ((const int*)ctx)[1];
  const int2* work = (const int2*)extra;
  const int g = grid_of(c);
#define OBA_BG(BT, TA, TB) bgemm_kernel<T, BT, TA, TB><<<g, 256, 0, s>>>(tab, work, nitems, ctr)
  if (c.var == 0) {
    if (!ta && !tb) OBA_BG(32, false, false);
    else if (!ta && tb) OBA_BG(32, false, true);
    else if (ta && !tb) OBA_BG(32, true, false);
    else OBA_BG(32, true, true);
  } else {
    if (!ta && !tb) OBA_BG(64, false, false);
    else if (!ta && tb) OBA_BG(64, false, true);
    else if (ta && !tb) OBA_BG(64, true, false);
    else OBA_BG(64, true, true);
  }
#undef OBA_BG
  OBA_HIP(hipGetLastError());
}

template <typename T>
void launch_getrf(const Cfg& c, const void* ctx, const BDesc* tab, const void*, int, int count, unsigned* ctr,
                  int* info, hipStream_t s) {
  const int rpt = ((const int*)ctx)[0];
  const int g = grid_of(c);
  // panel width: 8 or 16 columns for FP64 (wider FP64 panels exceed the
  // register file), 16 or 32 for FP32
  constexpr int W0 = sizeof(T) == 8 ? 8 : 16;
  if (rpt == 1) {
    if (c.var == 0) bgetrf_kernel<T, 1, W0><<<g, 256, 0, s>>>(tab, count, ctr, info);
    else bgetrf_kernel<T, 1, 2 * W0><<<g, 256, 0, s>>>(tab, count, ctr, info);
  } else {
    if (c.var == 0) bgetrf_kernel<T, 2, W0><<<g, 256, 0, s>>>(tab, count, ctr, info);
    else bgetrf_kernel<T, 2, 2 * W0><<<g, 256, 0, s>>>(tab, count, ctr, info);
  }
  OBA_HIP(hipGetLastError());
}

template <typename T>
void launch_potrf(const Cfg& c, const void* ctx, const BDesc* tab, const void*, int, int count, unsigned* ctr,
                  int* info, hipStream_t s) {
  const int rpt = ((const int*)ctx)[0], lower = ((const int*)ctx)[1];
  const int g = grid_of(c);
#define OBA_BP(R, W, L) bpotrf_kernel<T, R, W, L><<<g, 256, 0, s>>>(tab, count, ctr, info)
  if (lower) {
    if (rpt == 1) { if (c.var == 0) OBA_BP(1, 16, true); else OBA_BP(1, 32, true); }
    else { if (c.var == 0) OBA_BP(2, 16, true); else OBA_BP(2, 32, true); }
  } else {
    if (rpt == 1) { if (c.var == 0) OBA_BP(1, 16, false); else OBA_BP(1, 32, false); }
    else { if (c.var == 0) OBA_BP(2, 16, false); else OBA_BP(2, 32, false); }
  }
#undef OBA_BP
  OBA_HIP(hipGetLastError());
}

template <typename T>
void launch_getrs(const Cfg& c, const void* ctx, const BDesc* tab, const void*, int, int count, unsigned* ctr, int*,
                  hipStream_t s) {
  const int S = ((const int*)ctx)[0], trans = ((const int*)ctx)[1];
  const int g = grid_of(c);
#define OBA_BS(SS)                                                                        \
  if (trans) bgetrs_kernel<T, SS, true><<<g, 256, 0, s>>>(tab, count, ctr);             \
  else bgetrs_kernel<T, SS, false><<<g, 256, 0, s>>>(tab, count, ctr);
  if (S == 1) { OBA_BS(1) } else if (S == 2) { OBA_BS(2) } else if (S == 4) { OBA_BS(4) } else { OBA_BS(8) }
#undef OBA_BS
  OBA_HIP(hipGetLastError());
}

template <typename T> constexpr int tslot(int kind) { return kind * 2 + (sizeof(T) == 8 ? 0 : 1); }

// dinfo[p] is the info of table entry p; the caller's slot is probs[p].idx
void collect_info(const int* dinfo, const std::vector<BDesc>& probs, int* info_host) {
  if (probs.empty()) return;
  Runtime& r = rt();
  std::vector<int> h(probs.size(), 0);
  OBA_HIP(hipMemcpyAsync(h.data(), dinfo, sizeof(int) * h.size(), hipMemcpyDeviceToHost, r.stream));
  OBA_HIP(hipStreamSynchronize(r.stream));
  for (size_t p = 0; p < probs.size(); p++) info_host[probs[p].idx] = h[p];
}

}  // namespace

// ---- drivers -------------------------------------------------------------------
template <typename T>
void gemm_batched(int ta, int tb, std::vector<BDesc>& probs) {
  if (probs.empty()) return;
  // work list: every 64x64 tile of every C, problems in the given order
  std::vector<int> v;
  for (size_t p = 0; p < probs.size(); p++) {
    const long tm = (probs[p].m + 63) / 64, tn = (probs[p].n + 63) / 64;
    for (long t = 0; t < tm * tn; t++) { v.push_back((int)p); v.push_back((int)t); }
  }
  std::vector<char> wl(v.size() * sizeof(int));
  memcpy(wl.data(), v.data(), wl.size());
  const int nw = (int)(v.size() / 2);
  const int ctx[2] = {ta, tb};
  std::vector<Cfg> cands;
  for (int var = 0; var < 2; var++)
    for (int g : {2, 4, 8}) cands.push_back(Cfg{var, g});
  Cfg c = tune(signature(BK_GEMM, sizeof(T), ta | (tb << 1), probs), cands, launch_gemm<T>, ctx, probs, wl, nw, 2, 0,
               1, sizeof(T));
  if (const char* e = getenv("OPENBLAS_AMD_BATCH_GEMM_TILE")) c.var = atoi(e) == 64 ? 1 : 0;
  double fl = 0, by = 0;
  for (auto& d : probs) {
    fl += 2.0 * d.m * d.n * d.k;
    by += sizeof(T) * ((double)d.m * d.k + (double)d.k * d.n + (d.beta != 0.0 ? 2.0 : 1.0) * d.m * d.n);
  }
  KTimer kt("batched_gemm", rt().stream, fl, by);
  run_plan(tslot<T>(BK_GEMM), launch_gemm<T>, c, ctx, sizeof ctx, probs, wl, nw);
}

template <typename T>
void getrf_batched(std::vector<BDesc>& probs, int* info_host) {
  std::vector<BDesc> small, big;
  for (auto& d : probs) (d.m <= 512 ? small : big).push_back(d);
  if (!small.empty()) {
    int mx = 0;
    for (auto& d : small) mx = std::max(mx, d.m);
    const int ctx[1] = {mx <= 256 ? 1 : 2};
    std::vector<Cfg> cands;
    for (int v = 0; v < 2; v++)
      for (int g : {1, 2, 4}) cands.push_back(Cfg{v, g});
    const std::vector<char> none;
    Cfg c = tune(signature(BK_GETRF, sizeof(T), ctx[0], small), cands, launch_getrf<T>, ctx, small, none, 0, 0, 0, 1,
                 sizeof(T));
    if (const char* e = getenv("OPENBLAS_AMD_BATCH_LU_W")) c.var = atoi(e) == 32 ? 1 : 0;
    double fl = 0, by = 0;
    for (auto& d : small) {
      const double mm = d.m, nn = d.n, k = std::min(mm, nn);
      fl += mm * nn * k - (mm + nn) * k * k / 2.0 + k * k * k / 3.0;   // LU of m x n (LAPACK count, leading terms)
      by += 2.0 * sizeof(T) * mm * nn;
    }
    const int* di;
    {
      KTimer kt("batched_getrf", rt().stream, fl, by);
      di = run_plan(tslot<T>(BK_GETRF), launch_getrf<T>, c, ctx, sizeof ctx, small, none, 0);
    }
    collect_info(di, small, info_host);
  }
  for (auto& d : big)
    info_host[d.idx] = getrf<T>(d.m, d.n, (T*)d.a, d.lda, d.ipiv, rt().stream);
}

template <typename T>
void potrf_batched(bool lower, std::vector<BDesc>& probs, int* info_host) {
  std::vector<BDesc> small, big;
  for (auto& d : probs) (d.n <= 512 ? small : big).push_back(d);
  if (!small.empty()) {
    int mx = 0;
    for (auto& d : small) mx = std::max(mx, d.n);
    const int ctx[2] = {mx <= 256 ? 1 : 2, lower ? 1 : 0};
    std::vector<Cfg> cands;
    for (int v = 0; v < 2; v++)
      for (int g : {1, 2, 4}) cands.push_back(Cfg{v, g});
    const std::vector<char> none;
    Cfg c = tune(signature(BK_POTRF, sizeof(T), ctx[0] * 2 + ctx[1], small), cands, launch_potrf<T>, ctx, small, none,
                 0, 0, 1, 1, sizeof(T));
    double fl = 0, by = 0;
    for (auto& d : small) {
      fl += (double)d.n * d.n * d.n / 3.0;
      by += sizeof(T) * (double)d.n * (d.n + 1);   // the triangle read and written
    }
    const int* di;
    {
      KTimer kt("batched_potrf", rt().stream, fl, by);
      di = run_plan(tslot<T>(BK_POTRF), launch_potrf<T>, c, ctx, sizeof ctx, small, none, 0);
    }
    collect_info(di, small, info_host);
  }
  for (auto& d : big) info_host[d.idx] = potrf<T>(lower, d.n, (T*)d.a, d.lda, rt().stream);
}

template <typename T>
void getrs_batched(bool trans, std::vector<BDesc>& probs, int* info_host) {
  std::vector<BDesc> small, big;
  for (auto& d : probs) (d.n <= 512 ? small : big).push_back(d);
  for (auto& d : probs) info_host[d.idx] = 0;
  if (!small.empty()) {
    int mx = 0;
    for (auto& d : small) mx = std::max(mx, d.n);
    const int S = mx <= 64 ? 1 : mx <= 128 ? 2 : mx <= 256 ? 4 : 8;
    const int ctx[2] = {S, trans ? 1 : 0};
    std::vector<Cfg> cands;
    for (int g : {2, 4, 8, 16}) cands.push_back(Cfg{0, g});
    const std::vector<char> none;
    Cfg c = tune(signature(BK_GETRS, sizeof(T), ctx[0] * 2 + ctx[1], small), cands, launch_getrs<T>, ctx, small, none,
                 0, 1, 1, 2, sizeof(T));
    double fl = 0, by = 0;
    for (auto& d : small) {
      fl += 2.0 * d.n * d.n * std::max(d.k, 0);
      by += sizeof(T) * ((double)d.n * d.n + 2.0 * d.n * std::max(d.k, 0));
    }
    KTimer kt("batched_getrs", rt().stream, fl, by);
    run_plan(tslot<T>(BK_GETRS), launch_getrs<T>, c, ctx, sizeof ctx, small, none, 0);
  }
  for (auto& d : big) {   // large problems: the single-problem solve
    Runtime& r = rt();
    std::vector<int> ipv((size_t)d.n), perm;
    OBA_HIP(hipMemcpyAsync(ipv.data(), d.ipiv, sizeof(int) * (size_t)d.n, hipMemcpyDeviceToHost, r.stream));
    OBA_HIP(hipStreamSynchronize(r.stream));
    ipiv_to_perm(d.n, ipv.data(), perm);
    DevBuf pb(sizeof(int) * perm.size());
    OBA_HIP(hipMemcpyAsync(pb.p, perm.data(), sizeof(int) * perm.size(), hipMemcpyHostToDevice, r.stream));
    getrs_perm<T>(trans, d.n, d.k, (const T*)d.a, d.lda, pb.as<int>(), (T*)d.b, d.ldb, r.stream);
    OBA_HIP(hipStreamSynchronize(r.stream));
  }
}

template void gemm_batched<double>(int, int, std::vector<BDesc>&);
template void gemm_batched<float>(int, int, std::vector<BDesc>&);
template void getrf_batched<double>(std::vector<BDesc>&, int*);
template void getrf_batched<float>(std::vector<BDesc>&, int*);
template void getrs_batched<double>(bool, std::vector<BDesc>&, int*);
template void getrs_batched<float>(bool, std::vector<BDesc>&, int*);
template void potrf_batched<double>(bool, std::vector<BDesc>&, int*);
template void potrf_batched<float>(bool, std::vector<BDesc>&, int*);

}  // namespace oba
