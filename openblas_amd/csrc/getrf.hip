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
#include "lu_device.h"
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <utility>
#include <vector>

namespace oba {

constexpr int LU_W = 32;        // leaf panel width
constexpr int LU_PMAX = 128;    // max workgroups of the panel kernel
// solves with this many columns or fewer use trsv_lookback (one chain per
// column, all in one launch per triangle); more go to the TRSM recursion
constexpr int LU_TRSV_MAX_RHS = 16;
constexpr int LU_EB = 4;        // candidate record groups a wavefront polls at once


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

// ---- panel kernel ----------------------------------------------------------
// Factors A(j0:m, j0:j0+w) held as At(j0:j0+w, j0:m).  Workgroup g owns the
// panel rows [g*ROWS, (g+1)*ROWS) in registers, thread t rows g*ROWS+q*NT+t.
// Rows are never moved inside the kernel: each keeps its current position
// (LAPACK order after the interchanges so far) in a register, so an
// interchange is an O(1) update of two positions and the rows land at their
// final positions when written back.  Per column: a wavefront max-reduction
// (each wavefront's best row staged in LDS), one workgroup barrier, the
// workgroup's best row is published; with one workgroup (P == 1) it goes
// through LDS, otherwise wavefront 0 writes it through to memory as a
// record of 8-byte words, each word = {32 bits of the row, column tag}, plus
// one word {position, tag}.  An 8-byte store is seen whole or not at all, so
// a record is complete exactly when every word carries the tag: no store
// completion wait, no separate header, and the readers (all wavefronts, a
// share of the records each, coalesced loads) get the rows in the same round
// trip that shows they are there.  A second workgroup barrier hands the
// loaded rows and the per-wavefront winners to every thread.  No atomics, no
// cache maintenance.  The column loop runs at run time, but every register
// index is static: the
// update step of column j is
// selected by a uniform switch (a per-element runtime index costs GPR index
// mode, measured at ~10 us per column).
// threads per workgroup (1024-thread workgroups measured slower: the
// per-column barriers cost more with 16 wavefronts)
template <typename T> struct LuNt { static constexpr int v = 512; };
constexpr int LU_MV = 1 + 4 * LU_W;   // moves slot: count + 2*W (dst, src) pairs

// rows per thread (registers: RPT * W values of T per thread)
template <typename T> struct LuRpt { static constexpr int v = 1; };

// candidate records of one column: [2 buffers][LU_PMAX] records of
// LuRec::RW words (double-buffered by column parity: a workgroup can only
// publish column j+2 after every workgroup has read column j+1, hence
// column j).  The tag base + j + 1 never repeats within one factorization
// and the records start zeroed.
template <typename T, int W>
struct LuRec {
  static constexpr int DW = W * (int)sizeof(T) / 4;   // 32-bit words of a row
  static constexpr int RW = DW + 1;                   // + the position word
  static_assert(DW <= 64 && 64 % DW == 0, "a load instruction covers whole records");
};

// Cross-workgroup records: write-through stores (to memory, past the
// non-coherent L2) and loads that bypass the caches, so an exchange needs
// no L2 writeback or invalidation (both measured expensive on this part).
template <typename V>
__device__ inline void wt_store(V* p, V v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM); }
template <typename V>
__device__ inline V bypass_load(const V* p) {
  return __hip_atomic_load(const_cast<V*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// x - l*u with one rounding (the update of every element of the panel)
__device__ inline double lu_fms(double l, double u, double x) { return __builtin_fma(-l, u, x); }
__device__ inline float lu_fms(float l, float u, float x) { return __builtin_fmaf(-l, u, x); }


// x[k] -= l * u[k + OFF] for k in [K0, K1] with k <= kmax (kmax uniform;
// u in LDS): whole chunks of 4 under one uniform branch, their 4 reads
// issued together, selects only in the chunk that straddles kmax, no
// per-element branch or wait
template <int K0, int K1, int OFF, typename T, int W>
__device__ __attribute__((always_inline)) inline void lu_tail_update(T (&x)[W], T l, const T* u, int kmax) {
#pragma unroll
  for (int c0 = K0; c0 <= K1; c0 += 4) {
    if (c0 > kmax) break;
    T uc[4];
#pragma unroll
    for (int k = 0; k < 4; k++) uc[k] = c0 + k <= K1 ? u[c0 + k + OFF] : T(0);
    if (c0 + 3 <= kmax) {
#pragma unroll
      for (int k = c0; k < c0 + 4; k++)
        if (k <= K1) x[k] = lu_fms(l, uc[k - c0], x[k]);
    } else {
#pragma unroll
      for (int k = c0; k < c0 + 4; k++) {
        if (k <= K1) {
          const T f = lu_fms(l, uc[k - c0], x[k]);
          x[k] = k <= kmax ? f : x[k];
        }
      }
    }
  }
}

// rotate a register row left by one element (static indices only)
template <typename T, int W>
__device__ __attribute__((always_inline)) inline void lu_rotate(T (&r)[W]) {
  const T t0 = r[0];
#pragma unroll
  for (int k = 0; k < W - 1; k++) r[k] = r[k + 1];
  r[W - 1] = t0;
}

template <typename T, int W, int RPT, int NT>
__global__ __launch_bounds__(NT) void lu_panel_kernel(T* __restrict__ At, long ldt, long m, long j0, int w,
                                                      int* __restrict__ ipiv, int* __restrict__ info,
                                                      int* __restrict__ moves, unsigned long long* __restrict__ rec,
                                                      unsigned base, unsigned* err, int pre) {
  static_assert(W <= 64, "one pivot-row element per lane");
  constexpr int ROWS = NT * RPT;
  constexpr int NWAVE = NT / 64;
  const int P = gridDim.x, g = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const long mp = m - j0;
  constexpr int DW = LuRec<T, W>::DW, RW = LuRec<T, W>::RW;
  constexpr int RPI = 64 / DW;         // records per load instruction
  // LDS buffers that a wavefront may still read while a faster one starts
  // the next column are double-buffered by column parity
  __shared__ double s_v[NWAVE];
  __shared__ int s_p[NWAVE];
  __shared__ T s_cand[2][NWAVE][W];    // each wavefront's best row
  __shared__ T s_rows[2][LU_PMAX][W];  // the P candidate rows of the column (P > 1)
  __shared__ double s_wv[NWAVE];       // per wavefront: best record key,
  __shared__ int s_wp[NWAVE];          // its position,
  __shared__ int s_wr[NWAVE];          // its record (-2: timed out)
  __shared__ int s_ipiv[W];            // the panel's pivots, written out at the end
  int zinfo = 0;                       // first zero pivot (1-based, absolute)

  // Row q of this thread, rotated: at column step j, x[q][k] holds element
  // (j + k) mod W.  The current column is always x[q][0] and the column step
  // is one code path with static register indices (no per-column code, which
  // would not fit the instruction cache); rows in LDS and in the records are
  // kept in the same rotated order.
  T x[RPT][W];
  int pos[RPT];
  bool live[RPT];   // not yet a pivot row
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const long r = (long)g * ROWS + q * NT + t;
    live[q] = r < mp;
    pos[q] = (int)r;
    const T* src = At + j0 + (j0 + r) * ldt;
#pragma unroll
    for (int c = 0; c < W; c++) x[q][c] = (live[q] && c < w) ? src[c] : T(0);
  }
  // pre: this panel is the right half of a 2W-wide node whose left half
  // [j0-W, j0) was just factored and its interchanges applied here.  The
  // node's TRSM and GEMM happen in this kernel, before the column loop:
  // U12 = L11^{-1} A12 (every workgroup solves the W x w block in LDS; WG 0
  // stores it at the end, after every workgroup has read A12, which the
  // first column's exchange guarantees), then each row takes
  // x -= l(row) * U12 with its W multipliers of the left half.
  __shared__ T s_u12[W][W];
  if (pre) {
    T* stg = &s_rows[0][0][0];   // L11 | A12 staging, free before the loop
    // At column j0-W+k, rows j0-W .. j0+W: row k of L11 (first W), of A12;
    // all loads issued before the LDS writes (entries past the panel read
    // row j0 instead, in bounds, and are zeroed)
    static_assert((W * 2 * W) % NT == 0, "the staging divides over the workgroup");
    constexpr int SU = W * 2 * W / NT;
    T sv[SU];
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const int idx = t + u * NT, k = idx / (2 * W), e = idx % (2 * W);
      const int er = (e < W || e - W < w) ? e : W;
      sv[u] = At[(j0 - W + er) + (j0 - W + k) * ldt];
    }
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const int idx = t + u * NT, e = idx % (2 * W);
      stg[idx] = (e < W || e - W < w) ? sv[u] : T(0);
    }
    __syncthreads();
    if (t < W) {   // column t of U12: forward substitution with unit L11
      T u[W];
#pragma unroll
      for (int k = 0; k < W; k++) {
        T v = stg[k * 2 * W + W + t];
#pragma unroll
        for (int kk = 0; kk < k; kk++) v = lu_fms(stg[k * 2 * W + kk], u[kk], v);
        u[k] = v;
      }
#pragma unroll
      for (int k = 0; k < W; k++) s_u12[k][t] = u[k];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      if (live[q]) {
        const long r = (long)g * ROWS + q * NT + t;
        const T* lr = At + (j0 - W) + (j0 + r) * ldt;
#pragma unroll
        for (int k = 0; k < W; k++) {
          const T l = lr[k];
#pragma unroll
          for (int c = 0; c < W; c++) x[q][c] = lu_fms(l, s_u12[k][c], x[q][c]);
        }
      }
    }
  }
  // the row loads complete here, once: otherwise the compiler keeps a
  // vmcnt(0) wait inside the column loop, which then also waits for the
  // write-through stores of the previous column
  __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0) only (gfx9 encoding)
  // With P > 1 the tail of column j's update (elements j+2..) is deferred
  // into column j+1's exchange, where it overlaps the wait for the other
  // workgroups' records; the published candidate rows get the same update
  // (same operands, same rounding) on the way out.
  T lqp[RPT];        // the previous column's multipliers,
  bool updp[RPT];    // which rows took its update,
  int frp = -1;      // and its pivot record in s_rows[(j-1)&1]
#pragma unroll
  for (int q = 0; q < RPT; q++) { lqp[q] = T(0); updp[q] = false; }

#pragma unroll 1
  for (int j = 0; j < w; j++) {
    const int buf = j & 1;
    const int kmax = W - 1 - j;   // x[.][1..kmax]: the columns right of j
    // (a) this thread's, then this wavefront's best: max key, lowest position
    double bv = -2.0;
    int bp = INT_MAX;
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      if (live[q]) {
        const double v = pivot_key(x[q][0], pos[q], j);
        if (better(v, pos[q], bv, bp)) { bv = v; bp = pos[q]; }
      }
    }
    lu_wave_best(bv, bp);
#pragma unroll
    for (int q = 0; q < RPT; q++)
      if (live[q] && pos[q] == bp) {
#pragma unroll
        for (int c = 0; c < W; c++) s_cand[buf][wave][c] = x[q][c];
      }
    if (lane == 0) { s_v[wave] = bv; s_p[wave] = bp; }
    __syncthreads();
    // (b) the workgroup's best (every thread reduces the same values)
    double wv = s_v[0];
    int wp = s_p[0], wk = 0;
#pragma unroll
    for (int k = 1; k < NWAVE; k++)
      if (better(s_v[k], s_p[k], wv, wp)) { wv = s_v[k]; wp = s_p[k]; wk = k; }
    int gp;          // pivot position
    const T* urow;   // the pivot row in LDS (rotated like x)
    int fr = -1;
    if (P == 1) {
      gp = wp;
      urow = &s_cand[buf][wk][0];
    } else {
      const unsigned tag = base + (unsigned)(j + 1);
      const unsigned long long tg = (unsigned long long)tag << 32;
      unsigned long long* recs = rec + (size_t)buf * LU_PMAX * RW;
      // the previous pivot row, rotated for column j-1: its element k+1 is
      // column j+k
      const T* uprev = &s_rows[buf ^ 1][frp < 0 ? 0 : frp][0];
      // (c) publish: word e of the record = {word e of the row, tag}; the
      // last word {position (INT_MAX: no live row here), tag}.  The row gets
      // the deferred tail of column j-1 first (its multiplier of column j-1
      // is element W-1).
      if (wave == 0) {
        unsigned long long* dst = recs + (size_t)g * RW;
        if (lane < W) {
          T v = s_cand[buf][wk][lane];
          if (j > 0 && lane >= 1 && lane <= kmax) v = lu_fms(s_cand[buf][wk][W - 1], uprev[lane + 1], v);
          if constexpr (sizeof(T) == 8) {
            wt_store(dst + 2 * lane, tg | (unsigned)__double2loint(v));
            wt_store(dst + 2 * lane + 1, tg | (unsigned)__double2hiint(v));
          } else {
            wt_store(dst + lane, tg | (unsigned)__float_as_int(v));
          }
        }
        if (lane == 0) wt_store(dst + DW, tg | (unsigned)wp);
      }
      // (d) the deferred tail of column j-1 on this thread's rows (all
      // pivot-row reads first, then selects: no per-element branch waits)
      if (j > 0) {
#pragma unroll
        for (int q = 0; q < RPT; q++)
          if (updp[q]) lu_tail_update<1, W - 2, 1, T, W>(x[q], lqp[q], uprev, kmax);
      }
      // (e) wavefront v collects records v*RPI.., (v+NWAVE)*RPI.., ...:
      // lane = h*DW + d loads word d of record r0+h, and word DW when d == 0.
      // Up to LU_EB record groups are in flight at once (one round trip
      // for all of them; polling them one after the other costs one round
      // trip each).
      constexpr int GS = NWAVE * RPI;   // record stride between a wavefront's groups
      const int h = lane / DW, d = lane % DW;
      double bv2 = -2.0;
      int bp2 = INT_MAX, br2 = -1;
      bool ok = true;
      for (int c0 = wave * RPI; c0 < P; c0 += LU_EB * GS) {
        unsigned long long wd[LU_EB], wq[LU_EB];
        bool have[LU_EB];
#pragma unroll
        for (int k = 0; k < LU_EB; k++) {
          have[k] = c0 + k * GS + h >= P;
          wd[k] = 0;
          wq[k] = 0;
        }
        long spins = 0;
        while (true) {
#pragma unroll
          for (int k = 0; k < LU_EB; k++) {
            if (!have[k]) {
              const unsigned long long* src = recs + (size_t)(c0 + k * GS + h) * RW;
              wd[k] = bypass_load(src + d);
              if (d == 0) wq[k] = bypass_load(src + DW);
            }
          }
          bool all = true;
#pragma unroll
          for (int k = 0; k < LU_EB; k++) {
            if (!have[k])
              have[k] = (unsigned)(wd[k] >> 32) == tag && (d != 0 || (unsigned)(wq[k] >> 32) == tag);
            all = all && have[k];
          }
          if (__all(all)) break;
          __builtin_amdgcn_s_sleep(1);
          if (++spins > (1L << 22)) { ok = false; break; }
        }
        if (!ok) break;
#pragma unroll
        for (int k = 0; k < LU_EB; k++) {
          const int rb = c0 + k * GS;
          if (rb >= P) break;
          if (rb + h < P) reinterpret_cast<unsigned*>(&s_rows[buf][rb + h][0])[d] = (unsigned)wd[k];
          const int wdi = (int)(unsigned)wd[k], wqi = (int)(unsigned)wq[k];
#pragma unroll
          for (int hh = 0; hh < RPI; hh++) {
            if (rb + hh < P) {
              const int cp = __builtin_amdgcn_readlane(wqi, hh * DW);
              T key;   // element 0: column j
              if constexpr (sizeof(T) == 8) {
                const int lo = __builtin_amdgcn_readlane(wdi, hh * DW);
                const int hi = __builtin_amdgcn_readlane(wdi, hh * DW + 1);
                key = __hiloint2double(hi, lo);
              } else {
                key = __int_as_float(__builtin_amdgcn_readlane(wdi, hh * DW));
              }
              if (cp != INT_MAX) {
                const double vv = pivot_key(key, cp, j);
                if (better(vv, cp, bv2, bp2)) { bv2 = vv; bp2 = cp; br2 = rb + hh; }
              }
            }
          }
        }
      }
      if (lane == 0) { s_wv[wave] = bv2; s_wp[wave] = bp2; s_wr[wave] = ok ? br2 : -2; }
      __syncthreads();
      // (f) the column's pivot: every thread reduces the same values
      double fv = s_wv[0];
      int fp = s_wp[0];
      fr = s_wr[0];
      bool allok = fr != -2;
#pragma unroll
      for (int k = 1; k < NWAVE; k++) {
        const int kr = s_wr[k];
        allok = allok && kr != -2;
        if (better(s_wv[k], s_wp[k], fv, fp)) { fv = s_wv[k]; fp = s_wp[k]; fr = kr; }
      }
      if (!allok || fr < 0 || fr >= P || fp < j || fp >= mp) {   // timeout; the rest cannot happen
        if (t == 0) atomicExch(err, allok ? 2u : 1u);
        break;
      }
      gp = fp;
      urow = &s_rows[buf][fr][0];
    }
    const T piv = urow[0];
    const bool nz = piv != T(0);
    const bool doscale = nz && fabs(piv) >= MinNormal<T>::v;
    if (t == 0) s_ipiv[j] = (int)(j0 + gp + 1);
    if (!nz && zinfo == 0) zinfo = (int)(j0 + j + 1);
    const T rinv = doscale ? T(1) / piv : T(1);
    // interchange (positions only), multipliers, the update of column j+1
    // (and, with P == 1, of the rest), then the rotation to column j+1
    const bool defer = P > 1;
    T lq[RPT];
    bool upd[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      upd[q] = false;
      lq[q] = T(0);
      if (live[q]) {
        if (pos[q] == gp) {   // the pivot row: now a row of U
          pos[q] = j;
          live[q] = false;
        } else {
          if (pos[q] == j) pos[q] = gp;
          lq[q] = doscale ? x[q][0] * rinv : x[q][0];
          upd[q] = true;
        }
      }
      if (upd[q]) {
        x[q][0] = lq[q];
        if (kmax >= 1) x[q][1] = lu_fms(lq[q], urow[1], x[q][1]);
        if (!defer) lu_tail_update<2, W - 1, 0, T, W>(x[q], lq[q], urow, kmax);
      }
      lu_rotate<T, W>(x[q]);
    }
#pragma unroll
    for (int q = 0; q < RPT; q++) { lqp[q] = lq[q]; updp[q] = upd[q]; }
    frp = fr;
  }
  if (g == 0) {   // the pivots and the first zero pivot; U12
    __syncthreads();
    if (t < w) ipiv[j0 + t] = s_ipiv[t];
    if (t == 0 && zinfo != 0 && *info == 0) *info = zinfo;
    if (pre) {
      for (int idx = t; idx < W * w; idx += NT) {
        const int k = idx / w, c = idx % w;
        At[(j0 + c) + (j0 - W + k) * ldt] = s_u12[k][c];
      }
    }
  }
  // the last deferred tail only touches elements >= w + 1 (padding); turn
  // the rows back to column order
#pragma unroll 1
  for (int jj = w; jj < W; jj++) {
#pragma unroll
    for (int q = 0; q < RPT; q++) lu_rotate<T, W>(x[q]);
  }
  if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  // every row to its final position; list the rows that moved
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const long r = (long)g * ROWS + q * NT + t;
    if (r < mp) {
      T* dst = At + j0 + (j0 + pos[q]) * ldt;
#pragma unroll
      for (int c = 0; c < W; c++)
        if (c < w) dst[c] = x[q][c];
      if (pos[q] != (int)r) {
        const int k = atomicAdd(moves, 1);
        if (k < 2 * W) {   // at most 2w rows change position
          moves[1 + 2 * k] = pos[q];
          moves[2 + 2 * k] = (int)r;
        }
      }
    }
  }
}

// ---- panel kernel, second form ---------------------------------------------------
// Same contract as lu_panel_kernel, shaped for fewer instructions per column:
//  * 256-thread workgroups (one wavefront per SIMD) with RPT rows per thread,
//    so the per-column fixed work (reductions, barriers, the pivot-row
//    reads) is paid by 4 wavefronts instead of 8;
//  * pivot keys in the panel's own precision (NaN mapped so that it wins only
//    at the current position, kernel/arm/iamax.c:51-74): a wavefront max of
//    the key, then the lowest position among the lanes holding that key, each
//    a DPP butterfly within the 16-lane rows plus four readlanes;
//  * rows rotated by U columns once every U columns (column j0+u of a group
//    sits in register u of the group's rotation), so each of the U unrolled
//    column bodies has static register indices and the rotation costs W/U
//    moves per column instead of W;
//  * each wavefront keeps only the best record of its share in LDS, so the
//    LDS footprint does not grow with the workgroup count (P is bounded by
//    the record area and by residency, not by LDS).
// With P > 1 the tail of column j's row update is deferred into column j+1's
// exchange wait (as in lu_panel_kernel).
constexpr int L2_NT = 256;
constexpr int L2_NW = L2_NT / 64;
constexpr int L2_U = 4;           // columns per rotation group
constexpr int L2_PMAX = 1024;     // record slots per column buffer
constexpr int L2_EB = 8;          // records a wavefront polls at once

template <typename T>
__device__ inline T lu2_key(T x, int pos, int j) {
  return isnan(x) ? (pos == j ? T(INFINITY) : T(-1)) : fabs(x);
}

template <int CTRL>
__device__ inline float lu2_dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ inline double lu2_dpp(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ inline int lu2_dpp(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
__device__ inline float lu2_rl(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ inline double lu2_rl(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l), __builtin_amdgcn_readlane(__double2loint(v), l));
}

// (key, position) best of the wavefront, uniform in every lane: the largest
// key, ties to the lowest position (IAMAX: first maximum)
template <typename T>
__device__ __attribute__((always_inline)) inline void lu2_wave_best(T& key, int& pos) {
  T m = key;
  m = fmax(m, lu2_dpp<0xB1>(m));
  m = fmax(m, lu2_dpp<0x4E>(m));
  m = fmax(m, lu2_dpp<0x141>(m));
  m = fmax(m, lu2_dpp<0x140>(m));
  m = fmax(fmax(lu2_rl(m, 0), lu2_rl(m, 16)), fmax(lu2_rl(m, 32), lu2_rl(m, 48)));
  int p = key == m ? pos : INT_MAX;
  p = min(p, lu2_dpp<0xB1>(p));
  p = min(p, lu2_dpp<0x4E>(p));
  p = min(p, lu2_dpp<0x141>(p));
  p = min(p, lu2_dpp<0x140>(p));
  p = min(min(__builtin_amdgcn_readlane(p, 0), __builtin_amdgcn_readlane(p, 16)),
          min(__builtin_amdgcn_readlane(p, 32), __builtin_amdgcn_readlane(p, 48)));
  key = m;
  pos = p;
}

// rotate a register row left by U elements (static indices only)
template <int U, typename T, int W>
__device__ __attribute__((always_inline)) inline void lu2_rotate(T (&r)[W]) {
  T t[U];
#pragma unroll
  for (int k = 0; k < U; k++) t[k] = r[k];
#pragma unroll
  for (int k = 0; k < W - U; k++) r[k] = r[k + U];
#pragma unroll
  for (int k = 0; k < U; k++) r[W - U + k] = t[k];
}

// x[k] -= l * u[k + OFF] for k in [K0, W-1] with k <= kmax (kmax uniform):
// chunks of 4 under one uniform branch, selects only in the straddling chunk
template <int K0, int OFF, typename T, int W>
__device__ __attribute__((always_inline)) inline void lu2_tail(T (&x)[W], T l, const T* u, int kmax) {
#pragma unroll
  for (int c0 = K0; c0 < W; c0 += 4) {
    if (c0 > kmax) break;
    T uc[4];
#pragma unroll
    for (int k = 0; k < 4; k++) uc[k] = (c0 + k < W && c0 + k + OFF < W) ? u[c0 + k + OFF] : T(0);
    if (c0 + 3 <= kmax) {
#pragma unroll
      for (int k = c0; k < c0 + 4; k++)
        if (k < W) x[k] = lu_fms(l, uc[k - c0], x[k]);
    } else {
#pragma unroll
      for (int k = c0; k < c0 + 4; k++) {
        if (k < W) {
          const T f = lu_fms(l, uc[k - c0], x[k]);
          x[k] = k <= kmax ? f : x[k];
        }
      }
    }
  }
}

template <typename T, int W, int RPT>
__global__ __launch_bounds__(L2_NT) void lu_panel2_kernel(T* __restrict__ At, long ldt, long m, long j0, int w,
                                                          int* __restrict__ ipiv, int* __restrict__ info,
                                                          int* __restrict__ moves, unsigned long long* __restrict__ rec,
                                                          unsigned base, unsigned* err, int pre) {
  constexpr int NT = L2_NT, NW = L2_NW, U = L2_U;
  constexpr int ROWS = NT * RPT;
  static_assert(W % U == 0 && W <= 64, "rotation groups must tile the panel");
  const int P = gridDim.x, g = blockIdx.x, t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const long mp = m - j0;
  constexpr int DW = W * (int)sizeof(T) / 4;   // 32-bit words of a row
  constexpr int RW = DW + 1;                   // + the position word
  constexpr int RPI = 64 / DW;                 // records per load instruction
  static_assert(DW <= 64 && 64 % DW == 0, "a load instruction covers whole records");
  __shared__ T s_cand[2][NW][W];     // each wavefront's best local row
  __shared__ T s_key[2][NW];
  __shared__ int s_pos[2][NW];
  __shared__ T s_best[2][NW][W];     // each wavefront's best polled record row (P > 1)
  __shared__ T s_bkey[2][NW];
  __shared__ int s_bpos[2][NW];
  __shared__ int s_bok[2][NW];
  __shared__ int s_ipiv[W];
  __shared__ T s_u12[W][W];
  __shared__ T s_stage[W * 2 * W];   // L11 | A12 staging for `pre`
  int zinfo = 0;

  T x[RPT][W];
  int pos[RPT];
  bool live[RPT];
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const long r = (long)g * ROWS + q * NT + t;
    live[q] = r < mp;
    pos[q] = (int)r;
    const T* src = At + j0 + (j0 + r) * ldt;
#pragma unroll
    for (int c = 0; c < W; c++) x[q][c] = (live[q] && c < w) ? src[c] : T(0);
  }
  // pre: the node's U12 = L11^{-1} A12 and each row's x -= l(row) U12 (see
  // lu_panel_kernel); the k loop stays rolled to keep the code small
  if (pre) {
    // all loads before the LDS writes, as in lu_panel_kernel
    static_assert((W * 2 * W) % NT == 0, "the staging divides over the workgroup");
    constexpr int SU = W * 2 * W / NT;
    T sv[SU];
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const int idx = t + u * NT, k = idx / (2 * W), e = idx % (2 * W);
      const int er = (e < W || e - W < w) ? e : W;
      sv[u] = At[(j0 - W + er) + (j0 - W + k) * ldt];
    }
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const int idx = t + u * NT, e = idx % (2 * W);
      s_stage[idx] = (e < W || e - W < w) ? sv[u] : T(0);
    }
    __syncthreads();
    if (t < W) {
      T u[W];
#pragma unroll
      for (int k = 0; k < W; k++) {
        T v = s_stage[k * 2 * W + W + t];
#pragma unroll
        for (int kk = 0; kk < k; kk++) v = lu_fms(s_stage[k * 2 * W + kk], u[kk], v);
        u[k] = v;
      }
#pragma unroll
      for (int k = 0; k < W; k++) s_u12[k][t] = u[k];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      if (live[q]) {
        const long r = (long)g * ROWS + q * NT + t;
        const T* lr = At + (j0 - W) + (j0 + r) * ldt;
#pragma unroll 1
        for (int k = 0; k < W; k++) {
          const T l = lr[k];
#pragma unroll
          for (int c = 0; c < W; c++) x[q][c] = lu_fms(l, s_u12[k][c], x[q][c]);
        }
      }
    }
  }
  __builtin_amdgcn_s_waitcnt(0x0F70);   // the row loads complete here, once (vmcnt(0))

  T lqp[RPT];
  bool updp[RPT];
  int frp = 0;   // the previous column's pivot record slot in s_best[(j-1)&1] (P > 1)
#pragma unroll
  for (int q = 0; q < RPT; q++) { lqp[q] = T(0); updp[q] = false; }
  bool failed = false;
  int jg = 0;   // first column of the current rotation group
  // one column of the current group; u (static) is the column's register
  auto column = [&](auto u_tag) __attribute__((always_inline)) {
      constexpr int u = decltype(u_tag)::value;
      const int j = jg + u;
      if (j >= w || failed) return;
      const int buf = j & 1;
      const int kmax = W - 1 - j;    // elements right of column j: x[u+1 .. u+kmax]
      // (a) thread best over its rows, then the wavefront's
      T bk = T(-2);
      int bp = INT_MAX;
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        if (live[q]) {
          const T kq = lu2_key(x[q][u], pos[q], j);
          if (kq > bk || (kq == bk && pos[q] < bp)) { bk = kq; bp = pos[q]; }
        }
      }
      lu2_wave_best(bk, bp);
#pragma unroll
      for (int q = 0; q < RPT; q++)
        if (live[q] && pos[q] == bp) {
#pragma unroll
          for (int c = 0; c < W; c++) s_cand[buf][wave][c] = x[q][c];
        }
      if (lane == 0) { s_key[buf][wave] = bk; s_pos[buf][wave] = bp; }
      __syncthreads();
      // (b) the workgroup's best (every thread reduces the same values)
      T wk = s_key[buf][0];
      int wp = s_pos[buf][0], wv = 0;
#pragma unroll
      for (int v = 1; v < NW; v++) {
        const T kv = s_key[buf][v];
        const int pv = s_pos[buf][v];
        if (kv > wk || (kv == wk && pv < wp)) { wk = kv; wp = pv; wv = v; }
      }
      int gp;
      const T* urow;
      // the previous column's pivot row (group rotation of column j-1: for
      // u == 0 it was staged in the previous group's rotation, U further)
      constexpr int POFF = (u == 0) ? U : 0;
      if (P == 1) {
        gp = wp;
        urow = &s_cand[buf][wv][0];
      } else {
        const unsigned tag = base + (unsigned)(j + 1);
        const unsigned long long tg = (unsigned long long)tag << 32;
        unsigned long long* recs = rec + (size_t)buf * L2_PMAX * RW;
        const T* uprev = &s_best[buf ^ 1][frp][0];
        // (c) publish the workgroup's candidate, with column j-1's deferred
        // tail applied (its multiplier is element u-1, or W-1 when u == 0)
        if (wave == 0) {
          unsigned long long* dst = recs + (size_t)g * RW;
          constexpr int EPL = (int)sizeof(T) / 4;   // words per element
          const int e = lane / EPL;                 // element of this lane's word
          if (lane < DW) {
            T v = s_cand[buf][wv][e];
            const T lprev = s_cand[buf][wv][(u + W - 1) % W];
            if (j > 0 && e >= u + 1 && e <= u + kmax) v = lu_fms(lprev, uprev[e + POFF < W ? e + POFF : W - 1], v);
            unsigned word;
            if constexpr (sizeof(T) == 8) word = (lane & 1) ? (unsigned)__double2hiint(v) : (unsigned)__double2loint(v);
            else word = (unsigned)__float_as_int(v);
            wt_store(dst + lane, tg | word);
          }
          if (lane == 0) wt_store(dst + DW, tg | (unsigned)wp);
        }
        // (d) the deferred tail of column j-1 on this thread's rows
        if (j > 0) {
#pragma unroll
          for (int q = 0; q < RPT; q++)
            if (updp[q]) lu2_tail<u + 1, POFF, T, W>(x[q], lqp[q], uprev, u + kmax);
        }
        // (e) wavefront v polls records v*RPI.., (v+NW)*RPI.., ...; lane
        // h*DW + d holds word d of record (group base + h); the best record
        // of the share goes to s_best[buf][v]
        constexpr int GS = NW * RPI;
        const int h = lane / DW, d = lane % DW;
        T vbk = T(-2);
        int vbp = INT_MAX;
        bool ok = true;
        for (int c0 = wave * RPI; c0 < P && ok; c0 += L2_EB * GS) {
          unsigned long long wd[L2_EB], wq[L2_EB];
          bool have[L2_EB];
#pragma unroll
          for (int k = 0; k < L2_EB; k++) {
            have[k] = c0 + k * GS + h >= P;
            wd[k] = 0;
            wq[k] = 0;
          }
          long spins = 0;
          while (true) {
#pragma unroll
            for (int k = 0; k < L2_EB; k++) {
              if (!have[k]) {
                const unsigned long long* src = recs + (size_t)(c0 + k * GS + h) * RW;
                wd[k] = bypass_load(src + d);
                if (d == 0) wq[k] = bypass_load(src + DW);
              }
            }
            bool all = true;
#pragma unroll
            for (int k = 0; k < L2_EB; k++) {
              if (!have[k]) have[k] = (unsigned)(wd[k] >> 32) == tag && (d != 0 || (unsigned)(wq[k] >> 32) == tag);
              all = all && have[k];
            }
            if (__all(all)) break;
            __builtin_amdgcn_s_sleep(1);
            if (++spins > (1L << 22)) { ok = false; break; }
          }
          if (!ok) break;
#pragma unroll
          for (int k = 0; k < L2_EB; k++) {
            const int rb = c0 + k * GS;
            if (rb >= P) break;
            const int wdi = (int)(unsigned)wd[k], wqi = (int)(unsigned)wq[k];
#pragma unroll
            for (int hh = 0; hh < RPI; hh++) {
              if (rb + hh < P) {
                const int cp = __builtin_amdgcn_readlane(wqi, hh * DW);
                T key;   // element u of the record: column j
                if constexpr (sizeof(T) == 8) {
                  const int lo = __builtin_amdgcn_readlane(wdi, hh * DW + 2 * u);
                  const int hi = __builtin_amdgcn_readlane(wdi, hh * DW + 2 * u + 1);
                  key = __hiloint2double(hi, lo);
                } else {
                  key = __int_as_float(__builtin_amdgcn_readlane(wdi, hh * DW + u));
                }
                if (cp != INT_MAX) {
                  const T kv = lu2_key(key, cp, j);
                  if (kv > vbk || (kv == vbk && cp < vbp)) {
                    vbk = kv;
                    vbp = cp;
                    // stage this record's row (the lanes that hold it)
                    if (h == hh) reinterpret_cast<unsigned*>(&s_best[buf][wave][0])[d] = (unsigned)wd[k];
                  }
                }
              }
            }
          }
        }
        if (lane == 0) { s_bkey[buf][wave] = vbk; s_bpos[buf][wave] = vbp; s_bok[buf][wave] = ok ? 1 : 0; }
        __syncthreads();
        // (f) the column's pivot: every thread reduces the same values
        T fk = s_bkey[buf][0];
        int fp = s_bpos[buf][0], fv = 0;
        bool allok = s_bok[buf][0] != 0;
#pragma unroll
        for (int v = 1; v < NW; v++) {
          allok = allok && s_bok[buf][v] != 0;
          const T kv = s_bkey[buf][v];
          const int pv = s_bpos[buf][v];
          if (kv > fk || (kv == fk && pv < fp)) { fk = kv; fp = pv; fv = v; }
        }
        if (!allok || fp < j || fp >= mp) {   // a wait ran out; the rest cannot happen
          if (t == 0) atomicExch(err, allok ? 2u : 1u);
          failed = true;
          return;
        }
        gp = fp;
        urow = &s_best[buf][fv][0];
        frp = fv;
      }
      const T piv = urow[u];
      const bool nz = piv != T(0);
      const bool doscale = nz && fabs(piv) >= MinNormal<T>::v;
      if (t == 0) s_ipiv[j] = (int)(j0 + gp + 1);
      if (!nz && zinfo == 0) zinfo = (int)(j0 + j + 1);
      const T rinv = doscale ? T(1) / piv : T(1);
      const bool defer = P > 1;
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        bool upd = false;
        T lq = T(0);
        if (live[q]) {
          if (pos[q] == gp) {   // the pivot row: now a row of U
            pos[q] = j;
            live[q] = false;
          } else {
            if (pos[q] == j) pos[q] = gp;
            lq = doscale ? x[q][u] * rinv : x[q][u];
            upd = true;
          }
        }
        if (upd) {
          x[q][u] = lq;
          if (kmax >= 1) x[q][u + 1] = lu_fms(lq, urow[u + 1], x[q][u + 1]);
          if (!defer) lu2_tail<u + 2, 0, T, W>(x[q], lq, urow, u + kmax);
        }
        lqp[q] = lq;
        updp[q] = upd;
      }
  };
  static_assert(U == 4, "the group below calls four column bodies");
#pragma unroll 1
  for (; jg < w && !failed; jg += U) {
    column(std::integral_constant<int, 0>{});
    column(std::integral_constant<int, 1>{});
    column(std::integral_constant<int, 2>{});
    column(std::integral_constant<int, 3>{});
    if (failed) break;
    // next group: rotate by U (column jg+U moves to register 0)
    if (jg + U < w) {
#pragma unroll
      for (int q = 0; q < RPT; q++) lu2_rotate<U, T, W>(x[q]);
    }
  }
  if (g == 0) {   // the pivots and the first zero pivot; U12
    __syncthreads();
    if (t < w && !failed) ipiv[j0 + t] = s_ipiv[t];
    if (t == 0 && zinfo != 0 && *info == 0) *info = zinfo;
    if (pre) {
      for (int idx = t; idx < W * w; idx += NT) {
        const int k = idx / w, c = idx % w;
        At[(j0 + c) + (j0 - W + k) * ldt] = s_u12[k][c];
      }
    }
  }
  if (failed || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  // back to column order: the rotation so far is the start of the last group
  const int rot = (jg < w ? jg : (w - 1) / U * U) % W;
#pragma unroll 1
  for (int k = rot; k % W != 0; k += U) {
#pragma unroll
    for (int q = 0; q < RPT; q++) lu2_rotate<U, T, W>(x[q]);
  }
  // every row to its final position; list the rows that moved
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const long r = (long)g * ROWS + q * NT + t;
    if (r < mp) {
      T* dst = At + j0 + (j0 + pos[q]) * ldt;
#pragma unroll
      for (int c = 0; c < W; c++)
        if (c < w) dst[c] = x[q][c];
      if (pos[q] != (int)r) {
        const int k = atomicAdd(moves, 1);
        if (k < 2 * W) {
          moves[1 + 2 * k] = pos[q];
          moves[2 + 2 * k] = (int)r;
        }
      }
    }
  }
}

// Apply a panel's interchanges to the columns of A in [lo, hi) outside
// [ex_lo, ex_hi), i.e. to those rows of At: the moves (destination position,
// source row, both relative to the panel's first row j0) form a permutation
// of at most 2w columns of At, so every source element is read before any
// destination is written (no sequential dependence, coalesced along the rows
// of At).  The grid covers [lo, hi) in 64-row slabs.
template <typename T, int W>
__global__ __launch_bounds__(256) void lu_swap_kernel(T* __restrict__ At, long ldt, long lo, long hi, long ex_lo,
                                                      long ex_hi, long j0, long mp, const int* __restrict__ moves) {
  constexpr int RB = 64;                       // At rows per workgroup
  constexpr int MU = 2 * W * RB / 256;         // gathers per thread at most
  static_assert(2 * W <= 256 && (2 * W * RB) % 256 == 0, "one move per thread; whole gather rounds");
  __shared__ int sd[2 * W], ss[2 * W];
  // the count and this thread's move in one round trip
  const int t = threadIdx.x;
  const int c0 = moves[0];
  int d = 0, sr = 0;
  if (t < 2 * W) {
    d = moves[1 + 2 * t];
    sr = moves[2 + 2 * t];
  }
  const int cnt = min(max(c0, 0), 2 * W);
  if (t < 2 * W) {
    if (t >= cnt || d < 0 || d >= mp || sr < 0 || sr >= mp) d = sr = 0;   // never index outside the panel rows
    sd[t] = d;
    ss[t] = sr;
  }
  __syncthreads();
  const long r0 = lo + (long)blockIdx.x * RB;
  // gather every source first (coalesced along the rows of At; all of a
  // thread's loads in flight together) ...
  T v[MU];
#pragma unroll
  for (int u = 0; u < MU; u++) {
    const int idx = t + 256 * u, k = idx / RB, rr = idx % RB;
    const long r = r0 + rr;
    const bool ok = idx < cnt * RB && r < hi && !(r >= ex_lo && r < ex_hi);
    v[u] = ok ? At[r + (j0 + ss[k]) * ldt] : T(0);
  }
  __syncthreads();   // every read before any write: a move's destination can be another's source
  // ... then scatter: the moves are a permutation of the touched columns
#pragma unroll
  for (int u = 0; u < MU; u++) {
    const int idx = t + 256 * u, k = idx / RB, rr = idx % RB;
    const long r = r0 + rr;
    if (idx < cnt * RB && r < hi && !(r >= ex_lo && r < ex_hi)) At[r + (j0 + sd[k]) * ldt] = v[u];
  }
}

template <typename T>
static void lu_swap(T* At, long ldt, long lo, long hi, long ex_lo, long ex_hi, long j0, long mp, const int* mv,
                    hipStream_t s) {
  if (hi <= lo || (lo >= ex_lo && hi <= ex_hi)) return;   // nothing outside the excluded range
  lu_swap_kernel<T, LU_W><<<(unsigned)((hi - lo + 63) / 64), 256, 0, s>>>(At, ldt, lo, hi, ex_lo, ex_hi, j0, mp, mv);
  OBA_HIP(hipGetLastError());
}

// one factored leaf panel: its first row/column and its moves slot
struct LuLeaf {
  long j0;
  int* mv;
};

// the leaves of one block, in order, for lu_swap_block_kernel
constexpr int LU_MAXL = 16;
struct LuLeafSet {
  long j0[LU_MAXL];
  const int* mv[LU_MAXL];
  int n;
};

// The interchanges of all leaves of a block, in order, on the At rows in
// [lo, hi) outside [ex_lo, ex_hi): lu_swap_kernel's gather/scatter once per
// leaf inside one launch (each workgroup owns its 64-row slab throughout).
template <typename T, int W>
__global__ __launch_bounds__(256) void lu_swap_block_kernel(T* __restrict__ At, long ldt, long lo, long hi, long ex_lo,
                                                            long ex_hi, long m, LuLeafSet ls) {
  constexpr int RB = 64;
  __shared__ int sd[2 * W], ss[2 * W];
  __shared__ int s_cnt;
  __shared__ T vals[2 * W][RB];
  const long r0 = lo + (long)blockIdx.x * RB;
  for (int l = 0; l < ls.n; l++) {
    const long j0 = ls.j0[l], mp = m - j0;
    const int* moves = ls.mv[l];
    if (threadIdx.x == 0) s_cnt = min(max(moves[0], 0), 2 * W);
    __syncthreads();
    const int cnt = s_cnt;
    if ((int)threadIdx.x < cnt) {
      int d = moves[1 + 2 * threadIdx.x], sr = moves[2 + 2 * threadIdx.x];
      if (d < 0 || d >= mp || sr < 0 || sr >= mp) d = sr = 0;   // never index outside the panel rows
      sd[threadIdx.x] = d;
      ss[threadIdx.x] = sr;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < cnt * RB; idx += blockDim.x) {
      const int k = idx / RB, rr = idx % RB;
      const long r = r0 + rr;
      if (r < hi && !(r >= ex_lo && r < ex_hi)) vals[k][rr] = At[r + (j0 + ss[k]) * ldt];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < cnt * RB; idx += blockDim.x) {
      const int k = idx / RB, rr = idx % RB;
      const long r = r0 + rr;
      if (r < hi && !(r >= ex_lo && r < ex_hi)) At[r + (j0 + sd[k]) * ldt] = vals[k][rr];
    }
    __syncthreads();   // before the next leaf's moves overwrite s_cnt, sd, ss and vals
  }
}

template <typename T>
struct LuCtx {
  T* At;
  long ldt, m, nrows_t;
  int* ipiv;
  int* info;
  int* moves;          // LU_MV ints per panel: a count and up to 2*LU_W pairs
  long panel_no = 0;
  unsigned long long* rec;   // candidate records (LuRec)
  unsigned base = 0;
  unsigned* err;
  long sw_lo = 0, sw_hi = 0;     // At rows (columns of A) the leaf interchanges go to at once
  std::vector<LuLeaf> leaves;    // every leaf so far, in order
  hipStream_t s;
};

// rows per thread of lu_panel2_kernel
constexpr int L2_RPT = 2;

// Which panel kernel factors a leaf.  The first form (512-thread workgroups,
// one row per thread) is the faster one up to its LU_PMAX workgroups (65536
// rows): measured at N=16384 the FP32 factorization takes ~4.5 us per
// column with it against ~6.2 us with the second form.  Taller panels take
// the second form (256-thread workgroups, two rows per thread, up to
// L2_PMAX workgroups).  OPENBLAS_AMD_LU_PANEL=2 forces the second form.
static bool lu_force_panel2() {
  static const bool v = [] {
    const char* e = getenv("OPENBLAS_AMD_LU_PANEL");
    return e && atoi(e) == 2;
  }();
  return v;
}

// Workgroups of a panel kernel that the device can hold at once.  All of a
// launch must become resident together (they exchange candidates every
// column); workgroups that find no free CU start as soon as the trailing
// update's workgroups drain (that GEMM never waits for the panel), so the
// bound is the whole device's capacity.
template <typename K>
static long lu_capacity(K kernel, int threads) {
  int per_cu = 0;
  OBA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
  return (long)std::max(per_cu, 0) * rt().num_cu;
}

// The panel kernel for a leaf of mp rows (1: first form, 2: second form, 0:
// none can hold it resident) and its workgroup count *P.
template <typename T>
static int lu_panel_form(long mp, long* P) {
  constexpr int RPT = LuRpt<T>::v;
  constexpr int NT = LuNt<T>::v;
  constexpr long ROWS1 = (long)NT * RPT, ROWS2 = (long)L2_NT * L2_RPT;
  static const long cap1 = std::min<long>(LU_PMAX, lu_capacity(lu_panel_kernel<T, LU_W, RPT, NT>, NT));
  static const long cap2 = std::min<long>(L2_PMAX, lu_capacity(lu_panel2_kernel<T, LU_W, L2_RPT>, L2_NT));
  const long P1 = (mp + ROWS1 - 1) / ROWS1, P2 = (mp + ROWS2 - 1) / ROWS2;
  if (!lu_force_panel2() && P1 <= cap1) { *P = P1; return 1; }
  if (P2 <= cap2) { *P = P2; return 2; }
  *P = 0;
  return 0;
}

template <typename T>
static void lu_panel(LuCtx<T>& c, long j0, int w, bool pre = false) {
  const long mp = c.m - j0;
  int* mv = c.moves + c.panel_no * LU_MV;
  long P;
  // algorithmic work of the leaf: per column j, (mp-1-j) multipliers and a
  // rank-1 update of (mp-1-j) x (w-1-j); with `pre`, the node's W x w solve
  // and the mp x w x W update; bytes: the panel read and written once
  // (+ the W multipliers per row with `pre`)
  double lf = 0;
  for (int j = 0; j < w; j++) lf += (double)(mp - 1 - j) * (2.0 * (w - 1 - j) + 1.0);
  if (pre) lf += (double)LU_W * LU_W * w + 2.0 * mp * LU_W * w;
  {
  KTimer kt("lu_panel", c.s, lf, (double)sizeof(T) * mp * (2.0 * w + (pre ? LU_W : 0)));
  const int form = lu_panel_form<T>(mp, &P);
  if (form == 1) {
    lu_panel_kernel<T, LU_W, LuRpt<T>::v, LuNt<T>::v><<<(unsigned)P, LuNt<T>::v, 0, c.s>>>(
        c.At, c.ldt, c.m, j0, w, c.ipiv, c.info, mv, c.rec, c.base, c.err, pre ? 1 : 0);
  } else if (form == 2) {
    lu_panel2_kernel<T, LU_W, L2_RPT><<<(unsigned)P, L2_NT, 0, c.s>>>(c.At, c.ldt, c.m, j0, w, c.ipiv, c.info, mv,
                                                                      c.rec, c.base, c.err, pre ? 1 : 0);
  } else {
    fatal("getrf: more rows than the panel kernels can hold resident at once", nullptr);
  }
  OBA_HIP(hipGetLastError());
  }
  if (P > 1) c.base += (unsigned)w;   // record tags of this panel: base+1 .. base+w
  c.panel_no++;
  c.leaves.push_back(LuLeaf{j0, mv});
  lu_swap<T>(c.At, c.ldt, c.sw_lo, c.sw_hi, j0, j0 + w, j0, mp, mv, c.s);
}

template <typename T>
static void lu_rec(LuCtx<T>& c, long j0, long nc) {
  if (nc <= LU_W) { lu_panel<T>(c, j0, (int)nc); return; }
  if (nc <= 2 * LU_W) {   // two leaves: the node's TRSM + GEMM run inside the right leaf's kernel
    lu_panel<T>(c, j0, LU_W);
    lu_panel<T>(c, j0 + LU_W, (int)(nc - LU_W), true);
    return;
  }
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
  if (below > 0) {
    KTimer kt("lu_node_gemm", c.s, 2.0 * n2 * below * n1, (double)sizeof(T) * (n2 * n1 + n1 * below + 2.0 * n2 * below));
    gemm<T>(false, false, n2, below, n1, T(-1), At + (j0 + n1) + j0 * ldt, ldt, At + j0 + (j0 + n1) * ldt, ldt,
            T(1), At + (j0 + n1) + (j0 + n1) * ldt, ldt, c.s);
  }
  lu_rec<T>(c, j0 + n1, n2);
}

// Blocked right-looking LU with look-ahead, for large factorizations: the
// columns are cut into blocks of nb; block k+1 is factored (recursively, as
// above, its interchanges kept inside the block) on the caller's stream while
// the trailing update by block k -- the bulk of the flops, one TRSM and one
// large GEMM -- runs on a second stream that leaves some CUs free, so the
// latency-bound leaf panels and the throughput-bound GEMM overlap.  A
// block's interchanges reach the columns outside it after the trailing update
// that reads them has finished.  The trailing updates are sums over the
// blocks instead of over the recursion's halves: a different association of
// the same products (rounding-level differences, as with any blocking).
template <typename T>
static void lu_blocked(LuCtx<T>& c, long mn, long nb) {
  Runtime& r = rt();
  hipStream_t M = c.s, X = bulk_stream();
  T* At = c.At;
  const long ldt = c.ldt;
  DevBuf invb[2] = {DevBuf(trsm_right_inv_bytes<T>(nb)), DevBuf(trsm_right_inv_bytes<T>(nb))};
  invb[0].use_on(X);   // the trailing TRSMs read the inverses on the bulk stream
  invb[1].use_on(X);
  auto factor_block = [&](long jb, long w) {
    c.sw_lo = jb;
    c.sw_hi = jb + w;
    const size_t l0 = c.leaves.size();
    lu_rec<T>(c, jb, w);
    return l0;
  };
  size_t l0 = factor_block(0, std::min(nb, mn));
  bool pending = false;
  for (long jb = 0; jb < mn; jb += nb) {
    const long w = std::min(nb, mn - jb), jn = jb + w;
    if (pending) {   // the trailing update by the previous block has finished
      OBA_HIP(hipStreamWaitEvent(M, r.ev_b, 0));
      pending = false;
    }
    for (size_t l = l0; l < c.leaves.size(); l += LU_MAXL) {   // the block's interchanges outside it
      LuLeafSet ls;
      ls.n = (int)std::min<size_t>(LU_MAXL, c.leaves.size() - l);
      for (int i = 0; i < ls.n; i++) { ls.j0[i] = c.leaves[l + i].j0; ls.mv[i] = c.leaves[l + i].mv; }
      lu_swap_block_kernel<T, LU_W><<<(unsigned)((c.nrows_t + 63) / 64), 256, 0, M>>>(At, ldt, 0, c.nrows_t, jb, jn,
                                                                                    c.m, ls);
      OBA_HIP(hipGetLastError());
    }
    if (jn >= mn) break;
    const long w1 = std::min(nb, mn - jn);
    const long below = c.m - jn;
    T* inv = invb[(jb / nb) & 1].as<T>();
    const T* u = At + jb + jb * ldt;   // unit upper triangle of the block (At view)
    trsm_right_prepare<T>(false, false, true, w, u, ldt, inv, M);
    // next block: At rows [jn, jn+w1)
    trsm_right_prepared<T>(false, false, true, w1, w, u, ldt, At + jn + jb * ldt, ldt, inv, M);
    {
      KTimer kt("lu_next_gemm", M, 2.0 * w1 * below * w, (double)sizeof(T) * (w1 * w + w * below + 2.0 * w1 * below));
      gemm<T>(false, false, w1, below, w, T(-1), At + jn + jb * ldt, ldt, At + jb + jn * ldt, ldt, T(1),
              At + jn + jn * ldt, ldt, M);
    }
    // trailing columns [jn+w1, mn) on the bulk stream
    const long rest = mn - (jn + w1);
    if (rest > 0) {
      OBA_HIP(hipEventRecord(r.ev_a, M));
      OBA_HIP(hipStreamWaitEvent(X, r.ev_a, 0));
      T* xr = At + (jn + w1) + jb * ldt;
      trsm_right_prepared<T>(false, false, true, rest, w, u, ldt, xr, ldt, inv, X);
      {
        KTimer kt("lu_trailing_gemm", X, 2.0 * rest * below * w,
                  (double)sizeof(T) * (rest * w + w * below + 2.0 * rest * below));
        gemm<T>(false, false, rest, below, w, T(-1), xr, ldt, At + jb + jn * ldt, ldt, T(1),
                At + (jn + w1) + jn * ldt, ldt, X);
      }
      OBA_HIP(hipEventRecord(r.ev_b, X));
      pending = true;
    }
    l0 = factor_block(jn, w1);
  }
  if (pending) OBA_HIP(hipStreamWaitEvent(M, r.ev_b, 0));
}

// block width of the look-ahead driver (OPENBLAS_AMD_LU_NB; 0 turns it off);
// it is used from 4 blocks on
template <typename T>
static long lu_block_nb() {
  // swept at n = 16384 over 128..1024 (with 16..128 CUs kept free): the
  // FP32 factorization is bound by the panel chain, the FP64 one also by
  // the trailing GEMM, which prefers the deeper updates of wider blocks
  long v = sizeof(T) == 8 ? 512 : 256;
  if (const char* e = getenv("OPENBLAS_AMD_LU_NB")) v = atol(e);
  return v;
}

template <typename T>
int getrf_t(long m, long n, T* At, long ldt, int* ipiv_dev, bool track_perm, hipStream_t s) {
  const long mn = std::min(m, n);
  if (mn <= 0) return 0;
  long p_first = 0;   // the first leaf is the tallest: reject before any work is queued
  if (lu_panel_form<T>(m, &p_first) == 0)
    fatal("getrf: more rows than the panel kernels can hold resident at once", nullptr);
  const long npanels = (mn + LU_W - 1) / LU_W + 1;
  const size_t rec_bytes = sizeof(unsigned long long) * LuRec<T, LU_W>::RW * 2 * std::max(LU_PMAX, L2_PMAX);
  const size_t mv_bytes = sizeof(int) * npanels * LU_MV;
  DevBuf ws(4096 + rec_bytes + mv_bytes);
  char* p = ws.as<char>();
  LuCtx<T> c;
  c.At = At; c.ldt = ldt; c.m = m; c.nrows_t = n + (track_perm ? 1 : 0); c.ipiv = ipiv_dev; c.s = s;
  c.info = (int*)p;
  c.err = (unsigned*)(p + 64);
  c.rec = (unsigned long long*)(p + 4096);
  c.moves = (int*)(p + 4096 + rec_bytes);
  OBA_HIP(hipMemsetAsync(p, 0, 4096 + rec_bytes + mv_bytes, s));
  if (track_perm) {
    iota_row_kernel<T><<<(unsigned)((m + 255) / 256), 256, 0, s>>>(At + n, m, ldt);
    OBA_HIP(hipGetLastError());
  }
  const long nb = lu_block_nb<T>();
  if (nb >= 2 * LU_W && nb % LU_W == 0 && mn >= 4 * nb) {
    lu_blocked<T>(c, mn, nb);
  } else {
    c.sw_lo = 0;
    c.sw_hi = c.nrows_t;
    lu_rec<T>(c, 0, mn);
  }
  if (n > mn) trsm<T>(true, false, false, true, n - mn, mn, T(1), At, ldt, At + mn, ldt, s);
  int h[2] = {0, 0};
  OBA_HIP(hipMemcpyAsync(&h[0], c.info, sizeof(int), hipMemcpyDeviceToHost, s));
  OBA_HIP(hipMemcpyAsync(&h[1], c.err, sizeof(int), hipMemcpyDeviceToHost, s));
  OBA_HIP(hipStreamSynchronize(s));
  if (h[1]) fatal("getrf: panel candidate exchange timed out", nullptr);
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
void permute_rows_ws(bool inverse, long n, long nrhs, const int* perm_dev, T* b, long ldb, T* tmp, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return;
  const long total = n * nrhs;
  permute_rows_kernel<T><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(n, nrhs, perm_dev, b, ldb, tmp, n, inverse);
  OBA_HIP(hipGetLastError());
  OBA_HIP(hipMemcpy2DAsync(b, sizeof(T) * ldb, tmp, sizeof(T) * n, sizeof(T) * n, nrhs, hipMemcpyDeviceToDevice, s));
}

template <typename T>
void permute_rows(bool inverse, long n, long nrhs, const int* perm_dev, T* b, long ldb, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return;
  DevBuf tmp(sizeof(T) * (size_t)n * (size_t)nrhs);
  permute_rows_ws<T>(inverse, n, nrhs, perm_dev, b, ldb, tmp.as<T>(), s);
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
  const bool few = nrhs <= LU_TRSV_MAX_RHS;   // one launch per triangle
  if (!trans) {
    permute_rows<T>(false, n, nrhs, perm_dev, b, ldb, s);
    if (few) {
      trsv_lookback<T>(true, true, n, a, 1, lda, b, ldb, nrhs, s);
      trsv_lookback<T>(false, false, n, a, 1, lda, b, ldb, nrhs, s);
    } else {
      trsm<T>(false, true, false, true, n, nrhs, T(1), a, lda, b, ldb, s);
      trsm<T>(false, false, false, false, n, nrhs, T(1), a, lda, b, ldb, s);
    }
  } else {
    if (few) {   // U^T y = b (lower), then L^T x = y (upper, unit)
      trsv_lookback<T>(true, false, n, a, lda, 1, b, ldb, nrhs, s);
      trsv_lookback<T>(false, true, n, a, lda, 1, b, ldb, nrhs, s);
    } else {
      trsm<T>(false, false, true, false, n, nrhs, T(1), a, lda, b, ldb, s);
      trsm<T>(false, true, true, true, n, nrhs, T(1), a, lda, b, ldb, s);
    }
    permute_rows<T>(true, n, nrhs, perm_dev, b, ldb, s);
  }
}

// Same with the factors stored transposed (At = LU^T as left by getrf_t).
template <typename T>
void getrs_t(long n, long nrhs, const T* At, long ldt, const int* perm_dev, T* b, long ldb, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return;
  permute_rows<T>(false, n, nrhs, perm_dev, b, ldb, s);
  // L y = b: L = (unit upper part of At)^T ; U x = y: U = (lower part of At)^T
  if (nrhs <= LU_TRSV_MAX_RHS) {   // L(i,k) = At[k + i*ldt], U(i,k) = At[k + i*ldt]
    trsv_lookback<T>(true, true, n, At, ldt, 1, b, ldb, nrhs, s);
    trsv_lookback<T>(false, false, n, At, ldt, 1, b, ldb, nrhs, s);
    return;
  }
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
template void permute_rows_ws<double>(bool, long, long, const int*, double*, long, double*, hipStream_t);
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
