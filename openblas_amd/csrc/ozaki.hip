// FP64 GEMM emulated with INT8 slices on the INT8 MFMA (Ozaki-style
// splitting): no counterpart in the reference; it computes the same
// C <- alpha*op(A)*op(B) + beta*C as cblas_dgemm (interface/gemm.c:294).
//
// Each row i of op(A) is scaled by 2^-e_i (e_i from its largest magnitude)
// and cut into s signed 7-bit digits a_i^(t) (t = 0..s-1, truncation, so
// every digit lies in [-127, 127]); every column j of op(B) likewise with
// 2^-f_j and digits b_j^(u).  Then
//   C_ij = 2^(e_i+f_j) * sum_L 2^(-7(L+2)) * sum_{t+u=L} sum_k a_ik^(t) b_kj^(u)
// The inner sums are INT8 GEMMs with exact int32 accumulation
// (v_mfma_i32_16x16x64_i8); all pairs of one level L = t+u share an int32
// accumulator (exact while (L+1)*K*127^2 < 2^31, K is chunked otherwise),
// levels L >= s are dropped (the truncation error of the scheme), and the
// levels are combined in FP64 from the least significant up.
//
// Kernels: per-vector statistics (max |x| and sum |x|, coalesced in either
// layout), the split into k-contiguous int8 slices (64x64 tiles through
// LDS), and one fused GEMM kernel: a workgroup owns a 64x64 tile of C, four
// wavefronts own 32x32 each, all s slices of a 64-deep k slab of both
// operands sit in LDS (two slabs for s <= 9), every level's int32 accumulators and the FP64
// accumulator stay in registers, and C is read and written once.
#include "kernels.h"
#include "ozaki.h"
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

namespace oba {

typedef int i4v __attribute__((ext_vector_type(4)));
constexpr int OZ_SMAX = 12;     // slices supported by the fused kernel
constexpr int OZ_NONFINITE = -2;   // return code: NaN/Inf operand, native GEMM used
constexpr int OZ_BK = 64;       // k per slab (one i8 MFMA k-depth)
constexpr int OZ_T = 64;        // C tile edge

// Per-vector statistics: the largest |x| (it sets the vector's scale) and,
// over the nonzero entries, the sum of log2|x| and their count (their
// geometric mean says how far below the scale the typical entry sits, which
// is what the slice count has to cover).
// log2|x| to about 0.09 (exponent plus the linear mantissa term): enough for
// a slice count, and no transcendental call per element
__device__ inline double oz_log2_approx(double a) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(a);
  const int ex = (int)((b >> 52) & 0x7ff) - 1023;
  const double m = (double)(b & 0xFFFFFFFFFFFFFull) * 0x1p-52;
  return (double)ex + m;
}

// element (v, k) of an operand "vector" v at X[v*sv + k*sk]
// sv == 1: the k range is cut into gridDim.y chunks (one vector per thread
// and chunk, loads coalesced across the vectors); partial statistics per
// chunk, combined in chunk order by oz_stats_combine (deterministic)
// A NaN or infinity anywhere raises *nonfinite: the slices cannot carry it
// (the host then runs the native FP64 GEMM, which propagates it).
__global__ void oz_stats_vcontig(long V, long K, const double* __restrict__ X, long sk, double* pmax, double* plog,
                                 double* pcnt, int* nonfinite) {
  const long v = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (v >= V) return;
  const long kc = (K + gridDim.y - 1) / gridDim.y, k0 = blockIdx.y * kc, k1 = min(K, k0 + kc);
  double mx = 0.0, lg = 0.0, cn = 0.0;
  bool bad = false;
  for (long k = k0; k < k1; k++) {
    double a = fabs(X[v + k * sk]);
    bad |= !(a <= DBL_MAX);
    mx = a > mx ? a : mx;
    if (a > 0.0) { lg += oz_log2_approx(a); cn += 1.0; }
  }
  if (bad) atomicOr(nonfinite, 1);
  pmax[blockIdx.y * V + v] = mx;
  plog[blockIdx.y * V + v] = lg;
  pcnt[blockIdx.y * V + v] = cn;
}

__global__ void oz_stats_combine(long V, int parts, const double* __restrict__ pmax, const double* __restrict__ plog,
                                 const double* __restrict__ pcnt, double* vmax, double* vlog, double* vcnt) {
  const long v = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (v >= V) return;
  double mx = 0.0, lg = 0.0, cn = 0.0;
  for (int p = 0; p < parts; p++) {
    const double a = pmax[p * V + v];
    mx = a > mx ? a : mx;
    lg += plog[p * V + v];
    cn += pcnt[p * V + v];
  }
  vmax[v] = mx;
  vlog[v] = lg;
  vcnt[v] = cn;
}

__global__ void oz_stats_kcontig(long V, long K, const double* __restrict__ X, long sv, double* vmax, double* vlog,
                                 double* vcnt, int* nonfinite) {
  const long v = blockIdx.x * (long)(blockDim.x >> 6) + (threadIdx.x >> 6);   // sk == 1, one wave per vector
  const int lane = threadIdx.x & 63;
  if (v >= V) return;
  double mx = 0.0, lg = 0.0, cn = 0.0;
  bool bad = false;
  for (long k = lane; k < K; k += 64) {
    double a = fabs(X[v * sv + k]);
    bad |= !(a <= DBL_MAX);
    mx = a > mx ? a : mx;
    if (a > 0.0) { lg += oz_log2_approx(a); cn += 1.0; }
  }
  if (bad) atomicOr(nonfinite, 1);
  for (int off = 32; off > 0; off >>= 1) {
    mx = fmax(mx, __shfl_xor(mx, off));
    lg += __shfl_xor(lg, off);
    cn += __shfl_xor(cn, off);
  }
  if (lane == 0) { vmax[v] = mx; vlog[v] = lg; vcnt[v] = cn; }
}

// exponent e with max*2^-e < 1 (max = f*2^e, f in [0.5,1)); 0 for max == 0
__device__ inline int oz_exp(double mx) {
  int e = 0;
  if (mx > 0.0) frexp(mx, &e);
  return e;
}

// Splits a 64(v) x 64(k) tile into s slices S[t][v][k] (k contiguous, leading
// dimension Kp, Vp vectors per slice), zero outside V x K.
__global__ __launch_bounds__(256) void oz_split_kernel(long V, long K, const double* __restrict__ X, long sv, long sk,
                                                       const double* __restrict__ vmax, int s,
                                                       signed char* __restrict__ S, long Vp, long Kp) {
  __shared__ double tile[64][65];
  const long v0 = (long)blockIdx.x * 64, k0 = (long)blockIdx.y * 64;
  const bool vfast = (sv == 1);
  // all 16 loads of a thread in flight together, then the LDS stores
  double ld[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int idx = (int)threadIdx.x + i * 256;
    const int f = idx & 63, g = idx >> 6;   // f: along the contiguous index
    const int vv = vfast ? f : g, kk = vfast ? g : f;
    const long v = v0 + vv, k = k0 + kk;
    ld[i] = (v < V && k < K) ? X[v * sv + k * sk] : 0.0;
  }
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int idx = (int)threadIdx.x + i * 256;
    const int f = idx & 63, g = idx >> 6;
    tile[vfast ? f : g][vfast ? g : f] = ld[i];
  }
  __syncthreads();
  // each thread: one vector row of the tile, 16 consecutive k -> s x 16 bytes
  const int vv = threadIdx.x >> 2, kq = (threadIdx.x & 3) * 16;
  const long v = v0 + vv;
  const int e = (v < V) ? oz_exp(vmax[v]) : 0;
  double r[16];
#pragma unroll
  for (int c = 0; c < 16; c++) r[c] = ldexp(tile[vv][kq + c], -e);
  // digit t of every element, four digits per 32-bit word (byte c of the
  // 16-byte store = element c); static indices keep r and the words in
  // registers
#pragma unroll
  for (int t = 0; t < OZ_SMAX; t++) {
    if (t < s) {
      i4v w;
#pragma unroll
      for (int q4 = 0; q4 < 4; q4++) {
        unsigned word = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) {
          const int c = q4 * 4 + b;
          r[c] *= 128.0;
          const double q = trunc(r[c]);
          r[c] -= q;
          word |= ((unsigned)(int)q & 0xFFu) << (8 * b);
        }
        w[q4] = (int)word;
      }
      *(i4v*)(S + ((size_t)t * Vp + v) * Kp + k0 + kq) = w;
    }
  }
}

// LDS byte offset of 16-byte chunk c (0..3) of row r of a slab tile.  With
// two slab buffers (SW) rows are 64 bytes and the chunks are XOR-swizzled by
// row bits 1 and 3 (stores: 4 lanes per row; MFMA fragment reads: 16 rows x
// one chunk per 16 lanes; both touch every bank once per LDS cycle); with
// one buffer rows are padded to 80 bytes.
template <bool SW>
__device__ __attribute__((always_inline)) inline int oz_lds_off(int r, int c) {
  return SW ? r * 64 + ((c ^ (((r >> 1) ^ (r >> 3)) & 3)) << 4) : r * 80 + (c << 4);
}

// The fused slice GEMM.  LDS holds all s slices of a 64-deep slab for the
// tile's 64 rows of op(A) and 64 columns of op(B); for s <= 9 there are two
// such buffers (147 KB), so the next slab is stored while this one is read
// and each slab needs one barrier instead of two.
template <int NS>
__global__ __launch_bounds__(256) void oz_gemm_kernel(long M, long N, long K, long Kp, const signed char* __restrict__ SA,
                                                      long Mp, const signed char* __restrict__ SB, long Np,
                                                      const double* __restrict__ amax, const double* __restrict__ bmax,
                                                      double alpha, double beta, double* __restrict__ C, long ldc,
                                                      long kchunk) {
  constexpr bool DB = NS <= 9;
  constexpr int NBUF = DB ? 2 : 1, RS = DB ? 64 : 80;
  __shared__ __attribute__((aligned(16))) signed char la[NBUF][NS][OZ_T * RS];
  __shared__ __attribute__((aligned(16))) signed char lb[NBUF][NS][OZ_T * RS];
  // XCD-aware grouped tile order (workgroup b runs on XCD b%8): each XCD
  // takes a contiguous range of an order that walks 8 tile rows at a time,
  // so the slices of an A row block and a B column block are reused from
  // that XCD's L2 instead of being streamed from HBM once per tile
  const long tiles_m = (M + OZ_T - 1) / OZ_T, tiles_n = (N + OZ_T - 1) / OZ_T;
  const long T_all = tiles_m * tiles_n;
  long lin = blockIdx.x;
  if ((T_all & 7) == 0) lin = (long)(blockIdx.x & 7) * (T_all >> 3) + (blockIdx.x >> 3);
  constexpr long GM = 8;
  const long group = lin / (GM * tiles_n), first_m = group * GM;
  const long gsize = min(tiles_m - first_m, GM), in_group = lin - group * GM * tiles_n;
  const long tm = first_m + in_group % gsize, tn = in_group / gsize;
  const long m0 = tm * OZ_T, n0 = tn * OZ_T;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;

  i4v acc[NS][2][2];
  double dacc[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) dacc[a][b][r] = 0.0;

  for (long kc0 = 0; kc0 < K; kc0 += kchunk) {
    const long kc1 = min(K, kc0 + kchunk);
#pragma unroll
    for (int L = 0; L < NS; L++)
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[L][a][b] = i4v{0, 0, 0, 0};
    // software pipeline: the next slab's global loads are in flight in
    // registers while the current slab is multiplied out of LDS
    // thread's share of a slab: row tid>>2, bytes (tid&3)*16.. of every slice
    const int frow = threadIdx.x >> 2, fq = (threadIdx.x & 3) * 16;
    const signed char* ga = SA + ((size_t)m0 + frow) * Kp + fq;
    const signed char* gb = SB + ((size_t)n0 + frow) * Kp + fq;
    i4v pa[NS], pb[NS];
#pragma unroll
    for (int t = 0; t < NS; t++) {
      pa[t] = *(const i4v*)(ga + (size_t)t * Mp * Kp + kc0);
      pb[t] = *(const i4v*)(gb + (size_t)t * Np * Kp + kc0);
    }
    // stage a slab: NS slices x 64 rows x 64 bytes for each operand
    const int soff = oz_lds_off<DB>(frow, fq >> 4);
    int buf = 0;
    if (DB) {
#pragma unroll
      for (int t = 0; t < NS; t++) {
        *(i4v*)(&la[0][t][soff]) = pa[t];
        *(i4v*)(&lb[0][t][soff]) = pb[t];
      }
      __syncthreads();
    }
    for (long k0 = kc0; k0 < kc1; k0 += OZ_BK) {
      if (!DB) {
#pragma unroll
        for (int t = 0; t < NS; t++) {
          *(i4v*)(&la[0][t][soff]) = pa[t];
          *(i4v*)(&lb[0][t][soff]) = pb[t];
        }
        __syncthreads();
      }
      if (k0 + OZ_BK < kc1) {
#pragma unroll
        for (int t = 0; t < NS; t++) {
          pa[t] = *(const i4v*)(ga + (size_t)t * Mp * Kp + k0 + OZ_BK);
          pb[t] = *(const i4v*)(gb + (size_t)t * Np * Kp + k0 + OZ_BK);
        }
      }
      // fragments: lane l holds row (l%16) of a 16-row block, k bytes (l/16)*16..+15
      i4v fb[NS][2];
#pragma unroll
      for (int u = 0; u < NS; u++)
#pragma unroll
        for (int b = 0; b < 2; b++) fb[u][b] = *(const i4v*)(&lb[buf][u][oz_lds_off<DB>(wn + b * 16 + (lane & 15), lane >> 4)]);
#pragma unroll
      for (int t = 0; t < NS; t++) {
        i4v fa[2];
#pragma unroll
        for (int a = 0; a < 2; a++) fa[a] = *(const i4v*)(&la[buf][t][oz_lds_off<DB>(wm + a * 16 + (lane & 15), lane >> 4)]);
#pragma unroll
        for (int u = 0; u < NS - t; u++)
#pragma unroll
          for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
              acc[t + u][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fb[u][b], fa[a], acc[t + u][a][b], 0, 0, 0);
      }
      if (DB) {
        // the other buffer was last read before the previous barrier
        if (k0 + OZ_BK < kc1) {
#pragma unroll
          for (int t = 0; t < NS; t++) {
            *(i4v*)(&la[buf ^ 1][t][soff]) = pa[t];
            *(i4v*)(&lb[buf ^ 1][t][soff]) = pb[t];
          }
        }
        buf ^= 1;
      }
      __syncthreads();
    }
    // fold the exact int32 levels into FP64, least significant first
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          double sum = 0.0;
#pragma unroll
          for (int L = NS - 1; L >= 0; L--) sum += ldexp((double)acc[L][a][b][r], -7 * (L + 2));
          dacc[a][b][r] += sum;
        }
  }
  // D[j][i]: lane l, element r -> column j = 4*(l/16)+r of the 16-block, row i = l%16
  const bool bz = beta == 0.0;
#pragma unroll
  for (int a = 0; a < 2; a++) {
    const long i = m0 + wm + a * 16 + (lane & 15);
    if (i >= M) continue;
    const int ei = oz_exp(amax[i]);
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const long j = n0 + wn + b * 16 + 4 * (lane >> 4) + r;
        if (j >= N) continue;
        const int fj = oz_exp(bmax[j]);
        double v = alpha * ldexp(dacc[a][b][r], ei + fj);
        double* p = C + i + j * ldc;
        if (!bz) v += beta * *p;
        *p = v;
      }
  }
}

static void oz_stats(long V, long K, const double* X, long sv, long sk, double* vmax, double* vlog, double* vcnt,
                     int* nonfinite, hipStream_t s) {
  if (V <= 0) return;
  if (sv == 1) {
    // enough workgroups to fill the device: ~1024 (vector blocks x k chunks)
    const long vb = (V + 255) / 256;
    const int parts = (int)std::max(1L, std::min((K + 63) / 64, 1024 / vb));
    DevBuf part(sizeof(double) * 3 * (size_t)parts * V);
    double* pmax = part.as<double>();
    double* plog = pmax + (size_t)parts * V;
    double* pcnt = plog + (size_t)parts * V;
    oz_stats_vcontig<<<dim3((unsigned)vb, (unsigned)parts), 256, 0, s>>>(V, K, X, sk, pmax, plog, pcnt, nonfinite);
    OBA_HIP(hipGetLastError());
    oz_stats_combine<<<(unsigned)vb, 256, 0, s>>>(V, parts, pmax, plog, pcnt, vmax, vlog, vcnt);
  } else {
    oz_stats_kcontig<<<(unsigned)((V + 3) / 4), 256, 0, s>>>(V, K, X, sv, vmax, vlog, vcnt, nonfinite);
  }
  OBA_HIP(hipGetLastError());
}

int ozaki_slices_auto(long M, long N, long K, const double* amax, const double* alog, const double* acnt,
                      const double* bmax, const double* blog, const double* bcnt) {
  // Spread of a row (column): log2 of its largest |x| over the geometric
  // mean of its nonzero |x|.  The digits are cut at a fixed distance below
  // the vector's largest entry, so an entry that sits d bits below it keeps
  // d bits fewer; the typical entry of a row times the typical entry of a
  // column is what C's components are made of; 50 bits of it are kept,
  // which holds the result at the accuracy of a native FP64 product's own
  // rounding (8192^3, entries u*exp(phi*g): phi = 0.5 -> 9 slices, error
  // 7e-17 against native 5.5e-16; phi = 2 -> 12 slices, 3e-15 against
  // 1.2e-14; phi = 8 needs 26 and takes the native path).  The arithmetic
  // mean is dominated by the largest entries of a heavy-tailed vector and
  // underestimates the spread (it gave phi = 8 data 12 slices and a 2e-9
  // result).
  (void)K;
  auto spread = [](long V, const double* mx, const double* lg, const double* cn) {
    double worst = 0.0;
    for (long v = 0; v < V; v++) {
      if (mx[v] <= 0.0 || cn[v] <= 0.0) continue;
      const double r = std::log2(mx[v]) - lg[v] / cn[v];
      if (r > worst) worst = r;
    }
    return worst;
  };
  const double bits = 50.0 + spread(M, amax, alog, acnt) + spread(N, bmax, blog, bcnt);
  const double s = std::ceil(bits / 7.0);
  return (int)std::max(2.0, std::min(s, 1000.0));   // the required count, not clamped to OZ_SMAX
}

int ozaki_dgemm(bool ta, bool tb, long M, long N, long K, double alpha, const double* A, long lda, const double* B,
                long ldb, double beta, double* C, long ldc, int slices, hipStream_t s) {
  if (M <= 0 || N <= 0) return 0;
  if (K <= 0 || alpha == 0.0) { scale<double>(M, N, beta, C, ldc, s); return 0; }
  // op(A) row i, element k: A[i + k*lda] (N) or A[k + i*lda] (T)
  const long a_sv = ta ? lda : 1, a_sk = ta ? 1 : lda;
  // op(B) column j, element k: B[k + j*ldb] (N) or B[j + k*ldb] (T)
  const long b_sv = tb ? 1 : ldb, b_sk = tb ? ldb : 1;
  // statistics: amax, alog, acnt (M each), then bmax, blog, bcnt (N each)
  DevBuf st(sizeof(double) * 3 * (M + N) + 256);
  double* amax = st.as<double>();
  double* alog = amax + M;
  double* acnt = alog + M;
  double* bmax = acnt + M;
  double* blog = bmax + N;
  double* bcnt = blog + N;
  int* nonfinite = (int*)(bcnt + N);
  OBA_HIP(hipMemsetAsync(nonfinite, 0, sizeof(int), s));
  oz_stats(M, K, A, a_sv, a_sk, amax, alog, acnt, nonfinite, s);
  oz_stats(N, K, B, b_sv, b_sk, bmax, blog, bcnt, nonfinite, s);
  int ns = slices;
  int h_nonfinite = 0;
  std::vector<double> h(ns <= 0 ? 3 * (M + N) : 0);
  OBA_HIP(hipMemcpyAsync(&h_nonfinite, nonfinite, sizeof(int), hipMemcpyDeviceToHost, s));
  if (ns <= 0) OBA_HIP(hipMemcpyAsync(h.data(), amax, sizeof(double) * 3 * (M + N), hipMemcpyDeviceToHost, s));
  OBA_HIP(hipStreamSynchronize(s));
  if (h_nonfinite) {
    // NaN / Inf: the native FP64 GEMM gives cblas_dgemm's propagation
    gemm<double>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
    return OZ_NONFINITE;
  }
  if (ns <= 0) {
    const double* hb = h.data() + 3 * M;
    ns = ozaki_slices_auto(M, N, K, h.data(), h.data() + M, h.data() + 2 * M, hb, hb + N, hb + 2 * N);
    if (ns > OZ_SMAX) {
      // the exponent spread needs more digits than the fused kernel holds: a
      // 12-slice result would miss FP64 accuracy, so the native FP64 GEMM runs
      gemm<double>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
      return -ns;
    }
  }
  if (ns < 1 || ns > OZ_SMAX) fatal("ozaki: slice count out of range", nullptr);
  const long Kp = (K + OZ_BK - 1) / OZ_BK * OZ_BK;
  const long Mp = (M + OZ_T - 1) / OZ_T * OZ_T, Np = (N + OZ_T - 1) / OZ_T * OZ_T;
  DevBuf sa((size_t)ns * Mp * Kp), sb((size_t)ns * Np * Kp);
  {
    dim3 ga((unsigned)(Mp / 64), (unsigned)(Kp / 64)), gb((unsigned)(Np / 64), (unsigned)(Kp / 64));
    oz_split_kernel<<<ga, 256, 0, s>>>(M, K, A, a_sv, a_sk, amax, ns, sa.as<signed char>(), Mp, Kp);
    oz_split_kernel<<<gb, 256, 0, s>>>(N, K, B, b_sv, b_sk, bmax, ns, sb.as<signed char>(), Np, Kp);
    OBA_HIP(hipGetLastError());
  }
  // exact int32 accumulation: (pairs per level) * kchunk * 127^2 < 2^31
  long kchunk = (long)(2147483647.0 / ((double)ns * 127.0 * 127.0)) / OZ_BK * OZ_BK;
  if (kchunk < OZ_BK) kchunk = OZ_BK;
  const long tiles = (Mp / OZ_T) * (Np / OZ_T);
  if (tiles > 0x7fffffffL) fatal("ozaki: grid too large", nullptr);
#define OZ_L(NSV)                                                                                                  \
  case NSV:                                                                                                        \
    oz_gemm_kernel<NSV><<<(unsigned)tiles, 256, 0, s>>>(M, N, K, Kp, sa.as<signed char>(), Mp, sb.as<signed char>(), \
                                                        Np, amax, bmax, alpha, beta, C, ldc, kchunk);             \
    break;
  // algorithmic work: s(s+1)/2 slice products of 2MNK int8 ops each; bytes:
  // the slices of both operands (s bytes per element) and C read + written
  KTimer kt("ozaki_gemm", s, (double)ns * (ns + 1) / 2 * 2.0 * M * N * K,
            (double)ns * (M * K + K * N) + 8.0 * M * N * (beta == 0.0 ? 1 : 2));
  switch (ns) {
    OZ_L(1) OZ_L(2) OZ_L(3) OZ_L(4) OZ_L(5) OZ_L(6) OZ_L(7) OZ_L(8) OZ_L(9) OZ_L(10) OZ_L(11) OZ_L(12)
  }
#undef OZ_L
  OBA_HIP(hipGetLastError());
  return ns;
}

// ---- Scheme II: integer moduli and the Chinese remainder theorem -----------
//
// The published Ozaki scheme II (Ozaki, Uchino, Imamura, 2025): row i of
// op(A) is scaled to integers a'_ik = trunc(a_ik 2^(tA-e_i)) (|a'| < 2^tA,
// e_i the exponent of the row's largest |a|), column j of op(B) likewise
// with tB and f_j.  For NM pairwise coprime moduli p_m <= 256 the residues
// a'_ik mod p_m, b'_kj mod p_m are int8 (balanced, in [-128, 127]) and
// C'_m = A'_m B'_m is an INT8 GEMM with exact int32 accumulation (K * 128^2
// < 2^31 per launch); its value mod p_m is C' = A'B' mod p_m.  With P =
// prod p_m > 2 K 2^(tA+tB) > 2 |C'|, the Chinese remainder theorem gives C'
// exactly: C' = (sum_m c_m Q_m mod P) in the balanced range, Q_m = (P/p_m)
// ((P/p_m)^-1 mod p_m).  The sum is formed in 32-bit limbs, rounded once to
// FP64 and scaled by 2^(e_i+f_j-tA-tB).  The only error is the truncation
// of the operands to tA and tB bits (below each vector's largest entry);
// NM INT8 GEMMs replace the s(s+1)/2 of the slice scheme (FP64 accuracy:
// ~17 against 45).
constexpr int CRT_NMAX = 30;    // moduli in the table
constexpr int CRT_LIMBS = 8;    // 32-bit limbs of the sums: S < NM * 256 * P < 2^256
constexpr int CRT_T = 256;      // operand padding: a multiple of every GEMM tile edge
constexpr long CRT_KCHUNK = 131008;   // k per GEMM launch: 131008 * 128^2 < 2^31
constexpr double CRT_KEEP = 53.0;     // bits kept below the typical entry of each operand
// pairwise coprime, largest first (greedy from 256 down)
static const int kCrtModuli[CRT_NMAX] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197,
                                         193, 191, 181, 179, 173, 167, 163, 157, 151, 149, 139, 137, 131, 127, 113};

constexpr int CRT_PIECES = 14;   // 18-bit pieces of a scaled operand: t <= 252 bits
constexpr int CRT_L16 = 2 * CRT_LIMBS;   // 16-bit limbs of the CRT sums
struct CrtTab {
  int p[CRT_NMAX];
  double invp[CRT_NMAX];
  float invpf[CRT_NMAX];
  int c18[CRT_NMAX][CRT_PIECES];                     // 2^(18 i) mod p
  unsigned q[CRT_NMAX + 1][CRT_NMAX][CRT_LIMBS];     // Q_m for the first NM moduli (32-bit limbs)
  unsigned short q16[CRT_NMAX + 1][CRT_NMAX][CRT_L16];   // the same in 16-bit limbs
  unsigned short P16[CRT_NMAX + 1][CRT_L16];         // product of the first NM moduli
  unsigned short Ph16[CRT_NMAX + 1][CRT_L16];        // floor(P / 2)
  double Pd[CRT_NMAX + 1];                           // P rounded to FP64
  double log2P[CRT_NMAX + 1];
};

// host bignums: little-endian 32-bit limbs
typedef std::vector<unsigned> Big;
static void big_mul_small(Big& v, unsigned x) {
  unsigned long long c = 0;
  for (auto& l : v) { c += (unsigned long long)l * x; l = (unsigned)c; c >>= 32; }
  if (c) v.push_back((unsigned)c);
}
static unsigned big_divmod_small(Big& v, unsigned x) {   // v /= x, returns v mod x
  unsigned long long r = 0;
  for (size_t i = v.size(); i-- > 0;) {
    const unsigned long long cur = (r << 32) | v[i];
    v[i] = (unsigned)(cur / x);
    r = cur % x;
  }
  while (v.size() > 1 && v.back() == 0) v.pop_back();
  return (unsigned)r;
}
static long inv_mod(long a, long m) {   // a^-1 mod m (gcd 1)
  long t = 0, nt = 1, r = m, nr = ((a % m) + m) % m;
  while (nr) { const long q = r / nr; long tmp = t - q * nt; t = nt; nt = tmp; tmp = r - q * nr; r = nr; nr = tmp; }
  return t < 0 ? t + m : t;
}

static const CrtTab& crt_host_tab() {
  static CrtTab T;
  static bool built = false;
  if (built) return T;
  std::memset(static_cast<void*>(&T), 0, sizeof T);
  for (int m = 0; m < CRT_NMAX; m++) {
    const int p = kCrtModuli[m];
    T.p[m] = p;
    T.invp[m] = 1.0 / p;
    T.invpf[m] = 1.0f / (float)p;
    long v = 1 % p;
    for (int i = 0; i < CRT_PIECES; i++) { T.c18[m][i] = (int)v; v = (v * (1L << 18)) % p; }
  }
  for (int nm = 1; nm <= CRT_NMAX; nm++) {
    Big P{1u};
    for (int m = 0; m < nm; m++) big_mul_small(P, (unsigned)kCrtModuli[m]);
    if (P.size() > CRT_LIMBS) fatal("ozaki2: modulus product exceeds the limb count", nullptr);
    double pd = 0.0;
    for (size_t l = P.size(); l-- > 0;) pd = pd * 4294967296.0 + P[l];
    T.Pd[nm] = pd;
    T.log2P[nm] = std::log2(pd);
    for (size_t l = 0; l < P.size(); l++) { T.P16[nm][2 * l] = (unsigned short)P[l]; T.P16[nm][2 * l + 1] = (unsigned short)(P[l] >> 16); }
    Big h = P;
    unsigned carry = 0;
    for (size_t l = h.size(); l-- > 0;) { const unsigned nc = h[l] & 1u; h[l] = (h[l] >> 1) | (carry << 31); carry = nc; }
    for (size_t l = 0; l < h.size(); l++) { T.Ph16[nm][2 * l] = (unsigned short)h[l]; T.Ph16[nm][2 * l + 1] = (unsigned short)(h[l] >> 16); }
    for (int m = 0; m < nm; m++) {
      Big Mm = P;
      big_divmod_small(Mm, (unsigned)kCrtModuli[m]);    // P / p_m
      Big t = Mm;
      const unsigned r = big_divmod_small(t, (unsigned)kCrtModuli[m]);   // (P/p_m) mod p_m
      const long inv = inv_mod(r, kCrtModuli[m]);
      Big q = Mm;
      big_mul_small(q, (unsigned)inv);   // < P: inv < p_m
      if (q.size() > CRT_LIMBS) fatal("ozaki2: CRT weight exceeds the limb count", nullptr);
      for (size_t l = 0; l < q.size(); l++) {
        T.q[nm][m][l] = q[l];
        T.q16[nm][m][2 * l] = (unsigned short)q[l];
        T.q16[nm][m][2 * l + 1] = (unsigned short)(q[l] >> 16);
      }
    }
  }
  built = true;
  return T;
}

// the table in device memory, uploaded once per process
static const CrtTab* crt_dev_tab(hipStream_t s) {
  static void* d = nullptr;
  if (!d) {
    const CrtTab& h = crt_host_tab();
    OBA_HIP(hipMalloc(&d, sizeof(CrtTab)));
    OBA_HIP(hipMemcpyAsync(d, &h, sizeof(CrtTab), hipMemcpyHostToDevice, s));
    OBA_HIP(hipStreamSynchronize(s));
  }
  return (const CrtTab*)d;
}

// Residues of the scaled operand: a 64(v) x 64(k) tile through LDS, then
// every thread takes 16 consecutive k of one vector and writes their
// balanced residues for each of the NM moduli (R[m][v][k], k contiguous).
// x' = trunc(|x| 2^(t-e_v)) < 2^t is cut into 18-bit pieces x_i (exact FP64
// steps: the scaled double is exact, floor of it exact, and the difference
// of two doubles within 2^18 of each other exact), once per element; per
// modulus x' mod p = (sum_i x_i (2^(18 i) mod p)) mod p, a sum of 24-bit
// integer products below 2^30, reduced with an FP32 quotient estimate and
// one correction step each way.
template <int NP>
__global__ __launch_bounds__(256) void crt_split_kernel(long V, long K, const double* __restrict__ X, long sv, long sk,
                                                        const double* __restrict__ vmax, int t, int NM,
                                                        const CrtTab* __restrict__ tab, signed char* __restrict__ R,
                                                        long Vp, long Kp) {
  __shared__ double tile[64][65];
  const long v0 = (long)blockIdx.x * 64, k0 = (long)blockIdx.y * 64;
  const bool vfast = (sv == 1);
  double ld[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int idx = (int)threadIdx.x + i * 256;
    const int f = idx & 63, g = idx >> 6;
    const int vv = vfast ? f : g, kk = vfast ? g : f;
    const long v = v0 + vv, k = k0 + kk;
    ld[i] = (v < V && k < K) ? X[v * sv + k * sk] : 0.0;
  }
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int idx = (int)threadIdx.x + i * 256;
    const int f = idx & 63, g = idx >> 6;
    tile[vfast ? f : g][vfast ? g : f] = ld[i];
  }
  __syncthreads();
  const int vv = threadIdx.x >> 2, kq = (threadIdx.x & 3) * 16;
  const long v = v0 + vv;
  const int e = (v < V) ? oz_exp(vmax[v]) : 0;
  unsigned pc[16][NP];   // 18-bit pieces, least significant first
  unsigned neg = 0;
#pragma unroll
  for (int c = 0; c < 16; c++) {
    double y = ldexp(tile[vv][kq + c], t - e);
    if (y < 0.0) { neg |= 1u << c; y = -y; }
    y = trunc(y);
    // from the top piece down: hi = floor(y / 2^(18 i)), piece = hi - 2^18 floor(hi / 2^18)
#pragma unroll
    for (int i = NP - 1; i >= 0; i--) {
      const double hi = floor(ldexp(y, -18 * i));
      const double up = floor(hi * 0x1p-18);
      pc[c][i] = (unsigned)(hi - up * 0x1p18);
    }
  }
  for (int m = 0; m < NM; m++) {
    const int p = tab->p[m];
    const float ipf = tab->invpf[m];
    const int half = (p + 1) / 2;
    int cm[NP];
#pragma unroll
    for (int i = 0; i < NP; i++) cm[i] = tab->c18[m][i];
    i4v w;
#pragma unroll
    for (int q4 = 0; q4 < 4; q4++) {
      unsigned word = 0;
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int c = q4 * 4 + b;
        unsigned a = 0;
#pragma unroll
        for (int i = 0; i < NP; i++) a = __umul24(pc[c][i], (unsigned)cm[i]) + a;   // < NP 2^26
        // the FP32 quotient is within 2 of the true one for a < 2^30
        int r = (int)a - (int)floorf((float)a * ipf) * p;
        if (r < 0) r += p;
        if (r < 0) r += p;
        if (r >= p) r -= p;
        if (r >= p) r -= p;
        if (r >= half) r -= p;   // balanced: [-(p-1)/2, (p-1)/2], [-128, 127] for 256
        if ((neg >> c) & 1u) r = -r;
        if (r == 128) r = -128;   // p = 256: 128 = -128 (mod 256)
        word |= ((unsigned)r & 0xFFu) << (8 * b);
      }
      w[q4] = (int)word;
    }
    *(i4v*)(R + ((size_t)m * Vp + v) * Kp + k0 + kq) = w;
  }
}

// One INT8 GEMM per modulus (blockIdx.y): a (32 FA) x (32 FB) tile of the
// residue product, four wavefronts of (16 FA) x (16 FB) (FA x FB MFMA
// blocks each), 64-deep k slabs double-buffered in LDS (the slice kernel's
// swizzled layout), the next slab's loads in flight during the MFMAs.  The
// exact int32 sums are reduced mod p_m and stored as balanced int8
// residues of C (RC[m][j][i], i contiguous); with `accumulate` the launch
// adds to the residues of the previous k chunk.
template <int FA, int FB>
__global__ __launch_bounds__(256) void crt_gemm_kernel(long M, long N, long Kc, long Kp, long k_off,
                                                       const signed char* __restrict__ RA, long Mp,
                                                       const signed char* __restrict__ RB, long Np,
                                                       const CrtTab* __restrict__ tab, signed char* __restrict__ RC,
                                                       long ldrc, long plane, int accumulate) {
  constexpr int BM = 32 * FA, BN = 32 * FB;
  constexpr int LA = BM / 64, LB = BN / 64;   // 16-byte loads per thread and slab
  __shared__ __attribute__((aligned(16))) signed char la[2][BM * 64];
  __shared__ __attribute__((aligned(16))) signed char lb[2][BN * 64];
  const int mod = blockIdx.y;
  RA += (size_t)mod * Mp * Kp + k_off;
  RB += (size_t)mod * Np * Kp + k_off;
  const long tiles_m = Mp / BM, tiles_n = Np / BN;
  const long T_all = tiles_m * tiles_n;
  long lin = blockIdx.x;
  if ((T_all & 7) == 0) lin = (long)(blockIdx.x & 7) * (T_all >> 3) + (blockIdx.x >> 3);
  constexpr long GM = 8;
  const long group = lin / (GM * tiles_n), first_m = group * GM;
  const long gsize = min(tiles_m - first_m, GM), in_group = lin - group * GM * tiles_n;
  const long tm = first_m + in_group % gsize, tn = in_group / gsize;
  const long m0 = tm * BM, n0 = tn * BN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BN / 2);

  i4v acc[FA][FB];
#pragma unroll
  for (int a = 0; a < FA; a++)
#pragma unroll
    for (int b = 0; b < FB; b++) acc[a][b] = i4v{0, 0, 0, 0};
  // thread's share of a slab: rows frow + 64 q, bytes fq..fq+15
  const int frow = threadIdx.x >> 2, fq = (threadIdx.x & 3) * 16;
  const signed char* ga = RA + ((size_t)m0 + frow) * Kp + fq;
  const signed char* gb = RB + ((size_t)n0 + frow) * Kp + fq;
  const size_t step = (size_t)64 * Kp;
  i4v pa[LA], pb[LB];
#pragma unroll
  for (int q = 0; q < LA; q++) pa[q] = *(const i4v*)(ga + q * step);
#pragma unroll
  for (int q = 0; q < LB; q++) pb[q] = *(const i4v*)(gb + q * step);
#pragma unroll
  for (int q = 0; q < LA; q++) *(i4v*)(&la[0][oz_lds_off<true>(frow + 64 * q, fq >> 4)]) = pa[q];
#pragma unroll
  for (int q = 0; q < LB; q++) *(i4v*)(&lb[0][oz_lds_off<true>(frow + 64 * q, fq >> 4)]) = pb[q];
  __syncthreads();
  int buf = 0;
  for (long k0 = 0; k0 < Kc; k0 += OZ_BK) {
    const bool more = k0 + OZ_BK < Kc;
    if (more) {
#pragma unroll
      for (int q = 0; q < LA; q++) pa[q] = *(const i4v*)(ga + q * step + k0 + OZ_BK);
#pragma unroll
      for (int q = 0; q < LB; q++) pb[q] = *(const i4v*)(gb + q * step + k0 + OZ_BK);
    }
    // lane l holds row (l%16) of a 16-row block, k bytes (l/16)*16..+15
    i4v fb[FB];
#pragma unroll
    for (int b = 0; b < FB; b++) fb[b] = *(const i4v*)(&lb[buf][oz_lds_off<true>(wn + b * 16 + (lane & 15), lane >> 4)]);
#pragma unroll
    for (int a = 0; a < FA; a++) {
      const i4v fa = *(const i4v*)(&la[buf][oz_lds_off<true>(wm + a * 16 + (lane & 15), lane >> 4)]);
#pragma unroll
      for (int b = 0; b < FB; b++) acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fb[b], fa, acc[a][b], 0, 0, 0);
    }
    if (more) {
      // the other buffer was last read before the previous barrier
#pragma unroll
      for (int q = 0; q < LA; q++) *(i4v*)(&la[buf ^ 1][oz_lds_off<true>(frow + 64 * q, fq >> 4)]) = pa[q];
#pragma unroll
      for (int q = 0; q < LB; q++) *(i4v*)(&lb[buf ^ 1][oz_lds_off<true>(frow + 64 * q, fq >> 4)]) = pb[q];
    }
    buf ^= 1;
    __syncthreads();
  }
  // lane l, element r -> row i = l%16 of block a, column j = 4*(l/16)+r of block b
  const int p = tab->p[mod], hp = (p + 1) / 2;
  const double ip = tab->invp[mod];
  signed char* rc = RC + (size_t)mod * plane;
#pragma unroll
  for (int a = 0; a < FA; a++) {
    const long i = m0 + wm + a * 16 + (lane & 15);
    if (i >= M) continue;
#pragma unroll
    for (int b = 0; b < FB; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const long j = n0 + wn + b * 16 + 4 * (lane >> 4) + r;
        if (j >= N) continue;
        signed char* pc = rc + j * ldrc + i;
        double x = (double)acc[a][b][r];
        if (accumulate) x += (double)*pc;
        x -= (double)p * floor(x * ip);
        if (x < 0.0) x += p;
        if (x >= p) x -= p;
        int v = (int)x;
        if (v >= hp) v -= p;
        *pc = (signed char)v;
      }
  }
}

// C = alpha * 2^(e_i+f_j-tA-tB) * CRT(c_1..c_NM) + beta * C.  A thread takes
// 4 consecutive rows of one column (one 32-bit load per residue plane).
// Per element: S = sum c_m Q_m (c_m in [0, p_m)) in 16-bit limbs with
// 32-bit accumulators (24-bit products, at most 30 terms: below 2^29), S mod P
// with a quotient estimated in FP64 (S < NM 256 P: below 2^13) and
// corrected by one step, the balanced representative, rounded to FP64 once
// through the top four limbs.
template <int L16>
__global__ __launch_bounds__(256) void crt_combine_kernel(long M, long N, int NM, const CrtTab* __restrict__ tab,
                                                          const signed char* __restrict__ RC, long ldrc, long plane,
                                                          const double* __restrict__ amax,
                                                          const double* __restrict__ bmax, int tsum, double alpha,
                                                          double beta, double* __restrict__ C, long ldc) {
  const long quads = (M + 3) / 4;
  const long idx = blockIdx.x * 256L + threadIdx.x;
  if (idx >= quads * N) return;
  const long i0 = (idx % quads) * 4, j = idx / quads;
  const bool full = i0 + 4 <= M && (ldrc & 3) == 0;
  unsigned acc[4][L16];
#pragma unroll
  for (int e = 0; e < 4; e++)
#pragma unroll
    for (int l = 0; l < L16; l++) acc[e][l] = 0;
  const unsigned short* q = &tab->q16[NM][0][0];
  for (int m = 0; m < NM; m++) {
    const signed char* src = RC + (size_t)m * plane + j * ldrc + i0;
    unsigned word;
    if (full) word = *(const unsigned*)src;
    else {
      word = 0;
      for (int e = 0; e < 4; e++)
        if (i0 + e < M) word |= ((unsigned)(unsigned char)src[e]) << (8 * e);
    }
    const int p = tab->p[m];
#pragma unroll
    for (int e = 0; e < 4; e++) {
      int c = (int)(signed char)(word >> (8 * e));
      if (c < 0) c += p;
#pragma unroll
      for (int l = 0; l < L16; l++) acc[e][l] = __umul24((unsigned)c, (unsigned)q[m * CRT_L16 + l]) + acc[e][l];
    }
  }
  const unsigned short* P = tab->P16[NM];
  const unsigned short* H = tab->Ph16[NM];
  const double Pd = tab->Pd[NM];
  const int fj = oz_exp(bmax[j]);
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const long i = i0 + e;
    if (i >= M) break;
    unsigned s[L16];
    unsigned carry = 0;
#pragma unroll
    for (int l = 0; l < L16; l++) {
      const unsigned v = acc[e][l] + carry;
      s[l] = v & 0xFFFFu;
      carry = v >> 16;
    }
    double sd = 0.0;
#pragma unroll
    for (int l = L16 - 1; l >= 0; l--) sd = sd * 65536.0 + (double)s[l];
    const unsigned qe = (unsigned)floor(sd / Pd);   // < 2^13
    // s - qe P, borrow in 32-bit
    unsigned mc = 0;
    int br = 0;
#pragma unroll
    for (int l = 0; l < L16; l++) {
      const unsigned pr = __umul24(qe, (unsigned)P[l]) + mc;   // < 2^29 + 2^13
      mc = pr >> 16;
      int d = (int)s[l] - (int)(pr & 0xFFFFu) - br;
      br = d < 0;
      s[l] = (unsigned)(d + (br << 16));
    }
    if (br || mc) {   // qe one too large: add P back
      unsigned c2 = 0;
#pragma unroll
      for (int l = 0; l < L16; l++) {
        const unsigned v = s[l] + P[l] + c2;
        s[l] = v & 0xFFFFu;
        c2 = v >> 16;
      }
    }
    bool ge = true;   // s >= P: qe one too small
#pragma unroll
    for (int l = L16 - 1; l >= 0; l--) {
      if (s[l] != P[l]) { ge = s[l] > P[l]; break; }
    }
    if (ge) {
      int b2 = 0;
#pragma unroll
      for (int l = 0; l < L16; l++) {
        const int d = (int)s[l] - (int)P[l] - b2;
        b2 = d < 0;
        s[l] = (unsigned)(d + (b2 << 16));
      }
    }
    bool gt_half = false;   // balanced: s > P/2 stands for s - P
#pragma unroll
    for (int l = L16 - 1; l >= 0; l--) {
      if (s[l] != H[l]) { gt_half = s[l] > H[l]; break; }
    }
    if (gt_half) {
      int b2 = 0;
#pragma unroll
      for (int l = 0; l < L16; l++) {   // s = P - s
        const int d = (int)P[l] - (int)s[l] - b2;
        b2 = d < 0;
        s[l] = (unsigned)(d + (b2 << 16));
      }
    }
    // correctly rounded to FP64: the top 64 bits, normalized (leading bit at
    // 63), with every lower nonzero bit folded into bit 0 (sticky), then one
    // rounding in the 64-bit integer -> FP64 conversion
    int h = -1;
#pragma unroll
    for (int l = L16 - 1; l >= 0; l--)
      if (h < 0 && s[l] != 0) h = l;
    double v = 0.0;
    if (h >= 0) {
      auto limb = [&](int l) -> unsigned long long {
        unsigned long long r = 0;
#pragma unroll
        for (int q2 = 0; q2 < L16; q2++)
          if (q2 == l) r = s[q2];
        return r;
      };
      unsigned long long x = (limb(h) << 48) | (limb(h - 1) << 32) | (limb(h - 2) << 16) | limb(h - 3);
      const int lz = __clzll(x);   // < 16: limb h is nonzero
      const unsigned long long nxt = limb(h - 4);
      x = lz ? (x << lz) | (nxt >> (16 - lz)) : x;
      bool sticky = lz ? (nxt & ((1ull << (16 - lz)) - 1)) != 0 : nxt != 0;
#pragma unroll
      for (int l = 0; l < L16; l++)
        if (l < h - 4 && s[l] != 0) sticky = true;
      if (sticky) x |= 1ull;
      v = ldexp((double)x, 16 * (h - 3) - lz);
    }
    if (gt_half) v = -v;
    v = ldexp(v, oz_exp(amax[i]) + fj - tsum);
    double* pc = C + i + j * ldc;
    double out = alpha * v;
    if (beta != 0.0) out += beta * *pc;
    *pc = out;
  }
}

int ozaki2_dgemm(bool ta, bool tb, long M, long N, long K, double alpha, const double* A, long lda, const double* B,
                 long ldb, double beta, double* C, long ldc, int moduli, hipStream_t s) {
  if (M <= 0 || N <= 0) return 0;
  if (K <= 0 || alpha == 0.0) { scale<double>(M, N, beta, C, ldc, s); return 0; }
  const long a_sv = ta ? lda : 1, a_sk = ta ? 1 : lda;
  const long b_sv = tb ? 1 : ldb, b_sk = tb ? ldb : 1;
  DevBuf st(sizeof(double) * 3 * (M + N) + 256);
  double* amax = st.as<double>();
  double* alog = amax + M;
  double* acnt = alog + M;
  double* bmax = acnt + M;
  double* blog = bmax + N;
  double* bcnt = blog + N;
  int* nonfinite = (int*)(bcnt + N);
  OBA_HIP(hipMemsetAsync(nonfinite, 0, sizeof(int), s));
  oz_stats(M, K, A, a_sv, a_sk, amax, alog, acnt, nonfinite, s);
  oz_stats(N, K, B, b_sv, b_sk, bmax, blog, bcnt, nonfinite, s);
  int h_nonfinite = 0;
  std::vector<double> h(3 * (M + N));
  OBA_HIP(hipMemcpyAsync(&h_nonfinite, nonfinite, sizeof(int), hipMemcpyDeviceToHost, s));
  OBA_HIP(hipMemcpyAsync(h.data(), amax, sizeof(double) * 3 * (M + N), hipMemcpyDeviceToHost, s));
  OBA_HIP(hipStreamSynchronize(s));
  if (h_nonfinite) {
    gemm<double>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
    return OZ_NONFINITE;
  }
  const CrtTab& ht = crt_host_tab();
  // bits below each vector's largest entry: the typical entry (geometric
  // mean of the nonzero |x|) keeps CRT_KEEP bits
  auto spread = [](long V, const double* mx, const double* lg, const double* cn) {
    double worst = 0.0;
    for (long v = 0; v < V; v++) {
      if (mx[v] <= 0.0 || cn[v] <= 0.0) continue;
      worst = std::max(worst, std::log2(mx[v]) - lg[v] / cn[v]);
    }
    return worst;
  };
  const double* hb = h.data() + 3 * M;
  const double sa = spread(M, h.data(), h.data() + M, h.data() + 2 * M), sb = spread(N, hb, hb + N, hb + 2 * N);
  const double kbits = std::ceil(std::log2((double)K)) + 1.0;   // |C'| < K 2^(tA+tB) <= P/2
  int nm = moduli;
  if (nm <= 0) {
    const double need = 2.0 * CRT_KEEP + sa + sb + kbits;
    nm = 1;
    while (nm <= CRT_NMAX && ht.log2P[nm] < need) nm++;
    if (nm > CRT_NMAX) {
      // the spread needs a longer modulus product than the table holds
      gemm<double>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
      const int want = (int)std::ceil(need / 7.5);
      return -std::max(want, CRT_NMAX + 1);
    }
  }
  if (nm < 1 || nm > CRT_NMAX) fatal("ozaki2: modulus count out of range", nullptr);
  // split the available bits between the operands, the slack by their spreads
  const int total = (int)std::floor(ht.log2P[nm] - kbits - 1e-9);
  int tA = (int)std::floor((total + sa - sb) / 2.0);
  tA = std::max(1, std::min(total - 1, tA));
  const int tB = total - tA;
  if (tB < 1) fatal("ozaki2: too few moduli for this K", nullptr);
  const CrtTab* dt = crt_dev_tab(s);
  const long Kp = (K + OZ_BK - 1) / OZ_BK * OZ_BK;
  const long Mp = (M + CRT_T - 1) / CRT_T * CRT_T, Np = (N + CRT_T - 1) / CRT_T * CRT_T;
  DevBuf ra((size_t)nm * Mp * Kp), rb((size_t)nm * Np * Kp), rc((size_t)nm * M * N);
  {
    dim3 ga((unsigned)(Mp / 64), (unsigned)(Kp / 64)), gb((unsigned)(Np / 64), (unsigned)(Kp / 64));
    auto split = [&](dim3 g, long V, const double* X, long sv, long sk, const double* vmx, int t, signed char* R,
                     long Vp) {
      switch ((t + 17) / 18) {   // 18-bit pieces of a t-bit integer
#define CRT_S(NPV) \
  case NPV: crt_split_kernel<NPV><<<g, 256, 0, s>>>(V, K, X, sv, sk, vmx, t, nm, dt, R, Vp, Kp); break;
        CRT_S(1) CRT_S(2) CRT_S(3) CRT_S(4) CRT_S(5) CRT_S(6) CRT_S(7) CRT_S(8) CRT_S(9) CRT_S(10) CRT_S(11)
        CRT_S(12) CRT_S(13) CRT_S(14)
#undef CRT_S
        default: fatal("ozaki2: operand bits out of range", nullptr);
      }
    };
    split(ga, M, A, a_sv, a_sk, amax, tA, ra.as<signed char>(), Mp);
    split(gb, N, B, b_sv, b_sk, bmax, tB, rb.as<signed char>(), Np);
    OBA_HIP(hipGetLastError());
  }
  if ((Mp / 128) * (Np / 128) > 0x7fffffffL) fatal("ozaki2: grid too large", nullptr);
  {
    // algorithmic work: nm INT8 GEMMs of 2MNK ops; bytes: the residues of
    // both operands once per modulus and the residues of C written
    KTimer kt("ozaki_gemm", s, (double)nm * 2.0 * M * N * K, (double)nm * (M * K + K * N + M * N));
    // 256x256 tiles (a 128x128 register tile per wavefront: half the
    // operand traffic per product of the 128x128 tiles) when they still
    // give every CU a tile per modulus; 8192^3, 17 moduli: 15.1 vs 15.5 ms
    const bool big = (Mp / 256) * (Np / 256) >= rt().num_cu;
    for (long k0 = 0; k0 < Kp; k0 += CRT_KCHUNK) {
      const long kc = std::min(CRT_KCHUNK, Kp - k0);
      if (big)
        crt_gemm_kernel<8, 8><<<dim3((unsigned)((Mp / 256) * (Np / 256)), (unsigned)nm), 256, 0, s>>>(
            M, N, kc, Kp, k0, ra.as<signed char>(), Mp, rb.as<signed char>(), Np, dt, rc.as<signed char>(), M, M * N,
            k0 > 0 ? 1 : 0);
      else
        crt_gemm_kernel<4, 4><<<dim3((unsigned)((Mp / 128) * (Np / 128)), (unsigned)nm), 256, 0, s>>>(
            M, N, kc, Kp, k0, ra.as<signed char>(), Mp, rb.as<signed char>(), Np, dt, rc.as<signed char>(), M, M * N,
            k0 > 0 ? 1 : 0);
      OBA_HIP(hipGetLastError());
    }
  }
  // 16-bit limbs: S < nm * 256 * P
  const int l16 = std::min(CRT_L16, (int)std::ceil((ht.log2P[nm] + std::log2((double)nm) + 8.0 + 1.0) / 16.0));
  const unsigned blocks = (unsigned)(((M + 3) / 4 * N + 255) / 256);
#define CRT_C(L)                                                                                                    \
  case L:                                                                                                           \
    crt_combine_kernel<L><<<blocks, 256, 0, s>>>(M, N, nm, dt, rc.as<signed char>(), M, M * N, amax, bmax, tA + tB, \
                                                 alpha, beta, C, ldc);                                              \
    break;
  switch (std::max(l16, 4)) {
    CRT_C(4) CRT_C(5) CRT_C(6) CRT_C(7) CRT_C(8) CRT_C(9) CRT_C(10) CRT_C(11) CRT_C(12) CRT_C(13) CRT_C(14) CRT_C(15)
    CRT_C(16)
  }
#undef CRT_C
  OBA_HIP(hipGetLastError());
  return nm;
}

}  // namespace oba

using namespace oba;

// Extension entry point (no reference counterpart): cblas_dgemm semantics
// plus the slice count (0 = chosen from the operands' exponent spread).
// Returns the slice count used (0: nothing to multiply), -1 after an argument
// error, OZ_NONFINITE (-2) when an operand holds a NaN or infinity and the
// native FP64 GEMM ran, -s (s > 12) when the operands need s slices and the
// native FP64 GEMM ran, OBA_DEVICE_ERROR after a device error.
OBA_API int openblas_amd_dgemm_ozaki(int order, int TransA, int TransB, int m, int n, int k, double alpha,
                                     const double* a, int lda, const double* b, int ldb, double beta,
                                     double* c, int ldc, int slices) {
  int transa = -1, transb = -1;
  long M = 0, N = 0, K = k, LDA = 0, LDB = 0;
  const double *pa = nullptr, *pb = nullptr;
  auto tr = [](int t) { return (t == 111 || t == 114) ? 0 : (t == 112 || t == 113) ? 1 : -1; };
  int info = 0;
  if (order == 102) { M = m; N = n; pa = a; pb = b; LDA = lda; LDB = ldb; transa = tr(TransA); transb = tr(TransB); }
  else if (order == 101) { M = n; N = m; pa = b; pb = a; LDA = ldb; LDB = lda; transa = tr(TransB); transb = tr(TransA); }
  if (order == 101 || order == 102) {
    long nrowa = transa == 1 ? K : M, nrowb = transb == 1 ? N : K;
    info = -1;
    if (ldc < M) info = 13;
    if (LDB < nrowb) info = 10;
    if (LDA < nrowa) info = 8;
    if (K < 0) info = 5;
    if (N < 0) info = 4;
    if (M < 0) info = 3;
    if (transb < 0) info = 2;
    if (transa < 0) info = 1;
  }
  if (info >= 0) { report_xerbla("DGEMM ", info); return -1; }
  if (slices < 0 || slices > OZ_SMAX) { report_xerbla("DGEMM ", 15); return -1; }
  if (M == 0 || N == 0) return 0;
  int ns = OBA_DEVICE_ERROR;
  guarded([&] {
    Runtime& r = rt();
    Mat<double> A, B, C;
    const bool mul = K > 0 && alpha != 0.0;
    if (mul) {
      A.in(const_cast<double*>(pa), transa ? K : M, transa ? M : K, LDA, true);
      B.in(const_cast<double*>(pb), transb ? N : K, transb ? K : N, LDB, true);
    }
    C.in(c, M, N, ldc, beta != 0.0);
    ns = ozaki_dgemm(transa == 1, transb == 1, M, N, K, alpha, A.d, A.ld, B.d, B.ld, beta, C.d, C.ld, slices,
                     r.stream);
    C.out();
    end_call();
  });
  return ns;
}

// Extension entry point (no reference counterpart): cblas_dgemm semantics
// computed with scheme II (integer moduli, Chinese remainder theorem) plus
// the modulus count (0 = chosen from the operands' exponent spread so that
// the typical entry of each operand keeps 53 bits).  Returns the modulus
// count used (0: nothing to multiply), -1 after an argument error,
// OZ_NONFINITE (-2) when an operand holds a NaN or infinity and the native
// FP64 GEMM ran, -n (n > 30) when the spread needs more moduli than the
// table holds and the native FP64 GEMM ran, OBA_DEVICE_ERROR after a device
// error.
OBA_API int openblas_amd_dgemm_ozaki2(int order, int TransA, int TransB, int m, int n, int k,
                                      double alpha, const double* a, int lda, const double* b, int ldb,
                                      double beta, double* c, int ldc, int moduli) {
  int transa = -1, transb = -1;
  long M = 0, N = 0, K = k, LDA = 0, LDB = 0;
  const double *pa = nullptr, *pb = nullptr;
  auto tr = [](int t) { return (t == 111 || t == 114) ? 0 : (t == 112 || t == 113) ? 1 : -1; };
  int info = 0;
  if (order == 102) { M = m; N = n; pa = a; pb = b; LDA = lda; LDB = ldb; transa = tr(TransA); transb = tr(TransB); }
  else if (order == 101) { M = n; N = m; pa = b; pb = a; LDA = ldb; LDB = lda; transa = tr(TransB); transb = tr(TransA); }
  if (order == 101 || order == 102) {
    long nrowa = transa == 1 ? K : M, nrowb = transb == 1 ? N : K;
    info = -1;
    if (ldc < M) info = 13;
    if (LDB < nrowb) info = 10;
    if (LDA < nrowa) info = 8;
    if (K < 0) info = 5;
    if (N < 0) info = 4;
    if (M < 0) info = 3;
    if (transb < 0) info = 2;
    if (transa < 0) info = 1;
  }
  if (info >= 0) { report_xerbla("DGEMM ", info); return -1; }
  if (moduli < 0 || moduli > CRT_NMAX) { report_xerbla("DGEMM ", 15); return -1; }
  if (M == 0 || N == 0) return 0;
  int nm = OBA_DEVICE_ERROR;
  guarded([&] {
    Runtime& r = rt();
    Mat<double> A, B, C;
    const bool mul = K > 0 && alpha != 0.0;
    if (mul) {
      A.in(const_cast<double*>(pa), transa ? K : M, transa ? M : K, LDA, true);
      B.in(const_cast<double*>(pb), transb ? N : K, transb ? K : N, LDB, true);
    }
    C.in(c, M, N, ldc, beta != 0.0);
    nm = ozaki2_dgemm(transa == 1, transb == 1, M, N, K, alpha, A.d, A.ld, B.d, B.ld, beta, C.d, C.ld, moduli,
                      r.stream);
    C.out();
    end_call();
  });
  return nm;
}
