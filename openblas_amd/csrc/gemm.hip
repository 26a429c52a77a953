// Level-3 GEMM for gfx950: C <- alpha*op(A)*op(B) + beta*C, column-major,
// FP64 on v_mfma_f64_16x16x4f64 and FP32 on v_mfma_f32_16x16x4f32.
//
// Reference semantics followed (not its blocking): the β pre-pass and the
// k==0 / α==0 early exit of driver/level3/level3.c:229-264, β==0 writing
// exact zeros (kernel/generic/gemm_beta.c:52), all four NN/NT/TN/TT forms of
// interface/gemm.c:87-111.
//
// Design (MI355X, not the GotoBLAS 5-loop): one workgroup of 4 wavefronts
// owns a BM x BN tile of C held in MFMA accumulators; K advances in BK slabs
// that are loaded from HBM with 16-byte vector loads into registers one slab
// ahead and staged through a double-buffered k-major LDS tile (padded so the
// fragment reads are bank-conflict free).  The MFMA is issued with op(B) as
// its first operand so that the 16 consecutive lanes of an output row map to
// 16 consecutive rows of column-major C (128-byte coalesced stores).  Tiles
// are mapped XCD-aware: the 8 XCDs each get a contiguous, grouped block of C
// tiles so that the panels of A and B they share stay in their own L2.
#include "kernels.h"
#include <type_traits>

namespace oba {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct MF;
template <> struct MF<double> {
  typedef d4 acc;
  __device__ static inline acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // accumulator element r of lane l holds output row 4r + l/16 (measured
  // lane layout of the f64 16x16x4 instruction), column l%16.
  __device__ static inline int orow(int l, int r) { return 4 * r + (l >> 4); }
};
template <> struct MF<float> {
  typedef f4 acc;
  __device__ static inline acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int orow(int l, int r) { return 4 * (l >> 4) + r; }
};

template <typename T> struct Vec16;
template <> struct Vec16<double> { typedef double2 type; static constexpr int n = 2; };
template <> struct Vec16<float> { typedef float4 type; static constexpr int n = 4; };

// Loads one BK-deep slab of an operand tile (TILE along m or n) into
// registers.  KC: the operand is contiguous along k in memory (Trans A /
// NoTrans B).  VEC: 16-byte vector loads are legal (aligned base and ld).
// The slab is staged k-major in LDS, s[BK][TILE + PAD] (PAD: 128 bytes of
// doubles, 64 of floats, so that the four 16-lane rows of one fragment read
// do not start in the same bank).
template <typename T, int TILE, int BK, bool KC, bool VEC>
struct TileLoader {
  static constexpr int V = VEC ? Vec16<T>::n : 1;
  static constexpr int NV = TILE * BK / V / 256;  // vectors per thread
  static_assert(NV * V * 256 == TILE * BK, "tile must divide over 256 threads");
  static constexpr int PAD = 64 / sizeof(T) * (sizeof(T) == 8 ? 2 : 1);  // 16 doubles / 16 floats
  static constexpr int LD = TILE + PAD;
  static constexpr int SIZE = BK * LD;   // elements per buffer
  __device__ static inline int at(int k, int mn) { return k * LD + mn; }
  T r[NV * V];

  // GUARD: per-element bounds checks (edge tiles, K tail); otherwise the
  // caller guarantees the whole slab is in range and the loads are
  // unconditional 16-byte vector loads.
  template <bool GUARD>
  __device__ inline void load(const T* __restrict__ g, long ld, long mn0, long k0, long MN, long K) {
    const int tid = threadIdx.x & 255;   // thread within its 256-thread group
    constexpr bool full = !GUARD;
#pragma unroll
    for (int i = 0; i < NV; i++) {
      const int v = tid + i * 256;
      long mn, k;
      if (!KC) { k = v / (TILE / V); mn = (v % (TILE / V)) * V; }
      else { mn = v / (BK / V); k = (v % (BK / V)) * V; }
      mn += mn0; k += k0;
      if (full) {
        if constexpr (V > 1) {
          const T* p = KC ? g + k + mn * ld : g + mn + k * ld;
          typedef typename Vec16<T>::type vt;
          vt x = *(const vt*)p;
          const T* xs = (const T*)&x;
#pragma unroll
          for (int e = 0; e < V; e++) r[i * V + e] = xs[e];
        } else {
          r[i] = KC ? g[k + mn * ld] : g[mn + k * ld];
        }
      } else {
#pragma unroll
        for (int e = 0; e < V; e++) {
          long mm = KC ? mn : mn + e, kk = KC ? k + e : k;
          r[i * V + e] = (mm < MN && kk < K) ? g[KC ? kk + mm * ld : mm + kk * ld] : T(0);
        }
      }
    }
  }
  // writes the slab into its LDS buffer (layout above)
  __device__ inline void store(T* s) const {
    const int tid = threadIdx.x & 255;
#pragma unroll
    for (int i = 0; i < NV; i++) {
      const int v = tid + i * 256;
      if (!KC) {
        int k = v / (TILE / V), mn = (v % (TILE / V)) * V;
        if constexpr (V > 1) {
          typedef typename Vec16<T>::type vt;
          vt x;
          T* xs = (T*)&x;
#pragma unroll
          for (int e = 0; e < V; e++) xs[e] = r[i * V + e];
          *(vt*)(s + at(k, mn)) = x;
        } else {
          s[at(k, mn)] = r[i];
        }
      } else {
        int mn = v / (BK / V), k = (v % (BK / V)) * V;
#pragma unroll
        for (int e = 0; e < V; e++) s[at(k + e, mn)] = r[i * V + e];
      }
    }
  }
};

// KS > 1 (sliced K): the workgroup is KS groups of 4 wavefronts; group g
// multiplies the g-th of KS equal, BK-aligned k ranges into its own
// accumulators, and the KS partial tiles are summed in LDS in the fixed
// order 0..KS-1 before the epilogue (one launch, no workspace, the same
// result on every run).  It fills a CU with one workgroup when there are
// only about as many tiles as CUs.
template <typename T, int BM, int BN, int BK, bool TA, bool TB, bool VEC, int KS = 1>
__global__ __launch_bounds__(256 * KS, KS == 1 ? 2 : 1) void gemm_kernel(long M, long N, long K, T alpha,
                                                      const T* __restrict__ A, long lda,
                                                      const T* __restrict__ B, long ldb, T beta,
                                                      T* __restrict__ C, long ldc, int tiles_m,
                                                      int tiles_n, int tri, const int* abort_flag,
                                                      long kc, long cstride) {
  // split K (gridDim.y parts, kc deep): part y multiplies its k range into
  // its own C buffer (cstride apart); a second kernel sums the parts
  if (kc > 0) {
    const long y = blockIdx.y;
    A += y * kc * (TA ? 1 : lda);
    B += y * kc * (TB ? ldb : 1);
    K = min(kc, K - y * kc);
    C += y * cstride;
  }
  typedef TileLoader<T, BM, BK, TA, VEC> LA;
  typedef TileLoader<T, BN, BK, !TB, VEC> LB;
  constexpr int WTM = BM / 2, WTN = BN / 2;  // 2x2 waves
  constexpr int FM = WTM / 16, FN = WTN / 16;
  typedef MF<T> mf;
  typedef typename mf::acc acc_t;

  constexpr int STAGE = 2 * (LA::SIZE + LB::SIZE);   // one group's double-buffered slabs
  constexpr int SMEM = KS == 1 ? STAGE : (KS * STAGE > KS * BM * BN ? KS * STAGE : KS * BM * BN);
  __shared__ T smem[SMEM];
  const int grp = KS == 1 ? 0 : (int)(threadIdx.x >> 8);
  // buffer b of the A slab at st + b*LA::SIZE, of the B slab after both A's
  T* const st = smem + grp * STAGE;
#define OBA_AS(b) (st + (b) * LA::SIZE)
#define OBA_BS(b) (st + 2 * LA::SIZE + (b) * LB::SIZE)
  long kg = 0;   // sliced K: the groups' common k-range length, a multiple of BK
  const long k_total = K;
  if (KS > 1) {
    kg = ((K + KS - 1) / KS + BK - 1) / BK * BK;
    A += grp * kg * (TA ? 1 : lda);
    B += grp * kg * (TB ? ldb : 1);
    K = max(0L, min(kg, K - grp * kg));
  }

  // XCD-aware tile order: workgroup b runs on XCD b%8; give each XCD a
  // contiguous range of the grouped tile order.
  const int T_all = tiles_m * tiles_n;
  int lin = blockIdx.x;
  if ((T_all & 7) == 0) lin = (blockIdx.x & 7) * (T_all >> 3) + (blockIdx.x >> 3);
  constexpr int GM = 8;
  const int group = lin / (GM * tiles_n);
  const int first_m = group * GM;
  const int gsize = min(tiles_m - first_m, GM);
  const int in_group = lin - group * GM * tiles_n;
  const int tm = first_m + in_group % gsize;
  const int tn = in_group / gsize;
  const long m0 = (long)tm * BM, n0 = (long)tn * BN;
  // tri: 1 = update only C(i,j) with i >= j (lower), 2 = only i <= j (upper)
  if (tri == 1 && n0 > m0 + BM - 1) return;
  if (tri == 2 && m0 > n0 + BN - 1) return;
  if (abort_flag && *abort_flag) return;

  const int lane = threadIdx.x & 63;
  const int wave = (threadIdx.x >> 6) & 3;
  const int wm0 = (wave & 1) * WTM, wn0 = (wave >> 1) * WTN;

  LA la;
  LB lb;

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++) acc[i][j] = acc_t{};

  // compute one staged slab: BK/4 MFMA k-steps on the wave's FM x FN tiles
  auto compute = [&](int buf) __attribute__((always_inline)) {
    const T* as = OBA_AS(buf);
    const T* bs = OBA_BS(buf);
#pragma unroll
    for (int kk = 0; kk < BK / 4; kk++) {
      const int kr = kk * 4 + (lane >> 4);
      T fa[FM], fb[FN];
#pragma unroll
      for (int i = 0; i < FM; i++) fa[i] = as[LA::at(kr, wm0 + i * 16 + (lane & 15))];
#pragma unroll
      for (int j = 0; j < FN; j++) fb[j] = bs[LB::at(kr, wn0 + j * 16 + (lane & 15))];
#pragma unroll
      for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) acc[i][j] = mf::mma(fb[j], fa[i], acc[i][j]);
    }
  };
  // Full slabs run with unconditional loads on interior tiles (no control
  // flow around the prefetch, so its loads stay in flight across the MFMAs);
  // the K tail and edge tiles use guarded loads.
  // With sliced K every group runs the same kg/BK slabs (the same barriers);
  // a group whose range ends early loads zeros through the guarded loads.
  auto run = [&](auto guard_tag) __attribute__((always_inline)) {
    constexpr bool G = decltype(guard_tag)::value;
    const int nkf = KS > 1 ? (int)(kg / BK) : (int)(K / BK);
    const bool tail = KS == 1 && (K % BK) != 0;
    if (nkf == 0) {
      la.template load<true>(A, lda, m0, 0, M, K);
      lb.template load<true>(B, ldb, n0, 0, N, K);
      la.store(OBA_AS(0));
      lb.store(OBA_BS(0));
      __syncthreads();
      compute(0);
      return;
    }
    la.template load<G>(A, lda, m0, 0, M, K);
    lb.template load<G>(B, ldb, n0, 0, N, K);
    la.store(OBA_AS(0));
    lb.store(OBA_BS(0));
    __syncthreads();
    for (int kt = 0; kt < nkf - 1; kt++) {
      const int cur = kt & 1;
      la.template load<G>(A, lda, m0, (long)(kt + 1) * BK, M, K);
      lb.template load<G>(B, ldb, n0, (long)(kt + 1) * BK, N, K);
      compute(cur);
      la.store(OBA_AS(cur ^ 1));
      lb.store(OBA_BS(cur ^ 1));
      __syncthreads();
    }
    const int last = (nkf - 1) & 1;
    if (tail) {
      la.template load<true>(A, lda, m0, (long)nkf * BK, M, K);
      lb.template load<true>(B, ldb, n0, (long)nkf * BK, N, K);
    }
    compute(last);
    if (tail) {
      la.store(OBA_AS(last ^ 1));
      lb.store(OBA_BS(last ^ 1));
      __syncthreads();
      compute(last ^ 1);
    }
  };
  // Unguarded loads on interior tiles; with sliced K also only when every
  // group's range is full (K_total == KS*kg), so that the choice, and with
  // it the barrier sequence, is the same for the whole workgroup.
  const bool interior = (m0 + BM <= M) && (n0 + BN <= N) && (KS == 1 || k_total == KS * kg);
  if (interior) run(std::false_type{});
  else run(std::true_type{});

#undef OBA_AS
#undef OBA_BS
  const bool bz = (beta == T(0));
  if constexpr (KS > 1) {
    // partial tiles to LDS (column-major BM x BN per group), then every
    // thread sums its elements over the groups in order and writes C
    __syncthreads();
    T* pt = smem + grp * BM * BN;
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
      for (int j = 0; j < FN; j++)
#pragma unroll
        for (int r = 0; r < 4; r++)
          pt[(wn0 + j * 16 + mf::orow(lane, r)) * BM + wm0 + i * 16 + (lane & 15)] = acc[i][j][r];
    __syncthreads();
#pragma unroll
    for (int e = threadIdx.x; e < BM * BN; e += 256 * KS) {
      const long m = m0 + e % BM, n = n0 + e / BM;
      T sum = smem[e];
#pragma unroll
      for (int g = 1; g < KS; g++) sum += smem[g * BM * BN + e];
      if (m < M && n < N) {
        T* p = C + m + n * ldc;
        T v = alpha * sum;
        if (!bz) v += beta * *p;
        *p = v;
      }
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < FM; i++) {
    const long m = m0 + wm0 + i * 16 + (lane & 15);
#pragma unroll
    for (int j = 0; j < FN; j++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const long n = n0 + wn0 + j * 16 + mf::orow(lane, r);
        if (m < M && n < N && !(tri == 1 && n > m) && !(tri == 2 && m > n)) {
          T* p = C + m + n * ldc;
          T v = alpha * acc[i][j][r];
          if (!bz) v += beta * *p;
          *p = v;
        }
      }
    }
  }
}

template <typename T>
__global__ void scale_kernel(long m, long n, T beta, T* c, long ldc) {
  long total = m * n;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    long i = idx % m, j = idx / m;
    T* p = c + i + j * ldc;
    *p = (beta == T(0)) ? T(0) : beta * *p;
  }
}

template <typename T>
void scale(long m, long n, T beta, T* c, long ldc, hipStream_t s) {
  if (m <= 0 || n <= 0 || beta == T(1)) return;
  long total = m * n;
  int blocks = (int)std::min<long>((total + 255) / 256, 4096);
  scale_kernel<T><<<blocks, 256, 0, s>>>(m, n, beta, c, ldc);
  OBA_HIP(hipGetLastError());
}

// C = alpha*A*B + beta*C with few columns (N <= GS_NMAX, no transposes):
// an HBM-bound matrix-vector product, where 64x64 MFMA tiles leave most of
// the device idle (one tile column: 1.36 ms for 16384^2 x 1 FP64, 6x the
// time to stream A).  Thread = row of A (coalesced down the columns), the k
// range cut into gridDim.y chunks for enough workgroups; the chunk's B
// values staged in LDS; per-chunk partial sums combined in chunk order by a
// second kernel (deterministic).
constexpr int GS_NMAX = 4;
constexpr int GS_KC = 256;   // B rows staged per LDS round

template <typename T, int NR>
__global__ __launch_bounds__(256) void gemm_skinny_part_kernel(long M, long K, long kchunk, const T* __restrict__ A,
                                                               long lda, const T* __restrict__ B, long ldb,
                                                               T* __restrict__ part) {
  __shared__ T sb[NR][GS_KC];
  const long i = blockIdx.x * 256L + threadIdx.x;
  const long k0 = (long)blockIdx.y * kchunk, k1 = min(K, k0 + kchunk);
  T acc[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) acc[r] = T(0);
  for (long kb = k0; kb < k1; kb += GS_KC) {
    const int kn = (int)min((long)GS_KC, k1 - kb);
    __syncthreads();
    for (int idx = threadIdx.x; idx < NR * GS_KC; idx += 256) {
      const int r = idx / GS_KC, kk = idx % GS_KC;
      sb[r][kk] = kk < kn ? B[kb + kk + r * ldb] : T(0);
    }
    __syncthreads();
    if (i < M) {
      const T* ap = A + i + kb * lda;
      int kk = 0;
      for (; kk + 8 <= kn; kk += 8) {   // 8 loads in flight
        T a[8];
#pragma unroll
        for (int u = 0; u < 8; u++) a[u] = ap[(long)(kk + u) * lda];
#pragma unroll
        for (int u = 0; u < 8; u++)
#pragma unroll
          for (int r = 0; r < NR; r++) acc[r] += a[u] * sb[r][kk + u];
      }
      for (; kk < kn; kk++) {
        const T a = ap[(long)kk * lda];
#pragma unroll
        for (int r = 0; r < NR; r++) acc[r] += a * sb[r][kk];
      }
    }
  }
  if (i < M) {
#pragma unroll
    for (int r = 0; r < NR; r++) part[((long)blockIdx.y * NR + r) * M + i] = acc[r];
  }
}

template <typename T, int NR>
__global__ __launch_bounds__(256) void gemm_skinny_sum_kernel(long M, int parts, const T* __restrict__ part, T alpha,
                                                              T beta, T* __restrict__ C, long ldc) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= M) return;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    T s = T(0);
    for (int p = 0; p < parts; p++) s += part[((long)p * NR + r) * M + i];
    T* cp = C + i + r * ldc;
    *cp = beta == T(0) ? alpha * s : alpha * s + beta * *cp;   // beta == 0: C is not read
  }
}

template <typename T, int NR>
static void gemm_skinny(long m, long k, T alpha, const T* a, long lda, const T* b, long ldb, T beta, T* c, long ldc,
                        hipStream_t s) {
  const long mb = (m + 255) / 256;
  const int parts = (int)std::max(1L, std::min((k + GS_KC - 1) / GS_KC, (4L * rt().num_cu + mb - 1) / mb));
  const long kchunk = ((k + parts - 1) / parts + GS_KC - 1) / GS_KC * GS_KC;
  const int np = (int)((k + kchunk - 1) / kchunk);
  DevBuf part(sizeof(T) * (size_t)np * NR * m);
  gemm_skinny_part_kernel<T, NR><<<dim3((unsigned)mb, (unsigned)np), 256, 0, s>>>(m, k, kchunk, a, lda, b, ldb,
                                                                                part.as<T>());
  OBA_HIP(hipGetLastError());
  gemm_skinny_sum_kernel<T, NR><<<(unsigned)mb, 256, 0, s>>>(m, np, part.as<T>(), alpha, beta, c, ldc);
  OBA_HIP(hipGetLastError());
}

template <typename T, int BM, int BN, int BK, bool VEC, int KS = 1>
static void launch_cfg(bool ta, bool tb, long M, long N, long K, T alpha, const T* A, long lda,
                       const T* B, long ldb, T beta, T* C, long ldc, hipStream_t s, int tri,
                       const int* abort_flag, int ksplit = 1, long kc = 0, long cstride = 0) {
  long tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  if (tm * tn > 0x7fffffffL) fatal("gemm: grid too large", nullptr);
  dim3 grid((unsigned)(tm * tn), (unsigned)ksplit), block(256 * KS);
#define OBA_GEMM_L(TA_, TB_) \
  gemm_kernel<T, BM, BN, BK, TA_, TB_, VEC, KS><<<grid, block, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, (int)tm, (int)tn, tri, abort_flag, kc, cstride)
  if (!ta && !tb) OBA_GEMM_L(false, false);
  else if (!ta && tb) OBA_GEMM_L(false, true);
  else if (ta && !tb) OBA_GEMM_L(true, false);
  else OBA_GEMM_L(true, true);
#undef OBA_GEMM_L
  OBA_HIP(hipGetLastError());
}

// C = sum of the ksplit partial products (alpha applied) + beta*C
template <typename T>
__global__ __launch_bounds__(256) void gemm_ksum_kernel(long M, long N, int parts, const T* __restrict__ W, T beta,
                                                        T* __restrict__ C, long ldc) {
  const long total = M * N;
  for (long idx = blockIdx.x * 256L + threadIdx.x; idx < total; idx += (long)gridDim.x * 256) {
    T sum = T(0);
    for (int p = 0; p < parts; p++) sum += W[p * total + idx];
    const long i = idx % M, j = idx / M;
    T* cp = C + i + j * ldc;
    *cp = beta == T(0) ? sum : sum + beta * *cp;   // beta == 0: C is not read
  }
}

template <typename T>
void gemm(bool ta, bool tb, long m, long n, long k, T alpha, const T* a, long lda, const T* b,
          long ldb, T beta, T* c, long ldc, hipStream_t s, int tri, const int* abort_flag) {
  if (m <= 0 || n <= 0) return;
  if (k <= 0 || alpha == T(0)) {
    if (tri) fatal("gemm: triangular update with k==0 or alpha==0 is not supported", nullptr);
    scale<T>(m, n, beta, c, ldc, s);
    return;
  }
  // The split-K and few-column paths take a pool workspace.  Pool blocks are
  // ordered on the library stream only (a block freed there may be handed
  // out again before work queued on another stream has finished with it), so
  // a GEMM on any other stream (the look-ahead LU's trailing update) uses
  // the workspace-free tile kernels.
  const bool ws_ok = s == rt().stream;
  if (ws_ok && !ta && !tb && n <= GS_NMAX && tri == 0 && abort_flag == nullptr && m >= 2048 && k >= 2048) {
    switch (n) {
      case 1: gemm_skinny<T, 1>(m, k, alpha, a, lda, b, ldb, beta, c, ldc, s); return;
      case 2: gemm_skinny<T, 2>(m, k, alpha, a, lda, b, ldb, beta, c, ldc, s); return;
      case 3: gemm_skinny<T, 3>(m, k, alpha, a, lda, b, ldb, beta, c, ldc, s); return;
      default: gemm_skinny<T, 4>(m, k, alpha, a, lda, b, ldb, beta, c, ldc, s); return;
    }
  }
  constexpr int V = Vec16<T>::n;
  const bool vec = (lda % V == 0) && (ldb % V == 0) && (((uintptr_t)a & 15) == 0) && (((uintptr_t)b & 15) == 0);
  const long big_tiles = ((m + 127) / 128) * ((n + 127) / 128);
  const bool small = big_tiles < rt().num_cu;
  // k-slab depths from a sweep on MI355X (8192^3): shallow slabs keep the
  // LDS footprint small and the MFMA pipe fed (FP64 128x128: BK 16 -> 4 is
  // 60.6 -> 66.0 TF/s; FP32 BK 32 -> 8 is 118.5 -> 134.8 TF/s); the tile
  // loaders need TILE*BK >= 256 vectors of 16 bytes
  constexpr int BK_BIG = sizeof(T) == 8 ? 4 : 8;
  constexpr int BK_SMALL = sizeof(T) == 8 ? 8 : 16;
  const long tiles64 = ((m + 63) / 64) * ((n + 63) / 64);
  const bool sliced_deep = k >= 1024 && 8 * tiles64 >= 7 * rt().num_cu && tiles64 <= rt().num_cu;
  // few tiles and a moderate K (no split-K path below 1024): the four wave
  // groups per tile hide each other's latencies (the recursive TRSM's
  // updates, 256..512 x n x 256..512)
  const bool sliced_mid = k >= 256 && k < 1024 && 2 * tiles64 <= rt().num_cu;
  if (small && tri == 0 && abort_flag == nullptr && (sliced_deep || sliced_mid)) {
    // about one 64x64 tile per CU, deep K: one workgroup of 4 wave groups per
    // tile, each group a quarter of K, partials summed in LDS (4 wavefronts
    // per SIMD instead of 1; no workspace, so any stream).  MI355X, 1024^3:
    // DGEMM 68.7 -> 59.5 us, SGEMM 47.4 -> 40.3 us per call against the
    // split across workgroups; with 192 tiles the split is still ahead.
    if (vec) launch_cfg<T, 64, 64, BK_SMALL, true, 4>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s, 0, nullptr);
    else launch_cfg<T, 64, 64, BK_SMALL, false, 4>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s, 0, nullptr);
    return;
  }
  if (ws_ok && small && tri == 0 && abort_flag == nullptr && k >= 1024 && tiles64 <= rt().num_cu) {
    // fewer tiles than half the CUs, deep K: split K across workgroups so
    // that ~4 of them per CU hide each other's latencies
    const int ks = (int)std::min<long>({8L, std::max<long>(1L, 4L * rt().num_cu / tiles64), k / 256});
    if (ks > 1) {
      const long kc = ((k + ks - 1) / ks + 15) / 16 * 16;
      const int parts = (int)((k + kc - 1) / kc);
      DevBuf w(sizeof(T) * (size_t)parts * m * n);
      if (vec)
        launch_cfg<T, 64, 64, BK_SMALL, true>(ta, tb, m, n, k, alpha, a, lda, b, ldb, T(0), w.as<T>(), m, s, 0, nullptr,
                                              parts, kc, m * n);
      else
        launch_cfg<T, 64, 64, BK_SMALL, false>(ta, tb, m, n, k, alpha, a, lda, b, ldb, T(0), w.as<T>(), m, s, 0,
                                               nullptr, parts, kc, m * n);
      const long total = m * n;
      gemm_ksum_kernel<T><<<(unsigned)std::min<long>((total + 255) / 256, 4096), 256, 0, s>>>(m, n, parts, w.as<T>(),
                                                                                             beta, c, ldc);
      OBA_HIP(hipGetLastError());
      return;
    }
  }
  if (small) {
    if (vec) launch_cfg<T, 64, 64, BK_SMALL, true>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s, tri, abort_flag);
    else launch_cfg<T, 64, 64, BK_SMALL, false>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s, tri, abort_flag);
  } else {
    if (vec) launch_cfg<T, 128, 128, BK_BIG, true>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s, tri, abort_flag);
    else launch_cfg<T, 128, 128, BK_BIG, false>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s, tri, abort_flag);
  }
}

template void gemm<double>(bool, bool, long, long, long, double, const double*, long, const double*, long, double, double*, long, hipStream_t, int, const int*);
template void gemm<float>(bool, bool, long, long, long, float, const float*, long, const float*, long, float, float*, long, hipStream_t, int, const int*);
template void scale<double>(long, long, double, double*, long, hipStream_t);
template void scale<float>(long, long, float, float*, long, hipStream_t);

}  // namespace oba
