// Triangular solve with multiple right-hand sides (all 8 side/uplo/trans
// forms, unit or non-unit diagonal), device pointers, column-major.
//
// Reference semantics: B <- alpha*B first (driver/level3/trsm_L.c:103-107;
// alpha == 0 zeroes B and returns), then op(A)*X = B (left) or X*op(A) = B
// (right), the diagonal applied as a multiplication by its reciprocal
// (kernel/generic/trsm_ltcopy_2.c:43 INV(a) = 1/a, trsm_kernel_LT.c:95).
//
// Design: recursive blocking on the triangle (split in halves on a 64-row
// grid) so that almost all flops land in large MFMA GEMM updates.  Left-side
// leaves (up to 256 rows) are solved by wavefronts of right-hand-side
// vectors: lane i owns x_i and row i of the current 64x64 tile of the
// triangle, and x_k is broadcast with v_readlane after it is final (forward
// or backward column-oriented substitution).  Right-side leaves multiply by
// explicit inverses of the 64x64 diagonal blocks, or solve rows of B (the
// transposed system) one row per lane.
#include "kernels.h"

namespace oba {

constexpr int TRSM_NB = 64;

__device__ inline double bcast_lane(double v, int k) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
  return __hiloint2double(hi, lo);
}
__device__ inline float bcast_lane(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// Solves E x = b for nvec vectors, E = op(A) (left) or op(A)^T (right),
// nb <= 64.  Vector v, element i lives at B[i*inc_e + v*inc_v].
// eT: E(i,k) = A(k,i) (else A(i,k)); elower: E is lower triangular.
// right-hand-side vectors per wavefront (more workgroups beat interleaved
// chains: DTRSM 8192 x 1024 3.52 ms with 2, 3.21 ms with 1)
constexpr int TRSM_VPW = 1;

template <typename T>
__global__ __launch_bounds__(256) void trsm_leaf_kernel(int nb, long nvec, const T* __restrict__ A, long lda,
                                                        T* __restrict__ B, long inc_e, long inc_v, bool eT,
                                                        bool elower, bool unit, const int* abort_flag) {
  constexpr int LDE = TRSM_NB + 1;     // padded: conflict-free transposed stores
  __shared__ T E[TRSM_NB * LDE];       // E(i,k) at E[k*LDE + i]
  __shared__ T dinv[TRSM_NB];
  if (abort_flag && *abort_flag) return;
  // consecutive threads read consecutive addresses of A in both layouts;
  // every load issued before the first LDS write (clamped in-bounds
  // addresses, the padding selected after)
  constexpr int SU = TRSM_NB * TRSM_NB / 256;
  T sv[SU];
#pragma unroll
  for (int u = 0; u < SU; u++) {
    const int idx = threadIdx.x + 256 * u, f = idx & 63, s = idx >> 6;   // f: fast (contiguous) index of A
    sv[u] = A[min(f, nb - 1) + (long)min(s, nb - 1) * lda];
  }
#pragma unroll
  for (int u = 0; u < SU; u++) {
    const int idx = threadIdx.x + 256 * u, f = idx & 63, s = idx >> 6;
    // eT: E(i,k) = A(k,i) -> A[f + s*lda] = A(f, s) = E(s, f)
    const int i = eT ? s : f, k = eT ? f : s;
    E[k * LDE + i] = (f < nb && s < nb) ? sv[u] : T(0);
  }
  __syncthreads();
  if (threadIdx.x < TRSM_NB) {
    const int k = threadIdx.x;
    dinv[k] = (k < nb && !unit) ? T(1) / E[k * LDE + k] : T(1);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long v0 = ((long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * TRSM_VPW;
  // lane i holds row i of E and its reciprocal pivot in registers; the
  // wavefront's TRSM_VPW vectors (interleaved chains when more than one)
  // run through fully unrolled steps (static register indices; a runtime step loop reading E
  // from LDS inside the chain took ~36 us per 64 x 1024 leaf)
  T e[TRSM_NB];
#pragma unroll
  for (int k = 0; k < TRSM_NB; k++) e[k] = E[k * LDE + lane];
  const T di = dinv[lane];
  T x[TRSM_VPW];
#pragma unroll
  for (int vv = 0; vv < TRSM_VPW; vv++) {
    const long v = v0 + vv;
    x[vv] = (v < nvec && lane < nb) ? B[v * inc_v + lane * inc_e] : T(0);
  }
  if (elower) {
#pragma unroll
    for (int k = 0; k < TRSM_NB; k++) {
      if (k < nb) {
        T xk[TRSM_VPW];
#pragma unroll
        for (int vv = 0; vv < TRSM_VPW; vv++) {
          if (lane == k) x[vv] *= di;
          xk[vv] = bcast_lane(x[vv], k);
        }
#pragma unroll
        for (int vv = 0; vv < TRSM_VPW; vv++)
          if (lane > k) x[vv] -= e[k] * xk[vv];
      }
    }
  } else {
#pragma unroll
    for (int k = TRSM_NB - 1; k >= 0; k--) {
      if (k < nb) {
        T xk[TRSM_VPW];
#pragma unroll
        for (int vv = 0; vv < TRSM_VPW; vv++) {
          if (lane == k) x[vv] *= di;
          xk[vv] = bcast_lane(x[vv], k);
        }
#pragma unroll
        for (int vv = 0; vv < TRSM_VPW; vv++)
          if (lane < k) x[vv] -= e[k] * xk[vv];
      }
    }
  }
#pragma unroll
  for (int vv = 0; vv < TRSM_VPW; vv++) {
    const long v = v0 + vv;
    if (v < nvec && lane < nb) B[v * inc_v + lane * inc_e] = x[vv];
  }
}

// Left-side leaves of up to TRSM_LEAF_BLKS x 64 rows in one launch: the
// 64-row blocks are solved in dependency order (forward for lower E,
// backward for upper), each solved block's terms subtracted from the later
// blocks of the leaf in registers (x_k broadcast by v_readlane), so the
// recursion's two lowest GEMM levels (64 x n x 64 and 128 x n x 128 products
// with few output tiles) and their launches disappear.  The 64x64 tiles of
// the triangle pass through double-buffered LDS (one barrier per tile, the
// next tile's loads in flight during the current tile's work); each lane
// takes its row of the tile into registers.  Vectors as in trsm_leaf_kernel.
constexpr int TRSM_LEAF_BLKS = 4;
constexpr int TRSM_LEAF_MAX = TRSM_LEAF_BLKS * TRSM_NB;

template <typename T>
__global__ __launch_bounds__(256) void trsm_leafb_kernel(int nb, long nvec, const T* __restrict__ A, long lda,
                                                         T* __restrict__ B, long inc_e, long inc_v, bool eT,
                                                         bool elower, bool unit, const int* abort_flag) {
  constexpr int LDE = TRSM_NB + 1;
  constexpr int MB = TRSM_LEAF_BLKS;
  constexpr int PF = TRSM_NB * TRSM_NB / 256;   // tile elements per thread
  __shared__ T E[2][TRSM_NB * LDE];             // tile element (i, k) at E[par][k*LDE + i]
  // per wavefront: its vectors' x, block by block (written and read back by
  // the same wavefront only, no barrier; a register array indexed by the
  // run-time block number was placed in scratch memory)
  __shared__ T Xb[4][MB][TRSM_VPW][TRSM_NB];
  if (abort_flag && *abort_flag) return;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int nblk = (nb + TRSM_NB - 1) / TRSM_NB;
  const long v0 = ((long)blockIdx.x * (blockDim.x >> 6) + (t >> 6)) * TRSM_VPW;
  // block b of the leaf in processing order q: b = q (lower) or nblk-1-q
  auto blk_of = [&](int q) { return elower ? q : nblk - 1 - q; };
  // the tile (i-block bi, k-block bj) of E = op(A) into registers
  auto fetch = [&](int bi, int bj, T (&r)[PF]) {
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int idx = t + u * 256, f = idx & 63, s = idx >> 6;   // f: contiguous index of A
      const int ra = (eT ? bj : bi) * TRSM_NB + f, ca = (eT ? bi : bj) * TRSM_NB + s;
      r[u] = (ra < nb && ca < nb) ? A[ra + (long)ca * lda] : T(0);
    }
  };
  auto store = [&](int par, const T (&r)[PF]) {
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int idx = t + u * 256, f = idx & 63, s = idx >> 6;
      const int i = eT ? s : f, k = eT ? f : s;
      E[par][k * LDE + i] = r[u];
    }
  };
#pragma unroll
  for (int b = 0; b < MB; b++)
#pragma unroll
    for (int vv = 0; vv < TRSM_VPW; vv++) {
      const long v = v0 + vv;
      const int row = b * TRSM_NB + lane;
      Xb[w][b][vv][lane] = (b < nblk && v < nvec && row < nb) ? B[v * inc_v + (long)row * inc_e] : T(0);
    }
  // tiles in order: for each block row qi, the solved blocks qj < qi, then the diagonal
  int qi = 0, qj = 0, par = 0;
  T pf[PF];
  fetch(blk_of(0), blk_of(0), pf);
  T xt[TRSM_VPW];
#pragma unroll
  for (int vv = 0; vv < TRSM_VPW; vv++) xt[vv] = Xb[w][blk_of(0)][vv][lane];
  while (true) {
    store(par, pf);
    __syncthreads();
    const int bi = blk_of(qi), bj = blk_of(qj);
    int ni = qi, nj = qj + 1;   // the next tile
    if (nj > qi) { ni = qi + 1; nj = 0; }
    const bool more = ni < nblk;
    if (more) fetch(blk_of(ni), blk_of(nj), pf);
    // Every step is straight-line code (a branch per step kept each step's
    // loads from being issued early): rows of a partial block beyond its
    // end hold exact zeros in x (the diagonal solve never updates them) and
    // in the tile (the fetch pads with zeros), so steps past the end change
    // nothing, whatever Inf or NaN the real values hold.
    const T* el = &E[par][lane];
    if (qj < qi) {   // xt -= E(bi, bj) x(bj): the tile row and x(bj) read from LDS
#pragma unroll
      for (int k = 0; k < TRSM_NB; k++) {
#pragma unroll
        for (int vv = 0; vv < TRSM_VPW; vv++) xt[vv] -= el[k * LDE] * Xb[w][bj][vv][k];
      }
    } else {         // the diagonal block, as trsm_leaf_kernel, the tile row in registers
      const int nk = min(TRSM_NB, nb - bi * TRSM_NB);
      const bool live = lane < nk;
      const T di = (live && !unit) ? T(1) / el[lane * LDE] : T(1);
      T e[TRSM_NB];
#pragma unroll
      for (int k = 0; k < TRSM_NB; k++) e[k] = el[k * LDE];
      if (elower) {
#pragma unroll
        for (int k = 0; k < TRSM_NB; k++) {
          T xk[TRSM_VPW];
#pragma unroll
          for (int vv = 0; vv < TRSM_VPW; vv++) {
            if (lane == k) xt[vv] *= di;
            xk[vv] = bcast_lane(xt[vv], k);
          }
#pragma unroll
          for (int vv = 0; vv < TRSM_VPW; vv++)
            if (lane > k && live) xt[vv] -= e[k] * xk[vv];
        }
      } else {
#pragma unroll
        for (int k = TRSM_NB - 1; k >= 0; k--) {
          T xk[TRSM_VPW];
#pragma unroll
          for (int vv = 0; vv < TRSM_VPW; vv++) {
            if (lane == k) xt[vv] *= di;
            xk[vv] = bcast_lane(xt[vv], k);
          }
#pragma unroll
          for (int vv = 0; vv < TRSM_VPW; vv++)
            if (lane < k && live) xt[vv] -= e[k] * xk[vv];
        }
      }
#pragma unroll
      for (int vv = 0; vv < TRSM_VPW; vv++) Xb[w][bi][vv][lane] = xt[vv];
      if (more) {
#pragma unroll
        for (int vv = 0; vv < TRSM_VPW; vv++) xt[vv] = Xb[w][blk_of(ni)][vv][lane];
      }
    }
    if (!more) break;
    qi = ni;
    qj = nj;
    par ^= 1;
  }
#pragma unroll
  for (int b = 0; b < MB; b++)
#pragma unroll
    for (int vv = 0; vv < TRSM_VPW; vv++) {
      const long v = v0 + vv;
      const int row = b * TRSM_NB + lane;
      if (b < nblk && v < nvec && row < nb) B[v * inc_v + (long)row * inc_e] = Xb[w][b][vv][lane];
    }
}

// Right-side leaves: the vectors are rows of B (contiguous across vectors,
// stride ldb along a vector), so one lane owns one whole vector in
// registers and the loads/stores of a wavefront are coalesced; the triangle
// is read from LDS as broadcasts.
template <typename T>
__global__ __launch_bounds__(256) void trsm_leaf_rows_kernel(int nb, long nvec, const T* __restrict__ A, long lda,
                                                             T* __restrict__ B, long ldb, bool eT, bool elower,
                                                             bool unit, const int* abort_flag) {
  constexpr int LDE = TRSM_NB + 2;
  __shared__ T Ec[TRSM_NB * LDE];      // E(c,k) at Ec[k*LDE + c]: column k contiguous
  __shared__ T dinv[TRSM_NB];
  if (abort_flag && *abort_flag) return;
  for (int idx = threadIdx.x; idx < TRSM_NB * TRSM_NB; idx += blockDim.x) {
    const int f = idx & 63, s = idx >> 6;   // A(f, s), f contiguous
    const int c = eT ? s : f, k = eT ? f : s;
    T v = T(0);
    if (f < nb && s < nb) v = A[f + (long)s * lda];
    Ec[k * LDE + c] = v;
  }
  __syncthreads();
  if (threadIdx.x < TRSM_NB) {
    const int k = threadIdx.x;
    dinv[k] = (k < nb && !unit) ? T(1) / Ec[k * LDE + k] : T(1);
  }
  __syncthreads();
  const long v = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (v >= nvec) return;
  T x[TRSM_NB];
#pragma unroll
  for (int c = 0; c < TRSM_NB; c++) x[c] = (c < nb) ? B[v + c * ldb] : T(0);
  // column-oriented substitution: x_k is final, then an independent axpy
  // over the remaining entries (ILP instead of a dependent dot product)
  if (elower) {
#pragma unroll
    for (int k = 0; k < TRSM_NB; k++) {
      x[k] *= dinv[k];
#pragma unroll
      for (int c = k + 1; c < TRSM_NB; c++) x[c] -= Ec[k * LDE + c] * x[k];
    }
  } else {
#pragma unroll
    for (int k = TRSM_NB - 1; k >= 0; k--) {
      x[k] *= dinv[k];
#pragma unroll
      for (int c = 0; c < k; c++) x[c] -= Ec[k * LDE + c] * x[k];
    }
  }
#pragma unroll
  for (int c = 0; c < TRSM_NB; c++)
    if (c < nb) B[v + c * ldb] = x[c];
}

// Right-side leaves through explicit inverses of the 64x64 diagonal blocks:
// a fully unrolled per-row substitution is ~4000 straight-line instructions
// run once per wavefront and measured at 40-65 us per leaf whatever its
// size (instruction fetch); the two loop-based kernels below replace it.
//
// (1) One workgroup per diagonal block b of M = op(A) (upper or lower,
//     unit or not) computes M_bb^{-1} by doubling: the four 16x16 diagonal
//     blocks are inverted by substitution (one thread per column, the
//     diagonal applied as a multiplication by its reciprocal as
//     trsm_kernel_LT.c:95), then pairs are joined with two small products
//     each (upper: Y12 = -Y11 M12 Y22; lower: Y21 = -Y22 M21 Y11), 16 -> 32
//     -> 64.  Stored row-major at inv[b*64*64 + i*64 + j].
template <typename T>
__global__ __launch_bounds__(256) void trsm_inv_diag_kernel(long na, const T* __restrict__ A, long lda, bool mT,
                                                            bool mupper, bool unit, T* __restrict__ inv,
                                                            const int* abort_flag) {
  constexpr int LD = TRSM_NB + 1;
  __shared__ T Ms[TRSM_NB * LD];   // M(i,k) at Ms[i*LD + k]
  __shared__ T Ys[TRSM_NB * LD];   // M^{-1}(i,j) at Ys[i*LD + j]
  __shared__ T Ts[TRSM_NB * LD];   // products of the joins
  if (abort_flag && *abort_flag) return;
  const long o = (long)blockIdx.x * TRSM_NB;
  const int nb = (int)min((long)TRSM_NB, na - o);
  const int t = threadIdx.x;
  // staging: every load issued before the first LDS write (unconditional
  // loads from clamped, in-bounds addresses; the padding is selected after)
  constexpr int SU = TRSM_NB * TRSM_NB / 256;
  T sv[SU];
#pragma unroll
  for (int u = 0; u < SU; u++) {
    const int idx = t + 256 * u, f = idx & 63, g = idx >> 6;   // A(o+f, o+g), f contiguous
    sv[u] = A[(o + min(f, nb - 1)) + (o + min(g, nb - 1)) * lda];
  }
#pragma unroll
  for (int u = 0; u < SU; u++) {
    const int idx = t + 256 * u, f = idx & 63, g = idx >> 6;
    const T v = (f < nb && g < nb) ? sv[u] : (f == g ? T(1) : T(0));   // pad with the identity
    if (mT) Ms[g * LD + f] = v;  // M(g,f) = A(f,g)
    else Ms[f * LD + g] = v;
    Ys[f * LD + g] = T(0);
  }
  __syncthreads();
  // 16x16 diagonal blocks: lane j < 16 of wavefront blk computes column j
  // of block blk in registers, every step unrolled with static indices; the
  // terms outside the column's support are skipped by selects (so an
  // unreferenced Inf/NaN never meets a zero), the sums run in the order of
  // the substitution: k upward from the column (lower) or from i+1 (upper).
  if ((t & 63) < 16) {
    const int blk = t >> 6, j = t & 15, o16 = blk * 16;
    auto dinv = [&](int i) { return unit ? T(1) : T(1) / Ms[(o16 + i) * LD + o16 + i]; };
    T y[16];
    if (mupper) {
#pragma unroll
      for (int i = 15; i >= 0; i--) {
        T acc = T(0);
#pragma unroll
        for (int k = i + 1; k < 16; k++) {
          const T f = acc + Ms[(o16 + i) * LD + o16 + k] * y[k];
          acc = k <= j ? f : acc;
        }
        y[i] = i > j ? T(0) : (i == j ? dinv(i) : -acc * dinv(i));
      }
    } else {
#pragma unroll
      for (int i = 0; i < 16; i++) {
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < i; k++) {
          const T f = acc + Ms[(o16 + i) * LD + o16 + k] * y[k];
          acc = k >= j ? f : acc;
        }
        y[i] = i < j ? T(0) : (i == j ? dinv(i) : -acc * dinv(i));
      }
    }
#pragma unroll
    for (int i = 0; i < 16; i++) Ys[(o16 + i) * LD + o16 + j] = y[i];
  }
  __syncthreads();
  // joins 16 -> 32 -> 64 with the block size a compile-time constant, so
  // every loop is unrolled and a thread's LDS reads issue together (same
  // operations, same order)
  auto join = [&](auto hc) __attribute__((always_inline)) {
    constexpr int h = decltype(hc)::value;
    constexpr int npair = TRSM_NB / (2 * h);
    constexpr int PER = npair * h * h / 256;   // outputs per thread
    static_assert(PER * 256 == npair * h * h, "whole rounds");
    // T = M_off * Y_near   (upper: M12 Y22; lower: M21 Y11)
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int idx = t + 256 * u;
      const int p = idx / (h * h), r = (idx / h) % h, c = idx % h, po = p * 2 * h;
      const int mr = mupper ? po + r : po + h + r, mc0 = mupper ? po + h : po;   // M rows / first column
      const int yc = mupper ? po + h + c : po + c;                                  // Y column
      T acc = T(0);
#pragma unroll
      for (int k = 0; k < h; k++) acc += Ms[mr * LD + mc0 + k] * Ys[(mc0 + k) * LD + yc];
      Ts[(po + r) * LD + c] = acc;
    }
    __syncthreads();
    // Y_off = -Y_far * T   (upper: Y12 = -Y11 T; lower: Y21 = -Y22 T)
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int idx = t + 256 * u;
      const int p = idx / (h * h), r = (idx / h) % h, c = idx % h, po = p * 2 * h;
      const int yr = mupper ? po + r : po + h + r, yk0 = mupper ? po : po + h;
      T acc = T(0);
#pragma unroll
      for (int k = 0; k < h; k++) acc += Ys[yr * LD + yk0 + k] * Ts[(po + k) * LD + c];
      const int oc = mupper ? po + h + c : po + c;
      Ys[yr * LD + oc] = -acc;
    }
    __syncthreads();
  };
  join(std::integral_constant<int, 16>{});
  join(std::integral_constant<int, 32>{});
  T* out = inv + (long)blockIdx.x * TRSM_NB * TRSM_NB;
  for (int idx = t; idx < TRSM_NB * TRSM_NB; idx += 256) {
    const int i = idx >> 6, j = idx & 63;
    out[idx] = (i < nb && j < nb) ? Ys[i * LD + j] : T(0);
  }
}

// (2) X <- X * M_bb^{-1} for the leaf's nb columns of B, in place: a
//     workgroup stages 64 rows x nb columns of B and the inverse in LDS;
//     wavefront w computes columns 16w..16w+15 of all 64 rows (the inverse
//     row is a broadcast read, the B column a contiguous one).
template <typename T>
__global__ __launch_bounds__(256) void trsm_right_inv_kernel(int nb, long nvec, const T* __restrict__ inv,
                                                             T* __restrict__ B, long ldb, const int* abort_flag) {
  constexpr int LDX = TRSM_NB + 1;
  __shared__ T Xs[TRSM_NB * LDX];                              // X(r, k) at Xs[k*LDX + r]
  __shared__ __attribute__((aligned(16))) T Ys[TRSM_NB * TRSM_NB];   // inverse, row-major
  if (abort_flag && *abort_flag) return;
  const long v0 = (long)blockIdx.x * 64;
  const int t = threadIdx.x, r = t & 63, cg = t >> 6;
  // staging: every load issued before the first LDS write (unconditional
  // loads from clamped, in-bounds addresses; the padding is selected after)
  constexpr int SU = TRSM_NB * TRSM_NB / 256;
  T xv[SU], yv[SU];
  const long vmax = nvec - 1 - v0;
#pragma unroll
  for (int u = 0; u < SU; u++) {
    const int idx = t + 256 * u, rr = idx & 63, k = idx >> 6;
    xv[u] = B[(v0 + min((long)rr, vmax)) + (long)min(k, nb - 1) * ldb];
    yv[u] = inv[idx];
  }
#pragma unroll
  for (int u = 0; u < SU; u++) {
    const int idx = t + 256 * u, rr = idx & 63, k = idx >> 6;
    Xs[k * LDX + rr] = (k < nb && v0 + rr < nvec) ? xv[u] : T(0);
    Ys[idx] = yv[u];
  }
  __syncthreads();
  T acc[16];
#pragma unroll
  for (int c = 0; c < 16; c++) acc[c] = T(0);
  // a static trip count (X and the inverse are zero past nb) lets each
  // unrolled group of steps issue its LDS reads together
#pragma unroll 8
  for (int k = 0; k < TRSM_NB; k++) {
    const T xv = Xs[k * LDX + r];
    const T* yr = Ys + k * TRSM_NB + cg * 16;
#pragma unroll
    for (int c = 0; c < 16; c++) acc[c] += xv * yr[c];
  }
  if (v0 + r < nvec) {
#pragma unroll
    for (int c = 0; c < 16; c++)
      if (cg * 16 + c < nb) B[(v0 + r) + (long)(cg * 16 + c) * ldb] = acc[c];
  }
}

// inverses of op(A)'s diagonal blocks for a right-side solve (a0: the
// triangle's first element; the recursion splits on the 64-grid from a0)
template <typename T>
struct InvCtx {
  const T* a0 = nullptr;
  long lda = 0;
  const T* inv = nullptr;
};

template <typename T>
static void leaf(bool right, bool lower, bool trans, bool unit, long nb, long nrhs, const T* a, long lda, T* b,
                 long ldb, hipStream_t s, const int* abort_flag, const InvCtx<T>& ic) {
  if (nb <= 0 || nrhs <= 0) return;
  // E = op(A) for left, op(A)^T for right
  const bool eT = trans != right;
  const bool elower = lower != eT;
  if (right && ic.inv) {
    const long blk = (long)(a - ic.a0) / (ic.lda + 1) / TRSM_NB;
    trsm_right_inv_kernel<T><<<(unsigned)((nrhs + 63) / 64), 256, 0, s>>>(
        (int)nb, nrhs, ic.inv + blk * TRSM_NB * TRSM_NB, b, ldb, abort_flag);
    OBA_HIP(hipGetLastError());
    return;
  }
  if (right) {
    trsm_leaf_rows_kernel<T><<<(unsigned)((nrhs + 63) / 64), 64, 0, s>>>((int)nb, nrhs, a, lda, b, ldb, eT, elower,
                                                                        unit, abort_flag);
    OBA_HIP(hipGetLastError());
    return;
  }
  const long inc_e = right ? ldb : 1;
  const long inc_v = right ? 1 : ldb;
  const int waves = 4;
  const long per_block = (long)waves * TRSM_VPW;
  long blocks = (nrhs + per_block - 1) / per_block;
  if (nb > TRSM_NB)
    trsm_leafb_kernel<T><<<(unsigned)blocks, 64 * waves, 0, s>>>((int)nb, nrhs, a, lda, b, inc_e, inc_v, eT, elower,
                                                                 unit, abort_flag);
  else
    trsm_leaf_kernel<T><<<(unsigned)blocks, 64 * waves, 0, s>>>((int)nb, nrhs, a, lda, b, inc_e, inc_v, eT, elower,
                                                                unit, abort_flag);
  OBA_HIP(hipGetLastError());
}

// alpha already applied; solves in place.
template <typename T>
static void rec(bool right, bool lower, bool trans, bool unit, long m, long n, const T* a, long lda, T* b,
                long ldb, hipStream_t s, const int* abort_flag, const InvCtx<T>& ic) {
  const long na = right ? n : m;   // order of the triangle
  // left-side leaves take up to TRSM_LEAF_MAX rows (trsm_leafb_kernel);
  // right-side leaves stay on the 64-grid of the diagonal inverses
  if (na <= (right ? TRSM_NB : TRSM_LEAF_MAX)) {
    leaf<T>(right, lower, trans, unit, na, right ? m : n, a, lda, b, ldb, s, abort_flag, ic);
    return;
  }
  long n1 = ((na / 2) + TRSM_NB - 1) / TRSM_NB * TRSM_NB;
  if (n1 >= na) n1 = na - TRSM_NB;
  const long n2 = na - n1;
  const T* a11 = a;
  const T* a22 = a + n1 + n1 * lda;
  const T* a21 = a + n1;          // A(n1:, :n1)
  const T* a12 = a + n1 * lda;    // A(:n1, n1:)
  const bool eff_lower = (lower != trans);  // op(A) lower triangular
  if (!right) {
    T* b1 = b;
    T* b2 = b + n1;
    if (eff_lower) {
      rec<T>(right, lower, trans, unit, n1, n, a11, lda, b1, ldb, s, abort_flag, ic);
      // B2 -= op(A)21 * X1 ; op(A)21 = A21 (lower, N) or A12^T (upper, T)
      if (!trans) gemm<T>(false, false, n2, n, n1, T(-1), a21, lda, b1, ldb, T(1), b2, ldb, s, 0, abort_flag);
      else gemm<T>(true, false, n2, n, n1, T(-1), a12, lda, b1, ldb, T(1), b2, ldb, s, 0, abort_flag);
      rec<T>(right, lower, trans, unit, n2, n, a22, lda, b2, ldb, s, abort_flag, ic);
    } else {
      rec<T>(right, lower, trans, unit, n2, n, a22, lda, b2, ldb, s, abort_flag, ic);
      // B1 -= op(A)12 * X2 ; op(A)12 = A12 (upper, N) or A21^T (lower, T)
      if (!trans) gemm<T>(false, false, n1, n, n2, T(-1), a12, lda, b2, ldb, T(1), b1, ldb, s, 0, abort_flag);
      else gemm<T>(true, false, n1, n, n2, T(-1), a21, lda, b2, ldb, T(1), b1, ldb, s, 0, abort_flag);
      rec<T>(right, lower, trans, unit, n1, n, a11, lda, b1, ldb, s, abort_flag, ic);
    }
  } else {
    T* b1 = b;
    T* b2 = b + n1 * ldb;
    if (eff_lower) {
      // X = [X1 X2], op(A) = [A11 0; op21 A22]: X2 first
      rec<T>(right, lower, trans, unit, m, n2, a22, lda, b2, ldb, s, abort_flag, ic);
      if (!trans) gemm<T>(false, false, m, n1, n2, T(-1), b2, ldb, a21, lda, T(1), b1, ldb, s, 0, abort_flag);
      else gemm<T>(false, true, m, n1, n2, T(-1), b2, ldb, a12, lda, T(1), b1, ldb, s, 0, abort_flag);
      rec<T>(right, lower, trans, unit, m, n1, a11, lda, b1, ldb, s, abort_flag, ic);
    } else {
      rec<T>(right, lower, trans, unit, m, n1, a11, lda, b1, ldb, s, abort_flag, ic);
      if (!trans) gemm<T>(false, false, m, n2, n1, T(-1), b1, ldb, a12, lda, T(1), b2, ldb, s, 0, abort_flag);
      else gemm<T>(false, true, m, n2, n1, T(-1), b1, ldb, a21, lda, T(1), b2, ldb, s, 0, abort_flag);
      rec<T>(right, lower, trans, unit, m, n2, a22, lda, b2, ldb, s, abort_flag, ic);
    }
  }
}

template <typename T>
void trsm(bool right, bool lower, bool trans, bool unit, long m, long n, T alpha, const T* a, long lda, T* b,
          long ldb, hipStream_t s, const int* abort_flag) {
  if (m <= 0 || n <= 0) return;
  if (alpha != T(1)) scale<T>(m, n, alpha, b, ldb, s);
  if (alpha == T(0)) return;
  InvCtx<T> ic;
  DevBuf invb;
  if (right) {
    // M = op(A): upper when (upper, no transpose) or (lower, transpose)
    const long nblk = (n + TRSM_NB - 1) / TRSM_NB;
    invb = DevBuf(sizeof(T) * (size_t)nblk * TRSM_NB * TRSM_NB);
    trsm_inv_diag_kernel<T><<<(unsigned)nblk, 256, 0, s>>>(n, a, lda, trans, lower == trans, unit, invb.as<T>(),
                                                           abort_flag);
    OBA_HIP(hipGetLastError());
    ic.a0 = a;
    ic.lda = lda;
    ic.inv = invb.as<T>();
  }
  rec<T>(right, lower, trans, unit, m, n, a, lda, b, ldb, s, abort_flag, ic);
}

template <typename T>
size_t trsm_right_inv_bytes(long n) {
  return sizeof(T) * (size_t)((n + TRSM_NB - 1) / TRSM_NB) * TRSM_NB * TRSM_NB;
}

template <typename T>
void trsm_right_prepare(bool lower, bool trans, bool unit, long n, const T* a, long lda, T* inv, hipStream_t s) {
  if (n <= 0) return;
  const long nblk = (n + TRSM_NB - 1) / TRSM_NB;
  trsm_inv_diag_kernel<T><<<(unsigned)nblk, 256, 0, s>>>(n, a, lda, trans, lower == trans, unit, inv, nullptr);
  OBA_HIP(hipGetLastError());
}

template <typename T>
void trsm_right_prepared(bool lower, bool trans, bool unit, long m, long n, const T* a, long lda, T* b, long ldb,
                         const T* inv, hipStream_t s, const int* abort_flag) {
  if (m <= 0 || n <= 0) return;
  InvCtx<T> ic;
  ic.a0 = a;
  ic.lda = lda;
  ic.inv = inv;
  rec<T>(true, lower, trans, unit, m, n, a, lda, b, ldb, s, abort_flag, ic);
}

template size_t trsm_right_inv_bytes<double>(long);
template size_t trsm_right_inv_bytes<float>(long);
template void trsm_right_prepare<double>(bool, bool, bool, long, const double*, long, double*, hipStream_t);
template void trsm_right_prepare<float>(bool, bool, bool, long, const float*, long, float*, hipStream_t);
template void trsm_right_prepared<double>(bool, bool, bool, long, long, const double*, long, double*, long,
                                          const double*, hipStream_t, const int*);
template void trsm_right_prepared<float>(bool, bool, bool, long, long, const float*, long, float*, long,
                                         const float*, hipStream_t, const int*);
template void trsm<double>(bool, bool, bool, bool, long, long, double, const double*, long, double*, long, hipStream_t, const int*);
template void trsm<float>(bool, bool, bool, bool, long, long, float, const float*, long, float*, long, hipStream_t, const int*);

}  // namespace oba
