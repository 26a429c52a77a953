// Cholesky factorization (potrf, lower and upper), device pointers.
//
// Reference semantics (lapack/potf2/potf2_L.c:50-97, potrf_L_single.c:81-236,
// potrf_U_single.c:79-193): only the selected triangle is read or written;
// at column j, ajj = a_jj - sum_k l_jk^2; if ajj <= 0 the value is stored in
// a_jj and the routine returns info = j+1 without touching later columns
// (a NaN ajj is not caught, exactly like the reference's `ajj <= 0` test);
// otherwise a_jj = sqrt(ajj) and the column below is scaled by 1/a_jj.
//
// MI355X design: recursion on halves (64-column grid): L21 <- A21*L11^-T is a
// right-side TRSM, A22 -= L21*L21^T is an MFMA GEMM restricted to the lower
// triangle (the role of the reference's SYRK kernel); 64x64 diagonal leaves
// are factored by one wavefront with the rows in registers.  Large lower
// factorizations run blocked with look-ahead (potrf_blocked).  A failing leaf raises a device flag
// that turns every later kernel of the same call into a no-op, so the call
// stops where the reference stops without a host round trip per leaf.
#include "kernels.h"
#include <cmath>

namespace oba {

constexpr int PO_NB = 64;

__device__ inline float po_rl(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ inline double po_rl(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// Factor the nb x nb (nb <= 64) diagonal block at a (lower: E(i,k) =
// a[i + k*lda]; upper: E(i,k) = a[k + i*lda], i >= k); info and the failure
// flag take the global column offset j0.  One wavefront, no barrier: lane i
// loads row i of the block straight into registers (all loads in flight at
// once; a staging loop through LDS waited for each load in turn, and a
// four-wavefront form with two barriers per column took ~73 us FP32 per
// leaf), x[k] holding column 16q + k while the column loop is in its
// quarter q (the row is shifted down by 16 registers between quarters, so
// every register index in the unrolled 16-column body is static).  Column
// j: the diagonal comes from lane j (readlane); ajj <= 0 (not NaN) is
// stored as it is and stops the leaf; otherwise the column is scaled by
// 1/l_jj (potf2_L.c's SCAL_K), its multipliers go through LDS (written and
// read back by the same wavefront) and every lane subtracts l_ij l_kj from
// its entries right of j (entries above the diagonal take garbage and are
// never stored; LDS past column 63 reads zeros).  A quarter's 16 columns
// are final when it ends and are stored then.
template <typename T>
__global__ __launch_bounds__(64) void potrf_leaf1_kernel(int nb, T* a, long lda, bool upper, long j0, int* info) {
  __shared__ T col[2 * PO_NB];
  if (*info) return;
  const int lane = threadIdx.x;
  const int rl = min(lane, nb - 1);   // in-bounds row for the unconditional loads
  // row `lane` straight into registers: all 64 loads in flight at once
  // (lower: E(i,k) = a[i + k*lda], coalesced across lanes; upper: a[k + i*lda])
  T x[PO_NB];
#pragma unroll
  for (int k = 0; k < PO_NB; k++) {
    const int kc = min(k, nb - 1);
    const T v = upper ? a[kc + (long)rl * lda] : a[rl + (long)kc * lda];
    x[k] = (lane < nb && k < nb && k <= lane) ? v : T(0);
  }
  col[PO_NB + lane] = T(0);
  int fail = 0;
#pragma unroll 1
  for (int q = 0; q < 4; q++) {
    int done = 16;   // columns of this quarter that are final in x[0..15]
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
      const int j = 16 * q + jj;
      if (j >= nb || fail) { done = min(done, jj); continue; }   // uniform
      const T d = po_rl(x[jj], j);
      if (!(d > T(0)) && !isnan(d)) {   // ajj <= 0: stored as it is, stop
        fail = j + 1;
        done = jj + 1;                   // the failing column keeps its diagonal
        continue;
      }
      const T ljj = sqrt(d);
      const T rinv = T(1) / ljj;
      const T lij = lane == j ? ljj : x[jj] * rinv;
      if (lane >= j) x[jj] = lij;
      const T li = lane > j ? lij : T(0);
      col[lane] = li;
#pragma unroll
      for (int k = jj + 1; k < PO_NB; k++) x[k] -= li * col[16 * q + k];
    }
    // the quarter's final columns back to memory (of a failing column
    // only its diagonal)
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int c = 16 * q + k;
      const bool st = k < done && lane < nb && c <= lane && !(fail && c == fail - 1 && lane > c);
      if (st) {
        if (upper) a[c + (long)lane * lda] = x[k];
        else a[lane + (long)c * lda] = x[k];
      }
    }
    if (fail || 16 * (q + 1) >= nb) break;
#pragma unroll
    for (int k = 0; k < PO_NB - 16; k++) x[k] = x[k + 16];
#pragma unroll
    for (int k = PO_NB - 16; k < PO_NB; k++) x[k] = T(0);
  }
  if (lane == 0 && fail) *info = (int)(j0 + fail);
}

template <typename T>
static void potrf_rec(bool lower, long j0, long nc, T* a, long lda, int* info, hipStream_t s) {
  T* a11 = a + j0 + j0 * lda;
  if (nc <= PO_NB) {
    potrf_leaf1_kernel<T><<<1, 64, 0, s>>>((int)nc, a11, lda, !lower, j0, info);
    OBA_HIP(hipGetLastError());
    return;
  }
  long n1 = ((nc / 2) + PO_NB - 1) / PO_NB * PO_NB;
  if (n1 >= nc) n1 = nc - PO_NB;
  const long n2 = nc - n1;
  potrf_rec<T>(lower, j0, n1, a, lda, info, s);
  T* a22 = a + (j0 + n1) + (j0 + n1) * lda;
  if (lower) {
    T* a21 = a + (j0 + n1) + j0 * lda;
    trsm<T>(true, true, true, false, n2, n1, T(1), a11, lda, a21, lda, s, info);   // A21 L11^-T
    gemm<T>(false, true, n2, n2, n1, T(-1), a21, lda, a21, lda, T(1), a22, lda, s, 1, info);
  } else {
    T* a12 = a + j0 + (j0 + n1) * lda;
    trsm<T>(false, false, true, false, n1, n2, T(1), a11, lda, a12, lda, s, info);  // U11^-T A12
    gemm<T>(true, false, n2, n2, n1, T(-1), a12, lda, a12, lda, T(1), a22, lda, s, 2, info);
  }
  potrf_rec<T>(lower, j0 + n1, n2, a, lda, info, s);
}

// Blocked right-looking Cholesky (lower) with look-ahead, for large n, as
// the blocked LU: the columns are cut into blocks of nb.  On the caller's
// stream: block k's diagonal block (the recursion above), then the rows of
// its panel that belong to block k+1 (right-side TRSM with the 64x64
// diagonal inverses computed once) and block k+1's diagonal update; block
// k+1 is factored next.  Meanwhile the bulk stream (CU mask, all but the
// reserved CUs) solves the rest of the panel and updates the trailing
// matrix (one GEMM for block k+1's column below its diagonal block, one
// lower-restricted GEMM for the rest), and the caller's stream waits for it
// before it touches those rows.  Every kernel takes the failure flag: after
// a non-positive pivot the rest of the call is a no-op, as in the
// recursion.  The trailing sums run over blocks instead of recursion
// halves: a different association of the same products.
template <typename T>
static void potrf_blocked(long n, T* a, long lda, int* info, long nb, hipStream_t s) {
  Runtime& r = rt();
  hipStream_t M = s, X = bulk_stream();
  DevBuf invb[2] = {DevBuf(trsm_right_inv_bytes<T>(nb)), DevBuf(trsm_right_inv_bytes<T>(nb))};
  invb[0].use_on(X);   // read by the bulk stream's solves
  invb[1].use_on(X);
  auto at = [&](long i, long j) { return a + i + j * lda; };
  potrf_rec<T>(true, 0, std::min(nb, n), a, lda, info, M);
  bool pending = false;
  for (long jb = 0; jb < n; jb += nb) {
    const long w = std::min(nb, n - jb), jn = jb + w;
    if (pending) {   // the trailing update by the previous block has finished
      OBA_HIP(hipStreamWaitEvent(M, r.ev_b, 0));
      pending = false;
    }
    if (jn >= n) break;
    const long w1 = std::min(nb, n - jn);
    T* inv = invb[(jb / nb) & 1].as<T>();
    const T* l11 = at(jb, jb);
    trsm_right_prepare<T>(true, true, false, w, l11, lda, inv, M);
    trsm_right_prepared<T>(true, true, false, w1, w, l11, lda, at(jn, jb), lda, inv, M, info);
    gemm<T>(false, true, w1, w1, w, T(-1), at(jn, jb), lda, at(jn, jb), lda, T(1), at(jn, jn), lda, M, 1, info);
    const long rest = n - (jn + w1);
    if (rest > 0) {
      OBA_HIP(hipEventRecord(r.ev_a, M));
      OBA_HIP(hipStreamWaitEvent(X, r.ev_a, 0));
      trsm_right_prepared<T>(true, true, false, rest, w, l11, lda, at(jn + w1, jb), lda, inv, X, info);
      gemm<T>(false, true, rest, w1, w, T(-1), at(jn + w1, jb), lda, at(jn, jb), lda, T(1), at(jn + w1, jn), lda, X, 0,
              info);
      gemm<T>(false, true, rest, rest, w, T(-1), at(jn + w1, jb), lda, at(jn + w1, jb), lda, T(1),
              at(jn + w1, jn + w1), lda, X, 1, info);
      OBA_HIP(hipEventRecord(r.ev_b, X));
      pending = true;
    }
    potrf_rec<T>(true, jn, w1, a, lda, info, M);
  }
  if (pending) OBA_HIP(hipStreamWaitEvent(M, r.ev_b, 0));
}

// block width of the look-ahead driver (OPENBLAS_AMD_PO_NB; 0 turns it off);
// it is used from 4 blocks on
template <typename T>
static long potrf_block_nb() {
  // swept at n = 16384 with the one-wavefront leaves (the chain no longer
  // the bound, the trailing updates prefer a deeper K): 256 / 320 / 384 /
  // 448 / 512 / 640 / 768 gave dpotrf 70.2 / 65.0 / 61.8 / 62.0 / 71.8 /
  // 60.5 / 64.8 ms, spotrf 41.3 / 39.3 / 38.6 / 38.3 / 43.0 / 39.5 / 41.5
  long v = 384;
  if (const char* e = getenv("OPENBLAS_AMD_PO_NB")) v = atol(e);
  return v;
}

template <typename T>
int potrf(bool lower, long n, T* a, long lda, hipStream_t s) {
  if (n <= 0) return 0;
  DevBuf ib(sizeof(int) * 64);
  int* info = ib.as<int>();
  OBA_HIP(hipMemsetAsync(info, 0, sizeof(int), s));
  const long nb = potrf_block_nb<T>();
  if (lower && nb > 0 && n >= 4 * nb) potrf_blocked<T>(n, a, lda, info, nb, s);
  else potrf_rec<T>(lower, 0, n, a, lda, info, s);
  int h = 0;
  OBA_HIP(hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, s));
  OBA_HIP(hipStreamSynchronize(s));
  return h;
}

template int potrf<double>(bool, long, double*, long, hipStream_t);
template int potrf<float>(bool, long, float*, long, hipStream_t);

}  // namespace oba
