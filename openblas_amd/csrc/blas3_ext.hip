// The other level-3 routines on the GEMM tile engine: SYRK, GEMMT, SYMM,
// TRMM (device pointers, column-major, alpha/beta semantics of the
// reference's drivers).
//
//  * SYRK  C := alpha op(A) op(A)^T + beta C on one triangle
//    (driver/level3/level3_syrk.c:150-165: beta first, beta == 0 writes
//    zeros, then nothing more for k == 0 or alpha == 0) and GEMMT
//    C := alpha op(A) op(B) + beta C on one triangle (interface/gemmt.c,
//    per column: SCAL_K by beta, then the GEMV unless alpha == 0): the
//    triangle-restricted MFMA GEMM (tiles outside the triangle are skipped,
//    elements outside it inside diagonal tiles are neither read nor
//    written).
//  * SYMM  C := alpha A B + beta C (side L) or alpha B A + beta C (side R)
//    with A symmetric in one stored triangle (driver/level3/symm_k.c runs
//    GEMM with symmetric packing): the triangle is expanded to the full
//    matrix in a pool buffer, then one GEMM (the expansion moves n^2
//    elements, the product 2 m n k flops).
//  * TRMM  B := alpha op(A) B (side L) or alpha B op(A) (side R), A
//    triangular (driver/level3/trmm_L.c, trmm_R.c: B scaled by alpha first,
//    alpha == 0 writes zeros and stops): recursive halving on a 64 grid;
//    the off-diagonal products are GEMMs that update a half of B with the
//    other half before that half is overwritten, the 64-wide diagonal
//    blocks are one kernel that reads only the stored triangle (no product
//    with the unstored zeros: an Inf in B meets no 0 * Inf that the
//    reference does not form).
#include "kernels.h"

namespace oba {

constexpr int X3_B = 64;   // TRMM leaf size

// ---- triangle scale: C(i,j) *= beta on i >= j (lower) or i <= j (upper) ----
template <typename T>
__global__ void tri_scale_kernel(long n, bool lower, T beta, T* __restrict__ c, long ldc) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long j = blockIdx.y;
  if (i >= n || j >= n) return;
  if (lower ? i < j : i > j) return;
  T* p = c + i + j * ldc;
  *p = beta == T(0) ? T(0) : beta * *p;
}

template <typename T>
void tri_scale(bool lower, long n, T beta, T* c, long ldc, hipStream_t s) {
  if (n <= 0 || beta == T(1)) return;
  if (n > 65535) fatal("tri_scale: order beyond one launch's grid", nullptr);
  tri_scale_kernel<T><<<dim3((unsigned)((n + 255) / 256), (unsigned)n), 256, 0, s>>>(n, lower, beta, c, ldc);
  OBA_HIP(hipGetLastError());
}

template <typename T>
void syrk_device(bool lower, bool trans, long n, long k, T alpha, const T* a, long lda, T beta, T* c, long ldc,
                 hipStream_t s) {
  if (n <= 0) return;
  if (k <= 0 || alpha == T(0)) { tri_scale<T>(lower, n, beta, c, ldc, s); return; }
  // op(A) op(A)^T: N -> A (n x k) times A^T; T -> A^T (n x k) times A
  gemm<T>(trans, !trans, n, n, k, alpha, a, lda, a, lda, beta, c, ldc, s, lower ? 1 : 2);
}

template <typename T>
void gemmt_device(bool lower, bool ta, bool tb, long m, long k, T alpha, const T* a, long lda, const T* b, long ldb,
                  T beta, T* c, long ldc, hipStream_t s) {
  if (m <= 0) return;
  if (k <= 0 || alpha == T(0)) { tri_scale<T>(lower, m, beta, c, ldc, s); return; }
  gemm<T>(ta, tb, m, m, k, alpha, a, lda, b, ldb, beta, c, ldc, s, lower ? 1 : 2);
}

// ---- SYMM: the full symmetric matrix from one stored triangle --------------
template <typename T>
__global__ __launch_bounds__(256) void sym_expand_kernel(long n, bool lower, const T* __restrict__ a, long lda,
                                                         T* __restrict__ f, long ldf) {
  __shared__ T tile[32][33];
  const long r0 = (long)blockIdx.x * 32, c0 = (long)blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  // tiles on the stored side are copied; the others are the transposes of
  // their mirror tiles (read through LDS so both accesses stay coalesced)
  const bool stored = lower ? r0 >= c0 : r0 <= c0;
  if (r0 == c0) {
    for (int q = ty; q < 32; q += 8) {
      const long i = r0 + tx, j = c0 + q;
      if (i < n && j < n) {
        const bool in = lower ? i >= j : i <= j;
        f[i + j * ldf] = in ? a[i + j * lda] : a[j + i * lda];
      }
    }
    return;
  }
  if (stored) {
    for (int q = ty; q < 32; q += 8) {
      const long i = r0 + tx, j = c0 + q;
      if (i < n && j < n) f[i + j * ldf] = a[i + j * lda];
    }
    return;
  }
  // f(r0.., c0..) = a(c0.., r0..)^T
  for (int q = ty; q < 32; q += 8) {
    const long i = c0 + tx, j = r0 + q;   // element (i, j) of the mirror tile
    if (i < n && j < n) tile[q][tx] = a[i + j * lda];
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    const long i = r0 + tx, j = c0 + q;   // f(i, j) = a(j, i) = tile[i - r0][j - c0]
    if (i < n && j < n) f[i + j * ldf] = tile[tx][q];
  }
}

template <typename T>
void symm_device(bool right, bool lower, long m, long n, T alpha, const T* a, long lda, const T* b, long ldb, T beta,
                 T* c, long ldc, hipStream_t s) {
  if (m <= 0 || n <= 0) return;
  const long na = right ? n : m;
  if (alpha == T(0)) { scale<T>(m, n, beta, c, ldc, s); return; }
  const long ldf = (na + 15) & ~15L;
  DevBuf full(sizeof(T) * (size_t)ldf * (size_t)na);
  const unsigned g = (unsigned)((na + 31) / 32);
  sym_expand_kernel<T><<<dim3(g, g), 256, 0, s>>>(na, lower, a, lda, full.as<T>(), ldf);
  OBA_HIP(hipGetLastError());
  if (!right) gemm<T>(false, false, m, n, m, alpha, full.as<T>(), ldf, b, ldb, beta, c, ldc, s);
  else gemm<T>(false, false, m, n, n, alpha, b, ldb, full.as<T>(), ldf, beta, c, ldc, s);
}

// ---- TRMM ---------------------------------------------------------------------
// Leaf: B(:, slab) := op(T) B(:, slab) (left, T nb x nb at a) or B(slab, :) :=
// B(slab, :) op(T) (right), nb <= 64; one workgroup per 64-wide slab of B's
// other dimension.  T is read from the stored triangle only (the diagonal
// taken as 1 when unit); op(T)(i,k) = trans ? A(k,i) : A(i,k).
template <typename T>
__global__ __launch_bounds__(256) void trmm_leaf_kernel(bool right, bool lower, bool trans, bool unit, int nb,
                                                        long other, const T* __restrict__ a, long lda,
                                                        T* __restrict__ b, long ldb) {
  __shared__ T tt[X3_B][X3_B + 1];   // op(T)(i, k) at tt[i][k]
  __shared__ T bs[X3_B][X3_B + 1];   // left: bs[k][col] = B(k, col); right: bs[row][k] = B(row, k)
  const long s0 = (long)blockIdx.x * X3_B;
  const int t = threadIdx.x;
  // op(A) is lower iff (lower, no transpose) or (upper, transposed)
  const bool oplower = lower != trans;
  for (int idx = t; idx < X3_B * X3_B; idx += 256) {
    const int i = idx & 63, k = idx >> 6;   // i fastest: A(i,k) reads coalesced when !trans
    T v = T(0);
    if (i < nb && k < nb) {
      const bool in = oplower ? i >= k : i <= k;
      if (i == k && unit) v = T(1);
      else if (in) v = trans ? a[k + (long)i * lda] : a[i + (long)k * lda];
    }
    tt[i][k] = v;
  }
  for (int idx = t; idx < X3_B * X3_B; idx += 256) {
    const int f = idx & 63, g = idx >> 6;   // f along B's contiguous index
    if (!right) {   // B(f, s0 + g), f < nb
      const long col = s0 + g;
      bs[f][g] = (f < nb && col < other) ? b[f + col * ldb] : T(0);
    } else {        // B(s0 + f, g), g < nb
      const long row = s0 + f;
      bs[f][g] = (row < other && g < nb) ? b[row + (long)g * ldb] : T(0);
    }
  }
  __syncthreads();
  // thread: one output row index r (0..63) and 16 consecutive columns
  const int r = t & 63, c0 = (t >> 6) * 16;
  T acc[16];
#pragma unroll
  for (int q = 0; q < 16; q++) acc[q] = T(0);
  // only the stored triangle's terms are formed (op(T)(i, k) with k <= i
  // when op(T) is lower, k >= i when upper)
  if (!right) {   // out(r, col) = sum_k op(T)(r, k) B(k, col)
    const int k0 = oplower ? 0 : r, k1 = oplower ? min(r + 1, nb) : nb;
    for (int k = k0; k < k1; k++) {
      const T tv = tt[r][k];
#pragma unroll
      for (int q = 0; q < 16; q++) acc[q] += tv * bs[k][c0 + q];
    }
  } else {        // out(row r, col c) = sum_k B(r, k) op(T)(k, c)
    for (int k = 0; k < nb; k++) {
      const T bv = bs[r][k];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int c = c0 + q;
        if (oplower ? k >= c : k <= c) acc[q] += bv * tt[k][c];
      }
    }
  }
  __syncthreads();   // every read of bs happens before B is overwritten below
#pragma unroll
  for (int q = 0; q < 16; q++) bs[r][c0 + q] = acc[q];   // bs reused as the output tile (row r)
  __syncthreads();
  for (int idx = t; idx < X3_B * X3_B; idx += 256) {
    const int f = idx & 63, g = idx >> 6;
    if (!right) {   // out(f, g) -> B(f, s0 + g)
      const long col = s0 + g;
      if (f < nb && col < other) b[f + col * ldb] = bs[f][g];
    } else {        // out(f, g) -> B(s0 + f, g)
      const long row = s0 + f;
      if (row < other && g < nb) b[row + (long)g * ldb] = bs[f][g];
    }
  }
}

// B (m x n) := op(A) B (left, A m x m) or B op(A) (right, A n x n), alpha = 1
template <typename T>
static void trmm_rec(bool right, bool lower, bool trans, bool unit, long m, long n, const T* a, long lda, T* b,
                     long ldb, hipStream_t s) {
  const long na = right ? n : m;
  if (na <= 0 || (right ? m : n) <= 0) return;
  if (na <= X3_B) {
    const long other = right ? m : n;
    trmm_leaf_kernel<T><<<(unsigned)((other + X3_B - 1) / X3_B), 256, 0, s>>>(right, lower, trans, unit, (int)na,
                                                                             other, a, lda, b, ldb);
    OBA_HIP(hipGetLastError());
    return;
  }
  const long n1 = ((na / 2) + X3_B - 1) / X3_B * X3_B, n2 = na - n1;
  const T* a11 = a;
  const T* a22 = a + n1 + n1 * lda;
  // op(A) = [T11 T12; T21 T22]: T21 = trans ? A12^T : A21, T12 = trans ? A21^T : A12
  const T* a21 = a + n1;          // A(n1:, :n1)
  const T* a12 = a + n1 * lda;    // A(:n1, n1:)
  const bool oplower = lower != trans;
  if (!right) {
    T* b1 = b;
    T* b2 = b + n1;
    if (oplower) {   // B2 <- T22 B2 + T21 B1, then B1 <- T11 B1
      trmm_rec<T>(false, lower, trans, unit, n2, n, a22, lda, b2, ldb, s);
      if (!trans) gemm<T>(false, false, n2, n, n1, T(1), a21, lda, b1, ldb, T(1), b2, ldb, s);
      else gemm<T>(true, false, n2, n, n1, T(1), a12, lda, b1, ldb, T(1), b2, ldb, s);
      trmm_rec<T>(false, lower, trans, unit, n1, n, a11, lda, b1, ldb, s);
    } else {         // B1 <- T11 B1 + T12 B2, then B2 <- T22 B2
      trmm_rec<T>(false, lower, trans, unit, n1, n, a11, lda, b1, ldb, s);
      if (!trans) gemm<T>(false, false, n1, n, n2, T(1), a12, lda, b2, ldb, T(1), b1, ldb, s);
      else gemm<T>(true, false, n1, n, n2, T(1), a21, lda, b2, ldb, T(1), b1, ldb, s);
      trmm_rec<T>(false, lower, trans, unit, n2, n, a22, lda, b2, ldb, s);
    }
  } else {
    T* b1 = b;
    T* b2 = b + n1 * ldb;
    if (oplower) {   // [B1 B2] [T11 0; T21 T22]: B1 <- B1 T11 + B2 T21, then B2 <- B2 T22
      trmm_rec<T>(true, lower, trans, unit, m, n1, a11, lda, b1, ldb, s);
      if (!trans) gemm<T>(false, false, m, n1, n2, T(1), b2, ldb, a21, lda, T(1), b1, ldb, s);
      else gemm<T>(false, true, m, n1, n2, T(1), b2, ldb, a12, lda, T(1), b1, ldb, s);
      trmm_rec<T>(true, lower, trans, unit, m, n2, a22, lda, b2, ldb, s);
    } else {         // [B1 B2] [T11 T12; 0 T22]: B2 <- B2 T22 + B1 T12, then B1 <- B1 T11
      trmm_rec<T>(true, lower, trans, unit, m, n2, a22, lda, b2, ldb, s);
      if (!trans) gemm<T>(false, false, m, n2, n1, T(1), b1, ldb, a12, lda, T(1), b2, ldb, s);
      else gemm<T>(false, true, m, n2, n1, T(1), b1, ldb, a21, lda, T(1), b2, ldb, s);
      trmm_rec<T>(true, lower, trans, unit, m, n1, a11, lda, b1, ldb, s);
    }
  }
}

template <typename T>
void trmm_device(bool right, bool lower, bool trans, bool unit, long m, long n, T alpha, const T* a, long lda, T* b,
                 long ldb, hipStream_t s) {
  if (m <= 0 || n <= 0) return;
  if (alpha != T(1)) scale<T>(m, n, alpha, b, ldb, s);   // trmm_L.c: GEMM_BETA(alpha) first
  if (alpha == T(0)) return;
  trmm_rec<T>(right, lower, trans, unit, m, n, a, lda, b, ldb, s);
}

#define X3_INST(T)                                                                                                \
  template void tri_scale<T>(bool, long, T, T*, long, hipStream_t);                                            \
  template void syrk_device<T>(bool, bool, long, long, T, const T*, long, T, T*, long, hipStream_t);           \
  template void gemmt_device<T>(bool, bool, bool, long, long, T, const T*, long, const T*, long, T, T*, long,  \
                                hipStream_t);                                                                  \
  template void symm_device<T>(bool, bool, long, long, T, const T*, long, const T*, long, T, T*, long,         \
                               hipStream_t);                                                                   \
  template void trmm_device<T>(bool, bool, bool, bool, long, long, T, const T*, long, T*, long, hipStream_t);
X3_INST(double)
X3_INST(float)

}  // namespace oba
