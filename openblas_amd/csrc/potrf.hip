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
// are factored in LDS by one workgroup.  A failing leaf raises a device flag
// that turns every later kernel of the same call into a no-op, so the call
// stops where the reference stops without a host round trip per leaf.
#include "kernels.h"
#include <cmath>

namespace oba {

constexpr int PO_NB = 64;

// Factor the nb x nb diagonal block at a (lower: E(i,k) = a[i + k*lda];
// upper: E(i,k) = a[k + i*lda], i >= k).  info/flag: global column offset j0.
template <typename T>
__global__ __launch_bounds__(256) void potrf_leaf_kernel(int nb, T* a, long lda, bool upper, long j0, int* info) {
  __shared__ T E[PO_NB][PO_NB + 1];
  __shared__ int s_fail;
  if (*info) return;
  const int t = threadIdx.x;
  for (int idx = t; idx < PO_NB * PO_NB; idx += 256) {
    int i = idx & 63, k = idx >> 6;
    T v = T(0);
    if (i < nb && k < nb && i >= k) v = upper ? a[k + (long)i * lda] : a[i + (long)k * lda];
    E[i][k] = v;
  }
  if (t == 0) s_fail = 0;
  __syncthreads();
  int done = nb;
  for (int j = 0; j < nb; j++) {
    if (t == 0) {
      T ajj = E[j][j];
      if (!(ajj > T(0)) && !isnan(ajj)) {   // ajj <= 0: store and stop
        s_fail = j + 1;
      } else {
        E[j][j] = sqrt(ajj);
      }
    }
    __syncthreads();
    if (s_fail) { done = j; break; }
    const T rinv = T(1) / E[j][j];
    for (int i = j + 1 + t; i < nb; i += 256) E[i][j] *= rinv;
    __syncthreads();
    // trailing lower update: E(i,k) -= E(i,j)*E(k,j), j < k <= i
    const int rem = nb - j - 1;
    for (int idx = t; idx < rem * rem; idx += 256) {
      int i = j + 1 + idx % rem, k = j + 1 + idx / rem;
      if (i >= k) E[i][k] -= E[i][j] * E[k][j];
    }
    __syncthreads();
  }
  // write back columns [0, done] (the failing diagonal value included)
  const int ncols = s_fail ? s_fail : nb;
  for (int idx = t; idx < PO_NB * PO_NB; idx += 256) {
    int i = idx & 63, k = idx >> 6;
    if (i < nb && k < ncols && i >= k) {
      if (s_fail && k == ncols - 1 && i > k) continue;   // failing column: only the diagonal
      T v = E[i][k];
      if (upper) a[k + (long)i * lda] = v;
      else a[i + (long)k * lda] = v;
    }
  }
  (void)done;
  if (t == 0 && s_fail) *info = (int)(j0 + s_fail);
}

template <typename T>
static void potrf_rec(bool lower, long j0, long nc, T* a, long lda, int* info, hipStream_t s) {
  T* a11 = a + j0 + j0 * lda;
  if (nc <= PO_NB) {
    potrf_leaf_kernel<T><<<1, 256, 0, s>>>((int)nc, a11, lda, !lower, j0, info);
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

template <typename T>
int potrf(bool lower, long n, T* a, long lda, hipStream_t s) {
  if (n <= 0) return 0;
  DevBuf ib(sizeof(int) * 64);
  int* info = ib.as<int>();
  OBA_HIP(hipMemsetAsync(info, 0, sizeof(int), s));
  potrf_rec<T>(lower, 0, n, a, lda, info, s);
  int h = 0;
  OBA_HIP(hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, s));
  OBA_HIP(hipStreamSynchronize(s));
  return h;
}

template int potrf<double>(bool, long, double*, long, hipStream_t);
template int potrf<float>(bool, long, float*, long, hipStream_t);

}  // namespace oba
