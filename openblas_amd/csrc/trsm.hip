// Triangular solve with multiple right-hand sides (all 8 side/uplo/trans
// forms, unit or non-unit diagonal), device pointers, column-major.
//
// Reference semantics: B <- alpha*B first (driver/level3/trsm_L.c:103-107;
// alpha == 0 zeroes B and returns), then op(A)*X = B (left) or X*op(A) = B
// (right), the diagonal applied as a multiplication by its reciprocal
// (kernel/generic/trsm_ltcopy_2.c:43 INV(a) = 1/a, trsm_kernel_LT.c:95).
//
// Design: recursive blocking on the triangle (split in halves on a 64-row
// grid) so that almost all flops land in large MFMA GEMM updates; the 64x64
// diagonal leaves are solved by one wavefront per right-hand-side vector:
// lane i owns x_i, the effective triangle sits in LDS column-major, and
// x_k is broadcast with v_readlane after it is final (forward or backward
// column-oriented substitution).  Right-side leaves solve rows of B, i.e.
// the transposed system, with the same kernel.
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
constexpr int TRSM_VPW = 4;   // right-hand-side vectors per wavefront

template <typename T>
__global__ __launch_bounds__(256) void trsm_leaf_kernel(int nb, long nvec, const T* __restrict__ A, long lda,
                                                        T* __restrict__ B, long inc_e, long inc_v, bool eT,
                                                        bool elower, bool unit, const int* abort_flag) {
  constexpr int LDE = TRSM_NB + 1;     // padded: conflict-free transposed stores
  __shared__ T E[TRSM_NB * LDE];       // E(i,k) at E[k*LDE + i]
  __shared__ T dinv[TRSM_NB];
  if (abort_flag && *abort_flag) return;
  // consecutive threads read consecutive addresses of A in both layouts
  for (int idx = threadIdx.x; idx < TRSM_NB * TRSM_NB; idx += blockDim.x) {
    const int f = idx & 63, s = idx >> 6;       // f: fast (contiguous) index of A
    // eT: E(i,k) = A(k,i) -> A[f + s*lda] = A(f, s) = E(s, f)
    const int i = eT ? s : f, k = eT ? f : s;
    T v = T(0);
    if (f < nb && s < nb) v = A[f + (long)s * lda];
    E[k * LDE + i] = v;
  }
  __syncthreads();
  if (threadIdx.x < TRSM_NB) {
    const int k = threadIdx.x;
    dinv[k] = (k < nb && !unit) ? T(1) / E[k * LDE + k] : T(1);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long v0 = ((long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * TRSM_VPW;
  for (int vv = 0; vv < TRSM_VPW; vv++) {
    const long v = v0 + vv;
    if (v >= nvec) return;
    T* b = B + v * inc_v;
    T x = lane < nb ? b[lane * inc_e] : T(0);
    if (elower) {
      for (int k = 0; k < nb; k++) {
        if (lane == k) x *= dinv[k];
        T xk = bcast_lane(x, k);
        if (lane > k) x -= E[k * LDE + lane] * xk;
      }
    } else {
      for (int k = nb - 1; k >= 0; k--) {
        if (lane == k) x *= dinv[k];
        T xk = bcast_lane(x, k);
        if (lane < k) x -= E[k * LDE + lane] * xk;
      }
    }
    if (lane < nb) b[lane * inc_e] = x;
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

template <typename T>
static void leaf(bool right, bool lower, bool trans, bool unit, long nb, long nrhs, const T* a, long lda, T* b,
                 long ldb, hipStream_t s, const int* abort_flag) {
  if (nb <= 0 || nrhs <= 0) return;
  // E = op(A) for left, op(A)^T for right
  const bool eT = trans != right;
  const bool elower = lower != eT;
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
  trsm_leaf_kernel<T><<<(unsigned)blocks, 64 * waves, 0, s>>>((int)nb, nrhs, a, lda, b, inc_e, inc_v, eT, elower,
                                                              unit, abort_flag);
  OBA_HIP(hipGetLastError());
}

// alpha already applied; solves in place.
template <typename T>
static void rec(bool right, bool lower, bool trans, bool unit, long m, long n, const T* a, long lda, T* b,
                long ldb, hipStream_t s, const int* abort_flag) {
  const long na = right ? n : m;   // order of the triangle
  if (na <= TRSM_NB) { leaf<T>(right, lower, trans, unit, na, right ? m : n, a, lda, b, ldb, s, abort_flag); return; }
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
      rec<T>(right, lower, trans, unit, n1, n, a11, lda, b1, ldb, s, abort_flag);
      // B2 -= op(A)21 * X1 ; op(A)21 = A21 (lower, N) or A12^T (upper, T)
      if (!trans) gemm<T>(false, false, n2, n, n1, T(-1), a21, lda, b1, ldb, T(1), b2, ldb, s, 0, abort_flag);
      else gemm<T>(true, false, n2, n, n1, T(-1), a12, lda, b1, ldb, T(1), b2, ldb, s, 0, abort_flag);
      rec<T>(right, lower, trans, unit, n2, n, a22, lda, b2, ldb, s, abort_flag);
    } else {
      rec<T>(right, lower, trans, unit, n2, n, a22, lda, b2, ldb, s, abort_flag);
      // B1 -= op(A)12 * X2 ; op(A)12 = A12 (upper, N) or A21^T (lower, T)
      if (!trans) gemm<T>(false, false, n1, n, n2, T(-1), a12, lda, b2, ldb, T(1), b1, ldb, s, 0, abort_flag);
      else gemm<T>(true, false, n1, n, n2, T(-1), a21, lda, b2, ldb, T(1), b1, ldb, s, 0, abort_flag);
      rec<T>(right, lower, trans, unit, n1, n, a11, lda, b1, ldb, s, abort_flag);
    }
  } else {
    T* b1 = b;
    T* b2 = b + n1 * ldb;
    if (eff_lower) {
      // X = [X1 X2], op(A) = [A11 0; op21 A22]: X2 first
      rec<T>(right, lower, trans, unit, m, n2, a22, lda, b2, ldb, s, abort_flag);
      if (!trans) gemm<T>(false, false, m, n1, n2, T(-1), b2, ldb, a21, lda, T(1), b1, ldb, s, 0, abort_flag);
      else gemm<T>(false, true, m, n1, n2, T(-1), b2, ldb, a12, lda, T(1), b1, ldb, s, 0, abort_flag);
      rec<T>(right, lower, trans, unit, m, n1, a11, lda, b1, ldb, s, abort_flag);
    } else {
      rec<T>(right, lower, trans, unit, m, n1, a11, lda, b1, ldb, s, abort_flag);
      if (!trans) gemm<T>(false, false, m, n2, n1, T(-1), b1, ldb, a12, lda, T(1), b2, ldb, s, 0, abort_flag);
      else gemm<T>(false, true, m, n2, n1, T(-1), b1, ldb, a21, lda, T(1), b2, ldb, s, 0, abort_flag);
      rec<T>(right, lower, trans, unit, m, n2, a22, lda, b2, ldb, s, abort_flag);
    }
  }
}

template <typename T>
void trsm(bool right, bool lower, bool trans, bool unit, long m, long n, T alpha, const T* a, long lda, T* b,
          long ldb, hipStream_t s, const int* abort_flag) {
  if (m <= 0 || n <= 0) return;
  if (alpha != T(1)) scale<T>(m, n, alpha, b, ldb, s);
  if (alpha == T(0)) return;
  rec<T>(right, lower, trans, unit, m, n, a, lda, b, ldb, s, abort_flag);
}

template void trsm<double>(bool, bool, bool, bool, long, long, double, const double*, long, double*, long, hipStream_t, const int*);
template void trsm<float>(bool, bool, bool, bool, long, long, float, const float*, long, float*, long, hipStream_t, const int*);

}  // namespace oba
