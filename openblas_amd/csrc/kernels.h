// Device-pointer launchers shared between the modules of libopenblas_amd.
// All take device pointers, column-major storage, and enqueue on `s`.
#pragma once
#include "common.h"

namespace oba {

// C <- alpha*op(A)*op(B) + beta*C.  beta == 0 overwrites C without reading it
// (kernel/generic/gemm_beta.c:52); k == 0 or alpha == 0 only scales C
// (driver/level3/level3.c:229-264).
// tri = 1 / 2 restricts the update to the lower / upper triangle of C
// (the SYRK-kernel role of driver/level3/syrk_kernel.c:56 in potrf);
// a non-null abort_flag that reads nonzero on the device turns the launch
// into a no-op (used to stop potrf after a non-positive pivot).
template <typename T>
void gemm(bool ta, bool tb, long m, long n, long k, T alpha, const T* a, long lda, const T* b,
          long ldb, T beta, T* c, long ldc, hipStream_t s, int tri = 0,
          const int* abort_flag = nullptr);

// C <- beta*C on an m x n block (beta==0 writes exact zeros).
template <typename T>
void scale(long m, long n, T beta, T* c, long ldc, hipStream_t s);

// Triangular solve with multiple right-hand sides, all 8 side/uplo/trans/diag
// forms plus alpha (interface/trsm.c:213, driver/level3/trsm_L.c, trsm_R.c).
template <typename T>
void trsm(bool right, bool lower, bool trans, bool unit, long m, long n, T alpha, const T* a,
          long lda, T* b, long ldb, hipStream_t s, const int* abort_flag = nullptr);

// Right-side solve X <- X * op(A)^{-1} (alpha = 1) in two steps, so that the
// inverses of op(A)'s 64x64 diagonal blocks can be computed once and reused
// by several solves with the same triangle (on any stream, no allocation):
// trsm_right_inv_bytes(n) bytes for `inv`.
template <typename T>
size_t trsm_right_inv_bytes(long n);
template <typename T>
void trsm_right_prepare(bool lower, bool trans, bool unit, long n, const T* a, long lda, T* inv, hipStream_t s);
// abort_flag: a device word that, once nonzero, turns the solve's kernels
// into no-ops (a factorization that already failed).
template <typename T>
void trsm_right_prepared(bool lower, bool trans, bool unit, long m, long n, const T* a, long lda, T* b, long ldb,
                         const T* inv, hipStream_t s, const int* abort_flag = nullptr);

// LU with partial pivoting, in place; ipiv is a 1-based device int array.
// Returns info (0, or first zero pivot j>0) -- lapack/getrf/getrf_single.c.
template <typename T>
int getrf(long m, long n, T* a, long lda, int* ipiv_dev, hipStream_t s);

// Row interchanges k1..k2 (1-based, inclusive) from ipiv on n columns.
template <typename T>
void laswp(long n, T* a, long lda, long k1, long k2, const int* ipiv_dev, int incx, hipStream_t s);

// Cholesky, in place (lapack/potrf/potrf_L_single.c, potrf_U_single.c).
template <typename T>
int potrf(bool lower, long n, T* a, long lda, hipStream_t s);

}  // namespace oba
