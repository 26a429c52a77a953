// INT8-slice emulated FP64 GEMM (device pointers); see ozaki.hip.
#pragma once
#include "common.h"

namespace oba {

// Returns the slice count used (slices <= 0: chosen from the operands'
// exponent spread).
int ozaki_dgemm(bool ta, bool tb, long m, long n, long k, double alpha, const double* a, long lda, const double* b,
                long ldb, double beta, double* c, long ldc, int slices, hipStream_t s);

// Scheme II (integer moduli + Chinese remainder theorem); returns the
// modulus count used (moduli <= 0: chosen from the exponent spread).
int ozaki2_dgemm(bool ta, bool tb, long m, long n, long k, double alpha, const double* a, long lda, const double* b,
                 long ldb, double beta, double* c, long ldc, int moduli, hipStream_t s);

}  // namespace oba
