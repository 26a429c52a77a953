// INT8-slice emulated FP64 GEMM (device pointers); see ozaki.hip.
#pragma once
#include "common.h"

namespace oba {

// Returns the slice count used (slices <= 0: chosen from the operands'
// exponent spread).
int ozaki_dgemm(bool ta, bool tb, long m, long n, long k, double alpha, const double* a, long lda, const double* b,
                long ldb, double beta, double* c, long ldc, int slices, hipStream_t s);

}  // namespace oba
