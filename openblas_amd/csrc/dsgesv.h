// Mixed-precision solve entry shared by lapack_api.cpp and dsgesv.hip.
#pragma once
#include "common.h"

namespace oba {

struct DsgesvStats {
  bool sgetrf_done = false;
  bool fallback = false;
};

// Device pointers; a is overwritten only on the FP64 fallback.
void dsgesv_device(long n, long nrhs, double* a, long lda, int* ipiv_dev, const double* b, long ldb, double* x,
                   long ldx, int* iter, int* info, DsgesvStats* stats, hipStream_t s);

}  // namespace oba
