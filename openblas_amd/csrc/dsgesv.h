// Mixed-precision solve entries shared by lapack_api.cpp and dsgesv.hip.
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

// Cholesky-based mixed-precision solve (dsposv); only the `lower` (else
// upper) triangle of a is read, and written only on the FP64 fallback.
void dsposv_device(bool lower, long n, long nrhs, double* a, long lda, const double* b, long ldb, double* x, long ldx,
                   int* iter, int* info, DsgesvStats* stats, hipStream_t s);

// Solve with a Cholesky factor (potrs): lower L L^T, else U^T U; in place on b.
template <typename T>
void potrs_device(bool lower, long n, long nrhs, const T* f, long ldf, T* b, long ldb, hipStream_t s);

}  // namespace oba
