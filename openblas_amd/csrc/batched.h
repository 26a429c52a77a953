// Batched engine: many independent small problems in one persistent launch.
#pragma once
#include "common.h"
#include <vector>

namespace oba {

// One problem of a batch (device table entry).  Unused fields are zero.
struct BDesc {
  void* a;
  void* b;
  void* c;
  int* ipiv;
  int m, n, k;          // gemm: m, n, k; getrf: m, n; getrs: n, nrhs (in k); potrf: n
  int lda, ldb, ldc;
  int idx;              // position of the problem in the caller's arrays (info slot)
  int flags;            // gemm: bit0 op(A) transposed, bit1 op(B) transposed
  double alpha, beta;   // gemm scalars (FP32 batches convert exactly)
};

enum BatchKind { BK_GEMM = 0, BK_GETRF = 1, BK_GETRS = 2, BK_POTRF = 3 };

// Batched drivers on device-resident problem tables built by the callers in
// batched_api.cpp.  `probs` must already be in launch order (largest first);
// info (getrf/getrs/potrf) is returned per problem in info_host[idx].
template <typename T>
void gemm_batched(int ta, int tb, std::vector<BDesc>& probs);
template <typename T>
void getrf_batched(std::vector<BDesc>& probs, int* info_host);
template <typename T>
void getrs_batched(bool trans, std::vector<BDesc>& probs, int* info_host);
template <typename T>
void potrf_batched(bool lower, std::vector<BDesc>& probs, int* info_host);

// The problems of one checked cblas_?gemm_batch call (column-major view):
// stages host operands, runs small problems through gemm_batched and large
// ones through the single GEMM, copies host outputs back.
template <typename T>
void gemm_batch_run(std::vector<BDesc>& probs);

}  // namespace oba
