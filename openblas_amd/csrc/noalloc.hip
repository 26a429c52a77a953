// dgetrs_ for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// openblas_amd_dgetrs_noalloc has dgetrs_'s arguments, checks and result
// (interface/lapack/getrs.c:63-100 semantics; the same kernels as the
// library's dgetrs_ for NRHS <= 16: the row permutation gather and the
// look-back triangular solves), with two differences:
//  * A and B must already be in device memory (staging a host matrix would
//    need an allocation); IPIV may be host or device memory;
//  * the workspace of those kernels (the composed permutation, the staging
//    copy of B's permuted rows, the solves' ticket counters and published
//    words) lives in `scratch`, a file-scope device array of NA_BYTES,
//    instead of the library's memory pool.  N is limited to NA_MAX_N; the
//    right-hand sides are solved NA_CHUNK at a time.
// The entry point calls no hipMalloc/hipFree and no allocating routine.  The
// library's one-time set-up (streams, the pinned device-error word) happens
// on the first call of any entry point, this one included.  Every call on a
// device shares that device's array and leaves its contents behind; calls
// that overlap on one device (several host threads in asynchronous mode, or
// several replays of a captured call in flight) race on it.
#include "kernels.h"
#include "lu_internal.h"
#include <cctype>
#include <cstring>
#include <mutex>

using namespace oba;

namespace {

constexpr long NA_MAX_N = 32768;   // largest order N
constexpr long NA_CHUNK = 8;       // right-hand sides per solve launch
// scratch layout: [perm: NA_MAX_N ints][B staging: NA_MAX_N x NA_CHUNK
// doubles][solve workspace: tickets + NA_CHUNK columns of published words]
constexpr size_t NA_PERM_OFF = 0;
constexpr size_t NA_TMP_OFF = NA_PERM_OFF + sizeof(int) * NA_MAX_N;
constexpr size_t NA_WS_OFF = NA_TMP_OFF + sizeof(double) * NA_MAX_N * NA_CHUNK;
// one column's published words: one 8-byte word per 32 bits of x
constexpr size_t NA_PUB_COL = sizeof(unsigned long long) * NA_MAX_N * (sizeof(double) / 4);
constexpr size_t NA_WS_BYTES = 256 + NA_CHUNK * NA_PUB_COL;
constexpr size_t NA_BYTES = NA_WS_OFF + NA_WS_BYTES;   // 6.1 MiB
static_assert(NA_TMP_OFF % 256 == 0 && NA_WS_OFF % 256 == 0, "regions stay 256-byte aligned");

__device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEV = 64;
std::mutex g_na_mu;
void* g_na_ptr[NA_MAX_DEV] = {};
int g_na_perm[NA_MAX_N];   // host-side composition of IPIV (under the API lock)
int g_na_ipiv[NA_MAX_N];

// The device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).
void* na_scratch(int dev) {
  if (dev < 0 || dev >= NA_MAX_DEV) fatal("dgetrs_noalloc: device index beyond the scratch table", nullptr);
  std::lock_guard<std::mutex> g(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) fatal("dgetrs_noalloc: hipGetSymbolAddress(scratch) failed", hipGetErrorString(e));
    g_na_ptr[dev] = p;
  }
  return g_na_ptr[dev];
}

// zero fill (no hipMemsetAsync: a caller may capture this call in a graph)
__global__ void na_zero_kernel(unsigned long long* p, size_t words) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

void na_zero(void* p, size_t bytes, hipStream_t s) {
  const size_t words = (bytes + 7) / 8;
  const unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 1024);
  na_zero_kernel<<<blocks, 256, 0, s>>>((unsigned long long*)p, words);
  OBA_HIP(hipGetLastError());
}

}  // namespace

OBA_API void openblas_amd_dgetrs_noalloc(const char* TRANS, const int* N, const int* NRHS, const double* a,
                                         const int* LDA, const int* ipiv, double* b, const int* LDB,
                                         int* Info) {
  const bool ok = guarded([&] {
    const char tc = (char)toupper((unsigned char)*TRANS);
    int trans = -1;
    if (tc == 'N' || tc == 'R') trans = 0;
    if (tc == 'T' || tc == 'C') trans = 1;
    const long n = *N, nrhs = *NRHS, lda = *LDA, ldb = *LDB;
    int info = 0;
    if (ldb < std::max(1L, n)) info = 8;
    if (lda < std::max(1L, n)) info = 5;
    if (nrhs < 0) info = 3;
    if (n < 0) info = 2;
    if (trans < 0) info = 1;
    // like dgetrs_: argument errors go to xerbla and leave *Info untouched
    if (info) { report_xerbla("DGETRS", info); return; }
    // beyond the scratch array, or operands it cannot use without staging
    if (n > NA_MAX_N) info = 2;
    else if (n > 0 && nrhs > 0 && !is_device_ptr(a)) info = 4;
    else if (n > 0 && nrhs > 0 && !is_device_ptr(b)) info = 7;
    if (info) {
      report_xerbla("DGETRS_NOALLOC", info);
      *Info = -info;
      return;
    }
    *Info = 0;
    if (n == 0 || nrhs == 0) return;
    Runtime& r = rt();
    char* ws = (char*)na_scratch(r.device);
    int* perm = (int*)(ws + NA_PERM_OFF);
    double* tmp = (double*)(ws + NA_TMP_OFF);
    void* tws = ws + NA_WS_OFF;
    if (trsv_lookback_ws_bytes<double>(n, NA_CHUNK) > NA_WS_BYTES)
      fatal("dgetrs_noalloc: solve workspace exceeds its scratch region", nullptr);
    // IPIV (host or device) -> the composed row permutation in scratch
    const int* hp = ipiv;
    if (is_device_ptr(ipiv)) {
      OBA_HIP(hipMemcpyAsync(g_na_ipiv, ipiv, sizeof(int) * n, hipMemcpyDeviceToHost, r.stream));
      OBA_HIP(hipStreamSynchronize(r.stream));
      hp = g_na_ipiv;
    }
    for (long i = 0; i < n; i++) g_na_perm[i] = (int)i;
    for (long i = 0; i < n; i++) {
      const long p = hp[i] - 1;
      if (p != i && p >= 0 && p < n) std::swap(g_na_perm[i], g_na_perm[p]);
    }
    OBA_HIP(hipMemcpyAsync(perm, g_na_perm, sizeof(int) * n, hipMemcpyHostToDevice, r.stream));
    const size_t wsb = trsv_lookback_ws_bytes<double>(n, NA_CHUNK);
    for (long c0 = 0; c0 < nrhs; c0 += NA_CHUNK) {
      const long nc = std::min(NA_CHUNK, nrhs - c0);
      double* bc = b + c0 * ldb;
      if (trans == 0) {   // P^T: gather, then L (unit lower) and U (upper)
        permute_rows_ws<double>(false, n, nc, perm, bc, ldb, tmp, r.stream);
        na_zero(tws, wsb, r.stream);
        trsv_lookback_ws<double>(true, true, n, a, 1, lda, bc, ldb, nc, tws, r.stream);
        na_zero(tws, wsb, r.stream);
        trsv_lookback_ws<double>(false, false, n, a, 1, lda, bc, ldb, nc, tws, r.stream);
      } else {            // U^T (lower), L^T (unit upper), then P
        na_zero(tws, wsb, r.stream);
        trsv_lookback_ws<double>(true, false, n, a, lda, 1, bc, ldb, nc, tws, r.stream);
        na_zero(tws, wsb, r.stream);
        trsv_lookback_ws<double>(false, true, n, a, lda, 1, bc, ldb, nc, tws, r.stream);
        permute_rows_ws<double>(true, n, nc, perm, bc, ldb, tmp, r.stream);
      }
    }
    end_call();
  });
  if (!ok) *Info = OBA_DEVICE_ERROR;
}
