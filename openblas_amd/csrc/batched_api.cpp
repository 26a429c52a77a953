// C-ABI of the batched engine.
//
// cblas_?gemm_batch (the reference's group API, interface/gemm_batch.c:117)
// is routed here from blas3_api.cpp after its argument checks.  The batched
// LAPACK entry points below have no reference counterpart (the reference
// only has the single-problem dgetrf_/dgetrs_/dpotrf_); each problem gets
// exactly the argument checks, info numbering and results of that
// single-problem routine (interface/lapack/getrf.c:76-80, getrs.c:90-100,
// potrf.c:80-88), and `info` is a host array with one entry per problem.
//
// Matrices may be in host or device memory, per pointer.  Device operands
// are used in place; host operands of a call are packed into one pinned
// buffer and moved with one copy each way.
#include "batched.h"
#include "kernels.h"
#include <algorithm>
#include <cctype>
#include <cstring>
#include <map>
#include <vector>

namespace oba {
namespace {

// Host staging of the operands of one batched call.
class Stager {
 public:
  // Registers operand *slot (rows x cols, leading dimension *ld, elements of
  // es bytes).  Device pointers stay; host ones are redirected to the
  // staging buffer once finish() ran.
  void add(void** slot, int* ld, long rows, long cols, size_t es, bool in, bool out) {
    if (!*slot || rows <= 0 || cols <= 0) return;
    if (is_dev(*slot)) return;
    Ent e{slot, ld, *slot, rows, cols, ld ? *ld : rows, es, in, out, total_};
    total_ += ((size_t)rows * cols * es + 255) & ~(size_t)255;
    ents_.push_back(e);
  }
  // copies host inputs to the device and patches the descriptors' pointers
  void finish() {
    if (ents_.empty()) return;
    Runtime& r = rt();
    dev_ = DevBuf(total_);
    char* h = pinned(total_);
    for (auto& e : ents_) {
      if (e.in) pack(h + e.off, e);
      *e.slot = dev_.as<char>() + e.off;
      if (e.ld) *e.ld = (int)e.rows;
    }
    OBA_HIP(hipMemcpyAsync(dev_.p, h, total_, hipMemcpyHostToDevice, r.stream));
  }
  // copies host outputs back (after the compute was enqueued)
  void drain() {
    if (ents_.empty()) return;
    Runtime& r = rt();
    char* h = pinned(total_);
    OBA_HIP(hipMemcpyAsync(h, dev_.p, total_, hipMemcpyDeviceToHost, r.stream));
    OBA_HIP(hipStreamSynchronize(r.stream));
    for (auto& e : ents_)
      if (e.out) unpack(h + e.off, e);
  }

 private:
  struct Ent {
    void** slot;
    int* ld;
    void* host;
    long rows, cols, host_ld;
    size_t es;
    bool in, out;
    size_t off;
  };
  std::vector<Ent> ents_;
  size_t total_ = 0;
  DevBuf dev_;
  std::map<uintptr_t, uintptr_t> dev_ranges_;   // known device ranges [base, end)

  bool is_dev(const void* p) {
    const uintptr_t u = (uintptr_t)p;
    auto it = dev_ranges_.upper_bound(u);
    if (it != dev_ranges_.begin()) {
      --it;
      if (u < it->second) return true;
    }
    if (!is_device_ptr(p)) return false;
    void* base = nullptr;
    size_t sz = 0;
    if (hipMemGetAddressRange(&base, &sz, (void*)p) == hipSuccess && base && sz)
      dev_ranges_[(uintptr_t)base] = (uintptr_t)base + sz;
    else {
      (void)hipGetLastError();
      dev_ranges_[u] = u + 1;
    }
    return true;
  }
  static void pack(char* dst, const Ent& e) {
    const size_t cb = (size_t)e.rows * e.es;
    for (long j = 0; j < e.cols; j++) memcpy(dst + j * cb, (const char*)e.host + (size_t)j * e.host_ld * e.es, cb);
  }
  static void unpack(const char* src, const Ent& e) {
    const size_t cb = (size_t)e.rows * e.es;
    for (long j = 0; j < e.cols; j++) memcpy((char*)e.host + (size_t)j * e.host_ld * e.es, src + j * cb, cb);
  }
  // one pinned host buffer, grown on demand and kept for later calls
  static char* pinned(size_t bytes) {
    static char* buf = nullptr;
    static size_t cap = 0;
    if (bytes > cap) {
      if (buf) OBA_HIP(hipHostFree(buf));
      cap = std::max(bytes, cap * 2);
      OBA_HIP(hipHostMalloc((void**)&buf, cap, hipHostMallocDefault));
    }
    return buf;
  }
};

double cost(const BDesc& d) { return (double)std::max(d.m, 1) * std::max(d.n, 1) * std::max(std::max(d.k, 1), std::min(d.m, d.n)); }

void order_largest_first(std::vector<BDesc>& p) {
  std::stable_sort(p.begin(), p.end(), [](const BDesc& a, const BDesc& b) { return cost(a) > cost(b); });
}

template <typename T>
int getrf_batch_impl(const char* name, const int* m, const int* n, T* const* a, const int* lda,
                     int* const* ipiv, int* info, int count) {
  if (count < 0) { report_xerbla(name, 7); return -7; }
  if (count == 0) return 0;
  std::vector<BDesc> probs;
  bool reported = false;
  for (int i = 0; i < count; i++) {
    int e = 0;
    if (lda[i] < std::max(1, m[i])) e = 4;
    if (n[i] < 0) e = 2;
    if (m[i] < 0) e = 1;
    if (e) {
      info[i] = -e;
      if (!reported) { report_xerbla(name, e); reported = true; }
      continue;
    }
    info[i] = 0;
    if (m[i] == 0 || n[i] == 0) continue;
    BDesc d{};
    d.a = a[i]; d.ipiv = ipiv[i]; d.m = m[i]; d.n = n[i]; d.lda = lda[i]; d.idx = i;
    probs.push_back(d);
  }
  if (probs.empty()) return 0;
  (void)rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  Stager st;
  for (auto& d : probs) {
    st.add(&d.a, &d.lda, d.m, d.n, sizeof(T), true, true);
    st.add((void**)&d.ipiv, nullptr, std::min(d.m, d.n), 1, sizeof(int), false, true);
  }
  st.finish();
  order_largest_first(probs);
  getrf_batched<T>(probs, info);
  st.drain();
  end_call();
  return 0;
}

template <typename T>
int getrs_batch_impl(const char* name, char trans, const int* n, const int* nrhs, T* const* a,
                     const int* lda, int* const* ipiv, T* const* b, const int* ldb, int* info,
                     int count) {
  const char tc = (char)toupper((unsigned char)trans);
  int tr = -1;
  if (tc == 'N' || tc == 'R') tr = 0;
  if (tc == 'T' || tc == 'C') tr = 1;
  if (tr < 0) { report_xerbla(name, 1); return -1; }
  if (count < 0) { report_xerbla(name, 10); return -10; }
  if (count == 0) return 0;
  std::vector<BDesc> probs;
  bool reported = false;
  for (int i = 0; i < count; i++) {
    int e = 0;
    if (ldb[i] < std::max(1, n[i])) e = 8;
    if (lda[i] < std::max(1, n[i])) e = 5;
    if (nrhs[i] < 0) e = 3;
    if (n[i] < 0) e = 2;
    if (e) {
      info[i] = -e;
      if (!reported) { report_xerbla(name, e); reported = true; }
      continue;
    }
    info[i] = 0;
    if (n[i] == 0 || nrhs[i] == 0) continue;
    BDesc d{};
    d.a = a[i]; d.b = b[i]; d.ipiv = ipiv[i]; d.m = n[i]; d.n = n[i]; d.k = nrhs[i];
    d.lda = lda[i]; d.ldb = ldb[i]; d.idx = i;
    probs.push_back(d);
  }
  if (probs.empty()) return 0;
  (void)rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  Stager st;
  for (auto& d : probs) {
    st.add(&d.a, &d.lda, d.n, d.n, sizeof(T), true, false);
    st.add((void**)&d.ipiv, nullptr, d.n, 1, sizeof(int), true, false);
    st.add(&d.b, &d.ldb, d.n, d.k, sizeof(T), true, true);
  }
  st.finish();
  order_largest_first(probs);
  getrs_batched<T>(tr == 1, probs, info);
  st.drain();
  end_call();
  return 0;
}

template <typename T>
int potrf_batch_impl(const char* name, char uplo, const int* n, T* const* a, const int* lda, int* info,
                     int count) {
  const char uc = (char)toupper((unsigned char)uplo);
  int lower = -1;
  if (uc == 'U') lower = 0;
  if (uc == 'L') lower = 1;
  if (lower < 0) { report_xerbla(name, 1); return -1; }
  if (count < 0) { report_xerbla(name, 6); return -6; }
  if (count == 0) return 0;
  std::vector<BDesc> probs;
  bool reported = false;
  for (int i = 0; i < count; i++) {
    int e = 0;
    if (lda[i] < std::max(1, n[i])) e = 4;
    if (n[i] < 0) e = 2;
    if (e) {
      info[i] = -e;
      if (!reported) { report_xerbla(name, e); reported = true; }
      continue;
    }
    info[i] = 0;
    if (n[i] == 0) continue;
    BDesc d{};
    d.a = a[i]; d.m = n[i]; d.n = n[i]; d.lda = lda[i]; d.idx = i;
    probs.push_back(d);
  }
  if (probs.empty()) return 0;
  (void)rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  Stager st;
  for (auto& d : probs) st.add(&d.a, &d.lda, d.n, d.n, sizeof(T), true, true);
  st.finish();
  order_largest_first(probs);
  potrf_batched<T>(lower == 1, probs, info);
  st.drain();
  end_call();
  return 0;
}

}  // namespace

// Problems of a checked cblas_?gemm_batch call (column-major view, A/B
// already swapped for row-major): small ones go to the batched engine, one
// persistent launch per transpose pair; large ones to the single GEMM.
template <typename T>
void gemm_batch_run(std::vector<BDesc>& all) {
  if (all.empty()) return;
  Runtime& r = rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  Stager st;
  for (auto& d : all) {
    const bool ta = d.flags & 1, tb = d.flags & 2;
    if (d.alpha != 0.0 && d.k > 0) {
      st.add(&d.a, &d.lda, ta ? d.k : d.m, ta ? d.m : d.k, sizeof(T), true, false);
      st.add(&d.b, &d.ldb, tb ? d.n : d.k, tb ? d.k : d.n, sizeof(T), true, false);
    }
    st.add(&d.c, &d.ldc, d.m, d.n, sizeof(T), d.beta != 0.0, true);
  }
  st.finish();
  std::vector<BDesc> cls[4];
  for (auto& d : all) {
    if ((double)d.m * d.n * std::max(d.k, 1) > 512.0 * 512.0 * 512.0) {
      gemm<T>(d.flags & 1, d.flags & 2, d.m, d.n, d.k, (T)d.alpha, (const T*)d.a, d.lda, (const T*)d.b, d.ldb,
              (T)d.beta, (T*)d.c, d.ldc, r.stream);
    } else {
      cls[d.flags & 3].push_back(d);
    }
  }
  for (int f = 0; f < 4; f++) {
    order_largest_first(cls[f]);
    gemm_batched<T>(f & 1, (f >> 1) & 1, cls[f]);
  }
  st.drain();
  end_call();
}
template void gemm_batch_run<double>(std::vector<BDesc>&);
template void gemm_batch_run<float>(std::vector<BDesc>&);

}  // namespace oba

using namespace oba;

OBA_API int openblas_amd_dgetrf_batch(const int* m, const int* n, double* const* a, const int* lda,
                                      int* const* ipiv, int* info, int count) {
  int r = OBA_DEVICE_ERROR;
  guarded([&] { r = getrf_batch_impl<double>("DGETRF", m, n, a, lda, ipiv, info, count); });
  return r;
}
OBA_API int openblas_amd_sgetrf_batch(const int* m, const int* n, float* const* a, const int* lda,
                                      int* const* ipiv, int* info, int count) {
  int r = OBA_DEVICE_ERROR;
  guarded([&] { r = getrf_batch_impl<float>("SGETRF", m, n, a, lda, ipiv, info, count); });
  return r;
}
OBA_API int openblas_amd_dgetrs_batch(char trans, const int* n, const int* nrhs, double* const* a,
                                      const int* lda, int* const* ipiv, double* const* b, const int* ldb,
                                      int* info, int count) {
  int r = OBA_DEVICE_ERROR;
  guarded([&] { r = getrs_batch_impl<double>("DGETRS", trans, n, nrhs, a, lda, ipiv, b, ldb, info, count); });
  return r;
}
OBA_API int openblas_amd_sgetrs_batch(char trans, const int* n, const int* nrhs, float* const* a,
                                      const int* lda, int* const* ipiv, float* const* b, const int* ldb,
                                      int* info, int count) {
  int r = OBA_DEVICE_ERROR;
  guarded([&] { r = getrs_batch_impl<float>("SGETRS", trans, n, nrhs, a, lda, ipiv, b, ldb, info, count); });
  return r;
}
OBA_API int openblas_amd_dpotrf_batch(char uplo, const int* n, double* const* a, const int* lda, int* info,
                                      int count) {
  int r = OBA_DEVICE_ERROR;
  guarded([&] { r = potrf_batch_impl<double>("DPOTRF", uplo, n, a, lda, info, count); });
  return r;
}
OBA_API int openblas_amd_spotrf_batch(char uplo, const int* n, float* const* a, const int* lda, int* info,
                                      int count) {
  int r = OBA_DEVICE_ERROR;
  guarded([&] { r = potrf_batch_impl<float>("SPOTRF", uplo, n, a, lda, info, count); });
  return r;
}
