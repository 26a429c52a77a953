// Internal definitions shared by the host layer and the HIP kernels of
// libopenblas_amd.  Nothing here is part of the public C ABI (that lives in
// include/openblas_amd.h).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <mutex>

typedef int blasint;

#define OBA_API extern "C" __attribute__((visibility("default")))

namespace oba {

// ---- fatal / error plumbing ------------------------------------------------
[[noreturn]] void fatal(const char* what, const char* detail);
void hip_check(hipError_t e, const char* what);
#define OBA_HIP(x) ::oba::hip_check((x), #x)

// BLAS/LAPACK argument-error reporting (calls the weak, user-overridable
// xerbla_ exactly like the reference, driver/others/xerbla.c:56-64).
void report_xerbla(const char* name, blasint info);

// ---- runtime: one device, one main stream, one look-ahead stream ---------
struct Runtime {
  int device = -1;
  int num_cu = 0;
  hipStream_t stream = nullptr;   // all library work is ordered on this
  hipStream_t side = nullptr;     // second stream (same priority)
  hipStream_t bulk = nullptr;     // like side, restricted to all but a few CUs
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  bool sync_on_return = true;     // synchronous semantics like the reference
  std::recursive_mutex lock;      // library calls are serialized
};
Runtime& rt();                    // initializes on first use
// Stream for throughput-bound work that runs beside latency-bound work on
// rt().stream: its kernels use all CUs but OPENBLAS_AMD_RESERVED_CUS (32 by
// default), which stay free for the other stream (falls back to rt().side
// when the device refuses CU masks).
hipStream_t bulk_stream();

// Device scratch pool (the analogue of blas_memory_alloc,
// driver/others/memory.c:1161): cached blocks, reused across calls.
void* pool_alloc(size_t bytes);
void pool_free(void* p);

struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { if (bytes) p = pool_alloc(bytes); }
  ~DevBuf() { if (p) pool_free(p); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { if (p) pool_free(p); p = o.p; o.p = nullptr; }
    return *this;
  }
  template <typename T> T* as() const { return (T*)p; }
};

bool is_device_ptr(const void* p);

// A column-major matrix argument, staged to the device when the caller
// passed host memory.  `d` is always a device pointer with leading dim `ld`.
template <typename T>
struct Mat {
  T* d = nullptr;
  long ld = 0;
  T* host = nullptr;
  long host_ld = 0;
  long rows = 0, cols = 0;
  DevBuf buf;
  void in(T* p, long r, long c, long ldp, bool copy_in);   // stage in
  void out();                                              // copy back
};

void end_call();   // synchronize if the runtime is in synchronous mode

}  // namespace oba
