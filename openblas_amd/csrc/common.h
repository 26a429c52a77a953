// Internal definitions shared by the host layer and the HIP kernels of
// libopenblas_amd.  Nothing here is part of the public C ABI (that lives in
// include/openblas_amd.h).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <stdexcept>

// BLAS/LAPACK integers: 32-bit (LP64), or 64-bit in the ILP64 build
// (OPENBLAS_AMD_ILP64, the analogue of the reference's INTERFACE64=1;
// LAPACKE's lapack_int follows, as with LAPACK_ILP64).
#ifdef OPENBLAS_AMD_ILP64
typedef long blasint;
#else
typedef int blasint;
#endif
typedef blasint lapack_int;

#define OBA_API extern "C" __attribute__((visibility("default")))

namespace oba {

// ---- error plumbing ----------------------------------------------------------
// A device-side failure (no usable device, a failed HIP call, a bounded wait
// that ran out, an internal limit) never aborts the caller's process: it is
// thrown as DeviceError, caught at the exported entry point (guarded()),
// recorded for openblas_amd_last_error(), and reported to the caller as
// INFO = OBA_DEVICE_ERROR (LAPACK/LAPACKE) or a skipped call (BLAS).
struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
constexpr int OBA_DEVICE_ERROR = -1000;
[[noreturn]] void fatal(const char* what, const char* detail);   // throws DeviceError
void hip_check(hipError_t e, const char* what);
#define OBA_HIP(x) ::oba::hip_check((x), #x)

// Library calls are serialized by one recursive mutex (taken by guarded()).
std::recursive_mutex& api_mutex();
void note_error(const DeviceError& e);   // drains the device, records the message

// Runs one exported call under the library lock; returns false (after
// note_error) when it failed on the device.
template <typename F>
inline bool guarded(F&& f) noexcept {
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  try {
    f();
    return true;
  } catch (const DeviceError& e) {
    note_error(e);
    return false;
  } catch (const std::bad_alloc&) {
    note_error(DeviceError("host allocation failed"));
    return false;
  }
}

// BLAS/LAPACK argument-error reporting (calls the weak, user-overridable
// xerbla_ exactly like the reference, driver/others/xerbla.c:56-64).
void report_xerbla(const char* name, long info);

// ---- runtime: one device, one main stream, one look-ahead stream ---------
struct Runtime {
  int device = -1;
  int num_cu = 0;
  hipStream_t stream = nullptr;   // all library work is ordered on this
  hipStream_t side = nullptr;     // second stream (same priority)
  hipStream_t bulk = nullptr;     // like side, restricted to all but a few CUs
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  hipEvent_t ev_free = nullptr;   // pool_free's ordering event
  bool sync_on_return = true;     // synchronous semantics like the reference
  // Sticky device error word in pinned, device-mapped host memory: kernels
  // with a bounded wait set it (system-scope store) when the wait runs out;
  // the host reads it after its synchronizations (check_device_error).
  unsigned* d_err = nullptr;      // device address
  volatile unsigned* h_err = nullptr;
};
Runtime& rt();                    // initializes on first use
// Stream for throughput-bound work that runs beside latency-bound work on
// rt().stream: its kernels use all CUs but OPENBLAS_AMD_RESERVED_CUS (32 by
// default), which stay free for the other stream (falls back to rt().side
// when the device refuses CU masks).
hipStream_t bulk_stream();

// Device scratch pool (the analogue of blas_memory_alloc,
// driver/others/memory.c:1161): cached blocks, reused across calls.  Blocks
// are handed out for work on rt().stream, where stream order separates
// successive users.  A block that work on another stream also uses is
// marked (DevBuf::use_on); freeing it makes rt().stream wait for that
// stream's work queued so far, so the next user of the block, on
// rt().stream, runs after it.
void* pool_alloc(size_t bytes);
void pool_free(void* p, hipStream_t used_on = nullptr);

struct DevBuf {
  void* p = nullptr;
  hipStream_t other = nullptr;   // another stream that uses the block
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { if (bytes) p = pool_alloc(bytes); }
  ~DevBuf() { if (p) pool_free(p, other); }
  void use_on(hipStream_t s) { other = s; }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), other(o.other) { o.p = nullptr; o.other = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      if (p) pool_free(p, other);
      p = o.p; other = o.other;
      o.p = nullptr; o.other = nullptr;
    }
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

// Opt-in kernel timer (openblas_amd_ktimer_*): when enabled, named launches
// are bracketed by HIP events on the stream they run on, with the launch's
// algorithmic flops and HBM bytes; reading resolves the events.  Off by
// default: then a KTimer costs one branch.
bool ktimer_on();
struct KTimer {
  const char* name = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr;
  double flops = 0, bytes = 0;
  KTimer(const char* n, hipStream_t st, double fl, double by);
  ~KTimer();
  KTimer(const KTimer&) = delete;
  KTimer& operator=(const KTimer&) = delete;
};

void end_call();   // synchronize if the runtime is in synchronous mode; check_device_error
// throws DeviceError if a kernel has set the device error word (call after a
// synchronization that covers the kernels in question)
void check_device_error();

}  // namespace oba
