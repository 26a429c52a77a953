// Host runtime of libopenblas_amd: device/stream singleton, device scratch
// pool, host<->device staging of caller matrices, xerbla plumbing, config
// strings and the small extension API used by tests and benchmarks.
//
// Reference counterparts: driver/others/memory.c:1161,1328 (buffer pool),
// driver/others/xerbla.c:56-64 (weak xerbla_), driver/others/openblas_env.c
// (environment), interface/... thread-count knobs (openblas_set_num_threads,
// cblas.h:13-24).  There is no CPU fallback anywhere: a missing device is a
// device error reported to the caller (DeviceError, caught per entry point).
#include "common.h"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <map>
#include <vector>
#include <string>

namespace oba {

static Runtime* g_rt = nullptr;
static std::once_flag g_rt_once;

[[noreturn]] void fatal(const char* what, const char* detail) {
  std::string m = what;
  if (detail) { m += ": "; m += detail; }
  throw DeviceError(m);
}

std::recursive_mutex& api_mutex() {
  static std::recursive_mutex m;
  return m;
}

static std::mutex g_err_mu;
static std::string g_last_error;

void note_error(const DeviceError& e) {
  {
    std::lock_guard<std::mutex> g(g_err_mu);
    g_last_error = e.what();
  }
  fprintf(stderr, "libopenblas_amd: device error: %s (call skipped)\n", e.what());
  fflush(stderr);
  // nothing the failed call queued may still run on buffers that go back to
  // the pool; every wait in the library's kernels is bounded, so this ends
  int n = 0;
  if (hipGetDeviceCount(&n) == hipSuccess && n > 0) (void)hipDeviceSynchronize();
  (void)hipGetLastError();
  if (g_rt && g_rt->h_err) *g_rt->h_err = 0u;
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fatal(what, hipGetErrorString(e));
}


Runtime& rt() {
  std::call_once(g_rt_once, [] {
    Runtime* r = new Runtime();
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
      fatal("no HIP device available (libopenblas_amd has no CPU path)", e != hipSuccess ? hipGetErrorString(e) : nullptr);
    int dev = 0;
    if (const char* s = getenv("OPENBLAS_AMD_DEVICE")) dev = atoi(s);
    if (dev < 0 || dev >= n) fatal("OPENBLAS_AMD_DEVICE out of range", nullptr);
    OBA_HIP(hipSetDevice(dev));
    r->device = dev;
    hipDeviceProp_t prop;
    OBA_HIP(hipGetDeviceProperties(&prop, dev));
    r->num_cu = prop.multiProcessorCount;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      fatal("libopenblas_amd is built for gfx950 only; device is", prop.gcnArchName);
    OBA_HIP(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    OBA_HIP(hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking));
    OBA_HIP(hipEventCreateWithFlags(&r->ev_a, hipEventDisableTiming));
    OBA_HIP(hipEventCreateWithFlags(&r->ev_b, hipEventDisableTiming));
    OBA_HIP(hipEventCreateWithFlags(&r->ev_free, hipEventDisableTiming));
    void* he = nullptr;
    OBA_HIP(hipHostMalloc(&he, 256, hipHostMallocMapped));
    void* de = nullptr;
    OBA_HIP(hipHostGetDevicePointer(&de, he, 0));
    r->h_err = (volatile unsigned*)he;
    r->d_err = (unsigned*)de;
    *r->h_err = 0u;
    if (const char* s = getenv("OPENBLAS_AMD_ASYNC")) r->sync_on_return = (atoi(s) == 0);
    g_rt = r;
  });
  // make the library's device current once per host thread
  static thread_local bool t_dev_set = false;
  if (!t_dev_set) {
    OBA_HIP(hipSetDevice(g_rt->device));
    t_dev_set = true;
  }
  return *g_rt;
}

hipStream_t bulk_stream() {
  Runtime& r = rt();
  static std::once_flag once;
  std::call_once(once, [&r] {
    int reserve = 32;
    if (const char* e = getenv("OPENBLAS_AMD_RESERVED_CUS")) reserve = atoi(e);
    if (reserve > 0 && r.num_cu > 2 * reserve) {
      std::vector<uint32_t> mask((size_t)(r.num_cu + 31) / 32, 0u);
      for (int i = 0; i < r.num_cu - reserve; i++) mask[(size_t)i / 32] |= 1u << (i % 32);
      hipStream_t st = nullptr;
      if (hipExtStreamCreateWithCUMask(&st, (uint32_t)mask.size(), mask.data()) == hipSuccess) r.bulk = st;
      else (void)hipGetLastError();
    }
    if (r.bulk) {
      // drain and release the masked queue before the runtime tears down
      std::atexit([] {
        Runtime& q = *g_rt;
        if (q.bulk && q.bulk != q.side) {
          (void)hipStreamSynchronize(q.bulk);
          (void)hipStreamDestroy(q.bulk);
          q.bulk = q.side;
        }
      });
    } else {
      r.bulk = r.side;
    }
  });
  return r.bulk;
}

void check_device_error() {
  Runtime& r = rt();
  const unsigned e = *r.h_err;
  if (e) {
    *r.h_err = 0u;
    throw DeviceError(e == 1u ? "a cross-workgroup wait ran out (kernel exchange timed out)"
                              : "a kernel found inconsistent exchange data");
  }
}

void end_call() {
  Runtime& r = rt();
  if (r.sync_on_return) {
    OBA_HIP(hipStreamSynchronize(r.stream));
    check_device_error();
  }
}

// ---- kernel timer ------------------------------------------------------------
namespace {
struct KRec {
  hipEvent_t e0, e1;
  double flops, bytes;
};
struct KAgg {
  long count = 0;
  double ms = 0, flops = 0, bytes = 0;
  std::vector<KRec> pending;
};
std::mutex g_kt_mu;
bool g_kt_on = false;
std::map<std::string, KAgg> g_kt;
std::vector<hipEvent_t> g_kt_pool;

hipEvent_t kt_event() {
  if (!g_kt_pool.empty()) {
    hipEvent_t e = g_kt_pool.back();
    g_kt_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  OBA_HIP(hipEventCreate(&e));
  return e;
}
// resolve recorded pairs into the aggregates (events complete)
void kt_resolve(KAgg& a) {
  for (auto& r : a.pending) {
    OBA_HIP(hipEventSynchronize(r.e1));
    float ms = 0.f;
    OBA_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
    a.count++;
    a.ms += ms;
    a.flops += r.flops;
    a.bytes += r.bytes;
    g_kt_pool.push_back(r.e0);
    g_kt_pool.push_back(r.e1);
  }
  a.pending.clear();
}
}  // namespace

bool ktimer_on() { return g_kt_on; }

KTimer::KTimer(const char* n, hipStream_t st, double fl, double by) {
  if (!g_kt_on) return;
  std::lock_guard<std::mutex> g(g_kt_mu);
  name = n; s = st; flops = fl; bytes = by;
  e0 = kt_event();
  OBA_HIP(hipEventRecord(e0, s));
}

KTimer::~KTimer() {
  if (!e0) return;
  std::lock_guard<std::mutex> g(g_kt_mu);
  hipEvent_t e1 = nullptr;
  try {
    e1 = kt_event();
  } catch (const DeviceError&) {
    return;   // never throw from a destructor; the pair is dropped
  }
  if (hipEventRecord(e1, s) != hipSuccess) { (void)hipGetLastError(); return; }
  g_kt[name].pending.push_back(KRec{e0, e1, flops, bytes});
}

// ---- device pool -----------------------------------------------------------
namespace {
std::mutex g_pool_mu;
std::multimap<size_t, void*> g_free;       // size -> block
std::map<void*, size_t> g_size;            // every live or cached block
size_t g_cached_bytes = 0;

size_t round_size(size_t b) {
  if (b <= (1u << 20)) return (b + 255) & ~size_t(255);
  return (b + (2u << 20) - 1) & ~size_t((2u << 20) - 1);
}
}  // namespace

void* pool_alloc(size_t bytes) {
  rt();
  size_t sz = round_size(bytes ? bytes : 1);
  {
    std::lock_guard<std::mutex> g(g_pool_mu);
    auto it = g_free.lower_bound(sz);
    if (it != g_free.end() && it->first <= 2 * sz + (4u << 20)) {
      void* p = it->second;
      g_cached_bytes -= it->first;
      g_free.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    // release the cache (after draining all queued work) and retry once
    OBA_HIP(hipDeviceSynchronize());
    std::lock_guard<std::mutex> g(g_pool_mu);
    for (auto& kv : g_free) { OBA_HIP(hipFree(kv.second)); g_size.erase(kv.second); }
    g_free.clear();
    g_cached_bytes = 0;
    e = hipMalloc(&p, sz);
    if (e != hipSuccess) { (void)hipGetLastError(); fatal("device allocation failed", hipGetErrorString(e)); }
  }
  std::lock_guard<std::mutex> g(g_pool_mu);
  g_size[p] = sz;
  return p;
}

void pool_free(void* p, hipStream_t used_on) {
  if (!p) return;
  Runtime& r = *g_rt;
  if (used_on && used_on != r.stream) {
    // the block was used on another stream: everything queued later on the
    // library stream (where the block is handed out again) waits for that
    // stream's work so far.  Errors are left for the next checked call
    // (this runs in destructors).
    if (hipEventRecord(r.ev_free, used_on) != hipSuccess || hipStreamWaitEvent(r.stream, r.ev_free, 0) != hipSuccess)
      (void)hipGetLastError();
  }
  std::lock_guard<std::mutex> g(g_pool_mu);
  auto it = g_size.find(p);
  if (it == g_size.end()) return;   // not a pool block (cannot happen; never throw from a destructor)
  g_free.emplace(it->second, p);
  g_cached_bytes += it->second;
}

bool is_device_ptr(const void* p) {
  if (!p) return false;
  rt();
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) { (void)hipGetLastError(); return false; }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

template <typename T>
void Mat<T>::in(T* p, long r, long c, long ldp, bool copy_in) {
  rows = r; cols = c; host = nullptr; host_ld = ldp;
  if (r <= 0 || c <= 0) { d = p; ld = ldp; return; }
  if (is_device_ptr(p)) { d = p; ld = ldp; return; }
  host = p;
  ld = (r + 15) & ~15L;
  buf = DevBuf(sizeof(T) * (size_t)ld * (size_t)c);
  d = buf.as<T>();
  if (copy_in)
    OBA_HIP(hipMemcpy2DAsync(d, sizeof(T) * ld, p, sizeof(T) * ldp, sizeof(T) * r, c,
                             hipMemcpyHostToDevice, rt().stream));
}

template <typename T>
void Mat<T>::out() {
  if (!host || rows <= 0 || cols <= 0) return;
  OBA_HIP(hipMemcpy2DAsync(host, sizeof(T) * host_ld, d, sizeof(T) * ld, sizeof(T) * rows, cols,
                           hipMemcpyDeviceToHost, rt().stream));
  OBA_HIP(hipStreamSynchronize(rt().stream));
}

template struct Mat<double>;
template struct Mat<float>;
template struct Mat<int>;

// ---- xerbla ----------------------------------------------------------------
static thread_local char g_last_name[16] = {0};
static thread_local int g_last_info = 0;

typedef int (*xerbla_fn)(char*, blasint*, blasint);
}  // namespace oba
extern "C" int xerbla_(char* message, blasint* info, blasint len);
namespace oba {

void report_xerbla(const char* name, long info) {
  strncpy(g_last_name, name, sizeof(g_last_name) - 1);
  g_last_info = info;
  // Route through the xerbla_ that the program itself resolves (a user
  // override, as the reference's weak alias allows), else ours.  The lookup
  // goes through the main program's handle: this library is linked with
  // -Bsymbolic, and for such an object dlsym(RTLD_DEFAULT, ...) searches the
  // object itself first and would always return its own definition.
  char buf[16];
  strncpy(buf, name, sizeof(buf) - 1);
  buf[sizeof(buf) - 1] = 0;
  static void* const prog = dlopen(nullptr, RTLD_LAZY);
  xerbla_fn f = prog ? (xerbla_fn)dlsym(prog, "xerbla_") : nullptr;
  if (!f) f = &xerbla_;
  blasint i = info;
  f(buf, &i, (blasint)strlen(buf));
}

}  // namespace oba

using namespace oba;

// Weak so that applications may override it (driver/others/xerbla.c:56-64).
extern "C" __attribute__((visibility("default"), weak)) int xerbla_(char* message, blasint* info, blasint len) {
  (void)len;
  printf(" ** On entry to %6s parameter number %2d had an illegal value\n", message, (int)*info);
  return 0;
}

// ---- informational API (cblas.h:13-27) -------------------------------------
static char g_config[256];
static int g_threads = 1;

OBA_API char* openblas_get_config(void) {
  snprintf(g_config, sizeof(g_config), "OpenBLAS-AMD gfx950 MI355X HIP MFMA LP64 (device-resident level-3 + LAPACK LU/Cholesky)");
  return g_config;
}
OBA_API char* openblas_get_corename(void) { static char n[] = "gfx950"; return n; }
OBA_API void openblas_set_num_threads(int n) { g_threads = n > 0 ? n : 1; }
OBA_API int openblas_get_num_threads(void) { return g_threads; }
OBA_API int openblas_get_num_procs(void) {
  int n = 0;
  guarded([&] { n = rt().num_cu; });
  return n;
}

// ---- extension API (no reference counterpart; used by tests/benchmarks) ----
OBA_API void* openblas_amd_malloc(size_t bytes) {
  void* p = nullptr;
  guarded([&] {
    rt();
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) { (void)hipGetLastError(); p = nullptr; }
  });
  return p;
}
OBA_API void openblas_amd_free(void* p) {
  if (p) guarded([&] { rt(); OBA_HIP(hipFree(p)); });
}
OBA_API int openblas_amd_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  return guarded([&] {
    Runtime& r = rt();
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    OBA_HIP(hipMemcpyAsync(dst, src, bytes, k, r.stream));
    OBA_HIP(hipStreamSynchronize(r.stream));
  }) ? 0 : OBA_DEVICE_ERROR;
}
OBA_API int openblas_amd_memset(void* dst, int value, size_t bytes) {
  return guarded([&] {
    Runtime& r = rt();
    OBA_HIP(hipMemsetAsync(dst, value, bytes, r.stream));
    OBA_HIP(hipStreamSynchronize(r.stream));
  }) ? 0 : OBA_DEVICE_ERROR;
}
OBA_API int openblas_amd_synchronize(void) {
  return guarded([] {
    OBA_HIP(hipStreamSynchronize(rt().stream));
    check_device_error();
  }) ? 0 : OBA_DEVICE_ERROR;
}
OBA_API void openblas_amd_set_sync(int on) { guarded([on] { rt().sync_on_return = on != 0; }); }
// message of the last device error (empty if none); clears it
OBA_API int openblas_amd_last_error(char* msg, int len) {
  std::lock_guard<std::mutex> g(g_err_mu);
  const int had = g_last_error.empty() ? 0 : 1;
  if (msg && len > 0) { strncpy(msg, g_last_error.c_str(), len - 1); msg[len - 1] = 0; }
  g_last_error.clear();
  return had;
}
OBA_API int openblas_amd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) { (void)hipGetLastError(); return 0; }
  return n;
}
OBA_API int openblas_amd_device_info(char* name, int len, int* num_cu, int* clock_khz) {
  return guarded([&] {
    Runtime& r = rt();
    hipDeviceProp_t p;
    OBA_HIP(hipGetDeviceProperties(&p, r.device));
    if (name && len > 0) { strncpy(name, p.gcnArchName, len - 1); name[len - 1] = 0; }
    if (num_cu) *num_cu = p.multiProcessorCount;
    if (clock_khz) *clock_khz = p.clockRate;
  }) ? 0 : OBA_DEVICE_ERROR;
}
// HIP-event timer on the library stream: every kernel the library launches
// goes to this stream, so start/stop bracket exactly the library's work.
static hipEvent_t g_t0 = nullptr, g_t1 = nullptr;
OBA_API void openblas_amd_timer_start(void) {
  guarded([] {
    Runtime& r = rt();
    if (!g_t0) { OBA_HIP(hipEventCreate(&g_t0)); OBA_HIP(hipEventCreate(&g_t1)); }
    OBA_HIP(hipEventRecord(g_t0, r.stream));
  });
}
OBA_API float openblas_amd_timer_stop(void) {
  float ms = -1.f;
  guarded([&] {
    Runtime& r = rt();
    OBA_HIP(hipEventRecord(g_t1, r.stream));
    OBA_HIP(hipEventSynchronize(g_t1));
    OBA_HIP(hipEventElapsedTime(&ms, g_t0, g_t1));
  });
  return ms;
}
// Kernel timer (no reference counterpart; used by the benchmark): enable or
// disable (disabling keeps what was recorded), read one named aggregate
// (launch count, total ms, algorithmic flops and bytes summed over the
// launches; returns 0, or -1 if the name has no record), reset all.
OBA_API void openblas_amd_ktimer_enable(int on) {
  guarded([on] {
    rt();
    std::lock_guard<std::mutex> g(g_kt_mu);
    g_kt_on = on != 0;
  });
}
OBA_API int openblas_amd_ktimer_read(const char* name, long* count, double* ms, double* flops, double* bytes) {
  int r = -1;
  guarded([&] {
    std::lock_guard<std::mutex> g(g_kt_mu);
    auto it = g_kt.find(name ? name : "");
    if (it == g_kt.end()) return;
    kt_resolve(it->second);
    if (count) *count = it->second.count;
    if (ms) *ms = it->second.ms;
    if (flops) *flops = it->second.flops;
    if (bytes) *bytes = it->second.bytes;
    r = 0;
  });
  return r;
}
OBA_API void openblas_amd_ktimer_reset(void) {
  guarded([] {
    std::lock_guard<std::mutex> g(g_kt_mu);
    for (auto& kv : g_kt) kt_resolve(kv.second);
    g_kt.clear();
  });
}
OBA_API int openblas_amd_last_xerbla(char* name, int len) {
  if (name && len > 0) { strncpy(name, g_last_name, len - 1); name[len - 1] = 0; }
  int i = g_last_info;
  g_last_info = 0;
  g_last_name[0] = 0;
  return i;
}
