// LAPACK (Fortran ABI) and LAPACKE (C ABI) entry points of the level-3 hot
// path: ?getrf, ?getf2, ?getrs, ?gesv, ?potrf, ?potf2, ?laswp, dsgesv.
// Argument checks and info codes follow the reference's own overrides:
//   interface/lapack/getrf.c:53-80, getrs.c:63-100, gesv.c:63-100,
//   potrf.c:63-90, laswp.c:55-72 and netlib dsgesv.f:238-260;
// the LAPACKE layer follows lapack-netlib/LAPACKE/src/lapacke_dgesv.c and
// lapacke_dgesv_work.c (layout check, optional NaN check, row-major
// transpose, info shift of -1 for the Fortran argument numbering).
// Row-major LAPACKE arguments are transposed on the device.
#include "kernels.h"
#include "lu_internal.h"
#include "dsgesv.h"
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace oba;

namespace {

// The caller's integer array (IPIV) as the kernels' 32-bit device array.
// LP64: a device array is used in place, a host one staged.  ILP64: always
// a device int array, converted from / to the caller's 64-bit values
// (host or device) through the host.
struct IntDev {
  DevBuf buf;
  int* d = nullptr;
  blasint* host = nullptr;    // caller's host array (LP64 staging)
  blasint* user = nullptr;    // caller's array (ILP64 conversion back)
  bool user_dev = false;
  long n = 0;
  void in(blasint* p, long count, bool copy_in) {
    n = count;
    if (count <= 0) return;
    if (sizeof(blasint) == sizeof(int)) {
      if (is_device_ptr(p)) { d = (int*)p; return; }
      host = p;
      buf = DevBuf(sizeof(int) * count);
      d = buf.as<int>();
      if (copy_in) OBA_HIP(hipMemcpyAsync(d, p, sizeof(int) * count, hipMemcpyHostToDevice, rt().stream));
      return;
    }
    user = p;
    user_dev = is_device_ptr(p);
    buf = DevBuf(sizeof(int) * count);
    d = buf.as<int>();
    if (copy_in) {
      std::vector<blasint> w(count);
      read_ints(p, user_dev, count, w.data());
      std::vector<int> v(w.begin(), w.end());
      OBA_HIP(hipMemcpyAsync(d, v.data(), sizeof(int) * count, hipMemcpyHostToDevice, rt().stream));
      OBA_HIP(hipStreamSynchronize(rt().stream));
    }
  }
  void out() {
    if (n <= 0) return;
    if (sizeof(blasint) == sizeof(int)) {
      if (!host) return;
      OBA_HIP(hipMemcpyAsync(host, d, sizeof(int) * n, hipMemcpyDeviceToHost, rt().stream));
      OBA_HIP(hipStreamSynchronize(rt().stream));
      return;
    }
    if (!user) return;
    std::vector<int> v(n);
    OBA_HIP(hipMemcpyAsync(v.data(), d, sizeof(int) * n, hipMemcpyDeviceToHost, rt().stream));
    OBA_HIP(hipStreamSynchronize(rt().stream));
    std::vector<blasint> w(v.begin(), v.end());
    if (user_dev) {
      OBA_HIP(hipMemcpyAsync(user, w.data(), sizeof(blasint) * n, hipMemcpyHostToDevice, rt().stream));
      OBA_HIP(hipStreamSynchronize(rt().stream));
    } else {
      memcpy(user, w.data(), sizeof(blasint) * n);
    }
  }
  // count integers of the caller's array (host or device) into out
  static void read_ints(const blasint* p, bool dev, long count, blasint* out) {
    if (dev) {
      OBA_HIP(hipMemcpyAsync(out, p, sizeof(blasint) * count, hipMemcpyDeviceToHost, rt().stream));
      OBA_HIP(hipStreamSynchronize(rt().stream));
    } else {
      memcpy(out, p, sizeof(blasint) * count);
    }
  }
};

// host (or device) ipiv prefix -> device composed permutation
void perm_from_ipiv(const blasint* ipiv, long n, DevBuf& buf, int** perm_dev) {
  std::vector<int> hp(n), pm;
  {
    std::vector<blasint> w(n);
    IntDev::read_ints(ipiv, is_device_ptr(ipiv), n, w.data());
    for (long i = 0; i < n; i++) hp[i] = (int)w[i];
  }
  ipiv_to_perm(n, hp.data(), pm);
  buf = DevBuf(sizeof(int) * n);
  *perm_dev = buf.as<int>();
  OBA_HIP(hipMemcpyAsync(*perm_dev, pm.data(), sizeof(int) * n, hipMemcpyHostToDevice, rt().stream));
}

template <typename T> struct Names;
template <> struct Names<double> {
  static constexpr const char* getrf = "DGETRF";
  static constexpr const char* getrs = "DGETRS"; static constexpr const char* gesv = "DGESV ";
  static constexpr const char* potrf = "DPOTRF";
};
template <> struct Names<float> {
  static constexpr const char* getrf = "SGETRF";
  static constexpr const char* getrs = "SGETRS"; static constexpr const char* gesv = "SGESV ";
  static constexpr const char* potrf = "SPOTRF";
};

template <typename T>
void getrf_impl(const char* name, const blasint* M, const blasint* N, T* a, const blasint* LDA, blasint* ipiv,
                blasint* Info) {
  long m = *M, n = *N, lda = *LDA;
  blasint info = 0;
  if (lda < std::max(1L, m)) info = 4;
  if (n < 0) info = 2;
  if (m < 0) info = 1;
  if (info) { report_xerbla(name, info); *Info = -info; return; }
  *Info = 0;
  if (m == 0 || n == 0) return;
  Runtime& r = rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  Mat<T> A;
  A.in(a, m, n, lda, true);
  IntDev P;
  P.in(ipiv, std::min(m, n), false);
  *Info = getrf<T>(m, n, A.d, A.ld, P.d, r.stream);
  A.out();
  P.out();
  end_call();
}

template <typename T>
void getrs_impl(const char* name, const char* TRANS, const blasint* N, const blasint* NRHS, const T* a,
                const blasint* LDA, const blasint* ipiv, T* b, const blasint* LDB, blasint* Info) {
  char tc = (char)toupper((unsigned char)*TRANS);
  int trans = -1;
  if (tc == 'N' || tc == 'R') trans = 0;
  if (tc == 'T' || tc == 'C') trans = 1;
  long n = *N, nrhs = *NRHS, lda = *LDA, ldb = *LDB;
  blasint info = 0;
  if (ldb < std::max(1L, n)) info = 8;
  if (lda < std::max(1L, n)) info = 5;
  if (nrhs < 0) info = 3;
  if (n < 0) info = 2;
  if (trans < 0) info = 1;
  // the reference reports through xerbla only and leaves *Info untouched
  if (info) { report_xerbla(name, info); return; }
  *Info = 0;
  if (n == 0 || nrhs == 0) return;
  Runtime& r = rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  Mat<T> A, B;
  A.in(const_cast<T*>(a), n, n, lda, true);
  B.in(b, n, nrhs, ldb, true);
  DevBuf pb;
  int* perm = nullptr;
  perm_from_ipiv(ipiv, n, pb, &perm);
  getrs_perm<T>(trans == 1, n, nrhs, A.d, A.ld, perm, B.d, B.ld, r.stream);
  B.out();
  end_call();
}

template <typename T>
void gesv_impl(const char* name, const blasint* N, const blasint* NRHS, T* a, const blasint* LDA, blasint* ipiv, T* b,
               const blasint* LDB, blasint* Info) {
  long n = *N, nrhs = *NRHS, lda = *LDA, ldb = *LDB;
  blasint info = 0;
  if (ldb < std::max(1L, n)) info = 7;
  if (lda < std::max(1L, n)) info = 4;
  if (nrhs < 0) info = 2;
  if (n < 0) info = 1;
  if (info) { report_xerbla(name, info); *Info = -info; return; }
  *Info = 0;
  if (n == 0 || nrhs == 0) return;
  Runtime& r = rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  Mat<T> A, B;
  A.in(a, n, n, lda, true);
  B.in(b, n, nrhs, ldb, true);
  IntDev P;
  P.in(ipiv, n, false);
  DevBuf pb(sizeof(int) * n);
  int f = getrf_perm<T>(n, n, A.d, A.ld, P.d, pb.as<int>(), r.stream);
  if (f == 0) getrs_perm<T>(false, n, nrhs, A.d, A.ld, pb.as<int>(), B.d, B.ld, r.stream);
  *Info = f;
  A.out();
  B.out();
  P.out();
  end_call();
}

template <typename T>
void potrf_impl(const char* name, const char* UPLO, const blasint* N, T* a, const blasint* LDA, blasint* Info) {
  char uc = (char)toupper((unsigned char)*UPLO);
  int uplo = uc == 'U' ? 0 : uc == 'L' ? 1 : -1;
  long n = *N, lda = *LDA;
  blasint info = 0;
  if (lda < std::max(1L, n)) info = 4;
  if (n < 0) info = 2;
  if (uplo < 0) info = 1;
  if (info) { report_xerbla(name, info); *Info = -info; return; }
  *Info = 0;
  if (n == 0) return;
  Runtime& r = rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  Mat<T> A;
  A.in(a, n, n, lda, true);
  *Info = potrf<T>(uplo == 1, n, A.d, A.ld, r.stream);
  A.out();
  end_call();
}

template <typename T>
void laswp_impl(const blasint* N, T* a, const blasint* LDA, const blasint* K1, const blasint* K2, const blasint* ipiv,
                const blasint* INCX) {
  long n = *N, lda = *LDA, k1 = *K1, k2 = *K2;
  int incx = *INCX;
  if (incx == 0 || n <= 0) return;   // interface/lapack/laswp.c:72
  if (k2 < k1) return;
  Runtime& r = rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  const long inc = incx > 0 ? incx : -incx;
  const long npiv = (k1 - 1) + (k2 - k1) * inc + 1;
  std::vector<int> hp(npiv);
  {
    std::vector<blasint> w(npiv);
    IntDev::read_ints(ipiv, is_device_ptr(ipiv), npiv, w.data());
    for (long i = 0; i < npiv; i++) hp[i] = (int)w[i];
  }
  long rows = k2;
  for (long i = k1; i <= k2; i++) rows = std::max(rows, (long)hp[(k1 - 1) + (i - k1) * inc]);
  Mat<T> A;
  A.in(a, rows, n, lda, true);
  DevBuf pd(sizeof(int) * npiv);
  OBA_HIP(hipMemcpyAsync(pd.p, hp.data(), sizeof(int) * npiv, hipMemcpyHostToDevice, r.stream));
  laswp<T>(n, A.d, A.ld, k1, k2, pd.as<int>(), incx, r.stream);
  A.out();
  end_call();
}

// ---- LAPACKE helpers ------------------------------------------------------------
enum { LAPACK_ROW_MAJOR = 101, LAPACK_COL_MAJOR = 102 };

void lapacke_xerbla(const char* name, int info) {
  if (info < 0) printf("Wrong parameter %d in %s\n", -info, name);
}

bool lapacke_nancheck_on() {
  const char* e = getenv("LAPACKE_NANCHECK");
  return !(e && atoi(e) == 0);
}

template <typename T>
__global__ void nan_any_kernel(long m, long n, const T* a, long lda, int tri, int* flag);

// A caller matrix that the LAPACK layer sees column-major on the device:
// column-major host/device memory is staged as usual; row-major host memory
// is staged as its transpose and transposed on the device.
template <typename T>
struct LMat {
  Mat<T> raw;
  DevBuf colbuf;
  T* d = nullptr;
  long ld = 0, rows = 0, cols = 0;
  bool rowmajor = false;
  void in(int layout, T* p, long r, long c, long ldp, bool copy_in) {
    rows = r; cols = c; rowmajor = layout == LAPACK_ROW_MAJOR;
    if (!rowmajor) { raw.in(p, r, c, ldp, copy_in); d = raw.d; ld = raw.ld; return; }
    raw.in(p, c, r, ldp, true);   // the row-major r x c array is a column-major c x r array
    ld = std::max(1L, r);
    colbuf = DevBuf(sizeof(T) * (size_t)ld * (size_t)std::max(1L, c));
    d = colbuf.as<T>();
    if (r > 0 && c > 0) transpose<T, T>(c, r, raw.d, raw.ld, d, ld, nullptr, rt().stream);
  }
  void out() {
    if (rowmajor && rows > 0 && cols > 0) transpose<T, T>(rows, cols, d, ld, raw.d, raw.ld, nullptr, rt().stream);
    raw.out();
  }
  // NaN anywhere (tri 0), or in the lower (1) / upper (2) triangle with the
  // diagonal of the column-major view (LAPACKE_dtr_nancheck, diag 'N')
  bool has_nan(int tri = 0) {
    if (rows <= 0 || cols <= 0) return false;
    DevBuf f(sizeof(int));
    OBA_HIP(hipMemsetAsync(f.p, 0, sizeof(int), rt().stream));
    long total = rows * cols;
    nan_any_kernel<T><<<(unsigned)std::min<long>((total + 255) / 256, 65535), 256, 0, rt().stream>>>(rows, cols, d, ld,
                                                                                                      tri, f.as<int>());
    OBA_HIP(hipGetLastError());
    int h = 0;
    OBA_HIP(hipMemcpyAsync(&h, f.p, sizeof(int), hipMemcpyDeviceToHost, rt().stream));
    OBA_HIP(hipStreamSynchronize(rt().stream));
    return h != 0;
  }
};

}  // namespace

namespace {
template <typename T>
__global__ void nan_any_kernel(long m, long n, const T* a, long lda, int tri, int* flag) {
  long total = m * n;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    long i = idx % m, j = idx / m;
    if ((tri == 1 && i < j) || (tri == 2 && i > j)) continue;
    if (isnan(a[i + j * lda])) { atomicOr(flag, 1); return; }
  }
}

template <typename T>
lapack_int lapacke_gesv(const char* name, const char* wname, int layout, lapack_int n, lapack_int nrhs, T* a, lapack_int lda, lapack_int* ipiv, T* b,
                 lapack_int ldb, bool nancheck) {
  if (layout != LAPACK_COL_MAJOR && layout != LAPACK_ROW_MAJOR) { lapacke_xerbla(name, -1); return -1; }
  if (layout == LAPACK_ROW_MAJOR) {
    if (lda < n) { lapacke_xerbla(wname, -5); return -5; }
    if (ldb < nrhs) { lapacke_xerbla(wname, -8); return -8; }
  }
  (void)rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  LMat<T> A, B;
  A.in(layout, a, n, n, lda, true);
  B.in(layout, b, n, nrhs, ldb, true);
  if (nancheck && lapacke_nancheck_on()) {
    if (A.has_nan()) return -4;
    if (B.has_nan()) return -7;
  }
  blasint N = n, NRHS = nrhs, LDA = (blasint)A.ld, LDB = (blasint)B.ld, info = 0;
  gesv_impl<T>(Names<T>::gesv, &N, &NRHS, A.d, &LDA, ipiv, B.d, &LDB, &info);
  if (info < 0) info -= 1;
  A.out();
  B.out();
  end_call();
  return info;
}

template <typename T>
lapack_int lapacke_getrf(const char* name, const char* wname, int layout, lapack_int m, lapack_int n, T* a, lapack_int lda, lapack_int* ipiv,
                  bool nancheck) {
  if (layout != LAPACK_COL_MAJOR && layout != LAPACK_ROW_MAJOR) { lapacke_xerbla(name, -1); return -1; }
  if (layout == LAPACK_ROW_MAJOR && lda < n) { lapacke_xerbla(wname, -5); return -5; }
  (void)rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  LMat<T> A;
  A.in(layout, a, m, n, lda, true);
  if (nancheck && lapacke_nancheck_on() && A.has_nan()) return -5;
  blasint M = m, N = n, LDA = (blasint)std::max(1L, A.ld), info = 0;
  getrf_impl<T>(Names<T>::getrf, &M, &N, A.d, &LDA, ipiv, &info);
  if (info < 0) info -= 1;
  A.out();
  end_call();
  return info;
}

template <typename T>
lapack_int lapacke_getrs(const char* name, const char* wname, int layout, char trans, lapack_int n, lapack_int nrhs, const T* a, lapack_int lda,
                  const lapack_int* ipiv, T* b, lapack_int ldb, bool nancheck) {
  if (layout != LAPACK_COL_MAJOR && layout != LAPACK_ROW_MAJOR) { lapacke_xerbla(name, -1); return -1; }
  if (layout == LAPACK_ROW_MAJOR) {
    if (lda < n) { lapacke_xerbla(wname, -6); return -6; }
    if (ldb < nrhs) { lapacke_xerbla(wname, -9); return -9; }
  }
  (void)rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  LMat<T> A, B;
  A.in(layout, const_cast<T*>(a), n, n, lda, true);
  B.in(layout, b, n, nrhs, ldb, true);
  if (nancheck && lapacke_nancheck_on()) {
    if (A.has_nan()) return -5;
    if (B.has_nan()) return -8;
  }
  blasint N = n, NRHS = nrhs, LDA = (blasint)std::max(1L, A.ld), LDB = (blasint)std::max(1L, B.ld), info = 0;
  char t = trans;
  getrs_impl<T>(Names<T>::getrs, &t, &N, &NRHS, A.d, &LDA, ipiv, B.d, &LDB, &info);
  if (info < 0) info -= 1;
  B.out();
  end_call();
  return info;
}

template <typename T>
lapack_int lapacke_potrf(const char* name, const char* wname, int layout, char uplo, lapack_int n, T* a, lapack_int lda, bool nancheck) {
  if (layout != LAPACK_COL_MAJOR && layout != LAPACK_ROW_MAJOR) { lapacke_xerbla(name, -1); return -1; }
  if (layout == LAPACK_ROW_MAJOR && lda < n) { lapacke_xerbla(wname, -5); return -5; }
  (void)rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  LMat<T> A;
  A.in(layout, a, n, n, lda, true);
  // LAPACKE_dpo_nancheck (lapacke_dpotrf.c:43-45): the referenced triangle
  // only; after the row-major transpose the column-major view's uplo
  // triangle is the referenced one.  A bad uplo skips the check (as
  // LAPACKE_dtr_nancheck does) and potrf reports it.
  const char uc = (char)toupper((unsigned char)uplo);
  if (nancheck && lapacke_nancheck_on() && (uc == 'L' || uc == 'U') && A.has_nan(uc == 'L' ? 1 : 2)) return -4;
  blasint N = n, LDA = (blasint)std::max(1L, A.ld), info = 0;
  char u = uplo;
  potrf_impl<T>(Names<T>::potrf, &u, &N, A.d, &LDA, &info);
  if (info < 0) info -= 1;
  A.out();
  end_call();
  return info;
}
}  // namespace

// ---- Fortran LAPACK symbols ------------------------------------------------------
OBA_API int dgetrf_(const blasint* m, const blasint* n, double* a, const blasint* lda, blasint* ipiv, blasint* info) {
  if (!guarded([&] { getrf_impl<double>("DGETRF", m, n, a, lda, ipiv, info); })) *info = OBA_DEVICE_ERROR;
  return 0;
}
OBA_API int sgetrf_(const blasint* m, const blasint* n, float* a, const blasint* lda, blasint* ipiv, blasint* info) {
  if (!guarded([&] { getrf_impl<float>("SGETRF", m, n, a, lda, ipiv, info); })) *info = OBA_DEVICE_ERROR;
  return 0;
}
OBA_API int dgetf2_(const blasint* m, const blasint* n, double* a, const blasint* lda, blasint* ipiv, blasint* info) {
  if (!guarded([&] { getrf_impl<double>("DGETF2", m, n, a, lda, ipiv, info); })) *info = OBA_DEVICE_ERROR;
  return 0;
}
OBA_API int sgetf2_(const blasint* m, const blasint* n, float* a, const blasint* lda, blasint* ipiv, blasint* info) {
  if (!guarded([&] { getrf_impl<float>("SGETF2", m, n, a, lda, ipiv, info); })) *info = OBA_DEVICE_ERROR;
  return 0;
}
OBA_API int dgetrs_(const char* trans, const blasint* n, const blasint* nrhs, const double* a, const blasint* lda,
                    const blasint* ipiv, double* b, const blasint* ldb, blasint* info) {
  if (!guarded([&] { getrs_impl<double>("DGETRS", trans, n, nrhs, a, lda, ipiv, b, ldb, info); })) *info = OBA_DEVICE_ERROR;
  return 0;
}
OBA_API int sgetrs_(const char* trans, const blasint* n, const blasint* nrhs, const float* a, const blasint* lda,
                    const blasint* ipiv, float* b, const blasint* ldb, blasint* info) {
  if (!guarded([&] { getrs_impl<float>("SGETRS", trans, n, nrhs, a, lda, ipiv, b, ldb, info); })) *info = OBA_DEVICE_ERROR;
  return 0;
}
OBA_API int dgesv_(const blasint* n, const blasint* nrhs, double* a, const blasint* lda, blasint* ipiv, double* b,
                   const blasint* ldb, blasint* info) {
  if (!guarded([&] { gesv_impl<double>("DGESV ", n, nrhs, a, lda, ipiv, b, ldb, info); })) *info = OBA_DEVICE_ERROR;
  return 0;
}
OBA_API int sgesv_(const blasint* n, const blasint* nrhs, float* a, const blasint* lda, blasint* ipiv, float* b,
                   const blasint* ldb, blasint* info) {
  if (!guarded([&] { gesv_impl<float>("SGESV ", n, nrhs, a, lda, ipiv, b, ldb, info); })) *info = OBA_DEVICE_ERROR;
  return 0;
}
OBA_API int dpotrf_(const char* uplo, const blasint* n, double* a, const blasint* lda, blasint* info) {
  if (!guarded([&] { potrf_impl<double>("DPOTRF", uplo, n, a, lda, info); })) *info = OBA_DEVICE_ERROR;
  return 0;
}
OBA_API int spotrf_(const char* uplo, const blasint* n, float* a, const blasint* lda, blasint* info) {
  if (!guarded([&] { potrf_impl<float>("SPOTRF", uplo, n, a, lda, info); })) *info = OBA_DEVICE_ERROR;
  return 0;
}
OBA_API int dpotf2_(const char* uplo, const blasint* n, double* a, const blasint* lda, blasint* info) {
  if (!guarded([&] { potrf_impl<double>("DPOTF2", uplo, n, a, lda, info); })) *info = OBA_DEVICE_ERROR;
  return 0;
}
OBA_API int spotf2_(const char* uplo, const blasint* n, float* a, const blasint* lda, blasint* info) {
  if (!guarded([&] { potrf_impl<float>("SPOTF2", uplo, n, a, lda, info); })) *info = OBA_DEVICE_ERROR;
  return 0;
}
OBA_API int dlaswp_(const blasint* n, double* a, const blasint* lda, const blasint* k1, const blasint* k2,
                    const blasint* ipiv, const blasint* incx) {
  guarded([&] { laswp_impl<double>(n, a, lda, k1, k2, ipiv, incx); });
  return 0;
}
OBA_API int slaswp_(const blasint* n, float* a, const blasint* lda, const blasint* k1, const blasint* k2,
                    const blasint* ipiv, const blasint* incx) {
  guarded([&] { laswp_impl<float>(n, a, lda, k1, k2, ipiv, incx); });
  return 0;
}

// netlib dsgesv.f:193 / dsposv.f:197 signatures; WORK and SWORK are accepted
// and unused (the device pool holds the FP32 copy and the residual).
namespace {
void dsgesv_impl(const blasint* N, const blasint* NRHS, double* a, const blasint* LDA, blasint* ipiv, const double* b,
                 const blasint* LDB, double* x, const blasint* LDX, blasint* ITER, blasint* INFO) {
  long n = *N, nrhs = *NRHS, lda = *LDA, ldb = *LDB, ldx = *LDX;
  blasint info = 0;
  *ITER = 0;
  if (n < 0) info = -1;
  else if (nrhs < 0) info = -2;
  else if (lda < std::max(1L, n)) info = -4;
  else if (ldb < std::max(1L, n)) info = -7;
  else if (ldx < std::max(1L, n)) info = -9;
  if (info) { *INFO = info; report_xerbla("DSGESV", -info); return; }
  *INFO = 0;
  if (n == 0) return;   // dsgesv.f:264; with NRHS = 0 the factorization still runs and sets IPIV
  Runtime& r = rt();
  Mat<double> A, B, X;
  A.in(a, n, n, lda, true);
  B.in(const_cast<double*>(b), n, nrhs, ldb, true);
  X.in(x, n, nrhs, ldx, false);
  IntDev P;
  P.in(ipiv, n, false);
  int it = 0, inf = 0;
  dsgesv_device(n, nrhs, A.d, A.ld, P.d, B.d, B.ld, X.d, X.ld, &it, &inf, nullptr, r.stream);
  *ITER = it;
  *INFO = inf;
  if (it < 0) A.out();   // A holds the FP64 factors only after the fallback
  X.out();
  P.out();
  end_call();
}

void dsposv_impl(const char* UPLO, const blasint* N, const blasint* NRHS, double* a, const blasint* LDA,
                 const double* b, const blasint* LDB, double* x, const blasint* LDX, blasint* ITER, blasint* INFO) {
  const char uc = (char)toupper((unsigned char)*UPLO);
  long n = *N, nrhs = *NRHS, lda = *LDA, ldb = *LDB, ldx = *LDX;
  blasint info = 0;
  *ITER = 0;
  if (uc != 'U' && uc != 'L') info = -1;
  else if (n < 0) info = -2;
  else if (nrhs < 0) info = -3;
  else if (lda < std::max(1L, n)) info = -5;
  else if (ldb < std::max(1L, n)) info = -7;
  else if (ldx < std::max(1L, n)) info = -9;
  if (info) { *INFO = info; report_xerbla("DSPOSV", -info); return; }
  *INFO = 0;
  if (n == 0) return;   // dsposv.f:257
  Runtime& r = rt();
  Mat<double> A, B, X;
  A.in(a, n, n, lda, true);
  B.in(const_cast<double*>(b), n, nrhs, ldb, true);
  X.in(x, n, nrhs, ldx, false);
  int it = 0, inf = 0;
  dsposv_device(uc == 'L', n, nrhs, A.d, A.ld, B.d, B.ld, X.d, X.ld, &it, &inf, nullptr, r.stream);
  *ITER = it;
  *INFO = inf;
  if (it < 0) A.out();   // the FP64 Cholesky factor, only after the fallback
  X.out();
  end_call();
}

// LAPACKE_dsgesv_work / LAPACKE_dsposv_work (lapacke_dsgesv_work.c:35-110,
// lapacke_dsposv_work.c:35-115): layout check, row-major transposes, info
// shift; the NaN pre-check belongs to the high-level wrappers only.
lapack_int lapacke_dsxsv(bool chol, char uplo, const char* name, const char* wname, int layout, lapack_int n, lapack_int nrhs, double* a,
                  lapack_int lda, lapack_int* ipiv, double* b, lapack_int ldb, double* x, lapack_int ldx, lapack_int* iter, bool nancheck) {
  if (layout != LAPACK_COL_MAJOR && layout != LAPACK_ROW_MAJOR) { lapacke_xerbla(name, -1); return -1; }
  const int sh = chol ? 1 : 0;   // dsposv has UPLO in front
  if (layout == LAPACK_ROW_MAJOR) {
    if (lda < n) { lapacke_xerbla(wname, -5 - sh); return -5 - sh; }
    if (ldb < nrhs) { lapacke_xerbla(wname, -8); return -8; }
    if (ldx < nrhs) { lapacke_xerbla(wname, -10); return -10; }
  }
  LMat<double> A, B, X;
  A.in(layout, a, n, n, lda, true);
  B.in(layout, b, n, nrhs, ldb, true);
  X.in(layout, x, n, nrhs, ldx, layout == LAPACK_ROW_MAJOR);
  if (nancheck && lapacke_nancheck_on()) {
    if (chol) {
      const char uc = (char)toupper((unsigned char)uplo);
      if ((uc == 'L' || uc == 'U') && A.has_nan(uc == 'L' ? 1 : 2)) return -5;
    } else if (A.has_nan()) {
      return -4;
    }
    if (B.has_nan()) return -7;
  }
  blasint N = n, NRHS = nrhs, LDA = (blasint)std::max(1L, A.ld), LDB = (blasint)std::max(1L, B.ld),
          LDX = (blasint)std::max(1L, X.ld), it = 0, info = 0;
  if (chol) {
    char u = uplo;
    dsposv_impl(&u, &N, &NRHS, A.d, &LDA, B.d, &LDB, X.d, &LDX, &it, &info);
  } else {
    dsgesv_impl(&N, &NRHS, A.d, &LDA, ipiv, B.d, &LDB, X.d, &LDX, &it, &info);
  }
  if (info < 0) info -= 1;
  *iter = it;
  if (it < 0) A.out();
  X.out();
  end_call();
  return info;
}
}  // namespace

OBA_API int dsgesv_(const blasint* N, const blasint* NRHS, double* a, const blasint* LDA, blasint* ipiv,
                    const double* b, const blasint* LDB, double* x, const blasint* LDX, double* work, float* swork,
                    blasint* ITER, blasint* INFO) {
  (void)work; (void)swork;
  if (!guarded([&] { dsgesv_impl(N, NRHS, a, LDA, ipiv, b, LDB, x, LDX, ITER, INFO); })) *INFO = OBA_DEVICE_ERROR;
  return 0;
}
OBA_API int dsposv_(const char* uplo, const blasint* N, const blasint* NRHS, double* a, const blasint* LDA,
                    const double* b, const blasint* LDB, double* x, const blasint* LDX, double* work, float* swork,
                    blasint* ITER, blasint* INFO) {
  (void)work; (void)swork;
  if (!guarded([&] { dsposv_impl(uplo, N, NRHS, a, LDA, b, LDB, x, LDX, ITER, INFO); })) *INFO = OBA_DEVICE_ERROR;
  return 0;
}

// ---- LAPACKE symbols ----------------------------------------------------------------
OBA_API lapack_int LAPACKE_dgesv(int layout, lapack_int n, lapack_int nrhs, double* a, lapack_int lda, lapack_int* ipiv, double* b, lapack_int ldb) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_gesv<double>("LAPACKE_dgesv", "LAPACKE_dgesv_work", layout, n, nrhs, a, lda, ipiv, b, ldb, true); });
  return r;
}
OBA_API lapack_int LAPACKE_dgesv_work(int layout, lapack_int n, lapack_int nrhs, double* a, lapack_int lda, lapack_int* ipiv, double* b, lapack_int ldb) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_gesv<double>("LAPACKE_dgesv_work", "LAPACKE_dgesv_work", layout, n, nrhs, a, lda, ipiv, b, ldb, false); });
  return r;
}
OBA_API lapack_int LAPACKE_sgesv(int layout, lapack_int n, lapack_int nrhs, float* a, lapack_int lda, lapack_int* ipiv, float* b, lapack_int ldb) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_gesv<float>("LAPACKE_sgesv", "LAPACKE_sgesv_work", layout, n, nrhs, a, lda, ipiv, b, ldb, true); });
  return r;
}
OBA_API lapack_int LAPACKE_sgesv_work(int layout, lapack_int n, lapack_int nrhs, float* a, lapack_int lda, lapack_int* ipiv, float* b, lapack_int ldb) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_gesv<float>("LAPACKE_sgesv_work", "LAPACKE_sgesv_work", layout, n, nrhs, a, lda, ipiv, b, ldb, false); });
  return r;
}
OBA_API lapack_int LAPACKE_dgetrf(int layout, lapack_int m, lapack_int n, double* a, lapack_int lda, lapack_int* ipiv) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_getrf<double>("LAPACKE_dgetrf", "LAPACKE_dgetrf_work", layout, m, n, a, lda, ipiv, true); });
  return r;
}
OBA_API lapack_int LAPACKE_dgetrf_work(int layout, lapack_int m, lapack_int n, double* a, lapack_int lda, lapack_int* ipiv) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_getrf<double>("LAPACKE_dgetrf_work", "LAPACKE_dgetrf_work", layout, m, n, a, lda, ipiv, false); });
  return r;
}
OBA_API lapack_int LAPACKE_sgetrf(int layout, lapack_int m, lapack_int n, float* a, lapack_int lda, lapack_int* ipiv) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_getrf<float>("LAPACKE_sgetrf", "LAPACKE_sgetrf_work", layout, m, n, a, lda, ipiv, true); });
  return r;
}
OBA_API lapack_int LAPACKE_sgetrf_work(int layout, lapack_int m, lapack_int n, float* a, lapack_int lda, lapack_int* ipiv) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_getrf<float>("LAPACKE_sgetrf_work", "LAPACKE_sgetrf_work", layout, m, n, a, lda, ipiv, false); });
  return r;
}
OBA_API lapack_int LAPACKE_dgetrs(int layout, char trans, lapack_int n, lapack_int nrhs, const double* a, lapack_int lda, const lapack_int* ipiv, double* b,
                           lapack_int ldb) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_getrs<double>("LAPACKE_dgetrs", "LAPACKE_dgetrs_work", layout, trans, n, nrhs, a, lda, ipiv, b, ldb, true); });
  return r;
}
OBA_API lapack_int LAPACKE_dgetrs_work(int layout, char trans, lapack_int n, lapack_int nrhs, const double* a, lapack_int lda, const lapack_int* ipiv,
                                double* b, lapack_int ldb) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_getrs<double>("LAPACKE_dgetrs_work", "LAPACKE_dgetrs_work", layout, trans, n, nrhs, a, lda, ipiv, b,
                               ldb, false); });
  return r;
}
OBA_API lapack_int LAPACKE_sgetrs(int layout, char trans, lapack_int n, lapack_int nrhs, const float* a, lapack_int lda, const lapack_int* ipiv, float* b,
                           lapack_int ldb) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_getrs<float>("LAPACKE_sgetrs", "LAPACKE_sgetrs_work", layout, trans, n, nrhs, a, lda, ipiv, b, ldb, true); });
  return r;
}
OBA_API lapack_int LAPACKE_sgetrs_work(int layout, char trans, lapack_int n, lapack_int nrhs, const float* a, lapack_int lda, const lapack_int* ipiv,
                                float* b, lapack_int ldb) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_getrs<float>("LAPACKE_sgetrs_work", "LAPACKE_sgetrs_work", layout, trans, n, nrhs, a, lda, ipiv, b,
                              ldb, false); });
  return r;
}
OBA_API lapack_int LAPACKE_dpotrf(int layout, char uplo, lapack_int n, double* a, lapack_int lda) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_potrf<double>("LAPACKE_dpotrf", "LAPACKE_dpotrf_work", layout, uplo, n, a, lda, true); });
  return r;
}
OBA_API lapack_int LAPACKE_dpotrf_work(int layout, char uplo, lapack_int n, double* a, lapack_int lda) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_potrf<double>("LAPACKE_dpotrf_work", "LAPACKE_dpotrf_work", layout, uplo, n, a, lda, false); });
  return r;
}
OBA_API lapack_int LAPACKE_spotrf(int layout, char uplo, lapack_int n, float* a, lapack_int lda) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_potrf<float>("LAPACKE_spotrf", "LAPACKE_spotrf_work", layout, uplo, n, a, lda, true); });
  return r;
}
OBA_API lapack_int LAPACKE_spotrf_work(int layout, char uplo, lapack_int n, float* a, lapack_int lda) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] { r = lapacke_potrf<float>("LAPACKE_spotrf_work", "LAPACKE_spotrf_work", layout, uplo, n, a, lda, false); });
  return r;
}
OBA_API lapack_int LAPACKE_dsgesv(int layout, lapack_int n, lapack_int nrhs, double* a, lapack_int lda, lapack_int* ipiv, double* b, lapack_int ldb, double* x,
                           lapack_int ldx, lapack_int* iter) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] {
    r = lapacke_dsxsv(false, 'N', "LAPACKE_dsgesv", "LAPACKE_dsgesv_work", layout, n, nrhs, a, lda, ipiv, b, ldb, x,
                      ldx, iter, true);
  });
  return r;
}
OBA_API lapack_int LAPACKE_dsgesv_work(int layout, lapack_int n, lapack_int nrhs, double* a, lapack_int lda, lapack_int* ipiv, double* b, lapack_int ldb,
                                double* x, lapack_int ldx, double* work, float* swork, lapack_int* iter) {
  (void)work; (void)swork;
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] {
    r = lapacke_dsxsv(false, 'N', "LAPACKE_dsgesv_work", "LAPACKE_dsgesv_work", layout, n, nrhs, a, lda, ipiv, b, ldb,
                      x, ldx, iter, false);
  });
  return r;
}
OBA_API lapack_int LAPACKE_dsposv(int layout, char uplo, lapack_int n, lapack_int nrhs, double* a, lapack_int lda, double* b, lapack_int ldb, double* x,
                           lapack_int ldx, lapack_int* iter) {
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] {
    r = lapacke_dsxsv(true, uplo, "LAPACKE_dsposv", "LAPACKE_dsposv_work", layout, n, nrhs, a, lda, nullptr, b, ldb, x,
                      ldx, iter, true);
  });
  return r;
}
OBA_API lapack_int LAPACKE_dsposv_work(int layout, char uplo, lapack_int n, lapack_int nrhs, double* a, lapack_int lda, double* b, lapack_int ldb,
                                double* x, lapack_int ldx, double* work, float* swork, lapack_int* iter) {
  (void)work; (void)swork;
  lapack_int r = OBA_DEVICE_ERROR;
  guarded([&] {
    r = lapacke_dsxsv(true, uplo, "LAPACKE_dsposv_work", "LAPACKE_dsposv_work", layout, n, nrhs, a, lda, nullptr, b,
                      ldb, x, ldx, iter, false);
  });
  return r;
}
