// Level-3 BLAS C ABI beyond GEMM/TRSM: cblas_?syrk, ?syrk_, cblas_?gemmt,
// ?gemmt_, cblas_?symm, ?symm_, cblas_?trmm, ?trmm_.  Argument checks, info
// numbers, the row-major mapping and the quick returns follow the
// reference's interface layer:
//   interface/syrk.c:100-196 (Fortran), :198-330 (CBLAS), :332 (n == 0)
//   interface/gemmt.c:67-198 (Fortran), :200-397 (CBLAS), :467 (m == 0)
//   interface/symm.c:100-230 (Fortran), :232-376 (CBLAS), :376 (m|n == 0)
//   interface/trsm.c (TRMM instantiation): the same checks as TRSM
// Caller memory may be host or device; host matrices are staged through the
// device pool.  Calls are synchronous.
#include "kernels.h"
#include <cctype>

namespace oba {
template <typename T>
void syrk_device(bool lower, bool trans, long n, long k, T alpha, const T* a, long lda, T beta, T* c, long ldc,
                 hipStream_t s);
template <typename T>
void gemmt_device(bool lower, bool ta, bool tb, long m, long k, T alpha, const T* a, long lda, const T* b, long ldb,
                  T beta, T* c, long ldc, hipStream_t s);
template <typename T>
void symm_device(bool right, bool lower, long m, long n, T alpha, const T* a, long lda, const T* b, long ldb, T beta,
                 T* c, long ldc, hipStream_t s);
template <typename T>
void trmm_device(bool right, bool lower, bool trans, bool unit, long m, long n, T alpha, const T* a, long lda, T* b,
                 long ldb, hipStream_t s);
}  // namespace oba

using namespace oba;

namespace {

enum { RowMajor = 101, ColMajor = 102, NoTrans = 111, Trans = 112, ConjTrans = 113, ConjNoTrans = 114,
       Upper = 121, Lower = 122, NonUnit = 131, Unit = 132, Left = 141, Right = 142 };

int cblas_trans(int t) {
  if (t == NoTrans || t == ConjNoTrans) return 0;
  if (t == Trans || t == ConjTrans) return 1;
  return -1;
}
int fortran_trans(char c) {
  c = (char)toupper((unsigned char)c);
  if (c == 'N' || c == 'R') return 0;
  if (c == 'T' || c == 'C') return 1;
  return -1;
}
int up(const char* c) { return toupper((unsigned char)*c); }

// ---- SYRK -----------------------------------------------------------------------
// uplo: 0 upper / 1 lower; trans: 0 C = A A^T (A n x k) / 1 C = A^T A (A k x n)
template <typename T>
void syrk_run(int uplo, int trans, long n, long k, T alpha, const T* a, long lda, T beta, T* c, long ldc) {
  if (n == 0) return;
  const bool only_scale = k == 0 || alpha == T(0);
  if (only_scale && beta == T(1)) return;
  Runtime& r = rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  Mat<T> A, C;
  if (!only_scale) A.in(const_cast<T*>(a), trans ? k : n, trans ? n : k, lda, true);
  C.in(c, n, n, ldc, true);   // the other triangle travels unchanged
  {
    KTimer kt(sizeof(T) == 8 ? "dsyrk" : "ssyrk", r.stream, only_scale ? 0.0 : (double)n * (n + 1) * k,
              (double)sizeof(T) * ((only_scale ? 0.0 : (double)n * k) + (double)n * (n + 1)));
    syrk_device<T>(uplo == 1, trans == 1, n, k, alpha, A.d, A.ld, beta, C.d, C.ld, r.stream);
  }
  C.out();
  end_call();
}

template <typename T>
void f_syrk_impl(const char* name, const char* UPLO, const char* TRANS, const blasint* N, const blasint* K,
                 const T* alpha, const T* a, const blasint* lda, const T* beta, T* c, const blasint* ldc) {
  const int u = up(UPLO), tc = up(TRANS);
  const int uplo = u == 'U' ? 0 : u == 'L' ? 1 : -1;
  const int trans = tc == 'N' ? 0 : (tc == 'T' || tc == 'C') ? 1 : -1;
  const long n = *N, k = *K;
  const long nrowa = trans == 1 ? k : n;
  blasint info = 0;
  if (*ldc < std::max(1L, n)) info = 10;
  if (*lda < std::max(1L, nrowa)) info = 7;
  if (k < 0) info = 4;
  if (n < 0) info = 3;
  if (trans < 0) info = 2;
  if (uplo < 0) info = 1;
  if (info) { report_xerbla(name, info); return; }
  syrk_run<T>(uplo, trans, n, k, *alpha, a, *lda, *beta, c, *ldc);
}

template <typename T>
void cblas_syrk_impl(const char* name, int order, int Uplo, int Trans, blasint n, blasint k, T alpha, const T* a,
                     blasint lda, T beta, T* c, blasint ldc) {
  int uplo = -1, trans = -1;
  blasint info = 0;
  if (order == ColMajor) {
    if (Uplo == Upper) uplo = 0;
    if (Uplo == Lower) uplo = 1;
    trans = cblas_trans(Trans);
  } else if (order == RowMajor) {
    if (Uplo == Upper) uplo = 1;
    if (Uplo == Lower) uplo = 0;
    const int t = cblas_trans(Trans);
    trans = t < 0 ? -1 : 1 - t;
  }
  if (order == ColMajor || order == RowMajor) {
    const long nrowa = trans == 1 ? k : n;
    info = -1;
    if (ldc < std::max(1, (int)n)) info = 10;
    if (lda < std::max(1L, nrowa)) info = 7;
    if (k < 0) info = 4;
    if (n < 0) info = 3;
    if (trans < 0) info = 2;
    if (uplo < 0) info = 1;
  }
  if (info >= 0) { report_xerbla(name, info); return; }
  syrk_run<T>(uplo, trans, n, k, alpha, a, lda, beta, c, ldc);
}

// ---- GEMMT ----------------------------------------------------------------------
template <typename T>
void gemmt_run(int uplo, int ta, int tb, long m, long k, T alpha, const T* a, long lda, const T* b, long ldb, T beta,
               T* c, long ldc) {
  if (m == 0) return;
  const bool only_scale = k == 0 || alpha == T(0);
  if (only_scale && beta == T(1)) return;
  Runtime& r = rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  Mat<T> A, B, C;
  if (!only_scale) {
    A.in(const_cast<T*>(a), ta ? k : m, ta ? m : k, lda, true);
    B.in(const_cast<T*>(b), tb ? m : k, tb ? k : m, ldb, true);
  }
  C.in(c, m, m, ldc, true);
  {
    KTimer kt(sizeof(T) == 8 ? "dgemmt" : "sgemmt", r.stream, only_scale ? 0.0 : (double)m * (m + 1) * k,
              (double)sizeof(T) * ((only_scale ? 0.0 : 2.0 * m * k) + (double)m * (m + 1)));
    gemmt_device<T>(uplo == 1, ta == 1, tb == 1, m, k, alpha, A.d, A.ld, B.d, B.ld, beta, C.d, C.ld, r.stream);
  }
  C.out();
  end_call();
}

template <typename T>
void f_gemmt_impl(const char* name, const char* UPLO, const char* TA, const char* TB, const blasint* M,
                  const blasint* K, const T* alpha, const T* a, const blasint* lda, const T* b, const blasint* ldb,
                  const T* beta, T* c, const blasint* ldc) {
  const int u = up(UPLO);
  const int uplo = u == 'U' ? 0 : u == 'L' ? 1 : -1;
  const int ta = fortran_trans(*TA), tb = fortran_trans(*TB);
  const long m = *M, k = *K;
  const long nrowa = ta == 1 ? k : m, nrowb = tb == 1 ? m : k;
  blasint info = 0;
  if (*ldc < std::max(1L, m)) info = 13;
  if (*ldb < std::max(1L, nrowb)) info = 10;
  if (*lda < std::max(1L, nrowa)) info = 8;
  if (k < 0) info = 5;
  if (m < 0) info = 4;
  if (tb < 0) info = 3;
  if (ta < 0) info = 2;
  if (uplo < 0) info = 1;
  if (info) { report_xerbla(name, info); return; }
  gemmt_run<T>(uplo, ta, tb, m, k, *alpha, a, *lda, b, *ldb, *beta, c, *ldc);
}

template <typename T>
void cblas_gemmt_impl(const char* name, int order, int Uplo, int TransA, int TransB, blasint m, blasint k, T alpha,
                      const T* A, blasint LDA, const T* B, blasint LDB, T beta, T* c, blasint ldc) {
  int uplo = -1, ta = -1, tb = -1;
  const T *pa = nullptr, *pb = nullptr;
  long lda = 0, ldb = 0;
  blasint info = 0;
  if (order == ColMajor) {
    if (Uplo == Upper) uplo = 0;
    if (Uplo == Lower) uplo = 1;
    ta = cblas_trans(TransA);
    tb = cblas_trans(TransB);
    pa = A; pb = B; lda = LDA; ldb = LDB;
    const long nrowa = ta == 1 ? k : m, nrowb = tb == 1 ? m : k;
    info = -1;
    if (ldc < std::max(1, (int)m)) info = 13;
    if (ldb < std::max(1L, nrowb)) info = 10;
    if (lda < std::max(1L, nrowa)) info = 8;
    if (k < 0) info = 5;
    if (m < 0) info = 4;
    if (tb < 0) info = 3;
    if (ta < 0) info = 2;
    if (uplo < 0) info = 1;
  } else if (order == RowMajor) {
    // C^T = op(B)^T op(A)^T on the other triangle
    pa = B; pb = A; lda = LDB; ldb = LDA;
    if (Uplo == Upper) uplo = 1;
    if (Uplo == Lower) uplo = 0;
    ta = cblas_trans(TransB);
    tb = cblas_trans(TransA);
    const long ncola = ta == 1 ? k : m, ncolb = tb == 1 ? m : k;
    info = -1;
    if (ldc < std::max(1, (int)m)) info = 13;
    if (ldb < std::max(1L, ncolb)) info = 8;
    if (lda < std::max(1L, ncola)) info = 10;
    if (k < 0) info = 5;
    if (m < 0) info = 4;
    if (tb < 0) info = 2;
    if (ta < 0) info = 3;
    if (uplo < 0) info = 1;
  }
  if (info >= 0) { report_xerbla(name, info); return; }
  gemmt_run<T>(uplo, ta, tb, m, k, alpha, pa, lda, pb, ldb, beta, c, ldc);
}

// ---- SYMM -----------------------------------------------------------------------
// side: 0 left (C = A B, A m x m) / 1 right (C = B A, A n x n); uplo: 0 upper / 1 lower
template <typename T>
void symm_run(int side, int uplo, long m, long n, T alpha, const T* a, long lda, const T* b, long ldb, T beta, T* c,
              long ldc) {
  if (m == 0 || n == 0) return;
  if (alpha == T(0) && beta == T(1)) return;
  Runtime& r = rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  const long na = side ? n : m;
  Mat<T> A, B, C;
  if (alpha != T(0)) {
    A.in(const_cast<T*>(a), na, na, lda, true);
    B.in(const_cast<T*>(b), m, n, ldb, true);
  }
  C.in(c, m, n, ldc, beta != T(0));
  {
    KTimer kt(sizeof(T) == 8 ? "dsymm" : "ssymm", r.stream, alpha == T(0) ? 0.0 : 2.0 * m * n * na,
              (double)sizeof(T) * ((alpha == T(0) ? 0.0 : (double)na * na + (double)m * n) + 2.0 * m * n));
    symm_device<T>(side == 1, uplo == 1, m, n, alpha, A.d, A.ld, B.d, B.ld, beta, C.d, C.ld, r.stream);
  }
  C.out();
  end_call();
}

template <typename T>
void symm_checks(int side, int uplo, long m, long n, long lda, long ldb, long ldc, blasint* info) {
  if (ldc < std::max(1L, m)) *info = 12;
  if (!side) {
    if (ldb < std::max(1L, m)) *info = 9;
    if (lda < std::max(1L, m)) *info = 7;
  } else {
    if (ldb < std::max(1L, m)) *info = 9;
    if (lda < std::max(1L, n)) *info = 7;
  }
  if (n < 0) *info = 4;
  if (m < 0) *info = 3;
  if (uplo < 0) *info = 2;
  if (side < 0) *info = 1;
}

template <typename T>
void f_symm_impl(const char* name, const char* SIDE, const char* UPLO, const blasint* M, const blasint* N,
                 const T* alpha, const T* a, const blasint* lda, const T* b, const blasint* ldb, const T* beta, T* c,
                 const blasint* ldc) {
  const int s = up(SIDE), u = up(UPLO);
  const int side = s == 'L' ? 0 : s == 'R' ? 1 : -1;
  const int uplo = u == 'U' ? 0 : u == 'L' ? 1 : -1;
  blasint info = 0;
  symm_checks<T>(side, uplo, *M, *N, *lda, *ldb, *ldc, &info);
  if (info) { report_xerbla(name, info); return; }
  symm_run<T>(side, uplo, *M, *N, *alpha, a, *lda, b, *ldb, *beta, c, *ldc);
}

template <typename T>
void cblas_symm_impl(const char* name, int order, int Side, int Uplo, blasint m, blasint n, T alpha, const T* a,
                     blasint lda, const T* b, blasint ldb, T beta, T* c, blasint ldc) {
  int side = -1, uplo = -1;
  long M = 0, N = 0;
  blasint info = 0;
  if (order == ColMajor) {
    if (Side == Left) side = 0;
    if (Side == Right) side = 1;
    if (Uplo == Upper) uplo = 0;
    if (Uplo == Lower) uplo = 1;
    M = m; N = n;
  } else if (order == RowMajor) {
    if (Side == Left) side = 1;
    if (Side == Right) side = 0;
    if (Uplo == Upper) uplo = 1;
    if (Uplo == Lower) uplo = 0;
    M = n; N = m;
  }
  if (order == ColMajor || order == RowMajor) {
    info = -1;
    blasint i2 = 0;
    symm_checks<T>(side, uplo, M, N, lda, ldb, ldc, &i2);
    if (i2) info = i2;
  }
  if (info >= 0) { report_xerbla(name, info); return; }
  symm_run<T>(side, uplo, M, N, alpha, a, lda, b, ldb, beta, c, ldc);
}

// ---- TRMM -----------------------------------------------------------------------
// side: 0 left / 1 right; uplo: 0 upper / 1 lower; unit: 0 unit / 1 non-unit
template <typename T>
void trmm_run(int side, int uplo, int trans, int unit, long m, long n, T alpha, const T* a, long lda, T* b, long ldb) {
  if (m == 0 || n == 0) return;
  Runtime& r = rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  const long na = side ? n : m;
  Mat<T> A, B;
  B.in(b, m, n, ldb, alpha != T(0));
  if (alpha != T(0)) A.in(const_cast<T*>(a), na, na, lda, true);
  {
    KTimer kt(sizeof(T) == 8 ? "dtrmm" : "strmm", r.stream, (double)m * n * (double)na,
              (double)sizeof(T) * (0.5 * na * na + 2.0 * m * n));
    trmm_device<T>(side == 1, uplo == 1, trans == 1, unit == 0, m, n, alpha, A.d, A.ld, B.d, B.ld, r.stream);
  }
  B.out();
  end_call();
}

template <typename T>
void f_trmm_impl(const char* name, const char* SIDE, const char* UPLO, const char* TRANSA, const char* DIAG,
                 const blasint* m, const blasint* n, const T* alpha, const T* a, const blasint* lda, T* b,
                 const blasint* ldb) {
  const int s = up(SIDE), u = up(UPLO), d = up(DIAG);
  const int side = s == 'L' ? 0 : s == 'R' ? 1 : -1;
  const int uplo = u == 'U' ? 0 : u == 'L' ? 1 : -1;
  const int trans = fortran_trans(*TRANSA);
  const int unit = d == 'U' ? 0 : d == 'N' ? 1 : -1;
  const long M = *m, N = *n, nrowa = side == 1 ? N : M;
  blasint info = 0;
  if (*ldb < std::max(1L, M)) info = 11;
  if (*lda < std::max(1L, nrowa)) info = 9;
  if (N < 0) info = 6;
  if (M < 0) info = 5;
  if (unit < 0) info = 4;
  if (trans < 0) info = 3;
  if (uplo < 0) info = 2;
  if (side < 0) info = 1;
  if (info) { report_xerbla(name, info); return; }
  trmm_run<T>(side, uplo, trans, unit, M, N, *alpha, a, *lda, b, *ldb);
}

template <typename T>
void cblas_trmm_impl(const char* name, int order, int Side, int Uplo, int Trans, int Diag, blasint m, blasint n,
                     T alpha, const T* a, blasint lda, T* b, blasint ldb) {
  int side = -1, uplo = -1, trans = -1, unit = -1;
  long M = 0, N = 0;
  blasint info = 0;
  if (order == ColMajor) {
    M = m; N = n;
    if (Side == Left) side = 0;
    if (Side == Right) side = 1;
    if (Uplo == Upper) uplo = 0;
    if (Uplo == Lower) uplo = 1;
  } else if (order == RowMajor) {
    M = n; N = m;
    if (Side == Left) side = 1;
    if (Side == Right) side = 0;
    if (Uplo == Upper) uplo = 1;
    if (Uplo == Lower) uplo = 0;
  }
  if (order == ColMajor || order == RowMajor) {
    trans = cblas_trans(Trans);
    if (Diag == Unit) unit = 0;
    if (Diag == NonUnit) unit = 1;
    const long nrowa = side == 1 ? N : M;
    info = -1;
    if (ldb < std::max(1L, M)) info = 11;
    if (lda < std::max(1L, nrowa)) info = 9;
    if (N < 0) info = 6;
    if (M < 0) info = 5;
    if (unit < 0) info = 4;
    if (trans < 0) info = 3;
    if (uplo < 0) info = 2;
    if (side < 0) info = 1;
  }
  if (info >= 0) { report_xerbla(name, info); return; }
  trmm_run<T>(side, uplo, trans, unit, M, N, alpha, a, lda, b, ldb);
}

}  // namespace

// ---- exported symbols ------------------------------------------------------------
#define X3_SYRK(P, T, NM)                                                                                          \
  OBA_API void cblas_##P##syrk(int order, int uplo, int trans, blasint n, blasint k, T alpha, const T* a,           \
                               blasint lda, T beta, T* c, blasint ldc) {                                          \
    guarded([&] { cblas_syrk_impl<T>(NM, order, uplo, trans, n, k, alpha, a, lda, beta, c, ldc); });              \
  }                                                                                                               \
  OBA_API void P##syrk_(const char* uplo, const char* trans, const blasint* n, const blasint* k, const T* alpha,   \
                        const T* a, const blasint* lda, const T* beta, T* c, const blasint* ldc) {                 \
    guarded([&] { f_syrk_impl<T>(NM, uplo, trans, n, k, alpha, a, lda, beta, c, ldc); });                         \
  }
X3_SYRK(d, double, "DSYRK ")
X3_SYRK(s, float, "SSYRK ")

#define X3_GEMMT(P, T, NM)                                                                                         \
  OBA_API void cblas_##P##gemmt(int order, int uplo, int ta, int tb, blasint m, blasint k, T alpha, const T* a,     \
                                blasint lda, const T* b, blasint ldb, T beta, T* c, blasint ldc) {                \
    guarded([&] { cblas_gemmt_impl<T>(NM, order, uplo, ta, tb, m, k, alpha, a, lda, b, ldb, beta, c, ldc); });    \
  }                                                                                                               \
  OBA_API void P##gemmt_(const char* uplo, const char* ta, const char* tb, const blasint* m, const blasint* k,      \
                         const T* alpha, const T* a, const blasint* lda, const T* b, const blasint* ldb,           \
                         const T* beta, T* c, const blasint* ldc) {                                                \
    guarded([&] { f_gemmt_impl<T>(NM, uplo, ta, tb, m, k, alpha, a, lda, b, ldb, beta, c, ldc); });               \
  }
X3_GEMMT(d, double, "DGEMMT ")
X3_GEMMT(s, float, "SGEMMT ")

#define X3_SYMM(P, T, NM)                                                                                          \
  OBA_API void cblas_##P##symm(int order, int side, int uplo, blasint m, blasint n, T alpha, const T* a,           \
                               blasint lda, const T* b, blasint ldb, T beta, T* c, blasint ldc) {                 \
    guarded([&] { cblas_symm_impl<T>(NM, order, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc); });       \
  }                                                                                                               \
  OBA_API void P##symm_(const char* side, const char* uplo, const blasint* m, const blasint* n, const T* alpha,    \
                        const T* a, const blasint* lda, const T* b, const blasint* ldb, const T* beta, T* c,       \
                        const blasint* ldc) {                                                                     \
    guarded([&] { f_symm_impl<T>(NM, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc); });                  \
  }
X3_SYMM(d, double, "DSYMM ")
X3_SYMM(s, float, "SSYMM ")

#define X3_TRMM(P, T, NM)                                                                                          \
  OBA_API void cblas_##P##trmm(int order, int side, int uplo, int trans, int diag, blasint m, blasint n, T alpha,   \
                               const T* a, blasint lda, T* b, blasint ldb) {                                       \
    guarded([&] { cblas_trmm_impl<T>(NM, order, side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb); });        \
  }                                                                                                               \
  OBA_API void P##trmm_(const char* side, const char* uplo, const char* ta, const char* diag, const blasint* m,    \
                        const blasint* n, const T* alpha, const T* a, const blasint* lda, T* b,                   \
                        const blasint* ldb) {                                                                     \
    guarded([&] { f_trmm_impl<T>(NM, side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb); });                      \
  }
X3_TRMM(d, double, "DTRMM ")
X3_TRMM(s, float, "STRMM ")
