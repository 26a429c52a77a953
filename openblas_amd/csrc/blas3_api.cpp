// Level-3 BLAS C ABI: cblas_?gemm, ?gemm_, cblas_?trsm, ?trsm_,
// cblas_?gemm_batch.  Argument checking, info numbering, the row-major to
// column-major mapping and the quick returns follow the reference's
// interface layer exactly:
//   interface/gemm.c:181-246 (Fortran), :294-476 (CBLAS), :494 (m|n==0)
//   interface/trsm.c:114 (Fortran), :213-354 (CBLAS)
//   interface/gemm_batch.c:117-372 (group API)
// Caller memory may be host or device; host matrices are staged through the
// device pool and only outputs are copied back.  Calls are synchronous.
#include "kernels.h"
#include "batched.h"
#include <cctype>
#include <vector>

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

// Validated column-major GEMM on caller memory.
template <typename T>
void gemm_run(int transa, int transb, long m, long n, long k, T alpha, const T* a, long lda, const T* b,
              long ldb, T beta, T* c, long ldc) {
  if (m == 0 || n == 0) return;
  const bool only_scale = (k == 0 || alpha == T(0));
  if (only_scale && beta == T(1)) return;
  Runtime& r = rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  Mat<T> A, B, C;
  if (!only_scale) {
    A.in(const_cast<T*>(a), transa ? k : m, transa ? m : k, lda, true);
    B.in(const_cast<T*>(b), transb ? n : k, transb ? k : n, ldb, true);
  }
  C.in(c, m, n, ldc, beta != T(0));
  {
    KTimer kt(sizeof(T) == 8 ? "dgemm" : "sgemm", r.stream, only_scale ? 0.0 : 2.0 * m * n * k,
              (double)sizeof(T) * ((only_scale ? 0.0 : (double)m * k + (double)k * n) + (beta != T(0) ? 2.0 : 1.0) * m * n));
    gemm<T>(transa, transb, m, n, k, alpha, A.d, A.ld, B.d, B.ld, beta, C.d, C.ld, r.stream);
  }
  C.out();
  end_call();
}

template <typename T>
void cblas_gemm_impl(const char* name, int order, int TransA, int TransB, blasint m, blasint n, blasint k,
                     T alpha, const T* a, blasint lda, const T* b, blasint ldb, T beta, T* c, blasint ldc) {
  int transa = -1, transb = -1;
  long M = 0, N = 0, K = k, LDA = 0, LDB = 0;
  const T *pa = nullptr, *pb = nullptr;
  blasint info = 0;
  if (order == ColMajor) {
    M = m; N = n; pa = a; pb = b; LDA = lda; LDB = ldb;
    transa = cblas_trans(TransA); transb = cblas_trans(TransB);
  } else if (order == RowMajor) {
    M = n; N = m; pa = b; pb = a; LDA = ldb; LDB = lda;
    transa = cblas_trans(TransB); transb = cblas_trans(TransA);
  }
  if (order == ColMajor || order == RowMajor) {
    long nrowa = transa == 1 ? K : M;
    long nrowb = transb == 1 ? N : K;
    info = -1;
    if (ldc < M) info = 13;
    if (LDB < nrowb) info = 10;
    if (LDA < nrowa) info = 8;
    if (K < 0) info = 5;
    if (N < 0) info = 4;
    if (M < 0) info = 3;
    if (transb < 0) info = 2;
    if (transa < 0) info = 1;
  }
  if (info >= 0) { report_xerbla(name, info); return; }
  gemm_run<T>(transa, transb, M, N, K, alpha, pa, LDA, pb, LDB, beta, c, ldc);
}

template <typename T>
void f_gemm_impl(const char* name, const char* TA, const char* TB, const blasint* m, const blasint* n,
                 const blasint* k, const T* alpha, const T* a, const blasint* lda, const T* b,
                 const blasint* ldb, const T* beta, T* c, const blasint* ldc) {
  int transa = fortran_trans(*TA), transb = fortran_trans(*TB);
  long M = *m, N = *n, K = *k;
  long nrowa = transa == 1 ? K : M;
  long nrowb = transb == 1 ? N : K;
  blasint info = 0;
  if (*ldc < M) info = 13;
  if (*ldb < nrowb) info = 10;
  if (*lda < nrowa) info = 8;
  if (K < 0) info = 5;
  if (N < 0) info = 4;
  if (M < 0) info = 3;
  if (transb < 0) info = 2;
  if (transa < 0) info = 1;
  if (info) { report_xerbla(name, info); return; }
  gemm_run<T>(transa, transb, M, N, K, *alpha, a, *lda, b, *ldb, *beta, c, *ldc);
}

// ---- TRSM --------------------------------------------------------------------
template <typename T>
void trsm_run(int side, int uplo, int trans, int unit, long m, long n, T alpha, const T* a, long lda, T* b,
              long ldb) {
  if (m == 0 || n == 0) return;
  Runtime& r = rt();
  std::lock_guard<std::recursive_mutex> g(api_mutex());
  long na = side ? n : m;
  Mat<T> A, B;
  B.in(b, m, n, ldb, alpha != T(0));
  if (alpha == T(0)) {
    scale<T>(m, n, T(0), B.d, B.ld, r.stream);   // B <- 0 (trsm_L.c:103-107)
  } else {
    A.in(const_cast<T*>(a), na, na, lda, true);
    // side: 0 left / 1 right; uplo: 0 upper / 1 lower; unit: 0 unit / 1 non-unit
    KTimer kt(sizeof(T) == 8 ? "dtrsm" : "strsm", r.stream, (double)m * n * (double)na,
              (double)sizeof(T) * (0.5 * na * na + 2.0 * m * n));
    trsm<T>(side == 1, uplo == 1, trans == 1, unit == 0, m, n, alpha, A.d, A.ld, B.d, B.ld, r.stream);
  }
  B.out();
  end_call();
}

template <typename T>
void cblas_trsm_impl(const char* name, int order, int Side, int Uplo, int Trans, int Diag, blasint m, blasint n,
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
    long nrowa = side == 1 ? N : M;
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
  trsm_run<T>(side, uplo, trans, unit, M, N, alpha, a, lda, b, ldb);
}

template <typename T>
void f_trsm_impl(const char* name, const char* SIDE, const char* UPLO, const char* TRANSA, const char* DIAG,
                 const blasint* m, const blasint* n, const T* alpha, const T* a, const blasint* lda, T* b,
                 const blasint* ldb) {
  char s = (char)toupper((unsigned char)*SIDE), u = (char)toupper((unsigned char)*UPLO),
       d = (char)toupper((unsigned char)*DIAG);
  int side = s == 'L' ? 0 : s == 'R' ? 1 : -1;
  int uplo = u == 'U' ? 0 : u == 'L' ? 1 : -1;
  int trans = fortran_trans(*TRANSA);
  int unit = d == 'U' ? 0 : d == 'N' ? 1 : -1;
  long M = *m, N = *n;
  long nrowa = side == 1 ? N : M;
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
  trsm_run<T>(side, uplo, trans, unit, M, N, *alpha, a, *lda, b, *ldb);
}

// ---- GEMM batch (group API) ----------------------------------------------------
template <typename T>
void gemm_batch_impl(const char* name, int order, const int* ta_arr, const int* tb_arr, const blasint* m_arr,
                     const blasint* n_arr, const blasint* k_arr, const T* alpha_arr, const T** a_arr,
                     const blasint* lda_arr, const T** b_arr, const blasint* ldb_arr, const T* beta_arr,
                     T** c_arr, const blasint* ldc_arr, blasint group_count, const blasint* group_size) {
  std::vector<BDesc> probs;
  long idx = 0;
  for (blasint i = 0; i < group_count; idx += group_size[i], i++) {
    int ta = -1, tb = -1;
    long M = 0, N = 0, K = k_arr[i], LDA = 0, LDB = 0, LDC = ldc_arr[i];
    blasint info = 0;
    if (order == ColMajor) {
      M = m_arr[i]; N = n_arr[i]; LDA = lda_arr[i]; LDB = ldb_arr[i];
      ta = cblas_trans(ta_arr[i]); tb = cblas_trans(tb_arr[i]);
    } else if (order == RowMajor) {
      M = n_arr[i]; N = m_arr[i]; LDA = ldb_arr[i]; LDB = lda_arr[i];
      ta = cblas_trans(tb_arr[i]); tb = cblas_trans(ta_arr[i]);
    }
    if (order == ColMajor || order == RowMajor) {
      long nrowa = ta == 1 ? K : M, nrowb = tb == 1 ? N : K;
      info = -1;
      if (LDC < M) info = 13;
      if (LDB < nrowb) info = 10;
      if (LDA < nrowa) info = 8;
      if (K < 0) info = 5;
      if (N < 0) info = 4;
      if (M < 0) info = 3;
      if (tb < 0) info = 2;
      if (ta < 0) info = 1;
    }
    if (info >= 0) { report_xerbla(name, info); return; }
    if (M == 0 || N == 0) continue;
    for (blasint j = 0; j < group_size[i]; j++) {
      BDesc d{};
      d.a = (void*)(order == ColMajor ? a_arr[idx + j] : b_arr[idx + j]);
      d.b = (void*)(order == ColMajor ? b_arr[idx + j] : a_arr[idx + j]);
      d.c = (void*)c_arr[idx + j];
      d.m = (int)M; d.n = (int)N; d.k = (int)K;
      d.lda = (int)LDA; d.ldb = (int)LDB; d.ldc = (int)LDC;
      d.idx = (int)(idx + j);
      d.flags = ta | (tb << 1);
      d.alpha = (double)alpha_arr[i];
      d.beta = (double)beta_arr[i];
      probs.push_back(d);
    }
  }
  gemm_batch_run<T>(probs);
}

// ---- GEMM batch, strided (extension of the group API) ----------------------------
// One shape and one (alpha, beta) for batch_size problems; problem i takes
// A + i*stridea, B + i*strideb, C + i*stridec (strides in elements).  The
// single-GEMM checks of interface/gemm.c:423-471 keep their CBLAS numbers
// (parameters counted after Order: TransA 1 ... lda 8, ldb 11, ldc 15);
// negative strides (9, 12, 16), C matrices that would overlap (16) and a
// negative batch_size (17) are errors too.  A or B strides of 0 share one
// operand across the batch.  Row-major swaps A and B before the checks, and
// the numbers are those of the swapped call, as in the group API
// (interface/gemm_batch.c:248-290).
template <typename T>
void gemm_batch_strided_impl(const char* name, int order, int TransA, int TransB, blasint m, blasint n, blasint k,
                             T alpha, const T* a, blasint lda, long stridea, const T* b, blasint ldb, long strideb,
                             T beta, T* c, blasint ldc, long stridec, blasint batch) {
  int transa = -1, transb = -1;
  long M = 0, N = 0, K = k, LDA = 0, LDB = 0, SA = 0, SB = 0;
  const T *pa = nullptr, *pb = nullptr;
  blasint info = 0;
  if (order == ColMajor) {
    M = m; N = n; pa = a; pb = b; LDA = lda; LDB = ldb; SA = stridea; SB = strideb;
    transa = cblas_trans(TransA); transb = cblas_trans(TransB);
  } else if (order == RowMajor) {
    M = n; N = m; pa = b; pb = a; LDA = ldb; LDB = lda; SA = strideb; SB = stridea;
    transa = cblas_trans(TransB); transb = cblas_trans(TransA);
  }
  if (order == ColMajor || order == RowMajor) {
    long nrowa = transa == 1 ? K : M;
    long nrowb = transb == 1 ? N : K;
    info = -1;
    if (batch < 0) info = 17;
    if (stridec < 0 || (batch > 1 && M > 0 && N > 0 && stridec < (long)ldc * N)) info = 16;
    if (ldc < M) info = 15;
    if (strideb < 0) info = 12;
    if (LDB < nrowb) info = 11;
    if (stridea < 0) info = 9;
    if (LDA < nrowa) info = 8;
    if (K < 0) info = 5;
    if (N < 0) info = 4;
    if (M < 0) info = 3;
    if (transb < 0) info = 2;
    if (transa < 0) info = 1;
  }
  if (info >= 0) { report_xerbla(name, info); return; }
  if (M == 0 || N == 0 || batch == 0) return;
  std::vector<BDesc> probs((size_t)batch);
  for (blasint i = 0; i < batch; i++) {
    BDesc& d = probs[i];
    d = BDesc{};
    d.a = (void*)(pa + i * SA);
    d.b = (void*)(pb + i * SB);
    d.c = (void*)(c + i * stridec);
    d.m = (int)M; d.n = (int)N; d.k = (int)K;
    d.lda = (int)LDA; d.ldb = (int)LDB; d.ldc = (int)ldc;
    d.idx = (int)i;
    d.flags = transa | (transb << 1);
    d.alpha = (double)alpha;
    d.beta = (double)beta;
  }
  gemm_batch_run<T>(probs);
}

}  // namespace

// ---- exported symbols ----------------------------------------------------------
OBA_API void cblas_dgemm(int order, int ta, int tb, blasint m, blasint n, blasint k, double alpha, const double* a,
                         blasint lda, const double* b, blasint ldb, double beta, double* c, blasint ldc) {
  guarded([&] { cblas_gemm_impl<double>("DGEMM ", order, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc); });
}
OBA_API void cblas_sgemm(int order, int ta, int tb, blasint m, blasint n, blasint k, float alpha, const float* a,
                         blasint lda, const float* b, blasint ldb, float beta, float* c, blasint ldc) {
  guarded([&] { cblas_gemm_impl<float>("SGEMM ", order, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc); });
}
OBA_API void dgemm_(const char* ta, const char* tb, const blasint* m, const blasint* n, const blasint* k,
                    const double* alpha, const double* a, const blasint* lda, const double* b, const blasint* ldb,
                    const double* beta, double* c, const blasint* ldc) {
  guarded([&] { f_gemm_impl<double>("DGEMM ", ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc); });
}
OBA_API void sgemm_(const char* ta, const char* tb, const blasint* m, const blasint* n, const blasint* k,
                    const float* alpha, const float* a, const blasint* lda, const float* b, const blasint* ldb,
                    const float* beta, float* c, const blasint* ldc) {
  guarded([&] { f_gemm_impl<float>("SGEMM ", ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc); });
}
OBA_API void cblas_dtrsm(int order, int side, int uplo, int trans, int diag, blasint m, blasint n, double alpha,
                         const double* a, blasint lda, double* b, blasint ldb) {
  guarded([&] { cblas_trsm_impl<double>("DTRSM ", order, side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb); });
}
OBA_API void cblas_strsm(int order, int side, int uplo, int trans, int diag, blasint m, blasint n, float alpha,
                         const float* a, blasint lda, float* b, blasint ldb) {
  guarded([&] { cblas_trsm_impl<float>("STRSM ", order, side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb); });
}
OBA_API void dtrsm_(const char* side, const char* uplo, const char* ta, const char* diag, const blasint* m,
                    const blasint* n, const double* alpha, const double* a, const blasint* lda, double* b,
                    const blasint* ldb) {
  guarded([&] { f_trsm_impl<double>("DTRSM ", side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb); });
}
OBA_API void strsm_(const char* side, const char* uplo, const char* ta, const char* diag, const blasint* m,
                    const blasint* n, const float* alpha, const float* a, const blasint* lda, float* b,
                    const blasint* ldb) {
  guarded([&] { f_trsm_impl<float>("STRSM ", side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb); });
}
OBA_API void cblas_dgemm_batch(int order, const int* ta, const int* tb, const blasint* m, const blasint* n,
                               const blasint* k, const double* alpha, const double** a, const blasint* lda,
                               const double** b, const blasint* ldb, const double* beta, double** c,
                               const blasint* ldc, blasint group_count, const blasint* group_size) {
  guarded([&] { gemm_batch_impl<double>("DGEMM_BATCH ", order, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, group_count,
                          group_size); });
}
OBA_API void cblas_dgemm_batch_strided(int order, int ta, int tb, blasint m, blasint n, blasint k, double alpha,
                                       const double* a, blasint lda, long stridea, const double* b, blasint ldb,
                                       long strideb, double beta, double* c, blasint ldc, long stridec,
                                       blasint batch_size) {
  guarded([&] {
    gemm_batch_strided_impl<double>("DGEMM_STRIDED ", order, ta, tb, m, n, k, alpha, a, lda, stridea, b, ldb,
                                    strideb, beta, c, ldc, stridec, batch_size);
  });
}
OBA_API void cblas_sgemm_batch_strided(int order, int ta, int tb, blasint m, blasint n, blasint k, float alpha,
                                       const float* a, blasint lda, long stridea, const float* b, blasint ldb,
                                       long strideb, float beta, float* c, blasint ldc, long stridec,
                                       blasint batch_size) {
  guarded([&] {
    gemm_batch_strided_impl<float>("SGEMM_STRIDED ", order, ta, tb, m, n, k, alpha, a, lda, stridea, b, ldb,
                                   strideb, beta, c, ldc, stridec, batch_size);
  });
}
OBA_API void cblas_sgemm_batch(int order, const int* ta, const int* tb, const blasint* m, const blasint* n,
                               const blasint* k, const float* alpha, const float** a, const blasint* lda,
                               const float** b, const blasint* ldb, const float* beta, float** c,
                               const blasint* ldc, blasint group_count, const blasint* group_size) {
  guarded([&] { gemm_batch_impl<float>("SGEMM_BATCH ", order, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, group_count,
                         group_size); });
}
