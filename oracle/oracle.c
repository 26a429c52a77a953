/*
 * CPU oracle for the level-3 hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's algorithms (OpenBLAS 0.3.28.dev,
 * paths relative to its source tree), used by tests/, by the smoke check and
 * by bench.py's CPU-baseline leg as the checker; nothing in the product
 * (openblas_amd/) links, loads or calls it.
 *
 * Precision choices: GEMM and TRSM accumulate in long double (x87 80-bit) so
 * that the oracle is more accurate than any FP64 implementation it checks;
 * the factorizations (getf2, potf2) run in the working precision, literally
 * in the reference's loop order, so that their pivot choices and info codes
 * are the reference's.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/* ---- GEMM: driver/level3/level3.c:229-264 (beta pre-pass, then early exit
 * when k == 0 or alpha == 0), kernel/generic/gemm_beta.c:41-57 (beta == 0
 * writes exact zeros), C += alpha * op(A) op(B). */
#define GEMM_BODY(T)                                                              \
  for (int j = 0; j < n; j++)                                                     \
    for (int i = 0; i < m; i++) {                                                 \
      T* c = &C[IDX(i, j, ldc)];                                                  \
      if (beta != (T)1) *c = (beta == (T)0) ? (T)0 : beta * *c;                   \
    }                                                                             \
  if (k == 0 || alpha == (T)0) return;                                            \
  for (int j = 0; j < n; j++)                                                     \
    for (int i = 0; i < m; i++) {                                                 \
      long double s = 0.0L;                                                       \
      for (int l = 0; l < k; l++) {                                               \
        long double a = transa ? A[IDX(l, i, lda)] : A[IDX(i, l, lda)];           \
        long double b = transb ? B[IDX(j, l, ldb)] : B[IDX(l, j, ldb)];           \
        s += a * b;                                                               \
      }                                                                           \
      C[IDX(i, j, ldc)] = (T)((long double)C[IDX(i, j, ldc)] + (long double)alpha * s); \
    }

void or_dgemm(int transa, int transb, int m, int n, int k, double alpha, const double* A, int lda,
              const double* B, int ldb, double beta, double* C, int ldc) {
  GEMM_BODY(double)
}
void or_sgemm(int transa, int transb, int m, int n, int k, float alpha, const float* A, int lda,
              const float* B, int ldb, float beta, float* C, int ldc) {
  GEMM_BODY(float)
}

/* ---- TRSM: driver/level3/trsm_L.c:103-107 (B <- alpha*B, alpha == 0 zeroes
 * B and returns), kernel/generic/trsm_ltcopy_2.c:43 (the diagonal is stored
 * as its reciprocal) and trsm_kernel_LT.c:95 (substitution multiplies by it).
 * side 0 = left (op(A) X = B), 1 = right (X op(A) = B); uplo 0 = upper,
 * 1 = lower; diag 0 = unit, 1 = non-unit (the reference's table order). */
#define TRSM_BODY(T)                                                              \
  const int na = side ? n : m;                                                    \
  for (int j = 0; j < n; j++)                                                     \
    for (int i = 0; i < m; i++) {                                                 \
      T* b = &B[IDX(i, j, ldb)];                                                  \
      if (alpha != (T)1) *b = (alpha == (T)0) ? (T)0 : alpha * *b;                \
    }                                                                             \
  if (alpha == (T)0) return;                                                      \
  /* effective matrix E = op(A) (left) or op(A)^T (right): E x = b */             \
  const int eT = (transa != 0) != (side != 0);                                    \
  const int elower = (uplo == 1) != eT;                                           \
  const int nvec = side ? m : n;                                                  \
  long double* x = (long double*)malloc(sizeof(long double) * (na > 0 ? na : 1)); \
  for (int v = 0; v < nvec; v++) {                                                \
    for (int i = 0; i < na; i++) x[i] = side ? B[IDX(v, i, ldb)] : B[IDX(i, v, ldb)]; \
    for (int step = 0; step < na; step++) {                                       \
      const int i = elower ? step : na - 1 - step;                                \
      long double s = x[i];                                                       \
      if (elower) {                                                               \
        for (int l = 0; l < i; l++)                                               \
          s -= (long double)(eT ? A[IDX(l, i, lda)] : A[IDX(i, l, lda)]) * x[l];  \
      } else {                                                                    \
        for (int l = i + 1; l < na; l++)                                          \
          s -= (long double)(eT ? A[IDX(l, i, lda)] : A[IDX(i, l, lda)]) * x[l];  \
      }                                                                           \
      if (diag) s *= (long double)((T)1 / A[IDX(i, i, lda)]);                    \
      x[i] = s;                                                                   \
    }                                                                             \
    for (int i = 0; i < na; i++) {                                                \
      if (side) B[IDX(v, i, ldb)] = (T)x[i];                                      \
      else B[IDX(i, v, ldb)] = (T)x[i];                                           \
    }                                                                             \
  }                                                                               \
  free(x);

void or_dtrsm(int side, int uplo, int transa, int diag, int m, int n, double alpha, const double* A, int lda,
              double* B, int ldb) {
  TRSM_BODY(double)
}
void or_strsm(int side, int uplo, int transa, int diag, int m, int n, float alpha, const float* A, int lda,
              float* B, int ldb) {
  TRSM_BODY(float)
}

/* ---- IAMAX: kernel/arm/iamax.c:51-74 -- first index of max |x| (1-based).
 * A NaN never compares greater, so it is only returned when it is first. */
#define IAMAX_BODY(T)                                                             \
  if (n <= 0) return 0;                                                           \
  int best = 0;                                                                   \
  T maxf = (T)fabs((double)x[0]);                                                 \
  for (int i = 1; i < n; i++) {                                                   \
    T v = (T)fabs((double)x[i]);                                                  \
    if (v > maxf) { maxf = v; best = i; }                                         \
  }                                                                               \
  return best + 1;

static int iamax_d(int n, const double* x) { IAMAX_BODY(double) }
static int iamax_s(int n, const float* x) { IAMAX_BODY(float) }

/* ---- GETF2: lapack/getf2/getf2_k.c:46-125, literally: left-looking column
 * j: apply earlier interchanges to column j (:80-88), solve the U part with
 * dot products (:90-92), GEMV update of the rest (:95), IAMAX pivot (:97),
 * interchange and scale by 1/pivot only when |pivot| >= DBL_MIN / FLT_MIN
 * (:103-116), an exact zero pivot sets info = j+1 once (:119). */
#define GETF2_BODY(T, TMIN, IAMAX)                                                \
  int info = 0;                                                                   \
  for (int j = 0; j < n; j++) {                                                   \
    T* b = &A[IDX(0, j, lda)];                                                    \
    const int len = j < m ? j : m;                                                \
    for (int i = 0; i < len; i++) {                                               \
      int ip = ipiv[i] - 1;                                                       \
      if (ip != i) { T t1 = b[i]; b[i] = b[ip]; b[ip] = t1; }                     \
    }                                                                             \
    for (int i = 1; i < len; i++) {                                               \
      T s = 0;                                                                    \
      for (int l = 0; l < i; l++) s += A[IDX(i, l, lda)] * b[l];                  \
      b[i] -= s;                                                                  \
    }                                                                             \
    if (j < m) {                                                                  \
      for (int i = j; i < m; i++) {                                               \
        T s = 0;                                                                  \
        for (int l = 0; l < j; l++) s += A[IDX(i, l, lda)] * b[l];                \
        b[i] -= s;                                                                \
      }                                                                           \
      int jp = j + IAMAX(m - j, b + j);                                           \
      if (jp > m) jp = m;                                                         \
      ipiv[j] = jp;                                                               \
      jp--;                                                                       \
      T t1 = b[jp];                                                               \
      if (t1 != (T)0) {                                                           \
        if (fabs((double)t1) >= TMIN) {                                           \
          t1 = (T)1 / t1;                                                         \
          if (jp != j)                                                            \
            for (int l = 0; l <= j; l++) {                                        \
              T t2 = A[IDX(j, l, lda)];                                           \
              A[IDX(j, l, lda)] = A[IDX(jp, l, lda)];                             \
              A[IDX(jp, l, lda)] = t2;                                            \
            }                                                                     \
          for (int i = j + 1; i < m; i++) b[i] *= t1;                             \
        }                                                                         \
      } else if (!info) {                                                         \
        info = j + 1;                                                             \
      }                                                                           \
    }                                                                             \
  }                                                                               \
  return info;

int or_dgetf2(int m, int n, double* A, int lda, int* ipiv) { GETF2_BODY(double, DBL_MIN, iamax_d) }
int or_sgetf2(int m, int n, float* A, int lda, int* ipiv) { GETF2_BODY(float, FLT_MIN, iamax_s) }

/* ---- LASWP: lapack/laswp/generic/laswp_k_1.c:48 and the LAPACK contract
 * of interface/lapack/laswp.c:55 -- row i in k1..k2 is exchanged with row
 * IPIV(k1 + (i-k1)*|incx|); incx < 0 walks k2 down to k1. */
#define LASWP_BODY(T)                                                             \
  if (incx == 0 || n <= 0) return;                                                \
  const int inc = incx > 0 ? incx : -incx;                                        \
  for (int s = 0; s <= k2 - k1; s++) {                                            \
    const int i = incx > 0 ? k1 + s : k2 - s;                                     \
    const int ip = ipiv[(k1 - 1) + (i - k1) * inc] - 1;                           \
    if (ip != i - 1)                                                              \
      for (int c = 0; c < n; c++) {                                               \
        T t1 = A[IDX(i - 1, c, lda)];                                             \
        A[IDX(i - 1, c, lda)] = A[IDX(ip, c, lda)];                               \
        A[IDX(ip, c, lda)] = t1;                                                  \
      }                                                                           \
  }

void or_dlaswp(int n, double* A, int lda, int k1, int k2, const int* ipiv, int incx) { LASWP_BODY(double) }
void or_slaswp(int n, float* A, int lda, int k1, int k2, const int* ipiv, int incx) { LASWP_BODY(float) }

/* ---- GETRS: lapack/getrs/getrs_single.c:42-68 -- N: LASWP_PLUS on B, then
 * unit-lower and non-unit-upper solves; T: the transposed solves, then
 * LASWP_MINUS. */
void or_dgetrs(int trans, int n, int nrhs, const double* A, int lda, const int* ipiv, double* B, int ldb) {
  if (n <= 0 || nrhs <= 0) return;
  if (!trans) {
    or_dlaswp(nrhs, B, ldb, 1, n, ipiv, 1);
    or_dtrsm(0, 1, 0, 0, n, nrhs, 1.0, A, lda, B, ldb);
    or_dtrsm(0, 0, 0, 1, n, nrhs, 1.0, A, lda, B, ldb);
  } else {
    or_dtrsm(0, 0, 1, 1, n, nrhs, 1.0, A, lda, B, ldb);
    or_dtrsm(0, 1, 1, 0, n, nrhs, 1.0, A, lda, B, ldb);
    or_dlaswp(nrhs, B, ldb, 1, n, ipiv, -1);
  }
}
void or_sgetrs(int trans, int n, int nrhs, const float* A, int lda, const int* ipiv, float* B, int ldb) {
  if (n <= 0 || nrhs <= 0) return;
  if (!trans) {
    or_slaswp(nrhs, B, ldb, 1, n, ipiv, 1);
    or_strsm(0, 1, 0, 0, n, nrhs, 1.0f, A, lda, B, ldb);
    or_strsm(0, 0, 0, 1, n, nrhs, 1.0f, A, lda, B, ldb);
  } else {
    or_strsm(0, 0, 1, 1, n, nrhs, 1.0f, A, lda, B, ldb);
    or_strsm(0, 1, 1, 0, n, nrhs, 1.0f, A, lda, B, ldb);
    or_slaswp(nrhs, B, ldb, 1, n, ipiv, -1);
  }
}

/* ---- POTF2: lapack/potf2/potf2_L.c:50-97 and potf2_U.c -- ajj = a_jj -
 * dot(row/column j); ajj <= 0 stores ajj and returns j+1 (NaN passes the
 * test); a_jj = sqrt(ajj); GEMV update and scale by 1/ajj.  uplo 1 = lower,
 * 0 = upper; only that triangle is read or written. */
#define POTF2_BODY(T, SQRT)                                                       \
  for (int j = 0; j < n; j++) {                                                   \
    T dot = 0;                                                                    \
    for (int l = 0; l < j; l++) {                                                 \
      T v = uplo ? A[IDX(j, l, lda)] : A[IDX(l, j, lda)];                         \
      dot += v * v;                                                               \
    }                                                                             \
    T ajj = A[IDX(j, j, lda)] - dot;                                              \
    if (ajj <= 0) { A[IDX(j, j, lda)] = ajj; return j + 1; }                      \
    ajj = SQRT(ajj);                                                              \
    A[IDX(j, j, lda)] = ajj;                                                      \
    for (int i = j + 1; i < n; i++) {                                             \
      T s = 0;                                                                    \
      for (int l = 0; l < j; l++)                                                 \
        s += uplo ? A[IDX(i, l, lda)] * A[IDX(j, l, lda)] : A[IDX(l, i, lda)] * A[IDX(l, j, lda)]; \
      if (uplo) A[IDX(i, j, lda)] = (A[IDX(i, j, lda)] - s) * ((T)1 / ajj);       \
      else A[IDX(j, i, lda)] = (A[IDX(j, i, lda)] - s) * ((T)1 / ajj);            \
    }                                                                             \
  }                                                                               \
  return 0;

int or_dpotf2(int uplo, int n, double* A, int lda) { POTF2_BODY(double, sqrt) }
int or_spotf2(int uplo, int n, float* A, int lda) { POTF2_BODY(float, sqrtf) }

/* ---- DSGESV: lapack-netlib/SRC/dsgesv.f:193-430 restated: ITERMAX = 30,
 * BWDMAX = 1, CTE = ||A||_inf * eps * sqrt(N) (eps = DLAMCH('E') = 2^-53),
 * DLAG2S overflow -> ITER = -2, SGETRF info -> -3, no convergence -> -31,
 * and the FP64 DGETRF/DGETRS fallback on the original A.  The FP32 LU is
 * or_sgetf2 (the reference's SGETRF pivots), the residual is computed with
 * double accumulation like DGEMM. */
static int dlag2s(int m, int n, const double* a, int lda, float* s, int lds) {
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      double v = a[IDX(i, j, lda)];
      if (v < -(double)FLT_MAX || v > (double)FLT_MAX) return 1;
      s[IDX(i, j, lds)] = (float)v;
    }
  return 0;
}

static void residual(int n, int nrhs, const double* A, int lda, const double* X, int ldx, const double* B, int ldb,
                     double* R) {
  for (int c = 0; c < nrhs; c++)
    for (int i = 0; i < n; i++) {
      double s = 0;
      for (int l = 0; l < n; l++) s += A[IDX(i, l, lda)] * X[IDX(l, c, ldx)];
      R[IDX(i, c, n)] = B[IDX(i, c, ldb)] - s;
    }
}

static int converged(int n, int nrhs, const double* X, int ldx, const double* R, double cte) {
  for (int c = 0; c < nrhs; c++) {
    double xn = fabs(X[IDX(iamax_d(n, X + IDX(0, c, ldx)) - 1, c, ldx)]);
    double rn = fabs(R[IDX(iamax_d(n, R + IDX(0, c, n)) - 1, c, n)]);
    if (rn > xn * cte) return 0;
  }
  return 1;
}

int or_dsgesv(int n, int nrhs, double* A, int lda, int* ipiv, const double* B, int ldb, double* X, int ldx,
              int* iter) {
  *iter = 0;
  if (n <= 0) return 0;
  const int ITERMAX = 30;
  double anrm = 0;
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int j = 0; j < n; j++) s += fabs(A[IDX(i, j, lda)]);
    if (s > anrm || isnan(s)) anrm = s;
  }
  const double eps = ldexp(1.0, -53);
  const double cte = anrm * eps * sqrt((double)n) * 1.0;
  float* SA = (float*)malloc(sizeof(float) * (size_t)n * n);
  float* SX = (float*)malloc(sizeof(float) * (size_t)n * (nrhs > 0 ? nrhs : 1));
  double* R = (double*)malloc(sizeof(double) * (size_t)n * (nrhs > 0 ? nrhs : 1));
  int it = 0, fb = 0;
  if (dlag2s(n, nrhs, B, ldb, SX, n)) { it = -2; fb = 1; }
  if (!fb && dlag2s(n, n, A, lda, SA, n)) { it = -2; fb = 1; }
  if (!fb && or_sgetf2(n, n, SA, n, ipiv) != 0) { it = -3; fb = 1; }
  if (!fb) {
    or_sgetrs(0, n, nrhs, SA, n, ipiv, SX, n);
    for (int c = 0; c < nrhs; c++)
      for (int i = 0; i < n; i++) X[IDX(i, c, ldx)] = SX[IDX(i, c, n)];
    residual(n, nrhs, A, lda, X, ldx, B, ldb, R);
    if (converged(n, nrhs, X, ldx, R, cte)) { *iter = 0; goto done; }
    for (int k = 1; k <= ITERMAX; k++) {
      if (dlag2s(n, nrhs, R, n, SX, n)) { it = -2; fb = 1; break; }
      or_sgetrs(0, n, nrhs, SA, n, ipiv, SX, n);
      for (int c = 0; c < nrhs; c++)
        for (int i = 0; i < n; i++) X[IDX(i, c, ldx)] += (double)SX[IDX(i, c, n)];
      residual(n, nrhs, A, lda, X, ldx, B, ldb, R);
      if (converged(n, nrhs, X, ldx, R, cte)) { *iter = k; goto done; }
    }
    if (!fb) { it = -ITERMAX - 1; fb = 1; }
  }
  {
    *iter = it;
    int info = or_dgetf2(n, n, A, lda, ipiv);
    if (info) { free(SA); free(SX); free(R); return info; }
    for (int c = 0; c < nrhs; c++)
      for (int i = 0; i < n; i++) X[IDX(i, c, ldx)] = B[IDX(i, c, ldb)];
    or_dgetrs(0, n, nrhs, A, lda, ipiv, X, ldx);
  }
done:
  free(SA); free(SX); free(R);
  return 0;
}

/* ---- DSPOSV: lapack-netlib/SRC/dsposv.f:197-440 restated: the dsgesv loop
 * above with the Cholesky pieces: ANRM = DLANSY('I') of the symmetric matrix
 * held in the uplo triangle, DLAT2S (demote the triangle, overflow -> ITER =
 * -2), SPOTRF (here the unblocked or_spotf2; info -> ITER = -3), SPOTRS, the
 * residual with DSYMM (double accumulation over the symmetric matrix), and
 * the FP64 DPOTRF/DPOTRS fallback on the original A (uplo 1 lower / 0 upper). */
#define SYM(A, i, j, lda, lower) ((((lower) && (i) >= (j)) || (!(lower) && (i) <= (j))) ? (A)[IDX(i, j, lda)] : (A)[IDX(j, i, lda)])

static int dlat2s(int lower, int n, const double* a, int lda, float* s, int lds) {
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      if (lower ? i < j : i > j) { s[IDX(i, j, lds)] = 0.f; continue; }
      double v = a[IDX(i, j, lda)];
      if (v < -(double)FLT_MAX || v > (double)FLT_MAX) return 1;
      s[IDX(i, j, lds)] = (float)v;
    }
  return 0;
}

/* potrs: L L^T (lower) or U^T U (upper) in the stored triangle */
#define POTRS_BODY(T)                                                              \
  for (int c = 0; c < nrhs; c++) {                                                 \
    T* x = B + IDX(0, c, ldb);                                                     \
    for (int i = 0; i < n; i++) { /* first triangle: forward */                     \
      T v = x[i];                                                                  \
      for (int k = 0; k < i; k++) v -= (lower ? F[IDX(i, k, ldf)] : F[IDX(k, i, ldf)]) * x[k]; \
      x[i] = v / F[IDX(i, i, ldf)];                                                \
    }                                                                              \
    for (int i = n - 1; i >= 0; i--) { /* second triangle: backward */              \
      T v = x[i];                                                                  \
      for (int k = i + 1; k < n; k++) v -= (lower ? F[IDX(k, i, ldf)] : F[IDX(i, k, ldf)]) * x[k]; \
      x[i] = v / F[IDX(i, i, ldf)];                                                \
    }                                                                              \
  }
void or_dpotrs(int lower, int n, int nrhs, const double* F, int ldf, double* B, int ldb) { POTRS_BODY(double) }
void or_spotrs(int lower, int n, int nrhs, const float* F, int ldf, float* B, int ldb) { POTRS_BODY(float) }

static void sym_residual(int lower, int n, int nrhs, const double* A, int lda, const double* X, int ldx,
                         const double* B, int ldb, double* R) {
  for (int c = 0; c < nrhs; c++)
    for (int i = 0; i < n; i++) {
      double s = 0;
      for (int l = 0; l < n; l++) s += SYM(A, i, l, lda, lower) * X[IDX(l, c, ldx)];
      R[IDX(i, c, n)] = B[IDX(i, c, ldb)] - s;
    }
}

int or_dsposv(int lower, int n, int nrhs, double* A, int lda, const double* B, int ldb, double* X, int ldx,
              int* iter) {
  *iter = 0;
  if (n <= 0) return 0;
  const int ITERMAX = 30;
  double anrm = 0;
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int j = 0; j < n; j++) s += fabs(SYM(A, i, j, lda, lower));
    if (s > anrm || isnan(s)) anrm = s;
  }
  const double eps = ldexp(1.0, -53);
  const double cte = anrm * eps * sqrt((double)n) * 1.0;
  float* SA = (float*)malloc(sizeof(float) * (size_t)n * n);
  float* SX = (float*)malloc(sizeof(float) * (size_t)n * (nrhs > 0 ? nrhs : 1));
  double* R = (double*)malloc(sizeof(double) * (size_t)n * (nrhs > 0 ? nrhs : 1));
  int it = 0, fb = 0;
  if (dlag2s(n, nrhs, B, ldb, SX, n)) { it = -2; fb = 1; }
  if (!fb && dlat2s(lower, n, A, lda, SA, n)) { it = -2; fb = 1; }
  if (!fb && or_spotf2(lower, n, SA, n) != 0) { it = -3; fb = 1; }
  if (!fb) {
    or_spotrs(lower, n, nrhs, SA, n, SX, n);
    for (int c = 0; c < nrhs; c++)
      for (int i = 0; i < n; i++) X[IDX(i, c, ldx)] = SX[IDX(i, c, n)];
    sym_residual(lower, n, nrhs, A, lda, X, ldx, B, ldb, R);
    if (converged(n, nrhs, X, ldx, R, cte)) { *iter = 0; goto done; }
    for (int k = 1; k <= ITERMAX; k++) {
      if (dlag2s(n, nrhs, R, n, SX, n)) { it = -2; fb = 1; break; }
      or_spotrs(lower, n, nrhs, SA, n, SX, n);
      for (int c = 0; c < nrhs; c++)
        for (int i = 0; i < n; i++) X[IDX(i, c, ldx)] += (double)SX[IDX(i, c, n)];
      sym_residual(lower, n, nrhs, A, lda, X, ldx, B, ldb, R);
      if (converged(n, nrhs, X, ldx, R, cte)) { *iter = k; goto done; }
    }
    if (!fb) { it = -ITERMAX - 1; fb = 1; }
  }
  {
    *iter = it;
    int info = or_dpotf2(lower, n, A, lda);
    if (info) { free(SA); free(SX); free(R); return info; }
    for (int c = 0; c < nrhs; c++)
      for (int i = 0; i < n; i++) X[IDX(i, c, ldx)] = B[IDX(i, c, ldb)];
    or_dpotrs(lower, n, nrhs, A, lda, X, ldx);
  }
done:
  free(SA); free(SX); free(R);
  return 0;
}

/* ---- ctest checker: ctest/c_dblat3c.c:2956-3127 (dmmch).  Returns the
 * largest error ratio |ct - cc| / eps / g with the gauge
 * g = |alpha| sum |a||b| + |beta| |c| (ratio not scaled where g == 0). */
double or_dmmch_ratio(int transa, int transb, int m, int n, int k, double alpha, const double* A, int lda,
                      const double* B, int ldb, double beta, const double* C_in, int ldc, const double* C_out,
                      int ldcc, double eps) {
  double err = 0;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      double ct = 0, g = 0;
      for (int l = 0; l < k; l++) {
        double a = transa ? A[IDX(l, i, lda)] : A[IDX(i, l, lda)];
        double b = transb ? B[IDX(j, l, ldb)] : B[IDX(l, j, ldb)];
        ct += a * b;
        g += fabs(a) * fabs(b);
      }
      ct = alpha * ct + beta * C_in[IDX(i, j, ldc)];
      g = fabs(alpha) * g + fabs(beta) * fabs(C_in[IDX(i, j, ldc)]);
      double e = fabs(ct - C_out[IDX(i, j, ldcc)]) / eps;
      if (g != 0) e /= g;
      if (e > err) err = e;
    }
  return err;
}

/* ---- ctest data generator: ctest/c_dblat3c.c:3264 (dbeg): I <- I*891 mod
 * 1000 from I = 7, skipping one value in six, value (I - 500)/1001. */
static int g_beg_i = 7, g_beg_ic = 0;
double or_dbeg(int reset) {
  if (reset) { g_beg_i = 7; g_beg_ic = 0; }
  ++g_beg_ic;
  for (;;) {
    g_beg_i *= 891;
    g_beg_i -= g_beg_i / 1000 * 1000;
    if (g_beg_ic >= 5) { g_beg_ic = 0; continue; }
    break;
  }
  return (g_beg_i - 500) / 1001.0;
}
