/* CPU oracle for the level-3 hot path -- test infrastructure only (see
 * oracle.c).  Column-major, 0/1 flags: trans 1 = transposed; side 0 = left;
 * uplo 1 = lower; diag 1 = non-unit. */
#ifndef OBA_ORACLE_H
#define OBA_ORACLE_H
void or_dgemm(int transa, int transb, int m, int n, int k, double alpha, const double* A, int lda,
              const double* B, int ldb, double beta, double* C, int ldc);
void or_sgemm(int transa, int transb, int m, int n, int k, float alpha, const float* A, int lda,
              const float* B, int ldb, float beta, float* C, int ldc);
void or_dtrsm(int side, int uplo, int transa, int diag, int m, int n, double alpha, const double* A, int lda,
              double* B, int ldb);
void or_strsm(int side, int uplo, int transa, int diag, int m, int n, float alpha, const float* A, int lda,
              float* B, int ldb);
int or_dgetf2(int m, int n, double* A, int lda, int* ipiv);
int or_sgetf2(int m, int n, float* A, int lda, int* ipiv);
void or_dlaswp(int n, double* A, int lda, int k1, int k2, const int* ipiv, int incx);
void or_slaswp(int n, float* A, int lda, int k1, int k2, const int* ipiv, int incx);
void or_dgetrs(int trans, int n, int nrhs, const double* A, int lda, const int* ipiv, double* B, int ldb);
void or_sgetrs(int trans, int n, int nrhs, const float* A, int lda, const int* ipiv, float* B, int ldb);
int or_dpotf2(int uplo, int n, double* A, int lda);
int or_spotf2(int uplo, int n, float* A, int lda);
int or_dsgesv(int n, int nrhs, double* A, int lda, int* ipiv, const double* B, int ldb, double* X, int ldx,
              int* iter);
int or_dsposv(int lower, int n, int nrhs, double* A, int lda, const double* B, int ldb, double* X, int ldx,
              int* iter);
void or_dpotrs(int lower, int n, int nrhs, const double* F, int ldf, double* B, int ldb);
void or_spotrs(int lower, int n, int nrhs, const float* F, int ldf, float* B, int ldb);
double or_dmmch_ratio(int transa, int transb, int m, int n, int k, double alpha, const double* A, int lda,
                      const double* B, int ldb, double beta, const double* C_in, int ldc, const double* C_out,
                      int ldcc, double eps);
double or_dbeg(int reset);
#endif
