/*
 * libopenblas_amd -- MI355X (gfx950) implementation of OpenBLAS's level-3
 * hot path behind OpenBLAS's own C ABI.
 *
 * Every function below except the `openblas_amd_*` extensions has exactly the
 * name, argument list, argument meaning, info numbering and error behaviour
 * of the OpenBLAS entry point it replaces; the comment above each names that
 * reference interface (paths relative to the OpenBLAS source tree).
 *
 * Matrices are column-major (unless a CBLAS/LAPACKE layout argument says
 * otherwise) and may live in host memory or in device memory allocated with
 * hipMalloc/openblas_amd_malloc: host arguments are staged to the GPU and
 * back, device arguments are used in place.  Every call is synchronous.
 * blasint is a 32-bit int, or 64-bit with libopenblas_amd_ilp64.so and
 * OPENBLAS_AMD_ILP64 defined.  There is no CPU fallback.
 *
 * Device errors (no usable gfx950 device, a failed HIP call, a bounded
 * in-kernel wait that ran out) never abort the caller: the call is skipped,
 * LAPACK routines set INFO = OPENBLAS_AMD_DEVICE_ERROR, LAPACKE routines and
 * the batched extensions return it, and openblas_amd_last_error() returns
 * the message.  BLAS routines (void) leave their outputs unspecified then.
 */
#ifndef OPENBLAS_AMD_H
#define OPENBLAS_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* common.h:269-278 (INTERFACE64): blasint is 32-bit unless the ILP64 library
   (libopenblas_amd_ilp64.so) is used, which needs OPENBLAS_AMD_ILP64 defined
   here too; lapack_int follows (lapacke.h with LAPACK_ILP64).  The
   extensions without a reference counterpart take 32-bit int in both. */
#ifdef OPENBLAS_AMD_ILP64
typedef long blasint;
#else
typedef int blasint;
#endif
typedef blasint lapack_int;

/* enum values are ABI facts of cblas.h:62-66 */
typedef enum CBLAS_ORDER     {CblasRowMajor = 101, CblasColMajor = 102} CBLAS_ORDER;
typedef enum CBLAS_TRANSPOSE {CblasNoTrans = 111, CblasTrans = 112, CblasConjTrans = 113,
                              CblasConjNoTrans = 114} CBLAS_TRANSPOSE;
typedef enum CBLAS_UPLO      {CblasUpper = 121, CblasLower = 122} CBLAS_UPLO;
typedef enum CBLAS_DIAG      {CblasNonUnit = 131, CblasUnit = 132} CBLAS_DIAG;
typedef enum CBLAS_SIDE      {CblasLeft = 141, CblasRight = 142} CBLAS_SIDE;

#define LAPACK_ROW_MAJOR 101
#define LAPACK_COL_MAJOR 102

/* INFO / return value of a call that failed on the device (not an argument
   position: no routine here has 1000 arguments) */
#define OPENBLAS_AMD_DEVICE_ERROR (-1000)

/* ---- CBLAS level 3 -------------------------------------------------------- */
/* cblas.h:298-300, interface/gemm.c:294 (CNAME built with -DCBLAS) */
void cblas_dgemm(enum CBLAS_ORDER Order, enum CBLAS_TRANSPOSE TransA, enum CBLAS_TRANSPOSE TransB,
                 blasint M, blasint N, blasint K, double alpha, const double *A, blasint lda,
                 const double *B, blasint ldb, double beta, double *C, blasint ldc);
void cblas_sgemm(enum CBLAS_ORDER Order, enum CBLAS_TRANSPOSE TransA, enum CBLAS_TRANSPOSE TransB,
                 blasint M, blasint N, blasint K, float alpha, const float *A, blasint lda,
                 const float *B, blasint ldb, float beta, float *C, blasint ldc);
/* cblas.h:356-358, interface/trsm.c:213 */
void cblas_dtrsm(enum CBLAS_ORDER Order, enum CBLAS_SIDE Side, enum CBLAS_UPLO Uplo,
                 enum CBLAS_TRANSPOSE TransA, enum CBLAS_DIAG Diag, blasint M, blasint N,
                 double alpha, const double *A, blasint lda, double *B, blasint ldb);
void cblas_strsm(enum CBLAS_ORDER Order, enum CBLAS_SIDE Side, enum CBLAS_UPLO Uplo,
                 enum CBLAS_TRANSPOSE TransA, enum CBLAS_DIAG Diag, blasint M, blasint N,
                 float alpha, const float *A, blasint lda, float *B, blasint ldb);
/* cblas.h:419-422, interface/gemm_batch.c:117 (group API) */
void cblas_dgemm_batch(enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE *TransA_array,
                       const enum CBLAS_TRANSPOSE *TransB_array, const blasint *M_array,
                       const blasint *N_array, const blasint *K_array, const double *alpha_array,
                       const double **A_array, const blasint *lda_array, const double **B_array,
                       const blasint *ldb_array, const double *beta_array, double **C_array,
                       const blasint *ldc_array, blasint group_count, const blasint *group_size);
void cblas_sgemm_batch(enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE *TransA_array,
                       const enum CBLAS_TRANSPOSE *TransB_array, const blasint *M_array,
                       const blasint *N_array, const blasint *K_array, const float *alpha_array,
                       const float **A_array, const blasint *lda_array, const float **B_array,
                       const blasint *ldb_array, const float *beta_array, float **C_array,
                       const blasint *ldc_array, blasint group_count, const blasint *group_size);
/* Strided-batch extension of the group API above (no counterpart in the
   reference): batch_size problems of one shape, problem i at A + i*stridea,
   B + i*strideb, C + i*stridec (element strides; 0 shares an operand).
   Same argument checks as cblas_?gemm (numbers counted after Order:
   lda 8, ldb 11, ldc 15), plus stridea < 0 -> 9, strideb < 0 -> 12,
   stridec < 0 or overlapping C matrices -> 16, batch_size < 0 -> 17; xerbla
   names DGEMM_STRIDED / SGEMM_STRIDED.  Runs on the same persistent batched
   engine as cblas_?gemm_batch. */
void cblas_dgemm_batch_strided(enum CBLAS_ORDER Order, enum CBLAS_TRANSPOSE TransA, enum CBLAS_TRANSPOSE TransB,
                               blasint M, blasint N, blasint K, double alpha, const double *A, blasint lda,
                               long stridea, const double *B, blasint ldb, long strideb, double beta, double *C,
                               blasint ldc, long stridec, blasint batch_size);
void cblas_sgemm_batch_strided(enum CBLAS_ORDER Order, enum CBLAS_TRANSPOSE TransA, enum CBLAS_TRANSPOSE TransB,
                               blasint M, blasint N, blasint K, float alpha, const float *A, blasint lda,
                               long stridea, const float *B, blasint ldb, long strideb, float beta, float *C,
                               blasint ldc, long stridec, blasint batch_size);

/* cblas.h:313 (gemmt), :322 (symm), :331 (syrk), :349 (trmm); interface/syrk.c,
   symm.c, trsm.c (TRMM), gemmt.c built with -DCBLAS */
void cblas_dsyrk(enum CBLAS_ORDER Order, enum CBLAS_UPLO Uplo, enum CBLAS_TRANSPOSE Trans, blasint N, blasint K,
                 double alpha, const double *A, blasint lda, double beta, double *C, blasint ldc);
void cblas_ssyrk(enum CBLAS_ORDER Order, enum CBLAS_UPLO Uplo, enum CBLAS_TRANSPOSE Trans, blasint N, blasint K,
                 float alpha, const float *A, blasint lda, float beta, float *C, blasint ldc);
void cblas_dgemmt(enum CBLAS_ORDER Order, enum CBLAS_UPLO Uplo, enum CBLAS_TRANSPOSE TransA,
                  enum CBLAS_TRANSPOSE TransB, blasint M, blasint K, double alpha, const double *A, blasint lda,
                  const double *B, blasint ldb, double beta, double *C, blasint ldc);
void cblas_sgemmt(enum CBLAS_ORDER Order, enum CBLAS_UPLO Uplo, enum CBLAS_TRANSPOSE TransA,
                  enum CBLAS_TRANSPOSE TransB, blasint M, blasint K, float alpha, const float *A, blasint lda,
                  const float *B, blasint ldb, float beta, float *C, blasint ldc);
void cblas_dsymm(enum CBLAS_ORDER Order, enum CBLAS_SIDE Side, enum CBLAS_UPLO Uplo, blasint M, blasint N,
                 double alpha, const double *A, blasint lda, const double *B, blasint ldb, double beta, double *C,
                 blasint ldc);
void cblas_ssymm(enum CBLAS_ORDER Order, enum CBLAS_SIDE Side, enum CBLAS_UPLO Uplo, blasint M, blasint N,
                 float alpha, const float *A, blasint lda, const float *B, blasint ldb, float beta, float *C,
                 blasint ldc);
void cblas_dtrmm(enum CBLAS_ORDER Order, enum CBLAS_SIDE Side, enum CBLAS_UPLO Uplo, enum CBLAS_TRANSPOSE TransA,
                 enum CBLAS_DIAG Diag, blasint M, blasint N, double alpha, const double *A, blasint lda, double *B,
                 blasint ldb);
void cblas_strmm(enum CBLAS_ORDER Order, enum CBLAS_SIDE Side, enum CBLAS_UPLO Uplo, enum CBLAS_TRANSPOSE TransA,
                 enum CBLAS_DIAG Diag, blasint M, blasint N, float alpha, const float *A, blasint lda, float *B,
                 blasint ldb);

/* ---- Fortran BLAS level 3 (BLASFUNC(x) = x##_, common.h:194) -------------- */
/* interface/gemm.c:181 */
void dgemm_(const char *transa, const char *transb, const blasint *m, const blasint *n, const blasint *k,
            const double *alpha, const double *a, const blasint *lda, const double *b, const blasint *ldb,
            const double *beta, double *c, const blasint *ldc);
void sgemm_(const char *transa, const char *transb, const blasint *m, const blasint *n, const blasint *k,
            const float *alpha, const float *a, const blasint *lda, const float *b, const blasint *ldb,
            const float *beta, float *c, const blasint *ldc);
/* interface/trsm.c:114 */
void dtrsm_(const char *side, const char *uplo, const char *transa, const char *diag, const blasint *m,
            const blasint *n, const double *alpha, const double *a, const blasint *lda, double *b,
            const blasint *ldb);
void strsm_(const char *side, const char *uplo, const char *transa, const char *diag, const blasint *m,
            const blasint *n, const float *alpha, const float *a, const blasint *lda, float *b,
            const blasint *ldb);

/* interface/syrk.c, gemmt.c, symm.c, trsm.c (TRMM) */
void dsyrk_(const char *uplo, const char *trans, const blasint *n, const blasint *k, const double *alpha,
            const double *a, const blasint *lda, const double *beta, double *c, const blasint *ldc);
void ssyrk_(const char *uplo, const char *trans, const blasint *n, const blasint *k, const float *alpha,
            const float *a, const blasint *lda, const float *beta, float *c, const blasint *ldc);
void dgemmt_(const char *uplo, const char *transa, const char *transb, const blasint *m, const blasint *k,
             const double *alpha, const double *a, const blasint *lda, const double *b, const blasint *ldb,
             const double *beta, double *c, const blasint *ldc);
void sgemmt_(const char *uplo, const char *transa, const char *transb, const blasint *m, const blasint *k,
             const float *alpha, const float *a, const blasint *lda, const float *b, const blasint *ldb,
             const float *beta, float *c, const blasint *ldc);
void dsymm_(const char *side, const char *uplo, const blasint *m, const blasint *n, const double *alpha,
            const double *a, const blasint *lda, const double *b, const blasint *ldb, const double *beta, double *c,
            const blasint *ldc);
void ssymm_(const char *side, const char *uplo, const blasint *m, const blasint *n, const float *alpha,
            const float *a, const blasint *lda, const float *b, const blasint *ldb, const float *beta, float *c,
            const blasint *ldc);
void dtrmm_(const char *side, const char *uplo, const char *transa, const char *diag, const blasint *m,
            const blasint *n, const double *alpha, const double *a, const blasint *lda, double *b,
            const blasint *ldb);
void strmm_(const char *side, const char *uplo, const char *transa, const char *diag, const blasint *m,
            const blasint *n, const float *alpha, const float *a, const blasint *lda, float *b,
            const blasint *ldb);

/* ---- LAPACK (Fortran ABI) --------------------------------------------------- */
/* interface/lapack/getrf.c:53 */
int dgetrf_(const blasint *m, const blasint *n, double *a, const blasint *lda, blasint *ipiv, blasint *info);
int sgetrf_(const blasint *m, const blasint *n, float *a, const blasint *lda, blasint *ipiv, blasint *info);
/* interface/lapack/getf2.c (same contract as getrf, unblocked in the reference) */
int dgetf2_(const blasint *m, const blasint *n, double *a, const blasint *lda, blasint *ipiv, blasint *info);
int sgetf2_(const blasint *m, const blasint *n, float *a, const blasint *lda, blasint *ipiv, blasint *info);
/* interface/lapack/getrs.c:63 */
int dgetrs_(const char *trans, const blasint *n, const blasint *nrhs, const double *a, const blasint *lda,
            const blasint *ipiv, double *b, const blasint *ldb, blasint *info);
int sgetrs_(const char *trans, const blasint *n, const blasint *nrhs, const float *a, const blasint *lda,
            const blasint *ipiv, float *b, const blasint *ldb, blasint *info);
/* interface/lapack/gesv.c:63 */
int dgesv_(const blasint *n, const blasint *nrhs, double *a, const blasint *lda, blasint *ipiv, double *b,
           const blasint *ldb, blasint *info);
int sgesv_(const blasint *n, const blasint *nrhs, float *a, const blasint *lda, blasint *ipiv, float *b,
           const blasint *ldb, blasint *info);
/* interface/lapack/potrf.c:63 */
int dpotrf_(const char *uplo, const blasint *n, double *a, const blasint *lda, blasint *info);
int spotrf_(const char *uplo, const blasint *n, float *a, const blasint *lda, blasint *info);
/* interface/lapack/potf2.c */
int dpotf2_(const char *uplo, const blasint *n, double *a, const blasint *lda, blasint *info);
int spotf2_(const char *uplo, const blasint *n, float *a, const blasint *lda, blasint *info);
/* interface/lapack/laswp.c:55 */
int dlaswp_(const blasint *n, double *a, const blasint *lda, const blasint *k1, const blasint *k2,
            const blasint *ipiv, const blasint *incx);
int slaswp_(const blasint *n, float *a, const blasint *lda, const blasint *k1, const blasint *k2,
            const blasint *ipiv, const blasint *incx);
/* lapack-netlib/SRC/dsgesv.f:193 (WORK and SWORK are accepted and not used) */
int dsgesv_(const blasint *n, const blasint *nrhs, double *a, const blasint *lda, blasint *ipiv,
            const double *b, const blasint *ldb, double *x, const blasint *ldx, double *work,
            float *swork, blasint *iter, blasint *info);
/* lapack-netlib/SRC/dsposv.f:197 (Cholesky-based; WORK and SWORK are accepted and not used) */
int dsposv_(const char *uplo, const blasint *n, const blasint *nrhs, double *a, const blasint *lda,
            const double *b, const blasint *ldb, double *x, const blasint *ldx, double *work,
            float *swork, blasint *iter, blasint *info);

/* ---- LAPACKE (lapack-netlib/LAPACKE/include/lapacke.h) ---------------------- */
/* lapacke.h:872 and lapacke_dgesv_work.c:35 */
lapack_int LAPACKE_dgesv(int matrix_layout, lapack_int n, lapack_int nrhs, double *a, lapack_int lda,
                         lapack_int *ipiv, double *b, lapack_int ldb);
lapack_int LAPACKE_dgesv_work(int matrix_layout, lapack_int n, lapack_int nrhs, double *a, lapack_int lda,
                              lapack_int *ipiv, double *b, lapack_int ldb);
lapack_int LAPACKE_sgesv(int matrix_layout, lapack_int n, lapack_int nrhs, float *a, lapack_int lda,
                         lapack_int *ipiv, float *b, lapack_int ldb);
lapack_int LAPACKE_sgesv_work(int matrix_layout, lapack_int n, lapack_int nrhs, float *a, lapack_int lda,
                              lapack_int *ipiv, float *b, lapack_int ldb);
/* lapacke.h:883 */
lapack_int LAPACKE_dsgesv(int matrix_layout, lapack_int n, lapack_int nrhs, double *a, lapack_int lda,
                          lapack_int *ipiv, double *b, lapack_int ldb, double *x, lapack_int ldx,
                          lapack_int *iter);
/* lapacke.h:5845 (WORK and SWORK are accepted and not used) */
lapack_int LAPACKE_dsgesv_work(int matrix_layout, lapack_int n, lapack_int nrhs, double *a, lapack_int lda,
                               lapack_int *ipiv, double *b, lapack_int ldb, double *x, lapack_int ldx,
                               double *work, float *swork, lapack_int *iter);
/* lapacke.h:2974 and lapacke.h:8477 */
lapack_int LAPACKE_dsposv(int matrix_layout, char uplo, lapack_int n, lapack_int nrhs, double *a,
                          lapack_int lda, double *b, lapack_int ldb, double *x, lapack_int ldx,
                          lapack_int *iter);
lapack_int LAPACKE_dsposv_work(int matrix_layout, char uplo, lapack_int n, lapack_int nrhs, double *a,
                               lapack_int lda, double *b, lapack_int ldb, double *x, lapack_int ldx,
                               double *work, float *swork, lapack_int *iter);
/* lapacke.h:1062-1064 */
lapack_int LAPACKE_dgetrf(int matrix_layout, lapack_int m, lapack_int n, double *a, lapack_int lda,
                          lapack_int *ipiv);
lapack_int LAPACKE_dgetrf_work(int matrix_layout, lapack_int m, lapack_int n, double *a, lapack_int lda,
                               lapack_int *ipiv);
lapack_int LAPACKE_sgetrf(int matrix_layout, lapack_int m, lapack_int n, float *a, lapack_int lda,
                          lapack_int *ipiv);
lapack_int LAPACKE_sgetrf_work(int matrix_layout, lapack_int m, lapack_int n, float *a, lapack_int lda,
                               lapack_int *ipiv);
/* lapacke.h:1098 */
lapack_int LAPACKE_dgetrs(int matrix_layout, char trans, lapack_int n, lapack_int nrhs, const double *a,
                          lapack_int lda, const lapack_int *ipiv, double *b, lapack_int ldb);
lapack_int LAPACKE_dgetrs_work(int matrix_layout, char trans, lapack_int n, lapack_int nrhs, const double *a,
                               lapack_int lda, const lapack_int *ipiv, double *b, lapack_int ldb);
lapack_int LAPACKE_sgetrs(int matrix_layout, char trans, lapack_int n, lapack_int nrhs, const float *a,
                          lapack_int lda, const lapack_int *ipiv, float *b, lapack_int ldb);
lapack_int LAPACKE_sgetrs_work(int matrix_layout, char trans, lapack_int n, lapack_int nrhs, const float *a,
                               lapack_int lda, const lapack_int *ipiv, float *b, lapack_int ldb);
/* lapacke.h:3055-3057 */
lapack_int LAPACKE_dpotrf(int matrix_layout, char uplo, lapack_int n, double *a, lapack_int lda);
lapack_int LAPACKE_dpotrf_work(int matrix_layout, char uplo, lapack_int n, double *a, lapack_int lda);
lapack_int LAPACKE_spotrf(int matrix_layout, char uplo, lapack_int n, float *a, lapack_int lda);
lapack_int LAPACKE_spotrf_work(int matrix_layout, char uplo, lapack_int n, float *a, lapack_int lda);

/* ---- runtime / informational (cblas.h:13-27, driver/others/xerbla.c:56) ----- */
int xerbla_(char *name, blasint *info, blasint len);   /* weak: applications may override */
char *openblas_get_config(void);
char *openblas_get_corename(void);
void openblas_set_num_threads(int num_threads);
int openblas_get_num_threads(void);
int openblas_get_num_procs(void);

/* ---- extensions: no counterpart in the reference ---------------------------- */
/* FP64 GEMM emulated with INT8 slices on the INT8 MFMA (Ozaki-style): the
   cblas_dgemm contract plus the slice count (0 = chosen from the operands'
   exponent spread, else 1..12).  Returns the slice count used (0: nothing
   to multiply, only beta*C ran); -1 after an argument error (xerbla
   "DGEMM ", info 15 for a bad slice count); -2 when A or B holds a NaN or
   an infinity (the native FP64 GEMM computed C, so they propagate as in
   cblas_dgemm); -s (s > 12) when the automatic choice needs s slices, more
   than the kernel holds (the native FP64 GEMM computed C, so the result
   keeps FP64 accuracy); OPENBLAS_AMD_DEVICE_ERROR after a device error. */
int openblas_amd_dgemm_ozaki(enum CBLAS_ORDER Order, enum CBLAS_TRANSPOSE TransA, enum CBLAS_TRANSPOSE TransB,
                             int M, int N, int K, double alpha, const double *A, int lda,
                             const double *B, int ldb, double beta, double *C, int ldc, int slices);
/* The same product with Ozaki scheme II (Ozaki, Uchino, Imamura 2025):
   rows of op(A) and columns of op(B) scaled to integers, one exact INT8
   GEMM per modulus (pairwise coprime moduli <= 256), the product rebuilt
   by the Chinese remainder theorem and rounded once to FP64.  moduli = 0
   chooses the count from the operands' exponent spread (the typical entry
   of each operand keeps 53 bits), else 1..30.  Returns the modulus count
   used (0: nothing to multiply); -1 after an argument error (xerbla
   "DGEMM ", info 15 for a bad modulus count); -2 when A or B holds a NaN or
   an infinity (native FP64 GEMM); -n (n > 30) when the spread needs more
   moduli than the table holds (native FP64 GEMM, FP64 accuracy kept);
   OPENBLAS_AMD_DEVICE_ERROR after a device error. */
int openblas_amd_dgemm_ozaki2(enum CBLAS_ORDER Order, enum CBLAS_TRANSPOSE TransA, enum CBLAS_TRANSPOSE TransB,
                              int M, int N, int K, double alpha, const double *A, int lda,
                              const double *B, int ldb, double beta, double *C, int ldc, int moduli);
/* Batched LU / LU solve / Cholesky of many independent small problems in one
   persistent launch per call (the reference has no batched LAPACK; these
   are the batched forms of interface/lapack/getrf.c:53, getrs.c:63 and
   potrf.c:63).  Problem i has the arguments of the single routine at index
   i of each array (matrices host or device, per pointer); info is a host
   array receiving each problem's info exactly as the single routine would
   set it (a problem with an illegal argument gets -position and is
   skipped; the first one is reported through xerbla).  Returns 0, or -k
   when argument k of the batched call itself is illegal (trans/uplo,
   count < 0).  Problems up to 512 rows run in the batched kernels, larger
   ones through the single-problem path. */
int openblas_amd_dgetrf_batch(const int *m, const int *n, double *const *a, const int *lda,
                              int *const *ipiv, int *info, int count);
int openblas_amd_sgetrf_batch(const int *m, const int *n, float *const *a, const int *lda,
                              int *const *ipiv, int *info, int count);
int openblas_amd_dgetrs_batch(char trans, const int *n, const int *nrhs, double *const *a,
                              const int *lda, int *const *ipiv, double *const *b, const int *ldb,
                              int *info, int count);
int openblas_amd_sgetrs_batch(char trans, const int *n, const int *nrhs, float *const *a,
                              const int *lda, int *const *ipiv, float *const *b, const int *ldb,
                              int *info, int count);
int openblas_amd_dpotrf_batch(char uplo, const int *n, double *const *a, const int *lda, int *info,
                              int count);
int openblas_amd_spotrf_batch(char uplo, const int *n, float *const *a, const int *lda, int *info,
                              int count);
/* device memory and synchronisation for callers that keep matrices in HBM */
void *openblas_amd_malloc(size_t bytes);
void openblas_amd_free(void *p);
/* kind: 0 host->device, 1 device->host, 2 device->device */
int openblas_amd_memcpy(void *dst, const void *src, size_t bytes, int kind);
int openblas_amd_memset(void *dst, int value, size_t bytes);
int openblas_amd_synchronize(void);
void openblas_amd_set_sync(int on);
int openblas_amd_device_count(void);
int openblas_amd_device_info(char *arch, int len, int *num_cu, int *clock_khz);
/* HIP events on the library's stream: brackets exactly the library's kernels */
void openblas_amd_timer_start(void);
float openblas_amd_timer_stop(void);
/* kernel timer: when enabled, the library brackets its main kernels (LU
   panels and trailing updates, look-back solves, the GEMM/TRSM entry
   points, the INT8-slice product, the batched kernels) with HIP events on
   the stream each runs on and records their algorithmic flops and HBM
   bytes; read returns one named aggregate (launches, total ms, flops,
   bytes; 0, or -1 if nothing was recorded under that name) */
void openblas_amd_ktimer_enable(int on);
int openblas_amd_ktimer_read(const char *name, long *count, double *ms, double *flops, double *bytes);
void openblas_amd_ktimer_reset(void);
/* dgetrs_ that allocates nothing (no counterpart in the reference): for
   callers that cannot allocate device memory -- code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc.  Same
   arguments, checks and result as dgetrs_ (the same kernels), but A and B
   must be device memory (IPIV host or device), N <= 32768, and the kernels'
   workspace is a 6.1 MiB static device array per GPU instead of the memory
   pool.  Concurrent calls on one GPU race on that array.  A host A or B, or
   N > 32768, is reported through xerbla ("DGETRS_NOALLOC", parameter 4, 7
   or 2) with INFO = -4, -7 or -2. */
void openblas_amd_dgetrs_noalloc(const char *trans, const int *n, const int *nrhs, const double *a,
                                 const int *lda, const int *ipiv, double *b, const int *ldb,
                                 int *info);
/* last argument error reported through xerbla on this thread (0 if none);
   copies the routine name; clears the record */
int openblas_amd_last_xerbla(char *name, int len);
/* 1 if a device error was recorded since the last call (0 if none); copies
   its message; clears the record */
int openblas_amd_last_error(char *msg, int len);

#ifdef __cplusplus
}
#endif

#endif
