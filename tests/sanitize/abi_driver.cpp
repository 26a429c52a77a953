// Drives libopenblas_amd's C ABI in a host-sanitized build (AddressSanitizer
// and UndefinedBehaviorSanitizer on the host code).
//
//   abi_driver cpu   argument errors (xerbla numbers, through an overriding
//                    xerbla_ as the reference allows, driver/others/xerbla.c),
//                    quick returns and the informational calls; no device work
//   abi_driver gpu   the above, then real calls on host and device operands:
//                    dgemm in both layouts and all transposes against a naive
//                    product, the group and strided batch APIs, dgesv_ /
//                    dsgesv_ / dpotrf_ residuals, LAPACKE row-major, and four
//                    host threads calling dgemm at once (results bit-equal)
//
// Prints "OK" and exits 0 when every check passed; a sanitizer report
// aborts the process with a non-zero status.
#include <openblas_amd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

static std::string g_name;
static int g_info = 0;
static int g_fail = 0;

// overrides the library's weak xerbla_ (the reference's override contract)
extern "C" int xerbla_(char* name, blasint* info, blasint len) {
  std::string s(name, strnlen(name, (size_t)len));
  while (!s.empty() && s.back() == ' ') s.pop_back();
  g_name = s;
  g_info = *info;
  return 0;
}

#define CHECK(cond, ...)                                   \
  do {                                                     \
    if (!(cond)) {                                         \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);     \
      std::printf(__VA_ARGS__);                            \
      std::printf("\n");                                   \
      g_fail++;                                            \
    }                                                      \
  } while (0)

static void expect_xerbla(const char* name, int info, const char* what) {
  CHECK(g_name == name && g_info == info, "%s: xerbla(%s, %d), want (%s, %d)", what, g_name.c_str(), g_info, name,
        info);
  g_name.clear();
  g_info = 0;
}

static void cpu_checks() {
  std::vector<double> z(64, 0.0), c(64, 7.0);
  std::vector<float> zf(64, 0.0f);
  // cblas_dgemm (interface/gemm.c:423-471 numbering)
  cblas_dgemm(CblasColMajor, (CBLAS_TRANSPOSE)99, CblasNoTrans, 2, 2, 2, 1.0, z.data(), 2, z.data(), 2, 0.0, c.data(), 2);
  expect_xerbla("DGEMM", 1, "dgemm transa");
  cblas_dgemm(CblasColMajor, CblasNoTrans, CblasNoTrans, 2, 2, 2, 1.0, z.data(), 1, z.data(), 2, 0.0, c.data(), 2);
  expect_xerbla("DGEMM", 8, "dgemm lda");
  cblas_dgemm(CblasColMajor, CblasNoTrans, CblasNoTrans, 2, 2, 2, 1.0, z.data(), 2, z.data(), 2, 0.0, c.data(), 1);
  expect_xerbla("DGEMM", 13, "dgemm ldc");
  cblas_sgemm(CblasRowMajor, CblasNoTrans, CblasNoTrans, 2, 3, 2, 1.0f, zf.data(), 1, zf.data(), 3, 0.0f,
              (float*)c.data(), 3);
  expect_xerbla("SGEMM", 10, "sgemm row-major lda");
  // TRSM
  cblas_dtrsm(CblasColMajor, CblasLeft, CblasLower, CblasNoTrans, CblasNonUnit, 4, 2, 1.0, z.data(), 3, c.data(), 4);
  expect_xerbla("DTRSM", 9, "dtrsm lda");
  // group batch and the strided extension
  CBLAS_TRANSPOSE nt[2] = {CblasNoTrans, CblasNoTrans};
  blasint mm[2] = {2, 2}, ld[2] = {2, 2}, ldbad[2] = {2, 1}, gs[2] = {1, 1};
  double al[2] = {1, 1}, be[2] = {0, 0};
  const double* ap[2] = {z.data(), z.data()};
  double* cp[2] = {c.data(), c.data()};
  cblas_dgemm_batch(CblasColMajor, nt, nt, mm, mm, mm, al, ap, ld, ap, ldbad, be, cp, ld, 2, gs);
  expect_xerbla("DGEMM_BATCH", 10, "gemm_batch ldb");
  cblas_dgemm_batch_strided(CblasColMajor, CblasNoTrans, CblasNoTrans, 2, 2, 2, 1.0, z.data(), 2, 4, z.data(), 2, 4,
                            0.0, c.data(), 2, 3, 2);
  expect_xerbla("DGEMM_STRIDED", 16, "strided overlap");
  cblas_dgemm_batch_strided(CblasColMajor, CblasNoTrans, CblasNoTrans, 2, 2, 2, 1.0, z.data(), 2, 4, z.data(), 2, 4,
                            0.0, c.data(), 2, 4, -1);
  expect_xerbla("DGEMM_STRIDED", 17, "strided batch_size");
  // level-3 extensions
  cblas_dsyrk(CblasColMajor, CblasLower, CblasNoTrans, 4, 2, 1.0, z.data(), 3, 0.0, c.data(), 4);
  expect_xerbla("DSYRK", 7, "dsyrk lda");
  cblas_dtrmm(CblasColMajor, CblasLeft, CblasUpper, CblasNoTrans, CblasNonUnit, 4, 2, 1.0, z.data(), 3, c.data(), 4);
  expect_xerbla("DTRMM", 9, "dtrmm lda");
  // LAPACK (interface/lapack/*.c): INFO = -i and xerbla
  blasint n = 4, nrhs = 1, lda = 3, info = 0;
  std::vector<int> piv(8);
  dgesv_(&n, &nrhs, z.data(), &lda, piv.data(), c.data(), &n, &info);
  CHECK(info == -4, "dgesv info %d", info);
  expect_xerbla("DGESV", 4, "dgesv lda");
  blasint m = -1;
  dgetrf_(&m, &n, z.data(), &n, piv.data(), &info);
  CHECK(info == -1, "dgetrf info %d", info);
  expect_xerbla("DGETRF", 1, "dgetrf m");
  char up = 'X';
  dpotrf_(&up, &n, z.data(), &n, &info);
  CHECK(info == -1, "dpotrf info %d", info);
  expect_xerbla("DPOTRF", 1, "dpotrf uplo");
  // LAPACKE: bad layout -> -1 (LAPACKE_xerbla, not xerbla_)
  CHECK(LAPACKE_dgesv(7, 2, 1, z.data(), 2, piv.data(), c.data(), 2) == -1, "LAPACKE layout");
  // quick returns leave C untouched and report nothing
  g_name.clear();
  g_info = 0;
  std::vector<double> c7(4, 7.0);
  cblas_dgemm(CblasColMajor, CblasNoTrans, CblasNoTrans, 0, 2, 2, 1.0, z.data(), 1, z.data(), 2, 0.0, c7.data(), 1);
  cblas_dgemm(CblasColMajor, CblasNoTrans, CblasNoTrans, 2, 2, 0, 1.0, z.data(), 2, z.data(), 1, 1.0, c7.data(), 2);
  CHECK(c7[0] == 7.0 && c7[3] == 7.0 && g_info == 0, "quick returns");
  // informational calls
  CHECK(openblas_get_config() && std::strlen(openblas_get_config()) > 0, "config");
  openblas_set_num_threads(4);
  CHECK(openblas_get_num_threads() >= 1, "num_threads");
  char msg[256];
  (void)openblas_amd_last_error(msg, sizeof msg);
}

// ---- device-backed checks ---------------------------------------------------------
static double naive_err(bool rowmajor, bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int lda,
                        const double* B, int ldb, double beta, const double* C0, const double* C, int ldc) {
  // element (i, j) of a logical matrix in the given layout
  auto at = [&](const double* X, int ldx, int i, int j) { return rowmajor ? X[(long)i * ldx + j] : X[i + (long)j * ldx]; };
  double worst = 0;
  for (int i = 0; i < m; i++)
    for (int j = 0; j < n; j++) {
      double s = 0, g = 0;
      for (int p = 0; p < k; p++) {
        double a = ta ? at(A, lda, p, i) : at(A, lda, i, p);
        double b = tb ? at(B, ldb, j, p) : at(B, ldb, p, j);
        s += a * b;
        g += std::fabs(a * b);
      }
      double ref = alpha * s + beta * at(C0, ldc, i, j);
      double gauge = std::fabs(alpha) * g + std::fabs(beta * at(C0, ldc, i, j)) + 1e-300;
      worst = std::max(worst, std::fabs(at(C, ldc, i, j) - ref) / gauge);
    }
  return worst;
}

static void gpu_checks() {
  std::mt19937_64 rng(42);
  std::uniform_real_distribution<double> U(-1, 1);
  auto fill = [&](std::vector<double>& v) { for (auto& x : v) x = U(rng); };
  const double eps = 2.220446049250313e-16;
  // dgemm: both layouts, all transposes, host operands with padding
  const int m = 70, n = 53, k = 91;
  for (int layout : {CblasColMajor, CblasRowMajor})
    for (int ta = 0; ta < 2; ta++)
      for (int tb = 0; tb < 2; tb++) {
        bool row = layout == CblasRowMajor;
        int ar = ta ? k : m, ac = ta ? m : k, br = tb ? n : k, bc = tb ? k : n;
        int lda = (row ? ac : ar) + 3, ldb = (row ? bc : br) + 1, ldc = (row ? n : m) + 2;
        std::vector<double> A((size_t)lda * (row ? ar : ac)), B((size_t)ldb * (row ? br : bc)),
            C((size_t)ldc * (row ? m : n));
        fill(A); fill(B); fill(C);
        std::vector<double> C0 = C;
        cblas_dgemm((CBLAS_ORDER)layout, ta ? CblasTrans : CblasNoTrans, tb ? CblasTrans : CblasNoTrans, m, n, k, 0.7,
                    A.data(), lda, B.data(), ldb, -0.3, C.data(), ldc);
        double r = naive_err(row, ta, tb, m, n, k, 0.7, A.data(), lda, B.data(), ldb, -0.3, C0.data(), C.data(), ldc);
        CHECK(r < 16 * k * eps, "dgemm layout %d ta %d tb %d: err %g", layout, ta, tb, r);
      }
  // strided batch on device memory, against the group API (same bits)
  {
    const int cnt = 12, mm = 20, nn = 17, kk = 33;
    std::vector<double> A((size_t)mm * kk * cnt), B((size_t)kk * nn * cnt), C((size_t)mm * nn * cnt, 0.0), C2 = C;
    fill(A); fill(B);
    double *dA = (double*)openblas_amd_malloc(A.size() * 8), *dB = (double*)openblas_amd_malloc(B.size() * 8),
           *dC = (double*)openblas_amd_malloc(C.size() * 8);
    openblas_amd_memcpy(dA, A.data(), A.size() * 8, 0);
    openblas_amd_memcpy(dB, B.data(), B.size() * 8, 0);
    openblas_amd_memcpy(dC, C.data(), C.size() * 8, 0);
    cblas_dgemm_batch_strided(CblasColMajor, CblasNoTrans, CblasTrans, mm, nn, kk, 1.0, dA, mm, (long)mm * kk, dB, nn,
                              (long)kk * nn, 0.0, dC, mm, (long)mm * nn, cnt);
    openblas_amd_memcpy(C.data(), dC, C.size() * 8, 1);
    std::vector<const double*> ap(cnt), bp(cnt);
    std::vector<double*> cp(cnt);
    for (int i = 0; i < cnt; i++) {
      ap[i] = A.data() + (size_t)i * mm * kk;
      bp[i] = B.data() + (size_t)i * kk * nn;
      cp[i] = C2.data() + (size_t)i * mm * nn;
    }
    CBLAS_TRANSPOSE t1 = CblasNoTrans, t2 = CblasTrans;
    blasint M = mm, N = nn, K = kk, LA = mm, LB = nn, LC = mm, G = cnt;
    double al = 1.0, be = 0.0;
    cblas_dgemm_batch(CblasColMajor, &t1, &t2, &M, &N, &K, &al, ap.data(), &LA, bp.data(), &LB, &be, cp.data(), &LC,
                      1, &G);
    CHECK(std::memcmp(C.data(), C2.data(), C.size() * 8) == 0, "strided (device) != group (host)");
    openblas_amd_free(dA);
    openblas_amd_free(dB);
    openblas_amd_free(dC);
  }
  // dgesv_, dsgesv_, dpotrf_ on host operands
  {
    const int N = 300;
    std::vector<double> A((size_t)N * N), b(N), A0, b0;
    fill(A);
    for (int i = 0; i < N; i++) A[(size_t)i * N + i] += N;
    fill(b);
    A0 = A; b0 = b;
    std::vector<int> piv(N);
    blasint n = N, one = 1, info = -9, iter = -9;
    dgesv_(&n, &one, A.data(), &n, piv.data(), b.data(), &n, &info);
    CHECK(info == 0, "dgesv info %d", info);
    double rmax = 0, an = 0, xn = 0;
    for (int i = 0; i < N; i++) {
      double s = -b0[i], rs = 0;
      for (int j = 0; j < N; j++) { s += A0[i + (size_t)j * N] * b[j]; rs += std::fabs(A0[i + (size_t)j * N]); }
      rmax = std::max(rmax, std::fabs(s)); an = std::max(an, rs); xn = std::max(xn, std::fabs(b[i]));
    }
    CHECK(rmax / (an * xn) < 64 * N * eps, "dgesv residual %g", rmax / (an * xn));
    std::vector<double> x(N), work((size_t)N);
    std::vector<float> swork((size_t)N * (N + 1));
    A = A0;
    dsgesv_(&n, &one, A.data(), &n, piv.data(), b0.data(), &n, x.data(), &n, work.data(), swork.data(), &iter, &info);
    CHECK(info == 0 && iter >= 0, "dsgesv info %d iter %d", info, iter);
    // SPD: A0^T A0 + N I, lower
    std::vector<double> S((size_t)N * N, 0.0);
    for (int i = 0; i < N; i++)
      for (int j = 0; j <= i; j++) {
        double s = 0;
        for (int p = 0; p < N; p++) s += A0[p + (size_t)i * N] * A0[p + (size_t)j * N];
        S[i + (size_t)j * N] = s + (i == j ? N : 0);
      }
    char lo = 'L';
    dpotrf_(&lo, &n, S.data(), &n, &info);
    CHECK(info == 0, "dpotrf info %d", info);
    // LAPACKE row-major solve: the same system transposed in memory
    std::vector<double> R((size_t)N * N), rb(N, 1.0);
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++) R[(size_t)i * N + j] = A0[i + (size_t)j * N];
    CHECK(LAPACKE_dgesv(LAPACK_ROW_MAJOR, N, 1, R.data(), N, piv.data(), rb.data(), 1) == 0, "LAPACKE row-major");
  }
  // LAPACKE_dpotrf's NaN check of the referenced triangle (done on the
  // staged copy, so it needs the device) -> -4
  {
    std::vector<double> an(4, 1.0);
    an[0] = std::nan("");
    CHECK(LAPACKE_dpotrf(LAPACK_COL_MAJOR, 'L', 2, an.data(), 2) == -4, "LAPACKE_dpotrf NaN");
  }
  // four threads, identical 256^3 products: bit-equal results
  {
    const int S = 256;
    std::vector<double> A((size_t)S * S), B((size_t)S * S);
    fill(A); fill(B);
    std::vector<std::vector<double>> C(4, std::vector<double>((size_t)S * S, 0.0));
    std::vector<std::thread> th;
    for (int t = 0; t < 4; t++)
      th.emplace_back([&, t] {
        for (int r = 0; r < 3; r++)
          cblas_dgemm(CblasColMajor, CblasNoTrans, CblasNoTrans, S, S, S, 1.0, A.data(), S, B.data(), S, 0.0,
                      C[t].data(), S);
      });
    for (auto& t : th) t.join();
    for (int t = 1; t < 4; t++) CHECK(std::memcmp(C[0].data(), C[t].data(), C[0].size() * 8) == 0, "thread %d", t);
  }
}

int main(int argc, char** argv) {
  const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
  cpu_checks();
  if (gpu) {
    if (openblas_amd_device_count() < 1) {
      std::printf("FAIL: no device\n");
      return 2;
    }
    gpu_checks();
  }
  if (g_fail) {
    std::printf("%d checks failed\n", g_fail);
    return 1;
  }
  std::printf("OK\n");
  return 0;
}
