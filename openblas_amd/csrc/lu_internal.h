// Internal LU helpers shared by getrf.hip, lapack_api.cpp and dsgesv.hip.
#pragma once
#include "common.h"
#include <vector>

namespace oba {

// out(c, r) = (To) in(r, c); sets *overflow when a value does not fit To.
template <typename Ti, typename To>
void transpose(long rows, long cols, const Ti* in, long ldi, To* out, long ldo, int* overflow, hipStream_t s);

// LU of A (m x n) given as its transpose At (n x m, leading dim ldt).  With
// track_perm, At must have ldt >= n+1: row n then carries the original row
// numbers through the interchanges (read them with perm_from_tracked).
template <typename T>
int getrf_t(long m, long n, T* At, long ldt, int* ipiv_dev, bool track_perm, hipStream_t s);

// getrf on column-major A (transposes in and out); with perm_dev the
// composed row permutation is returned as well (perm[i] = original row).
template <typename T>
int getrf_perm(long m, long n, T* a, long lda, int* ipiv_dev, int* perm_dev, hipStream_t s);

template <typename T>
void perm_from_tracked(const T* At, long ldt, long m, long n, int* perm_dev, hipStream_t s);

template <typename T>
void permute_rows(bool inverse, long n, long nrhs, const int* perm_dev, T* b, long ldb, hipStream_t s);
// the same with a caller-provided n x nrhs staging buffer (no allocation)
template <typename T>
void permute_rows_ws(bool inverse, long n, long nrhs, const int* perm_dev, T* b, long ldb, T* tmp, hipStream_t s);

void ipiv_to_perm(long n, const int* ipiv_host, std::vector<int>& perm);

template <typename T>
void getrs_perm(bool trans, long n, long nrhs, const T* a, long lda, const int* perm_dev, T* b, long ldb,
                hipStream_t s);

// M X = B in a single launch for the ncols columns X = x + c*ldx, M(i,k) =
// a[i*rs + k*cs] (rs == 1 or cs == 1), lower -> forward, upper -> backward
// substitution; one dependency chain per column, all chains side by side.
template <typename T>
void trsv_lookback(bool lower, bool unit, long n, const T* a, long rs, long cs, T* x, long ldx, long ncols,
                   hipStream_t s);

// The same on caller-provided workspace of trsv_lookback_ws_bytes(n, ncols)
// zeroed bytes; no allocation.
template <typename T>
size_t trsv_lookback_ws_bytes(long n, long ncols);
template <typename T>
void trsv_lookback_ws(bool lower, bool unit, long n, const T* a, long rs, long cs, T* x, long ldx, long ncols,
                      void* ws, hipStream_t s);

template <typename T>
void getrs_t(long n, long nrhs, const T* At, long ldt, const int* perm_dev, T* b, long ldb, hipStream_t s);

}  // namespace oba
