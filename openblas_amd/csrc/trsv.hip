// Triangular solves with few right-hand sides in one launch (the few-RHS
// case of the LU solves, lapack/getrs/getrs_single.c, which the reference
// runs as TRSV for NRHS = 1).
//
// M x = b with M(i,k) = A[i*rs + k*cs] (lower: forward substitution, upper:
// backward), unit or non-unit diagonal.  The rows are cut into 64-row
// blocks; one workgroup per block, blocks handed out by a ticket counter in
// dependency order (a workgroup only waits for blocks handed out before its
// own, which are running or done, so the grid cannot deadlock).  A block's
// workgroup first stages its diagonal block, then accumulates the
// contributions of the solved blocks as their solutions appear (each
// wavefront a quarter of every tile, loads coalesced along whichever index
// of A is contiguous), then wavefront 0 solves the diagonal block with
// v_readlane broadcasts and publishes its part of x as tagged words: every
// 32 bits of the solution paired with a nonzero tag in one 8-byte
// write-through store (seen whole or not at all).  Readers poll those words
// with cache-bypassing loads until every tag is set, so one round trip both
// shows that a block is solved and delivers its values; no flag, no wait for
// store completion, no cache writeback or invalidation.
// The recursion-based TRSM needs n/64 dependent launches for the same job.
#include "kernels.h"
#include "lu_internal.h"

namespace oba {

constexpr int TV_B = 64;

__device__ inline double tv_bcast(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
__device__ inline float tv_bcast(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }

// ROWLANES: rs == 1 (rows contiguous: lanes over rows); else cs == 1 (lanes
// over columns).
// published x of a block: TV_B values as 32-bit words, each {word, tag 1}
template <typename T> struct TvPub { static constexpr int WPV = sizeof(T) / 4; };

// Several right-hand sides: column c = blockIdx.y is x + c*ldx with its own
// ticket counter ticket[c] and published words pub + c*pstride, i.e. an
// independent chain per column; the chains run side by side in one launch.
template <typename T, bool ROWLANES>
__global__ __launch_bounds__(256) void trsv_lb_kernel(long n, const T* __restrict__ A, long rs, long cs, bool lower,
                                                      bool unit, T* __restrict__ x, long ldx,
                                                      unsigned long long* __restrict__ pub, long pstride,
                                                      unsigned* __restrict__ ticket, unsigned* __restrict__ err) {
  constexpr int WPV = TvPub<T>::WPV;
  x += (long)blockIdx.y * ldx;
  pub += (long)blockIdx.y * pstride;
  ticket += blockIdx.y;
  constexpr int LD = TV_B + 1;
  __shared__ T Md[TV_B * LD];      // M(r0+i, r0+k) at Md[i*LD + k]
  __shared__ T part[4][TV_B];
  __shared__ T s_red[ROWLANES ? 1 : TV_B * LD];   // column partial sums, transposed
  __shared__ T dinv[TV_B];
  __shared__ int s_blk;
  const int t = threadIdx.x, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const long nblk = (n + TV_B - 1) / TV_B;
  if (t == 0) s_blk = (int)atomicAdd(ticket, 1u);
  __syncthreads();
  const long b = s_blk;
  const long bi = lower ? b : nblk - 1 - b;
  const long r0 = bi * TV_B;
  const int nr = (int)min((long)TV_B, n - r0);
  for (int idx = t; idx < TV_B * TV_B; idx += 256) {
    const int f = idx & 63, g = idx >> 6;     // f along the contiguous index
    const int i = ROWLANES ? f : g, k = ROWLANES ? g : f;
    Md[i * LD + k] = (i < nr && k < nr) ? A[(r0 + i) * rs + (r0 + k) * cs] : T(0);
  }
  __syncthreads();
  if (t < TV_B) dinv[t] = (t < nr && !unit) ? T(1) / Md[t * LD + t] : T(1);
  // wavefront 0's inputs of the diagonal solve (its own right-hand side, the
  // diagonal block's row, the reciprocal pivot) do not depend on earlier
  // blocks: fetched now, off the dependency chain
  T v = T(0), myd = T(1);
  T md[TV_B];
  if (w == 0) {
    v = lane < nr ? x[r0 + lane] : T(0);
    myd = dinv[lane];   // written by this lane above
#pragma unroll
    for (int k = 0; k < TV_B; k++) md[k] = Md[lane * LD + k];
  }
  T acc = T(0);
  T accv[16];
#pragma unroll
  for (int q = 0; q < 16; q++) accv[q] = T(0);
  bool bad = false;
  for (long p = 0; p < b; p++) {
    const long kb = lower ? p : nblk - 1 - p;
    const long k0 = kb * TV_B;
    const int nk = (int)min((long)TV_B, n - k0);
    // this tile of A does not depend on block kb: load it before waiting
    T av[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      if (ROWLANES) {
        const int k = w * 16 + q;
        av[q] = (lane < nr && k < nk) ? A[(r0 + lane) * rs + (k0 + k) * cs] : T(0);
      } else {
        const int i = w * 16 + q;
        av[q] = (i < nr && lane < nk) ? A[(r0 + i) * rs + (k0 + lane) * cs] : T(0);
      }
    }
    // block kb's solution: poll its tagged words (lane e: value e)
    const unsigned long long* src = pub + (size_t)kb * TV_B * WPV;
    unsigned long long wv[WPV];
    bool have = lane >= nk;
#pragma unroll
    for (int h = 0; h < WPV; h++) wv[h] = 0;
    long spins = 0;
    bool ok = true;
    while (true) {
      if (!have) {
#pragma unroll
        for (int h = 0; h < WPV; h++) wv[h] = __hip_atomic_load(const_cast<unsigned long long*>(src + lane * WPV + h),
                                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        have = true;
#pragma unroll
        for (int h = 0; h < WPV; h++) have = have && (wv[h] >> 32) != 0;
      }
      if (__all(have)) break;
      __builtin_amdgcn_s_sleep(1);   // back off: every block's wavefronts poll
      ++spins;
      if (spins > (1L << 22) ||
          ((spins & 1023) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))) {
        if (lane == 0) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        ok = false;
        break;
      }
    }
    if (!ok) { bad = true; break; }
    T yl = T(0);
    if (lane < nk) {
      if constexpr (sizeof(T) == 8) yl = __hiloint2double((int)(unsigned)wv[1], (int)(unsigned)wv[0]);
      else yl = __int_as_float((int)(unsigned)wv[0]);
    }
    if (ROWLANES) {
#pragma unroll
      for (int q = 0; q < 16; q++) acc += av[q] * tv_bcast(yl, w * 16 + q);   // av = 0 outside the tile
    } else {
#pragma unroll
      for (int q = 0; q < 16; q++) accv[q] += av[q] * yl;
    }
  }
  // per-lane partial sums: by rows (ROWLANES) or, transposed through LDS,
  // by columns, summed along each row by one lane
  if (ROWLANES) {
    part[w][lane] = acc;
  } else {
#pragma unroll
    for (int q = 0; q < 16; q++) s_red[(w * 16 + q) * LD + lane] = accv[q];
  }
  __syncthreads();
  if (w != 0 || bad) return;
  if (ROWLANES) {
    v -= (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
  } else {
    T sacc[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int l = 0; l < TV_B; l++) sacc[l & 3] += s_red[lane * LD + l];
    v -= (sacc[0] + sacc[1]) + (sacc[2] + sacc[3]);
  }
  // the diagonal block's row of this lane is in registers; static steps
  if (lower) {
#pragma unroll
    for (int k = 0; k < TV_B; k++) {
      if (k < nr) {
        if (lane == k) v *= myd;
        const T xk = tv_bcast(v, k);
        if (lane > k) v -= md[k] * xk;
      }
    }
  } else {
#pragma unroll
    for (int k = TV_B - 1; k >= 0; k--) {
      if (k < nr) {
        if (lane == k) v *= myd;
        const T xk = tv_bcast(v, k);
        if (lane < k) v -= md[k] * xk;
      }
    }
  }
  // the tagged words for the later blocks, then the result itself
  if (lane < nr) {
    unsigned long long* dst = pub + (size_t)bi * TV_B * WPV + lane * WPV;
    const unsigned long long tg = 1ull << 32;
    if constexpr (sizeof(T) == 8) {
      __hip_atomic_store(dst, tg | (unsigned)__double2loint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(dst + 1, tg | (unsigned)__double2hiint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    } else {
      __hip_atomic_store(dst, tg | (unsigned)__float_as_int(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    x[r0 + lane] = v;
  }
}

// published words of one column's chain (bytes)
template <typename T>
static size_t tv_pub_bytes(long n) {
  return sizeof(unsigned long long) * (size_t)((n + TV_B - 1) / TV_B) * TV_B * TvPub<T>::WPV;
}
static size_t tv_ticket_bytes(long ncols) { return ((size_t)ncols * sizeof(unsigned) + 255) / 256 * 256; }

template <typename T>
size_t trsv_lookback_ws_bytes(long n, long ncols) {
  return tv_ticket_bytes(ncols) + (size_t)ncols * tv_pub_bytes<T>(n);
}

template <typename T>
void trsv_lookback(bool lower, bool unit, long n, const T* a, long rs, long cs, T* x, long ldx, long ncols,
                   hipStream_t s) {
  if (n <= 0 || ncols <= 0) return;
  const size_t bytes = trsv_lookback_ws_bytes<T>(n, ncols);
  DevBuf ws(bytes);
  OBA_HIP(hipMemsetAsync(ws.p, 0, bytes, s));
  trsv_lookback_ws<T>(lower, unit, n, a, rs, cs, x, ldx, ncols, ws.p, s);
}

// The launch on caller-provided, zeroed workspace of trsv_lookback_ws_bytes(n,
// ncols) bytes (the ticket words, then each column's published words).  A
// wait that runs out sets the runtime's device error word, which the API
// layer checks after its next synchronization (the kernel itself never
// blocks the host).
template <typename T>
void trsv_lookback_ws(bool lower, bool unit, long n, const T* a, long rs, long cs, T* x, long ldx, long ncols,
                      void* ws, hipStream_t s) {
  if (n <= 0 || ncols <= 0) return;
  if (rs != 1 && cs != 1) fatal("trsv_lookback: one index of A must be contiguous", nullptr);
  if (ncols > 65535) fatal("trsv_lookback: too many right-hand sides for one launch", nullptr);
  const long nblk = (n + TV_B - 1) / TV_B;
  unsigned* ticket = (unsigned*)ws;
  unsigned long long* pub = (unsigned long long*)((char*)ws + tv_ticket_bytes(ncols));
  const long pstride = (long)(tv_pub_bytes<T>(n) / sizeof(unsigned long long));
  unsigned* err = rt().d_err;
  KTimer kt("trsv_lookback", s, (double)n * n * ncols, (double)sizeof(T) * (0.5 * n * n + 2.0 * n * ncols));
  const dim3 grid((unsigned)nblk, (unsigned)ncols);
  if (rs == 1)
    trsv_lb_kernel<T, true><<<grid, 256, 0, s>>>(n, a, rs, cs, lower, unit, x, ldx, pub, pstride, ticket, err);
  else
    trsv_lb_kernel<T, false><<<grid, 256, 0, s>>>(n, a, rs, cs, lower, unit, x, ldx, pub, pstride, ticket, err);
  OBA_HIP(hipGetLastError());
}

template size_t trsv_lookback_ws_bytes<double>(long, long);
template size_t trsv_lookback_ws_bytes<float>(long, long);
template void trsv_lookback_ws<double>(bool, bool, long, const double*, long, long, double*, long, long, void*,
                                       hipStream_t);
template void trsv_lookback_ws<float>(bool, bool, long, const float*, long, long, float*, long, long, void*,
                                      hipStream_t);
template void trsv_lookback<double>(bool, bool, long, const double*, long, long, double*, long, long, hipStream_t);
template void trsv_lookback<float>(bool, bool, long, const float*, long, long, float*, long, long, hipStream_t);

}  // namespace oba
