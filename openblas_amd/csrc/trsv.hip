// Triangular solve with one right-hand side in one launch (the few-RHS case
// of the LU solves, lapack/getrs/getrs_single.c, which the reference runs as
// TRSV for NRHS = 1).
//
// M x = b with M(i,k) = A[i*rs + k*cs] (lower: forward substitution, upper:
// backward), unit or non-unit diagonal.  The rows are cut into 64-row
// blocks; one workgroup per block, blocks handed out by a ticket counter in
// dependency order (a workgroup only waits for blocks handed out before its
// own, which are running or done, so the grid cannot deadlock).  A block's
// workgroup first stages its diagonal block, then accumulates the
// contributions of the solved blocks as their ready flags appear (each
// wavefront a quarter of every tile, loads coalesced along whichever index
// of A is contiguous), then wavefront 0 solves the diagonal block with
// v_readlane broadcasts and publishes its part of x: write-through stores,
// a wait for their completion, then a write-through flag; readers poll the
// flag and read x with cache-bypassing loads, so no cache writeback or
// invalidation (expensive on this part) happens per tile.
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

template <typename T>
__device__ inline T tv_wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// ROWLANES: rs == 1 (rows contiguous: lanes over rows); else cs == 1 (lanes
// over columns).
template <typename T, bool ROWLANES>
__global__ __launch_bounds__(256) void trsv_lb_kernel(long n, const T* __restrict__ A, long rs, long cs, bool lower,
                                                      bool unit, T* __restrict__ x, int* __restrict__ flags,
                                                      unsigned* __restrict__ ticket, unsigned* __restrict__ err) {
  constexpr int LD = TV_B + 1;
  __shared__ T Md[TV_B * LD];      // M(r0+i, r0+k) at Md[i*LD + k]
  __shared__ T part[4][TV_B];
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
    int ok = 1;
    if (lane == 0) {
      long spins = 0;
      while (__hip_atomic_load(&flags[kb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) == 0) {
        __builtin_amdgcn_s_sleep(1);   // back off: every block's wavefronts poll
        ++spins;
        if (spins > (1L << 22) ||
            ((spins & 1023) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
          atomicExch(err, 1u);
          ok = 0;
          break;
        }
      }
    }
    if (!__builtin_amdgcn_readfirstlane(ok)) { bad = true; break; }
    // issued after the flag was seen; reads memory, not a stale cache line
    const T yl = lane < nk ? __hip_atomic_load(&x[k0 + lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) : T(0);
    if (ROWLANES) {
#pragma unroll
      for (int q = 0; q < 16; q++) acc += av[q] * tv_bcast(yl, w * 16 + q);   // av = 0 outside the tile
    } else {
#pragma unroll
      for (int q = 0; q < 16; q++) accv[q] += av[q] * yl;
    }
  }
  if (ROWLANES) {
    part[w][lane] = acc;
  } else {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const T v = tv_wave_sum(accv[q]);
      if (lane == 0) part[0][w * 16 + q] = v;
    }
  }
  __syncthreads();
  if (w != 0 || bad) return;
  T v = lane < nr ? x[r0 + lane] : T(0);
  v -= ROWLANES ? (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]) : part[0][lane];
  // the diagonal block's row of this lane in registers; static steps
  T md[TV_B];
#pragma unroll
  for (int k = 0; k < TV_B; k++) md[k] = Md[lane * LD + k];
  const T myd = dinv[lane];
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
  // write-through stores of this block's solution, wait for their
  // completion, then the (write-through) flag: no cache writeback needed
  if (lane < nr) __hip_atomic_store(&x[r0 + lane], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  if (lane == 0) __hip_atomic_store(&flags[bi], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <typename T>
void trsv_lookback(bool lower, bool unit, long n, const T* a, long rs, long cs, T* x, hipStream_t s) {
  if (n <= 0) return;
  if (rs != 1 && cs != 1) fatal("trsv_lookback: one index of A must be contiguous", nullptr);
  const long nblk = (n + TV_B - 1) / TV_B;
  DevBuf ws(sizeof(int) * (size_t)(nblk + 64));
  int* flags = ws.as<int>();
  unsigned* ticket = (unsigned*)(flags + nblk);
  unsigned* err = ticket + 16;
  OBA_HIP(hipMemsetAsync(ws.p, 0, sizeof(int) * (size_t)(nblk + 64), s));
  if (rs == 1)
    trsv_lb_kernel<T, true><<<(unsigned)nblk, 256, 0, s>>>(n, a, rs, cs, lower, unit, x, flags, ticket, err);
  else
    trsv_lb_kernel<T, false><<<(unsigned)nblk, 256, 0, s>>>(n, a, rs, cs, lower, unit, x, flags, ticket, err);
  OBA_HIP(hipGetLastError());
  unsigned h = 0;
  OBA_HIP(hipMemcpyAsync(&h, err, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  OBA_HIP(hipStreamSynchronize(s));
  if (h) fatal("trsv: block exchange timed out", nullptr);
}

template void trsv_lookback<double>(bool, bool, long, const double*, long, long, double*, hipStream_t);
template void trsv_lookback<float>(bool, bool, long, const float*, long, long, float*, hipStream_t);

}  // namespace oba
