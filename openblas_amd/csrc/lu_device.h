// Device helpers shared by the LU kernels (single-problem panel and batched).
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>

namespace oba {

template <typename T> struct MinNormal;
template <> struct MinNormal<double> { static constexpr double v = DBL_MIN; };
template <> struct MinNormal<float> { static constexpr float v = FLT_MIN; };

// a panel row lives in registers as one vector value; dynamic element
// indices go through static selects so the row never spills to scratch
template <typename T, int W> struct RowVec { typedef T type __attribute__((ext_vector_type(W))); };

template <typename V, int W>
__device__ __attribute__((always_inline)) inline auto pick(const V& x, int j) {
  auto v = x[0];
#pragma unroll
  for (int c = 1; c < W; c++) v = (c == j) ? x[c] : v;
  return v;
}

// A panel row in registers with element access by a uniform runtime index,
// which the compiler maps to GPR index mode (s_set_gpr_idx_on + v_mov)
// instead of selects or scratch.  FP64 rows are kept as their low and high
// 32-bit halves: index mode handles vectors of 32 dwords, not of 32 qwords.
template <typename T, int W> struct PRow;
template <int W> struct PRow<float, W> {
  typedef float V __attribute__((ext_vector_type(W)));
  V v;
  __device__ __attribute__((always_inline)) float get(int c) const { return v[c]; }
  __device__ __attribute__((always_inline)) void set(int c, float x) { v[c] = x; }
};
template <int W> struct PRow<double, W> {
  typedef int V __attribute__((ext_vector_type(W)));
  V lo, hi;
  __device__ __attribute__((always_inline)) double get(int c) const { return __hiloint2double(hi[c], lo[c]); }
  __device__ __attribute__((always_inline)) void set(int c, double x) {
    lo[c] = __double2loint(x);
    hi[c] = __double2hiint(x);
  }
};

// pivot key: |x|; a NaN only wins when it sits at the current position j
// (IAMAX returns the first element when that one is NaN, and otherwise never
// selects a NaN -- kernel/arm/iamax.c:51-74)
template <typename T>
__device__ inline double pivot_key(T xv, int pos, int j) {
  return isnan(xv) ? (pos == j ? (double)INFINITY : -1.0) : fabs((double)xv);
}

// larger key wins; equal keys -> lower position (IAMAX: first maximum)
__device__ inline bool better(double v, int p, double bv, int bp) { return v > bv || (v == bv && p < bp); }

// one step of a wavefront max-reduction of (key, position) through DPP lane
// moves (CTRL: a DPP control word whose source lanes are all in the row)
template <int CTRL>
__device__ __attribute__((always_inline)) inline void lu_dpp_best(double& bv, int& bp) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(bv), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(bv), CTRL, 0xF, 0xF, false);
  const int op = __builtin_amdgcn_update_dpp(0, bp, CTRL, 0xF, 0xF, false);
  const double ov = __hiloint2double(hi, lo);
  if (better(ov, op, bv, bp)) { bv = ov; bp = op; }
}

// the wavefront's best (key, position), returned in every lane: butterfly
// within each row of 16 lanes (quad_perm xor 1, xor 2, half-row mirror, row
// mirror), then the four row results combined as uniform values
__device__ __attribute__((always_inline)) inline void lu_wave_best(double& bv, int& bp) {
  lu_dpp_best<0xB1>(bv, bp);
  lu_dpp_best<0x4E>(bv, bp);
  lu_dpp_best<0x141>(bv, bp);
  lu_dpp_best<0x140>(bv, bp);
  const int blo = __double2loint(bv), bhi = __double2hiint(bv);
  double v = __hiloint2double(__builtin_amdgcn_readlane(bhi, 0), __builtin_amdgcn_readlane(blo, 0));
  int p = __builtin_amdgcn_readlane(bp, 0);
#pragma unroll
  for (int r = 16; r < 64; r += 16) {
    const double ov = __hiloint2double(__builtin_amdgcn_readlane(bhi, r), __builtin_amdgcn_readlane(blo, r));
    const int op = __builtin_amdgcn_readlane(bp, r);
    if (better(ov, op, v, p)) { v = ov; p = op; }
  }
  bv = v;
  bp = p;
}

}  // namespace oba
