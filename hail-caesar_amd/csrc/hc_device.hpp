// Device helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hcdev {

__device__ __forceinline__ double dmax_ref(double a, double b) {
  return (a < b) ? b : a;  // std::max(a, b)
}

// Correctly rounded a*a*a: a^2 = p + e and p*a = q + f exactly (fma), so
// a^3 = q + (f + e*a) with the bracket far below half an ulp of q.
__device__ __forceinline__ double cube_rn(double a) {
  const double p = a * a;
  const double e = __builtin_fma(a, a, -p);
  const double q = p * a;
  const double f = __builtin_fma(p, a, -q);
  const double t = __builtin_fma(e, a, f);
  return q + t;
}

__device__ __forceinline__ bool mask_bit(const uint64_t* __restrict__ m, int W, int x, int y) {
  return (m[static_cast<size_t>(x) * W + (y >> 6)] >> (y & 63)) & 1ull;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const double w = __shfl_xor(v, o, 64);
    v = (v < w) ? w : v;
  }
  return v;
}

__device__ __forceinline__ void atomic_max_pos(unsigned long long* dst, double v) {
  // v >= 0: the bit patterns of non-negative doubles order like the values
  if (v > 0) atomicMax(dst, static_cast<unsigned long long>(__double_as_longlong(v)));
}

// Unique edge cells: row 1, row imax, then columns 1 and jmax of rows
// 2..imax-1.  Every cell drains once, as in the reference where a corner's
// second visit already finds it at the threshold.
__device__ __forceinline__ void edge_cell(int k, int imax, int jmax, int& x, int& y) {
  if (k < jmax) { x = 1; y = k + 1; return; }
  k -= jmax;
  if (imax > 1) {
    if (k < jmax) { x = imax; y = k + 1; return; }
    k -= jmax;
  }
  const int nside = (jmax > 1) ? 2 : 1;
  x = 2 + k / nside;
  y = (k % nside == 0) ? 1 : jmax;
}


}  // namespace hcdev
