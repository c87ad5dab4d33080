// Device helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hcdev {

// A pointer read from device memory (a table of buffers) is generic to the
// compiler, which then emits flat loads and stores; those also count against
// the LDS wait counter, so every cross-lane shuffle waits for them.  Reading
// the pointer value itself as a global-address-space pointer lets the
// compiler use global_* accesses through it.
template <typename T, typename F>
__device__ __forceinline__ T* table_ptr(const F& field) {
  static_assert(sizeof(F) == sizeof(uintptr_t), "a table field holds one pointer");
  using GP = __attribute__((address_space(1))) T*;
  const uintptr_t bits = *reinterpret_cast<const uintptr_t*>(&field);
  return (T*)(GP)bits;
}

// Element i of a field addressed by a 32-bit byte offset from its base, so
// the access is one global instruction with a scalar base and a vector
// offset (the launcher checks that every offset fits in 32 bits).
template <typename T>
__device__ __forceinline__ T ld_elem(const T* base, uint32_t i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                     i * static_cast<uint32_t>(sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void st_elem(T* base, uint32_t i, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * static_cast<uint32_t>(sizeof(T))) = v;
}

// The same with the nontemporal hint: data streamed through once per launch
// (the fused kernel's outputs, and the inputs no neighbouring tile reads
// again), kept from displacing the rows that are read twice.
template <typename T>
__device__ __forceinline__ T ld_nt(const T* base, uint32_t i) {
  return __builtin_nontemporal_load(reinterpret_cast<const T*>(
      reinterpret_cast<const char*>(base) + i * static_cast<uint32_t>(sizeof(T))));
}
template <typename T>
__device__ __forceinline__ void st_nt(T* base, uint32_t i, T v) {
  __builtin_nontemporal_store(
      v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * static_cast<uint32_t>(sizeof(T))));
}

// A value materialised in a register at this point of the program (the
// compiler may not recompute it at, or sink it into, each use after here).
__device__ __forceinline__ float pin_value(float v) {
  asm volatile("" : "+v"(v));
  return v;
}
__device__ __forceinline__ double pin_value(double v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Zero in a vector register, opaque to the compiler.  Added to the index of
// a load at a wave-uniform address it makes a vector-memory load: those
// complete in order, so waiting for one leaves younger loads in flight,
// whereas scalar loads complete out of order and a wait for any of them is
// a wait for all (which empties a prefetch queue).
__device__ __forceinline__ uint32_t opaque_vzero() {
  uint32_t z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

// Lane shifts by one across the whole wavefront (DPP wave_shr:1 / wave_shl:1):
// lane_up1 gives lane l the value of lane l-1 (lane 0 keeps its own),
// lane_down1 the value of lane l+1 (lane 63 keeps its own).
__device__ __forceinline__ int dpp_up1(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int dpp_down1(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ float lane_up1(float v) {
  return __int_as_float(dpp_up1(__float_as_int(v)));
}
__device__ __forceinline__ float lane_down1(float v) {
  return __int_as_float(dpp_down1(__float_as_int(v)));
}
__device__ __forceinline__ int32_t lane_up1(int32_t v) { return dpp_up1(v); }
__device__ __forceinline__ double lane_up1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_up1(static_cast<int>(b));
  const int hi = dpp_up1(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<uint32_t>(lo));
}
__device__ __forceinline__ double lane_down1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_down1(static_cast<int>(b));
  const int hi = dpp_down1(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<uint32_t>(lo));
}

// The same shifts with the vacated lane (lane 0 for lane_up1, lane 63 for
// lane_down1) taking `edge` instead: with bound_ctrl off, a lane whose source
// lies outside the wavefront keeps the instruction's old value.
__device__ __forceinline__ int dpp_up1(int v, int edge) {
  return __builtin_amdgcn_update_dpp(edge, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int dpp_down1(int v, int edge) {
  return __builtin_amdgcn_update_dpp(edge, v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ float lane_up1(float v, float edge) {
  return __int_as_float(dpp_up1(__float_as_int(v), __float_as_int(edge)));
}
__device__ __forceinline__ float lane_down1(float v, float edge) {
  return __int_as_float(dpp_down1(__float_as_int(v), __float_as_int(edge)));
}
__device__ __forceinline__ int32_t lane_up1(int32_t v, int32_t edge) { return dpp_up1(v, edge); }
__device__ __forceinline__ double lane_up1(double v, double edge) {
  const long long b = __double_as_longlong(v), e = __double_as_longlong(edge);
  const int lo = dpp_up1(static_cast<int>(b), static_cast<int>(e));
  const int hi = dpp_up1(static_cast<int>(b >> 32), static_cast<int>(e >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<uint32_t>(lo));
}
__device__ __forceinline__ double lane_down1(double v, double edge) {
  const long long b = __double_as_longlong(v), e = __double_as_longlong(edge);
  const int lo = dpp_down1(static_cast<int>(b), static_cast<int>(e));
  const int hi = dpp_down1(static_cast<int>(b >> 32), static_cast<int>(e >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<uint32_t>(lo));
}

// Lane k's value, uniform over the wavefront (v_readlane; k uniform).
__device__ __forceinline__ float lane_value(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
__device__ __forceinline__ int32_t lane_value(int32_t v, int k) {
  return __builtin_amdgcn_readlane(v, k);
}
__device__ __forceinline__ double lane_value(double v, int k) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), k);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), k);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<uint32_t>(lo));
}

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
  // v >= 0: the bit patterns of non-negative doubles order like the values.
  // Read first and only contend when this value would raise the maximum:
  // thousands of blocks hitting one address otherwise serialise.
  if (v > 0) {
    const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned long long cur = __atomic_load_n(dst, __ATOMIC_RELAXED);
    if (bits > cur) atomicMax(dst, bits);
  }
}

// Max over a 256-thread block (4 wavefronts) of non-negative values, then one
// conditional atomic per block.  Every thread of the block must call it.
__device__ __forceinline__ void block_max_pos(unsigned long long* dst, double v) {
  __shared__ double s_wave_max[4];
  v = wave_max(v);
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  if ((t & 63) == 0) s_wave_max[t >> 6] = v;
  __syncthreads();
  if (t == 0) {
    double m = s_wave_max[0];
    for (int k = 1; k < 4; ++k) m = (m < s_wave_max[k]) ? s_wave_max[k] : m;
    atomic_max_pos(dst, m);
  }
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


__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Rows per scan work item: one wavefront takes one 64-column mask word over
// a strip of rows.  With the row above and the row below, a strip is 64 rows
// of 4-byte depths or 32 of 8-byte ones: the side columns are then fetched
// one row per lane, and the strip's row loads all stay in flight together.
__host__ __device__ constexpr int scan_rows(int esz) { return esz == 4 ? 62 : 30; }

__host__ __device__ inline int scan_items(int imax, int W, int esz) {
  const int r = scan_rows(esz);
  return W * ((imax + r - 1) / r);
}

// scan_area (LSDCatchmentModel.cpp:3821-3849) for one work item: a cell is
// on the active mask when any cell of its 3x3 neighbourhood holds water.
// Each lane reads its column over the strip plus one row either side, ORs
// vertical triples, and the row of bits is dilated sideways by one column;
// the edge lanes also read the columns just outside the word.  Columns
// 1..jmax of rows 1..imax are written; rows 0 and imax+1 are cleared.
template <typename T>
__device__ __forceinline__ void scan_item(int item, int imax, int jmax, int P, int W,
                                          const T* __restrict__ h,
                                          uint64_t* __restrict__ mask) {
  constexpr int SR = scan_rows(static_cast<int>(sizeof(T)));
  static_assert(SR + 2 <= 64, "a strip's rows are indexed by lanes");
  const int lane = threadIdx.x & 63;
  item = __builtin_amdgcn_readfirstlane(item);  // one item per wavefront
  const int word = item % W;
  const int strip = item / W;
  const int x0 = 1 + strip * SR;
  const int c = word * 64 + lane;  // < P = 64 W
  T hv[SR + 2];
#pragma unroll
  for (int k = 0; k < SR + 2; ++k) {
    const int x = min(x0 - 1 + k, imax + 1);  // rows past the grid repeat the last
    // a uniform row base and the lane's column: one vector register serves
    // every row's address
    hv[k] = ld_elem(h + static_cast<size_t>(x) * P, static_cast<uint32_t>(c));
  }
  // The columns either side of the word, lane k taking row x0-1+k, as bit
  // masks over the strip's rows (bit k: that column is wet in row x0-1+k).
  // Word 0 has no column to its west, the last word none to its east.
  const uint32_t xk = static_cast<uint32_t>(min(x0 - 1 + min(lane, SR + 1), imax + 1));
  const int cw = word * 64 - 1, ce = word * 64 + 64;
  const T* hs = h + static_cast<size_t>(x0 - 1) * P;  // row x0-1 (the lanes' offsets from it)
  const uint32_t ok = (xk - static_cast<uint32_t>(x0 - 1)) * static_cast<uint32_t>(P);
  T hw = T(0), he = T(0);
  if (cw >= 0) hw = ld_elem(hs + cw, ok);  // uniform branches
  if (ce < P) he = ld_elem(hs + ce, ok);
  constexpr uint64_t rows = SR + 2 == 64 ? ~0ull : ((1ull << (SR + 2)) - 1);
  const uint64_t wside = __ballot(hw > T(0)) & rows;
  const uint64_t eside = __ballot(he > T(0)) & rows;
  const int lo = word * 64;
  uint64_t keep = ~0ull;
  if (lo == 0) keep &= ~1ull;
  const int hi = jmax + 1 - lo;  // first column past jmax, relative to the word
  if (hi <= 0) keep = 0;
  else if (hi < 64) keep &= (1ull << hi) - 1;
  // The 3x3 dilation on the vector unit: per lane the largest depth of rows
  // k-1..k+1 (wet when > 0; the largest of three is > 0 exactly when one of
  // them is, and a NaN counts as dry either way), then the same over the
  // lane and its two neighbours through lane shifts, the side columns
  // entering as the value shifted into lanes 0 and 63.  Row k's mask word is
  // kept in lane k (both halves) and the strip's words go out in one store.
  uint32_t out_lo = 0, out_hi = 0;
  int nrows = 0;
#pragma unroll
  for (int k = 1; k <= SR; ++k) {
    const int x = x0 - 1 + k;
    if (x > imax) break;  // uniform per wavefront
    const T vm = fmax(fmax(hv[k - 1], hv[k]), hv[k + 1]);
    const T ws = ((wside >> (k - 1)) & 7ull) ? T(1) : T(0);  // column lo-1 wet
    const T es = ((eside >> (k - 1)) & 7ull) ? T(1) : T(0);  // column lo+64 wet
    const T hm = fmax(fmax(vm, lane_up1(vm, ws)), lane_down1(vm, es));
    const uint64_t dil = __ballot(hm > T(0)) & keep;
    out_lo = lane == k ? static_cast<uint32_t>(dil) : out_lo;
    out_hi = lane == k ? static_cast<uint32_t>(dil >> 32) : out_hi;
    nrows = k;
  }
  if (lane >= 1 && lane <= nrows)
    mask[static_cast<size_t>(x0 - 1 + lane) * W + word] =
        (static_cast<uint64_t>(out_hi) << 32) | out_lo;
  if (lane == 0) {
    if (strip == 0) mask[word] = 0ull;
    if (x0 + SR > imax) mask[static_cast<size_t>(imax + 1) * W + word] = 0ull;
  }
}

// water_flux_out's edge drain (LSDCatchmentModel.cpp:3005-3047) for the
// block's share of the unique edge cells; returns the block sum in thread 0.
template <typename T>
__device__ __forceinline__ double flux_block(int imax, int jmax, int P, double DX, double thr,
                                             double ts, T* __restrict__ h) {
  const int nedge = jmax + (imax > 1 ? jmax : 0) + (imax > 2 ? (imax - 2) * (jmax > 1 ? 2 : 1) : 0);
  double acc = 0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nedge; k += gridDim.x * blockDim.x) {
    int x, y;
    edge_cell(k, imax, jmax, x, y);
    T* hp = &h[static_cast<size_t>(x) * P + y];
    const double hv = static_cast<double>(*hp);
    if (hv > thr) {
      // the depth left behind is thr as stored: in FP32 float(thr), so the
      // volume counted out is exactly the volume taken (thr itself in FP64)
      acc += (hv - static_cast<double>(static_cast<T>(thr))) * DX * DX / ts;
      *hp = static_cast<T>(thr);
    }
  }
  __shared__ double s[256];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  return s[0];
}

// Deterministic grid sum: every block deposits its partial; the last block to
// arrive adds them in a fixed tree order.  Returns true (in all threads of
// that block) with *total set; the counter is re-armed for the next launch.
__device__ __forceinline__ bool grid_sum_last(double block_sum, double* __restrict__ partials,
                                              unsigned int* __restrict__ count, double* total) {
  __shared__ bool s_last;
  __shared__ double s[256];
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = block_sum;
    __threadfence();
    const unsigned prev = atomicAdd(count, 1u);
    s_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  double v = 0;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
    v += __longlong_as_double(
        __atomic_load_n(reinterpret_cast<const long long*>(&partials[b]), __ATOMIC_RELAXED));
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  *total = s[0];
  if (threadIdx.x == 0) *count = 0;
  return true;
}

}  // namespace hcdev
