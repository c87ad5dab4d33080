// Erosion and deposition ("erodepo"): LSDCatchmentModel::erode
// (LSDCatchmentModel.cpp:4211-4845) with its grain helpers sort_active,
// addGS, sand_fraction, d50 and slide_GS (3874-4171).  FP64 arithmetic in
// the reference's operation order; the sediment planes are stored in FP64 or,
// in the FP32 storage form, in FP32 (elevations and bedrock stay FP64).
//
// Layout: dense per-cell grain (11 planes), strata (10 layers x 9 fractions
// = 90 planes, plane z*9+n), a has-grain flag standing for index != -9999
// with a creation key per cell that recovers the reference's row numbers
// (hc::row_key), and the directional bedload su/sr/sd/sl as 4 x 10 planes.
// Where a row is made, the kernel records the key of the reference's loop
// position that makes it: the entrainment pass by atomic minimum over the
// cells that would make it (the first in the reference's order does), the
// serial hillslope sweeps and the reach points as they make it, and the
// deposition and y-slide sweeps, which only make the row of the cell they
// visit, by a stamp pass after the sweep.
//
// Phases and their parallel structure:
//   entrainment (4236-4565)   per-cell map over the active mask, y < jmax;
//                             grain rows a cell allocates for a neighbour are
//                             flagged and initialised by a second pass
//                             (addGS is deterministic, so order is moot);
//                             tempbmax is a max-reduction; repeated on the
//                             host while tempbmax > ERODEFACTOR.
//   deposition + x slide      one wavefront per column y walks x in the
//   (4570-4682)               reference's order: every write stays in the
//                             column, and each cell sees its upper
//                             neighbour's slide, exactly as the serial loop.
//                             The cells around the sweep position are held
//                             in registers (see the sweep section below).
//   y slide (4684-4747)       one wavefront per row x walks y in order
//                             (writes stay in the row).
//   edge outputs (4750-4827)  nine lanes, one per fraction, accumulate in
//                             float in the reference order (from LDS chunks
//                             staged by the whole block): a zeroed private
//                             sum per loop, added to the total afterwards, as
//                             GCC's OpenMP array reduction does.
// Transcendentals (pow, exp, log, sqrt) are the device library's, so values
// agree with the libm reference to a few ulp rather than bitwise.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "hc_device.hpp"
#include "hc_internal.hpp"

using hc::ErodeArgs;
using hc::kGravity;
using namespace hcdev;

namespace {

constexpr int kG = 10;  // G_MAX

// E: storage of the grain planes and every other sediment plane (double,
// or float in the FP32 storage form); all arithmetic is FP64, rounded to E
// where a value is stored.
template <typename E>
struct GrainT {
  E* g;            // 11 planes
  E* s;            // 90 planes
  int32_t* has;
  size_t nc;       // plane stride (cells)
  uint64_t* key;   // creation key of each cell's row (hc::kRowNone: none)
  __device__ E& G(size_t c, int n) const { return g[n * nc + c]; }
  __device__ E& S(size_t c, int z, int n) const { return s[(z * 9 + n) * nc + c]; }
};

// A cell's grain row as the grain helpers use it: get(n), and set(n, x),
// which keeps x as the storage type holds it.  MemRow works on the planes in
// memory; RegRow holds the row in registers (values rounded to the storage
// type at each set, so a run on registers and one through memory agree bit
// for bit).  The strata stay in memory in both.
// Both also answer whether the cell has a row (index != -9999) and record
// that it now has one: MemRow through the flags in memory, RegRow from a
// register copy it keeps in step with them.
template <typename E>
struct MemRow {
  E* g;
  int32_t* hasp;
  size_t nc, c;
  __device__ double get(int n) const { return g[n * nc + c]; }
  __device__ void set(int n, double x) const { g[n * nc + c] = static_cast<E>(x); }
  __device__ bool has() const { return hasp[c] != 0; }
  __device__ void set_has() const { hasp[c] = 1; }
};
template <typename E>
struct RegRow {
  double v[kG + 1];
  int32_t h;  // the flag as loaded (tested where a slide needs it)
  int32_t* hasp;
  size_t c;
  __device__ double get(int n) const { return v[n]; }
  __device__ void set(int n, double x) { v[n] = static_cast<double>(static_cast<E>(x)); }
  __device__ bool has() const { return h != 0; }
  __device__ void set_has() {
    h = 1;
    hasp[c] = 1;
  }
};
template <typename E>
__device__ __forceinline__ MemRow<E> mem_row(const GrainT<E>& gr, size_t c) {
  return MemRow<E>{gr.g, gr.has, gr.nc, c};
}
// a cell's row and flag into registers
template <typename E>
__device__ __forceinline__ void reg_row_load(RegRow<E>& r, const GrainT<E>& gr, size_t c) {
  r.hasp = gr.has;
  r.c = c;
  r.h = gr.has[c];
#pragma unroll
  for (int n = 0; n <= kG; ++n) r.v[n] = gr.G(c, n);
}

// LCM.cpp:3874-3947
template <typename E, typename R>
__device__ __forceinline__ void sort_active_r(const GrainT<E>& gr, size_t c, R& row,
                                              const ErodeArgs& a) {
  double total = 0.0;
#pragma unroll
  for (int n = 0; n <= kG; ++n) total += row.get(n);
  // (the rare layer shifts stay loops: unrolled, their loads would all be
  // live at once)
  if (total > (a.active * 1.5)) {
#pragma unroll 1
    for (int z = 9; z >= 1; --z)
#pragma unroll 1
      for (int n = 0; n <= kG - 2; ++n) gr.S(c, z, n) = gr.S(c, z - 1, n);
    const double coeff = a.active / total;
#pragma unroll
    for (int n = 1; n <= kG - 1; ++n) {
      const double gv = row.get(n);
      if (gv > 0.0) {
        const double amount = coeff * (gv);
        gr.S(c, 0, n - 1) = amount;
        row.set(n, gv - amount);
      }
    }
  }
  if (total < (a.active / 4)) {
#pragma unroll
    for (int n = 1; n <= kG - 1; ++n)  // in the storage type, as G += S adds there
      row.set(n, static_cast<E>(row.get(n)) + gr.S(c, 0, n - 1));
#pragma unroll 1
    for (int z = 0; z <= 8; ++z)
#pragma unroll 1
      for (int n = 0; n <= kG - 2; ++n) gr.S(c, z, n) = gr.S(c, z + 1, n);
    const double amount = a.active;
#pragma unroll 1
    for (int n = 1; n <= kG - 1; ++n) gr.S(c, 9, n - 1) = amount * a.dprop[n];
  }
}

// LCM.cpp:3949-3983 (sort_active afterwards, as addGS does)
template <typename E, typename R>
__device__ __forceinline__ void add_gs_r(const GrainT<E>& gr, size_t c, R& row, double elev,
                                         double bedrock, const ErodeArgs& a) {
  row.set_has();
  row.set(0, 0);
#pragma unroll
  for (int n = 1; n <= kG - 1; ++n) row.set(n, a.active * a.dprop[n]);
  row.set(kG, 0);
#pragma unroll 1
  for (int z = 0; z <= 9; ++z) {
    const bool zero = elev - (a.active * (z + 1)) < (bedrock - a.active);
#pragma unroll 1
    for (int n2 = 0; n2 <= kG - 2; ++n2) gr.S(c, z, n2) = zero ? 0.0 : (a.active) * a.dprop[n2 + 1];
  }
  sort_active_r(gr, c, row, a);
}

// sort_active as called from outside addGS: allocates first if needed
template <typename E, typename R>
__device__ __forceinline__ void sort_active_alloc_r(const GrainT<E>& gr, size_t c, R& row,
                                                    const double* elev, const double* bedrock,
                                                    const ErodeArgs& a) {
  if (!row.has()) add_gs_r(gr, c, row, elev[c], bedrock[c], a);
  sort_active_r(gr, c, row, a);
}

template <typename E>
__device__ void sort_active(const GrainT<E>& gr, size_t c, const ErodeArgs& a) {
  MemRow<E> r = mem_row(gr, c);
  sort_active_r(gr, c, r, a);
}
template <typename E>
__device__ void add_gs(const GrainT<E>& gr, size_t c, double elev, double bedrock, const ErodeArgs& a) {
  MemRow<E> r = mem_row(gr, c);
  add_gs_r(gr, c, r, elev, bedrock, a);
}
template <typename E>
__device__ void sort_active_alloc(const GrainT<E>& gr, size_t c, const double* elev,
                                  const double* bedrock, const ErodeArgs& a) {
  MemRow<E> r = mem_row(gr, c);
  sort_active_alloc_r(gr, c, r, elev, bedrock, a);
}

// LCM.cpp:3985-4009
template <typename E>
__device__ double sand_fraction(const GrainT<E>& gr, size_t c) {
  double at = 0, st = 0;
  for (int n = 1; n <= kG; ++n) at += (gr.G(c, n));
  for (int n = 1; n <= 2; ++n) st += (gr.G(c, n));
  if (at < 0.0001) return 0.0;
  return st / at;
}

// LCM.cpp:4038-4075
template <typename E>
__device__ double d50(const GrainT<E>& gr, size_t c, const ErodeArgs& a) {
  double at = 0, mx = 0, mn = 0;
  double cum[kG + 1];
  cum[0] = 0;
#pragma unroll
  for (int n = 1; n <= kG; ++n) {
    at += (gr.G(c, n));
    cum[n] = at;
  }
  // the reference's scan, i = the first of 1..9 with cum[i] >= at / 2 (not
  // below it), else 10, and the values it picks, by selects over constant
  // indices (a run-time index would put cum in scratch memory)
  int i = 10;
#pragma unroll
  for (int q = 9; q >= 1; --q)
    if (!(cum[q] < (at * 0.5))) i = q;
  double ci = cum[10], cim1 = cum[9];
#pragma unroll
  for (int q = 1; q <= 9; ++q) {
    if (i == q) {
      ci = cum[q];
      cim1 = cum[q - 1];
      mn = a.logd[(q == 1) ? 1 : q - 1];
      mx = a.logd[q];
    }
  }
  double D = exp(mx - ((mx - mn) * ((ci - (at * 0.5)) / (ci - cim1))));
  if (at < 0.0000001) D = 0;
  return D;
}

// LCM.cpp:4077-4171, material moves from `src` to `dst`; a row made for
// dst gets the creation key dst_key
template <typename E, typename RS, typename RD>
__device__ __forceinline__ void slide_gs_r(const GrainT<E>& gr, size_t src, RS& rs, size_t dst,
                                           RD& rd, double amount, const double* elev,
                                           const double* bedrock, const ErodeArgs& a,
                                           uint64_t dst_key) {
  double total = 0;
  if (rs.has() && rd.has()) {
#pragma unroll
    for (int n = 1; n <= kG - 1; ++n)
      if (rs.get(n) > 0) total += rs.get(n);
    if (amount > total) {
#pragma unroll
      for (int n = 1; n <= kG - 1; ++n) rd.set(n, rd.get(n) + (amount - total) * a.dprop[n]);
      amount = total;
    }
    if (total > 0) {
#pragma unroll
      for (int n = 1; n <= kG - 1; ++n) {
        const double tr = amount * (rs.get(n) / total);
        rd.set(n, rd.get(n) + tr);
        rs.set(n, rs.get(n) - tr);
        if (rs.get(n) < 0) rs.set(n, 0);
      }
    }
    sort_active_alloc_r(gr, src, rs, elev, bedrock, a);
    sort_active_alloc_r(gr, dst, rd, elev, bedrock, a);
    return;
  }
  if (!rs.has() && rd.has()) {
#pragma unroll
    for (int n = 1; n <= kG - 1; ++n) rd.set(n, rd.get(n) + (amount) * a.dprop[n]);
    sort_active_alloc_r(gr, dst, rd, elev, bedrock, a);
    return;
  }
  if (rs.has() && !rd.has()) {
    add_gs_r(gr, dst, rd, elev[dst], bedrock[dst], a);
    gr.key[dst] = dst_key;
    if (amount > a.active) {
#pragma unroll
      for (int n = 1; n <= kG - 1; ++n) rd.set(n, rd.get(n) + amount * a.dprop[n]);
      amount = a.active;
    }
#pragma unroll
    for (int n = 1; n <= kG - 1; ++n)
      if (rs.get(n) > 0) total += rs.get(n);
#pragma unroll
    for (int n = 1; n <= kG - 1; ++n) {
      if (total > 0) {
        rd.set(n, rd.get(n) + amount * (rs.get(n) / total));
        if (rs.get(n) > 0.0001) rs.set(n, rs.get(n) - amount * (rs.get(n) / total));
        if (rs.get(n) < 0) rs.set(n, 0);
      }
    }
    sort_active_alloc_r(gr, src, rs, elev, bedrock, a);
    sort_active_alloc_r(gr, dst, rd, elev, bedrock, a);
    return;
  }
}

template <typename E>
__device__ void slide_gs(const GrainT<E>& gr, size_t src, size_t dst, double amount, const double* elev,
                         const double* bedrock, const ErodeArgs& a, uint64_t dst_key) {
  MemRow<E> rs = mem_row(gr, src), rd = mem_row(gr, dst);
  slide_gs_r(gr, src, rs, dst, rd, amount, elev, bedrock, a, dst_key);
}

// direction p (1,3,5,7) -> (dx, dy) of LCM.hpp:445-446 and bedload plane
__device__ __forceinline__ int dxp(int p) { return p == 3 ? 1 : (p == 7 ? -1 : 0); }
__device__ __forceinline__ int dyp(int p) { return p == 1 ? -1 : (p == 5 ? 1 : 0); }
// su (p=1) -> 0, sr (p=3) -> 1, sd (p=5) -> 2, sl (p=7) -> 3
__device__ __forceinline__ int sdir_of(int p) { return (p - 1) / 2; }

// -------------------------------------------------------- entrainment
// Built for three wavefronts per SIMD (165 registers; unconstrained it takes
// 245 and holds two); with d50's scan free of scratch memory, 268 against
// 317 us per call at 1024^2 in FP64, 251 against 300 in FP32.
template <typename E>
__global__ __launch_bounds__(256, 3) void k_entrain(ErodeArgs a, double time_step, double mult,
                                                 const E* __restrict__ h,
                                                 const double* __restrict__ elev,
                                                 double* __restrict__ bedrock,
                                                 const E* __restrict__ nman,
                                                 const uint64_t* __restrict__ mask,
                                                 const E* __restrict__ V,
                                                 const E* __restrict__ vd,
                                                 E* __restrict__ Vel, E* __restrict__ Tau,
                                                 E* __restrict__ ss, E* __restrict__ sdir,
                                                 GrainT<E> gr, int32_t* __restrict__ need,
                                                 unsigned long long* __restrict__ red,
                                                 uint32_t ep) {
  const int y = blockIdx.x * 64 + threadIdx.x;
  const int x = blockIdx.y * blockDim.y + threadIdx.y;
  double tbmax = 0;
  if (x >= 1 && x <= a.imax && y >= 1 && y < a.jmax && mask_bit(mask, a.mwords, x, y)) {
    const size_t P = a.pitch, nc = gr.nc;
    const size_t c = static_cast<size_t>(x) * P + y;
    Vel[c] = 0;
    Tau[c] = 0;
    for (int d = 0; d < 4; ++d)
      for (int n = 0; n < kG; ++n) sdir[(d * 10 + n) * nc + c] = 0;
    ss[c] = 0;
    const double hc = h[c];
    if (hc > a.wdet) {
      const double rho = 1000.0;
      double veltot = 0, vel = 0, qtot = 0, tau = 0, slopetot = 0;
      double tempdir[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      double temp_dist[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      const double mann = a.spatial_n ? static_cast<double>(nman[c]) : a.mannings;
      if (!gr.has[c]) {  // LCM.cpp:4290, at loop position (y, x)
        add_gs(gr, c, elev[c], bedrock[c], a);
        atomicMin(reinterpret_cast<unsigned long long*>(&gr.key[c]),
                  static_cast<unsigned long long>(hc::row_key(ep, y, x, 0)));
      }
      for (int p = 1; p <= 7; p += 2) {
        const size_t c2 = c + static_cast<long>(dxp(p)) * P + dyp(p);
        if (h[c2] > a.wdet) {
          const double vdp = vd[sdir_of(p) * nc + c];
          if (vdp > 0) {
            vel = vdp;
            if (vel > a.max_vel) vel = a.max_vel;
            tempdir[p] = vel * vel;
            veltot += tempdir[p];
            qtot += (vel * vel);
            slopetot += ((elev[c] - elev[c2]) / a.DX) * vel;
          }
        }
      }
      if (qtot > 0) {
        vel = (sqrt(qtot));
        Vel[c] = vel;
        if (vel > a.max_vel) vel = a.max_vel;
        const double ci = kGravity * (mann * mann) * pow(hc, -0.33);
        if (slopetot > 0) slopetot = 0;
        tau = 1000 * ci * vel * vel * (1 + (1 * (slopetot / vel)));
        Tau[c] = tau;
      }
      if (tau > 0) {
        double d_50 = 0, Fs = 0, graintot = 0;
        if (a.wilcock == 1) {
          d_50 = d50(gr, c, a);
          if (d_50 < a.d[1]) d_50 = a.d[1];
          Fs = sand_fraction(gr, c);
          for (int n = 1; n <= kG; ++n) graintot += (gr.G(c, n));
        }
        double temptot1 = 0;
        for (int n = 1; n <= kG - 1; ++n) {
          const double Di = a.d[n];
          const double gn = gr.G(c, n);
          if (a.wilcock == 1) {
            const double tau_ri = (0.021 + (0.015 * exp(-20 * Fs))) * (rho * kGravity * d_50) *
                                  pow((Di / d_50), (0.67 / (1 + exp(1.5 - (Di / d_50)))));
            const double U_star = pow(tau / rho, 0.5);
            const double Fi = gn / graintot;
            double Wi_star;
            if ((tau / tau_ri) < 1.35)
              Wi_star = 0.002 * pow(tau / tau_ri, 7.5);
            else
              Wi_star = 14 * pow(1 - (0.894 / pow(tau / tau_ri, 0.5)), 4.5);
            temp_dist[n] = mult * time_step *
                           ((Fi * (U_star * U_star * U_star)) / ((2.65 - 1) * kGravity)) * Wi_star / a.DX;
          }
          if (a.einstein == 1) {
            temp_dist[n] = mult * time_step *
                           (40 * pow((1 / (((2650 - 1000) * Di) / (tau / kGravity))), 3)) /
                           sqrt(1000 / ((2250 - 1000) * kGravity * (Di * Di * Di))) / a.DX;
          }
          if (temp_dist[n] > gn) temp_dist[n] = gn;
          if (n == 1) {
            if ((temp_dist[n] + V[c]) / hc > a.Csuspmax) temp_dist[n] = (hc * a.Csuspmax) - V[c];
          }
          if (temp_dist[n] < 0) temp_dist[n] = 0;
          temptot1 += temp_dist[n];
        }
        if (elev[c] - temptot1 <= bedrock[c]) {
          const double elevdiff = elev[c] - bedrock[c];
          const double temptot3 = temptot1;
          temptot1 = 0;
          for (int n = 1; n <= kG - 1; ++n) {
            if (elev[c] <= bedrock[c]) {
              temp_dist[n] = 0;
            } else {
              temp_dist[n] = elevdiff * (temp_dist[n] / temptot3);
              if (temp_dist[n] < 0) temp_dist[n] = 0;
            }
            temptot1 += temp_dist[n];
          }
          if (tau > a.bedrock_thr) {
            const double amount =
                pow(a.bedrock_rate * tau, 1.5) * time_step * mult * 0.000000317;
            bedrock[c] -= amount;
            for (int n2 = 1; n2 <= kG - 1; ++n2) gr.G(c, n2) += amount * a.dprop[n2];
          }
        }
        if (temptot1 > tbmax) tbmax = temptot1;
        if (temptot1 > 0) {
          for (int p = 1; p <= 7; p += 2) {
            const size_t c2 = c + static_cast<long>(dxp(p)) * P + dyp(p);
            if (h[c2] > a.wdet) {
              // LCM.cpp:4506: a neighbour without a row when the pass began
              // (its key is none, or of this pass) gets one; the earliest
              // position that asks for it numbers it
              if (gr.key[c2] >= hc::row_epoch_base(ep)) {
                need[c2] = 1;
                atomicMin(reinterpret_cast<unsigned long long*>(&gr.key[c2]),
                          static_cast<unsigned long long>(hc::row_key(ep, y, x, 1 + (p - 1) / 2)));
              }
              double factor = 0;
              if (vd[sdir_of(p) * nc + c] > 0) factor += 0.75 * tempdir[p] / veltot;
              for (int n = 1; n <= kG - 1; ++n) {
                if (temp_dist[n] > 0) {
                  if (n == 1)
                    ss[c] = temp_dist[n];
                  else
                    sdir[(sdir_of(p) * 10 + n) * nc + c] = temp_dist[n] * factor;
                }
              }
            }
          }
        }
      }
    }
  }
  block_max_pos(red, tbmax);
}

// addGS for the neighbours the entrainment pass flagged
template <typename E>
__global__ __launch_bounds__(256) void k_alloc(ErodeArgs a, const double* __restrict__ elev,
                                               const double* __restrict__ bedrock, GrainT<E> gr,
                                               int32_t* __restrict__ need, size_t n) {
  for (size_t c = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<size_t>(gridDim.x) * blockDim.x) {
    if (need[c]) {
      if (!gr.has[c]) add_gs(gr, c, elev[c], bedrock[c], a);
      need[c] = 0;
    }
  }
}

// Key of every row a sweep made (has a row, no key yet): the sweeps that use
// this make only the row of the cell they visit, at the reference's loop
// position (y, x).
__global__ __launch_bounds__(256) void k_row_stamp(int imax, int jmax, int P,
                                                   const int32_t* __restrict__ has,
                                                   uint64_t* __restrict__ key, uint32_t ep) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  const int x = blockIdx.y;
  if (y > jmax + 1 || x > imax + 1) return;
  const size_t c = static_cast<size_t>(x) * P + y;
  if (has[c] && key[c] == hc::kRowNone) key[c] = hc::row_key(ep, y, x, 0);
}

// ------------------------------------------ register-window sweeps
// The deposition and y-slide sweeps are sequential along a column (row) in
// the reference, so one wavefront walks one column (row).  Each cell's
// grain row, elevation, bedrock, depth and allocation flag are held in
// registers from one cell before the sweep reaches it until one cell after
// (the only cells a step reads or changes are the current one and its two
// neighbours along the sweep), so the per-cell work runs on registers; the
// next cell's inputs are issued one step ahead.  The strata (10 layers x 9
// fractions) stay in memory and are moved only on the rare layer shifts and
// allocations, spread over the lanes: element i = z*9+n lives in lane i % 64,
// slot i / 64, and every load and store of an element is made by that lane,
// so program order within the lane orders all accesses to it.  All other
// values are uniform: every lane computes them alike.

// A cell within one step of the sweep position.  Its grain row is held one
// fraction per lane (lane n holds fraction n, n <= 10; other lanes hold 0),
// so per-fraction work is one vector operation; sums over fractions are
// taken in the reference's order through v_readlane.  (No member leaves
// padding: copies of padding bytes keep the compiler from holding the cell
// in registers.)
struct CellW {
  double g;
  double elev, bedrock, h;
  uint64_t has;
  size_t c;
};

__device__ __forceinline__ double bcast(double v, int l) {
  const long long b = __double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | lo);
}

// Lane n's value in every lane of its 16-lane row (DPP row_newbcast:n, one
// 64-bit move); for n < 16 row 0 receives lane n itself.
template <int N>
__device__ __forceinline__ double row_bcast(double v) {
  static_assert(N >= 0 && N < 16, "row_newbcast selects a lane of the row");
  return __longlong_as_double(
      __builtin_amdgcn_mov_dpp(__double_as_longlong(v), 0x150 + N, 0xf, 0xf, false));
}
template <int N, int HI>
__device__ __forceinline__ double row_sum_from(double t, double v) {
  if constexpr (N > HI) {
    return t;
  } else {
    return row_sum_from<N + 1, HI>(t + row_bcast<N>(v), v);
  }
}
// Sum of lanes LO .. HI (< 16) of v, added in lane order from 0.0 (as the
// reference's loops over n add), the same in every lane: row 0's sum is
// taken through DPP broadcasts, one move and one add per lane, then made
// uniform from lane 0.
template <int LO, int HI>
__device__ __forceinline__ double lane_sum(double v) {
  const double t = row_sum_from<LO, HI>(0.0, v);
  return bcast(t, 0);
}

// A cell's state as fetched; becomes a CellW when the cell enters the
// window, so the fetch does not wait for its loads.
struct CellRaw {
  double gv, elev, bedrock, h;
  size_t c;
  int64_t has;
};
// (vz: opaque_vzero(), so that every load is a vector-memory load)
template <typename E>
__device__ __forceinline__ void cell_fetch(CellRaw& r, const GrainT<E>& gr, size_t c, const E* h,
                                           const double* elev, const double* bedrock, int lane,
                                           uint32_t vz) {
  r.c = c;
  const size_t cv = c + vz;
  r.gv = (lane <= kG) ? gr.g[static_cast<size_t>(lane) * gr.nc + c] : 0.0;
  r.elev = elev[cv];
  r.bedrock = bedrock[cv];
  r.h = h[cv];
  r.has = gr.has[cv];
}
__device__ __forceinline__ void cell_take(CellW& w, const CellRaw& r) {
  w.c = r.c;
  w.g = r.gv;
  w.elev = r.elev;
  w.bedrock = r.bedrock;
  w.h = r.h;
  w.has = r.has != 0 ? 1u : 0u;
}

// cell state back to memory (fraction n by lane n, the rest by lane 0)
template <typename E>
__device__ __forceinline__ void cell_store(const CellW& w, const GrainT<E>& gr, double* elev,
                                           int lane) {
  if (lane <= kG) gr.g[static_cast<size_t>(lane) * gr.nc + w.c] = w.g;
  if (lane == 0) {
    elev[w.c] = w.elev;
    gr.has[w.c] = w.has ? 1 : 0;
  }
}

// strata of a cell into the lanes (slot 0: element lane, slot 1: lane + 64)
template <typename E>
__device__ __forceinline__ void strata_load(const GrainT<E>& gr, size_t c, int lane, double& s0,
                                            double& s1) {
  s0 = gr.s[static_cast<size_t>(lane) * gr.nc + c];
  s1 = (lane + 64 < 90) ? gr.s[static_cast<size_t>(lane + 64) * gr.nc + c] : 0.0;
}
template <typename E>
__device__ __forceinline__ void strata_store(const GrainT<E>& gr, size_t c, int lane, double s0,
                                             double s1) {
  gr.s[static_cast<size_t>(lane) * gr.nc + c] = s0;
  if (lane + 64 < 90) gr.s[static_cast<size_t>(lane + 64) * gr.nc + c] = s1;
}
// new element i = old element i + d (i + d in [0, 90)), for both slots
__device__ __forceinline__ void strata_move(double& s0, double& s1, int lane, int d, int lo,
                                            int hi) {
  const int i0 = lane, i1 = lane + 64;
  const int j0 = min(max(i0 + d, 0), 89), j1 = min(max(i1 + d, 0), 89);
  const double a0 = __shfl(s0, j0 & 63, 64), b0 = __shfl(s1, j0 & 63, 64);
  const double a1 = __shfl(s0, j1 & 63, 64), b1 = __shfl(s1, j1 & 63, 64);
  const double n0 = j0 < 64 ? a0 : b0;
  const double n1 = j1 < 64 ? a1 : b1;
  if (i0 >= lo && i0 < hi) s0 = n0;
  if (i1 >= lo && i1 < hi) s1 = n1;
}

// Per-lane constants of the grain helpers: this lane's fraction, whether it
// is one of the moving fractions 1..9, its default proportion, and the
// proportion of fraction lane-16 (the lane that holds that fraction's
// bottom strata layer).
struct LaneG {
  double dprop;   // dprop[lane] for lanes 0..10, else 0
  double dprop9;  // dprop[lane - 16] for lanes 17..25, else 0
  double dz0;     // dprop[lane % 9 + 1]: a new strata element lane's default (slot 0)
  double dz1;     // dprop[(lane + 64) % 9 + 1] (slot 1)
  int lane;
  int frac;       // 1 <= lane <= 9
};
__device__ __forceinline__ LaneG make_lane(const ErodeArgs& a, int lane) {
  LaneG L;
  L.lane = lane;
  L.frac = (lane >= 1 && lane <= kG - 1) ? 1 : 0;
  double d = 0.0, d9 = 0.0, z0 = 0.0, z1 = 0.0;
  const int f0 = lane % 9 + 1, f1 = (lane + 64) % 9 + 1;
#pragma unroll
  for (int n = 0; n <= kG; ++n) {
    d = (lane == n) ? a.dprop[n] : d;
    if (n >= 1 && n <= kG - 1) {
      d9 = (lane == 16 + n) ? a.dprop[n] : d9;
      z0 = (f0 == n) ? a.dprop[n] : z0;
      z1 = (f1 == n) ? a.dprop[n] : z1;
    }
  }
  L.dprop = pin_value(d);
  L.dprop9 = pin_value(d9);
  L.dz0 = pin_value(z0);
  L.dz1 = pin_value(z1);
  return L;
}

// The sweeps' copy of the erosion constants: the floating-point ones held in
// vector registers for the whole sweep (scalar registers are scarce there,
// and a constant the compiler reloads from the argument block inside the
// chain costs a scalar-memory round trip per use).
__device__ __forceinline__ ErodeArgs sweep_args(const ErodeArgs& in) {
  ErodeArgs a = in;
  a.DX = pin_value(a.DX);
  a.wdet = pin_value(a.wdet);
  a.ERODEFACTOR = pin_value(a.ERODEFACTOR);
  a.active = pin_value(a.active);
  a.lateral_constant = pin_value(a.lateral_constant);
  a.chann_lateral = pin_value(a.chann_lateral);
  a.veg_lat = pin_value(a.veg_lat);
  return a;
}

// sort_active (LCM.cpp:3874-3947) on a register cell
template <typename E>
__device__ __forceinline__ void sort_active_w(const GrainT<E>& gr, CellW& w, const ErodeArgs& a,
                                              const LaneG& L) {
  const double total = lane_sum<0, kG>(w.g);
  if (total > (a.active * 1.5)) {
    double s0, s1;
    strata_load(gr, w.c, L.lane, s0, s1);
    strata_move(s0, s1, L.lane, -9, 9, 90);  // layers 9..1 take the layer above (z-1)
    const double coeff = a.active / total;
    const bool pos = L.frac && w.g > 0.0;
    const double amount = coeff * (w.g);
    // layer 0, fraction n-1 lives in lane n-1: it takes lane n's amount
    const double am_next = __shfl_down(amount, 1, 64);
    const bool pos_next = __shfl_down(pos ? 1 : 0, 1, 64) != 0;
    if (L.lane <= kG - 2 && pos_next) s0 = am_next;
    if (pos) w.g = w.g - amount;
    strata_store(gr, w.c, L.lane, s0, s1);
  }
  if (total < (a.active / 4)) {
    double s0, s1;
    strata_load(gr, w.c, L.lane, s0, s1);
    const double s_prev = __shfl_up(s0, 1, 64);  // element (0, n-1) for lane n
    if (L.frac) w.g += s_prev;
    strata_move(s0, s1, L.lane, 9, 0, 81);  // layers 0..8 take the layer below (z+1)
    // layer 9, fraction n-1: element 80 + n, lane 16 + n slot 1
    const int n9 = L.lane - 16;
    if (n9 >= 1 && n9 <= kG - 1) s1 = a.active * L.dprop9;
    strata_store(gr, w.c, L.lane, s0, s1);
  }
}

// addGS (LCM.cpp:3949-3983) on a register cell
template <typename E>
__device__ __forceinline__ void add_gs_w(const GrainT<E>& gr, CellW& w, const ErodeArgs& a,
                                         const LaneG& L) {
  w.has = 1;
  w.g = L.frac ? a.active * L.dprop : 0.0;  // fractions 0 and 10 start empty
  auto val = [&](int i) {
    const int z = i / 9;
    const bool zero = w.elev - (a.active * (z + 1)) < (w.bedrock - a.active);
    return zero ? 0.0 : (a.active) * (i < 64 ? L.dz0 : L.dz1);  // dprop[i % 9 + 1]
  };
  strata_store(gr, w.c, L.lane, val(L.lane), (L.lane + 64 < 90) ? val(L.lane + 64) : 0.0);
  sort_active_w(gr, w, a, L);
}

template <typename E>
__device__ __forceinline__ void sort_active_alloc_w(const GrainT<E>& gr, CellW& w, const ErodeArgs& a,
                                                    const LaneG& L) {
  if (!w.has) add_gs_w(gr, w, a, L);
  sort_active_w(gr, w, a, L);
}

// slide_GS (LCM.cpp:4077-4171) between register cells: `src` to `dst`, one
// fraction per lane (lanes 1..9)
template <typename E>
__device__ __forceinline__ void slide_gs_w(const GrainT<E>& gr, CellW& src, CellW& dst, double amount,
                                           const ErodeArgs& a, const LaneG& L) {
  const bool f = L.frac != 0;
  // sum over n of G(src, n) where positive, in n order (adding 0.0 elsewhere
  // leaves a non-negative running sum unchanged)
  auto pos_total = [&]() { return lane_sum<1, kG - 1>((f && src.g > 0) ? src.g : 0.0); };
  if (src.has && dst.has) {
    const double total = pos_total();
    if (amount > total) {
      if (f) dst.g += (amount - total) * L.dprop;
      amount = total;
    }
    if (total > 0 && f) {
      const double tr = amount * (src.g / total);
      dst.g += tr;
      src.g -= tr;
      if (src.g < 0) src.g = 0;
    }
    sort_active_alloc_w(gr, src, a, L);
    sort_active_alloc_w(gr, dst, a, L);
    return;
  }
  if (!src.has && dst.has) {
    if (f) dst.g += (amount) * L.dprop;
    sort_active_alloc_w(gr, dst, a, L);
    return;
  }
  if (src.has && !dst.has) {
    add_gs_w(gr, dst, a, L);
    if (amount > a.active) {
      if (f) dst.g += amount * L.dprop;
      amount = a.active;
    }
    const double total = pos_total();
    if (total > 0 && f) {
      dst.g += amount * (src.g / total);
      if (src.g > 0.0001) src.g -= amount * (src.g / total);
      if (src.g < 0) src.g = 0;
    }
    sort_active_alloc_w(gr, src, a, L);
    sort_active_alloc_w(gr, dst, a, L);
    return;
  }
}

// one lateral slide from neighbour `nb` into cell `w` (LCM.cpp:4619-4664,
// 4689-4742; veg and edge are zero in this build).  As in the reference, once
// the first amount is positive the slide is carried out even when the bedrock
// or domain-edge test has set it back to zero: slide_GS with a zero amount
// still clamps negative fractions and re-sorts both active layers.
template <typename E>
__device__ __forceinline__ void lateral_slide_w(const ErodeArgs& a, const GrainT<E>& gr, double mult,
                                                double time_step, double tau_c, double e3,
                                                CellW& nb, CellW& w, bool at_edge,
                                                double& elev_update, const LaneG& L) {
  double amt = 0;
  if (nb.h < a.wdet)
    amt = mult * a.lateral_constant * tau_c * 0.0 * time_step / a.DX;
  else
    amt = a.chann_lateral * e3 * (nb.elev - w.elev) / a.DX * 0.1;
  if (amt > 0) {
    amt *= 1 - (0.0 * (1 - a.veg_lat));
    if ((nb.elev - amt) < nb.bedrock || at_edge) amt = 0;
    if (amt > a.ERODEFACTOR * 0.1) amt = a.ERODEFACTOR * 0.1;
    elev_update += amt;
    nb.elev -= amt;
    slide_gs_w(gr, nb, w, amt, a, L);
  }
}

// ------------------------------------------------- deposition pre-pass
// The parts of the deposition loop (LCM.cpp:4578-4617) that no other cell's
// step changes, for every cell the sweep will visit (active, deeper than the
// erosion threshold, 1 < x < imax, 1 < y < jmax), in parallel: the
// suspended-sediment entrainment and settling (V is final here), each
// bedload fraction's val2 - val1, and erodetot and erodetot3 as the
// reference accumulates them.  The sweep then adds these to the grain rows
// and elevations in order and runs the ordered part (sort_active, the
// slides).  Planes of dp (FP64, the sweep's arithmetic): 0 erodetot,
// 1 erodetot3, 2 the settled amount, 3 + k fraction 2 + k's val2 - val1.
constexpr int kDepPlanes = 11;
template <typename E>
__global__ __launch_bounds__(256) void k_deposit_pre(ErodeArgs a, double time_step,
                                                     const E* __restrict__ h,
                                                     const uint64_t* __restrict__ mask,
                                                     E* __restrict__ V, const E* __restrict__ ss,
                                                     const E* __restrict__ sdir, size_t nc,
                                                     E* __restrict__ etot, E* __restrict__ etot3,
                                                     double* __restrict__ dp) {
  const int y = blockIdx.x * 64 + threadIdx.x;
  const int x = blockIdx.y * blockDim.y + threadIdx.y;
  if (x < 2 || x >= a.imax || y < 2 || y >= a.jmax) return;
  const size_t P = a.pitch;
  const size_t c = static_cast<size_t>(x) * P + y;
  const double hc = h[c];
  if (!mask_bit(mask, a.mwords, x, y) || !(hc > a.wdet)) return;
  double et = 0, et3 = 0;
  {  // fraction 1: suspended load
    const int n = 1;
    const double s1 = ss[c];
    const double v = static_cast<double>(V[c]) + s1;
    et -= s1;
    double coeff = (a.fall[n] * time_step) / hc;
    if (coeff > 1) coeff = 1;
    double drop = coeff * v;
    if (drop > 0.001) drop = 0.001;
    et += drop;
    V[c] = static_cast<E>(v - drop);
    dp[2 * nc + c] = drop;
  }
  auto sd = [&](int d, int n, size_t cc) { return static_cast<double>(sdir[(static_cast<size_t>(d) * 10 + n) * nc + cc]); };
#pragma unroll
  for (int n = 2; n <= kG - 1; ++n) {
    // su sr sd sl of the cell; su of y+1, sd of y-1, sl of x+1, sr of x-1
    const double val1 = ((sd(0, n, c) + sd(1, n, c)) + sd(2, n, c)) + sd(3, n, c);
    const double val2 = ((sd(0, n, c + 1) + sd(2, n, c - 1)) + sd(3, n, c + P)) + sd(1, n, c - P);
    const double dv = val2 - val1;
    et += dv;
    et3 += val1;
    dp[static_cast<size_t>(1 + n) * nc + c] = dv;
  }
  dp[c] = et;
  dp[nc + c] = et3;
  etot[c] = static_cast<E>(et);
  etot3[c] = static_cast<E>(et3);
}

// ------------------------------------------------- sweep input staging
// A sweep's per-step inputs are the records of the cells it walks, in walk
// order (record i: the sweep's cell i + 1).  They come from memory 64 at a
// time, one record per lane (64 independent loads per field in flight), and
// go through LDS: a ring of two 64-record chunks.  The chunk that the walk
// needs next is loaded into registers a whole chunk ahead and written to the
// ring as the walk enters it, so the steps themselves only read LDS.  The
// ring holds the records as they were when the sweep started; a cell's
// record is taken into the register window before any step changes the
// cell, and nothing outside this wavefront changes the cells of its walk.
constexpr int kStageSlots = 128;           // ring slots (two chunks)
constexpr int kStagePitch = kStageSlots + 1;  // LDS row pitch: rows read by lanes at once fall in
                                              // different banks

// ------------------------------------------- deposition + x-lateral slide
// One wavefront per column y, walking x = 2 .. imax-1 (LCM.cpp:4570-4682),
// on the pre-pass's per-cell terms: lane n of a step holds fraction n's
// addition (lane 1 the suspended fraction's entrainment and settling).
// Record i is cell x = i + 1 of the column: its grain row, elevation,
// bedrock, depth, row flag, active-mask bit, ss, Tau and pre-pass terms.
template <typename E>
__global__ __launch_bounds__(64) void k_deposit(ErodeArgs a_in, double time_step, double mult,
                                                const E* __restrict__ h,
                                                double* __restrict__ elev,
                                                const double* __restrict__ bedrock,
                                                const uint64_t* __restrict__ mask,
                                                const E* __restrict__ Tau,
                                                const E* __restrict__ ss, GrainT<E> gr,
                                                const double* __restrict__ dp) {
  const int y = 2 + blockIdx.x;
  if (y >= a_in.jmax || a_in.imax < 3) return;
  const ErodeArgs a = sweep_args(a_in);
  time_step = pin_value(time_step);
  mult = pin_value(mult);
  const int lane = threadIdx.x;
  const LaneG L = make_lane(a, lane);
  const size_t P = a.pitch, nc = gr.nc;
  constexpr int NV = 7;  // elevation, bedrock, depth, ss, Tau, erodetot, erodetot3
  __shared__ double s_g[kG + 1][kStagePitch];  // fractions 0..10
  __shared__ double s_t[kG - 1][kStagePitch];  // fraction n's pre-pass term (n = 1..9)
  __shared__ double s_v[NV][kStagePitch];
  __shared__ int s_f[kStagePitch];             // row flag | active bit << 1
  struct Chunk {  // (the flag and mask word are combined as the chunk is stored, so
                  // that nothing waits for a load the step after it is issued)
    double g[kG + 1], t[kG - 1], v[NV];
    uint64_t mw;
    int has;
  };
  auto load_chunk = [&](int m, Chunk& ch) {  // lane r: record 64 m + r
    const int x = min(64 * m + lane + 1, a.imax);
    const size_t c = static_cast<size_t>(x) * P + y;
#pragma unroll
    for (int n = 0; n <= kG; ++n) ch.g[n] = gr.g[static_cast<size_t>(n) * nc + c];
    // fraction 1: the settled amount (plane 2); fraction n >= 2: val2 - val1
    // (plane n + 1)
#pragma unroll
    for (int k = 0; k < kG - 1; ++k) ch.t[k] = dp[static_cast<size_t>(k == 0 ? 2 : k + 2) * nc + c];
    ch.v[0] = elev[c];
    ch.v[1] = bedrock[c];
    ch.v[2] = h[c];
    ch.v[3] = ss[c];
    ch.v[4] = Tau[c];
    ch.v[5] = dp[c];
    ch.v[6] = dp[nc + c];
    ch.mw = mask[static_cast<size_t>(x) * a.mwords + (y >> 6)];
    ch.has = gr.has[c];
  };
  auto store_chunk = [&](int m, const Chunk& ch) {
    const int sl = (64 * m + lane) & (kStageSlots - 1);
#pragma unroll
    for (int n = 0; n <= kG; ++n) s_g[n][sl] = ch.g[n];
#pragma unroll
    for (int k = 0; k < kG - 1; ++k) s_t[k][sl] = ch.t[k];
#pragma unroll
    for (int k = 0; k < NV; ++k) s_v[k][sl] = ch.v[k];
    s_f[sl] = (ch.has != 0 ? 1 : 0) | (static_cast<int>((ch.mw >> (y & 63)) & 1ull) << 1);
    __syncthreads();  // (one wavefront: orders the lanes' writes before the reads)
  };
  // the window cell of record i
  auto take = [&](CellW& w, int i) {
    const int sl = i & (kStageSlots - 1);
    const double gv = s_g[min(lane, kG)][sl];
    w.g = lane <= kG ? gv : 0.0;
    w.elev = s_v[0][sl];
    w.bedrock = s_v[1][sl];
    w.h = s_v[2][sl];
    w.has = (s_f[sl] & 1) ? 1u : 0u;
    w.c = static_cast<size_t>(i + 1) * P + y;
  };
  // a step's own inputs (record of cell x)
  const int tl = min(max(lane - 1, 0), kG - 2);  // lane n's term row (fractions 1..9)
  struct In {
    double h, ss, tau, et, et3, term;
    int f;
  };
  auto read_in = [&](int i, In& in) {
    const int sl = i & (kStageSlots - 1);
    in.f = s_f[sl];
    in.h = s_v[2][sl];
    in.ss = s_v[3][sl];
    in.tau = s_v[4][sl];
    in.et = s_v[5][sl];
    in.et3 = s_v[6][sl];
    in.term = s_t[tl][sl];
  };
  Chunk pend;
  load_chunk(0, pend);
  store_chunk(0, pend);
  load_chunk(1, pend);
  CellW wm, wc, wp;
  take(wm, 0);
  take(wc, 1);
  take(wp, 2);
  In in;
  read_in(1, in);  // cell 2's
  for (int x = 2; x < a.imax; ++x) {
    // the far cell of this step (x + 2, record x + 1) opens chunk m: write
    // the chunk loaded a chunk ago and start loading the next
    if (((x + 1) & 63) == 0) {  // uniform
      const int m = (x + 1) >> 6;
      store_chunk(m, pend);
      load_chunk(m + 1, pend);
    }
    // LDS reads for later, issued first so that the step's arithmetic
    // covers their latency: the far cell (taken at the end of this step) and
    // the next step's inputs
    CellW far;
    take(far, x + 1);
    In nin;
    read_in(x, nin);
    if ((in.f & 2) && in.h > a.wdet) {
      if (!wc.has) add_gs_w(gr, wc, a, L);
      // fraction 1: entrained, then settled; fractions 2..9: val2 - val1
      if (lane == 1) wc.g = (wc.g - in.ss) + in.term;
      if (lane >= 2 && lane <= kG - 1) wc.g += in.term;
      const double et = in.et, et3 = in.et3;
      wc.elev += et;
      if (et < 0) sort_active_alloc_w(gr, wc, a, L);
      if (et3 > 0) {
        const double tau = in.tau;
        double elev_update = 0;
        // (the neighbours go through the slide by value: two calls on
        // references to different cells are merged by the compiler into one
        // call on a selected address, which puts the cells in scratch memory)
        if (wm.elev > wc.elev && x > 2) {
          CellW nb = wm;
          lateral_slide_w(a, gr, mult, time_step, tau, et3, nb, wc, x - 1 == 1, elev_update, L);
          wm = nb;
        }
        if (wp.elev > wc.elev && x < a.imax - 1) {
          CellW nb = wp;
          lateral_slide_w(a, gr, mult, time_step, tau, et3, nb, wc, x + 1 == a.imax, elev_update,
                          L);
          wp = nb;
        }
        wc.elev += elev_update;
      }
    }
    cell_store(wm, gr, elev, lane);  // row x-1 is final
    wm = wc;
    wc = wp;
    if (x + 2 <= a.imax) wp = far;  // cell x + 2
    in = nin;
  }
  // rows imax-1 and imax (the last two of the window)
  cell_store(wm, gr, elev, lane);
  if (a.imax >= 3) cell_store(wc, gr, elev, lane);
}

// --------------------------------------------------------- y-lateral slide
// One wavefront per row x, walking y = 2 .. jmax-1 (LCM.cpp:4684-4747), its
// records (cell y = i + 1: grain row, elevation, bedrock, depth, row flag,
// active-mask bit, Tau, erodetot3) staged as in the deposition sweep.
template <typename E>
__global__ __launch_bounds__(64) void k_lateral_y(ErodeArgs a_in, double time_step, double mult,
                                                  const E* __restrict__ h,
                                                  double* __restrict__ elev,
                                                  const double* __restrict__ bedrock,
                                                  const uint64_t* __restrict__ mask,
                                                  const E* __restrict__ Tau, GrainT<E> gr,
                                                  const E* __restrict__ etot3) {
  const int x = 1 + blockIdx.x;
  if (x > a_in.imax || a_in.jmax < 3) return;
  const ErodeArgs a = sweep_args(a_in);
  time_step = pin_value(time_step);
  mult = pin_value(mult);
  const int lane = threadIdx.x;
  const LaneG L = make_lane(a, lane);
  const size_t row = static_cast<size_t>(x) * a.pitch;
  const size_t nc = gr.nc;
  constexpr int NV = 5;  // elevation, bedrock, depth, Tau, erodetot3
  __shared__ double s_g[kG + 1][kStagePitch];
  __shared__ double s_v[NV][kStagePitch];
  __shared__ int s_f[kStagePitch];  // row flag | active bit << 1
  struct Chunk {  // (flag and mask word combined at the store, as in k_deposit)
    double g[kG + 1], v[NV];
    uint64_t mw;
    int has;
  };
  auto load_chunk = [&](int m, Chunk& ch) {  // lane r: record 64 m + r
    const int y = min(64 * m + lane + 1, a.jmax);
    const size_t c = row + y;
#pragma unroll
    for (int n = 0; n <= kG; ++n) ch.g[n] = gr.g[static_cast<size_t>(n) * nc + c];
    ch.v[0] = elev[c];
    ch.v[1] = bedrock[c];
    ch.v[2] = h[c];
    ch.v[3] = Tau[c];
    ch.v[4] = etot3[c];
    ch.mw = mask[static_cast<size_t>(x) * a.mwords + (y >> 6)];
    ch.has = gr.has[c];
  };
  auto store_chunk = [&](int m, const Chunk& ch) {
    const int sl = (64 * m + lane) & (kStageSlots - 1);
#pragma unroll
    for (int n = 0; n <= kG; ++n) s_g[n][sl] = ch.g[n];
#pragma unroll
    for (int k = 0; k < NV; ++k) s_v[k][sl] = ch.v[k];
    const int yc = min(64 * m + lane + 1, a.jmax);  // the record's column
    s_f[sl] = (ch.has != 0 ? 1 : 0) | (static_cast<int>((ch.mw >> (yc & 63)) & 1ull) << 1);
    __syncthreads();
  };
  auto take = [&](CellW& w, int i) {
    const int sl = i & (kStageSlots - 1);
    const double gv = s_g[min(lane, kG)][sl];
    w.g = lane <= kG ? gv : 0.0;
    w.elev = s_v[0][sl];
    w.bedrock = s_v[1][sl];
    w.h = s_v[2][sl];
    w.has = (s_f[sl] & 1) ? 1u : 0u;
    w.c = row + static_cast<size_t>(i + 1);
  };
  struct In {
    double tau, e3;
    int f;
  };
  auto read_in = [&](int i, In& in) {
    const int sl = i & (kStageSlots - 1);
    in.f = s_f[sl];
    in.tau = s_v[3][sl];
    in.e3 = s_v[4][sl];
  };
  Chunk pend;
  load_chunk(0, pend);
  store_chunk(0, pend);
  load_chunk(1, pend);
  CellW wl, wc, wr;
  take(wl, 0);
  take(wc, 1);
  take(wr, 2);
  In in;
  read_in(1, in);
  for (int y = 2; y < a.jmax; ++y) {
    if (((y + 1) & 63) == 0) {  // uniform: the far record opens chunk m
      const int m = (y + 1) >> 6;
      store_chunk(m, pend);
      load_chunk(m + 1, pend);
    }
    CellW far;  // LDS reads for later first (k_deposit)
    take(far, y + 1);
    In nin;
    read_in(y, nin);
    const double e3 = in.e3;
    if (e3 > 0 && (in.f & 2)) {
      const double tau = in.tau;
      double elev_update = 0;
      if (wl.elev > wc.elev) {  // (by value: as in the deposition sweep)
        CellW nb = wl;
        lateral_slide_w(a, gr, mult, time_step, tau, e3, nb, wc, y - 1 == 1, elev_update, L);
        wl = nb;
      }
      if (wr.elev > wc.elev) {
        CellW nb = wr;
        lateral_slide_w(a, gr, mult, time_step, tau, e3, nb, wc, y + 1 == a.jmax, elev_update, L);
        wr = nb;
      }
      wc.elev += elev_update;
    }
    cell_store(wl, gr, elev, lane);  // column y-1 is final
    wl = wc;
    wc = wr;
    if (y + 2 <= a.jmax) wr = far;  // column y + 2
    in = nin;
  }
  cell_store(wl, gr, elev, lane);
  if (a.jmax >= 3) cell_store(wc, gr, elev, lane);
}

// ------------------------------------------------------------ edge outputs
// The reference's two loops visit the edge cells in a fixed order (entry
// e: y or x = 2 + e/2, the far edge for even e, the near edge for odd e),
// and each fraction's float total must add them in that order.  The two
// loops' totals are independent (added together at the end): wavefront 0
// sums the rows' entries and wavefront 1 the columns'.  Chunks of entries
// are staged in LDS by the whole block (an entry's V, or 0 where the
// reference skips it, and its eight bedload terms), then lane n of the
// wavefront adds fraction n from LDS in entry order.  A skipped entry adds 0,
// which leaves a float total unchanged (it starts at +0 and 0 + -0 = +0), so
// the sum runs without branches.  The edge cells are distinct and each is
// read before it is cleared, so the clearing follows in a last pass.
template <typename E>
__global__ __launch_bounds__(512) void k_edges(ErodeArgs a, const E* __restrict__ h,
                                                E* __restrict__ V,
                                                const E* __restrict__ sdir, size_t nc,
                                                float* __restrict__ gtot) {
  constexpr int C = 256;  // entries of one loop per chunk (blockDim = 2C)
  __shared__ double s_val[2][kG - 1][C];  // [loop][fraction n - 1][entry]
  const int t = threadIdx.x;
  const int I = a.imax, J = a.jmax;
  const size_t P = a.pitch;
  const int ne[2] = {2 * max(J - 2, 0), 2 * max(I - 2, 0)};
  const int wv = t >> 6, ln = t & 63;
  const bool sums = wv < 2 && ln >= 1 && ln <= kG - 1;  // lane ln of wavefront wv adds fraction ln
  float acc = 0.0f;
  const int nchunk = (max(ne[0], ne[1]) + C - 1) / C;
  for (int k = 0; k < nchunk; ++k) {
    const int ph = t / C, e = k * C + (t % C);  // threads 0..C-1 stage loop 0, C..2C-1 loop 1
    if (e < ne[ph]) {
      size_t c, src;
      int d;
      const int i = 2 + e / 2;
      if (ph == 0) {  // rows imax (sr of the row above) and 1 (sl of the row below)
        c = (e % 2 == 0) ? static_cast<size_t>(I) * P + i : P + i;
        src = (e % 2 == 0) ? c - P : c + P;
        d = (e % 2 == 0) ? 1 : 3;
      } else {  // columns jmax (sd of the column left) and 1 (su of the column right)
        c = (e % 2 == 0) ? static_cast<size_t>(i) * P + J : static_cast<size_t>(i) * P + 1;
        src = (e % 2 == 0) ? c - 1 : c + 1;
        d = (e % 2 == 0) ? 2 : 0;
      }
      const double v = V[c];
      const bool cond = h[c] > a.wdet || v > 0;
      s_val[ph][0][t % C] = cond ? v : 0.0;
#pragma unroll
      for (int n = 2; n <= kG - 1; ++n)
        s_val[ph][n - 1][t % C] = cond ? static_cast<double>(sdir[(d * 10 + n) * nc + src]) : 0.0;
    }
    __syncthreads();
    if (sums) {
      const int cnt = min(C, ne[wv] - k * C);
      const double* sv = s_val[wv][ln - 1];
      int q = 0;
      for (; q + 8 <= cnt; q += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = static_cast<float>(static_cast<double>(acc) + sv[q + u]);
      }
      for (; q < cnt; ++q) acc = static_cast<float>(static_cast<double>(acc) + sv[q]);
    }
    __syncthreads();  // the chunk's LDS is rewritten by the next one
  }
  __shared__ float s_part[2][kG];
  if (sums) s_part[wv][ln] = acc;
  for (int y = 2 + t; y < J; y += blockDim.x) {
    const size_t cA = static_cast<size_t>(I) * P + y, cB = P + y;
    if (h[cA] > a.wdet || V[cA] > 0) V[cA] = 0;
    if (h[cB] > a.wdet || V[cB] > 0) V[cB] = 0;
  }
  for (int x = 2 + t; x < I; x += blockDim.x) {
    const size_t cA = static_cast<size_t>(x) * P + J, cB = static_cast<size_t>(x) * P + 1;
    if (h[cA] > a.wdet || V[cA] > 0) V[cA] = 0;
    if (h[cB] > a.wdet || V[cB] > 0) V[cB] = 0;
  }
  __syncthreads();
  if (t < 20) gtot[t] = (t >= 1 && t <= kG - 1) ? (s_part[0][t] + s_part[1][t]) : 0.0f;
}

ErodeArgs erode_args(const hc_ctx* c) {
  ErodeArgs a;
  std::memset(&a, 0, sizeof(a));
  a.imax = c->g.imax;
  a.jmax = c->g.jmax;
  a.pitch = c->P;
  a.mwords = c->W;
  a.DX = c->g.DX;
  a.wdet = c->prm.water_depth_erosion_threshold;
  a.ERODEFACTOR = c->prm.ERODEFACTOR;
  a.active = c->prm.active;
  a.Csuspmax = c->prm.Csuspmax;
  a.max_vel = c->prm.max_vel;
  a.mannings = c->prm.mannings;
  a.spatial_n = c->prm.spatially_var_mannings;
  a.wilcock = c->prm.transport_law == HC_WILCOCK ? 1 : 0;
  a.einstein = c->prm.transport_law == HC_EINSTEIN ? 1 : 0;
  for (int k = 0; k < 10; ++k) {
    a.d[k] = c->prm.d[k];
    a.logd[k] = (k >= 1) ? std::log(c->prm.d[k]) : 0.0;  // host libm, as the reference
  }
  for (int k = 0; k < 11; ++k) {
    a.dprop[k] = c->prm.dprop[k];
    a.fall[k] = c->prm.fallVelocity[k];
  }
  a.bedrock_thr = c->prm.bedrock_erosion_threshold;
  a.bedrock_rate = c->prm.bedrock_erosion_rate;
  a.lateral_constant = c->prm.lateral_constant;
  a.chann_lateral = c->prm.chann_lateral_erosion;
  a.veg_lat = c->prm.veg_lat_restriction;
  return a;
}

// reach_water_and_sediment_input's updates (LCM.cpp:3301-3452) at the few
// reach points, by one thread in the reference's point order: with
// sediment, a point wetter than the threshold gets a grain row if it has
// none (addGS) and each fraction's amount -- fraction 1 into Vsusptot, the
// others into the grain row and the elevation, in fraction order; then
// every point takes its water.  The amounts are formed on the host.
template <class T>
__global__ void k_reach(ErodeArgs a, GrainT<T> gr, int n, const int* __restrict__ idx,
                        const double* __restrict__ sed, const double* __restrict__ water,
                        double thr, T* __restrict__ h, double* __restrict__ elev,
                        const double* __restrict__ bedrock, T* __restrict__ V, uint32_t ep) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (sed) {
    for (int k = 0; k < n; ++k) {
      const size_t c = static_cast<size_t>(idx[k]);
      if (!(static_cast<double>(h[c]) > thr)) continue;
      if (!gr.has[c]) {  // LCM.cpp:3333, point k of the list
        add_gs(gr, c, elev[c], bedrock[c], a);
        gr.key[c] = hc::row_key(ep, static_cast<uint32_t>(k) >> 15, k & 0x7fff, 0);
      }
      for (int f = 1; f <= kG - 1; ++f) {
        const double amt = sed[static_cast<size_t>(k) * (kG - 1) + (f - 1)];
        if (f == 1) {
          V[c] += amt;
        } else {
          gr.G(c, f) += amt;
          elev[c] += amt;
        }
      }
    }
  }
  for (int k = 0; k < n; ++k) {
    const size_t c = static_cast<size_t>(idx[k]);
    h[c] = static_cast<T>(static_cast<double>(h[c]) + water[k]);
  }
}

// Whether any cell of rows 2..xmax, columns 2..jmax-1 would move material into a neighbour under the
// present elevations: the drop to the neighbour (over 1.41 on the
// diagonals) above the cell's threshold (wet_low where the cell is at or
// below bedrock + active, factor elsewhere), the neighbour above -9999 --
// the test both landsliding sweeps apply.  The sweeps change elevations only
// by such moves, so where no cell passes it a sweep in any order moves
// nothing: *any stays 0 and the sweep kernels return at once.  Neighbour
// order does not matter here.
__global__ __launch_bounds__(256) void k_slide_candidates(int xmax, int jmax, int P, double factor,
                                                          double wet_low, double active,
                                                          const double* __restrict__ elev,
                                                          const double* __restrict__ bedrock,
                                                          unsigned long long* __restrict__ any) {
  const int y = 2 + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int x = 2 + static_cast<int>(blockIdx.y);
  if (y > jmax - 1 || x > xmax) return;
  const size_t c = static_cast<size_t>(x) * P + y;
  const double e = elev[c];
  const double wet = (e <= bedrock[c] + active) ? wet_low : factor;
  bool found = false;
#pragma unroll
  for (int dx = -1; dx <= 1; ++dx) {
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      if (dx == 0 && dy == 0) continue;
      const double en = elev[static_cast<size_t>(static_cast<long>(c) + dx * P + dy)];
      const double drop = (dx != 0 && dy != 0) ? (e - en) / 1.41 : (e - en);
      if (drop > wet && en > -9999) found = true;
    }
  }
  if (found) *any = 1ull;  // every writer stores the same value
}

// global_landsliding (LCM.cpp:5529-5654).  The reference sweeps the
// interior row-major (x = 2..imax-1 outer, y = 2..jmax-1 inner); a cell
// reads its 8 neighbours and moves material into them in place, so two cells
// interact only within Chebyshev distance 2.  Cells on one wavefront
// t = 3x + y are at least 3 columns apart, and every cell within distance 2
// that precedes (x, y) in row-major order lies on a smaller t: visiting the
// wavefronts in order, the cells of each in parallel, reproduces the
// reference's sweep exactly.  One block walks them with a barrier between
// wavefronts.  Each applied move is kept in its cell's neighbour slot (0
// where nothing moved), so that a sweep whose moves are not all positive (a
// cell held at bedrock + active moves negative amounts) sums its total in
// the reference's order; otherwise the total is positive exactly when some
// move is.  Sweeps repeat while the total is positive, at most 200.
__global__ __launch_bounds__(1024) void k_global_landslide(int imax, int jmax, int P, double DX,
                                                           double factor, double active,
                                                           double EF, double* __restrict__ elev,
                                                           const double* __restrict__ bedrock,
                                                           double* __restrict__ moves,
                                                           size_t plane, int* __restrict__ sweeps,
                                                           const unsigned long long* __restrict__ any) {
  // neighbour order of the reference's checks
  constexpr int ox[8] = {1, 0, -1, -1, -1, 0, 1, 1};
  constexpr int oy[8] = {1, 1, 1, 0, -1, -1, -1, 0};
  __shared__ int s_pos, s_neg, s_again;
  const int tid = threadIdx.x;
  const int t0 = 3 * 2 + 2, t1 = 3 * (imax - 1) + (jmax - 1);
  if (*any == 0) {  // no cell can move: one sweep, total 0
    if (tid == 0) *sweeps = 1;
    return;
  }
  int inc = 0;
  for (;;) {
    ++inc;
    if (tid == 0) {
      s_pos = 0;
      s_neg = 0;
    }
    __syncthreads();
    int pos = 0, neg = 0;
    for (int t = t0; t <= t1; ++t) {
      const int xlo = max(2, (t - (jmax - 1) + 2) / 3);  // y = t - 3x <= jmax - 1
      const int xhi = min(imax - 1, (t - 2) / 3);        // y >= 2
      for (int x = xlo + tid; x <= xhi; x += blockDim.x) {
        const int y = t - 3 * x;
        const size_t c = static_cast<size_t>(x) * P + y;
        double e = elev[c];
        const double bpa = bedrock[c] + active;
        const double wet = (e <= bpa) ? 10 * DX : factor;
        // no other cell of this wavefront touches the 3x3 block: its
        // elevations are read at once and written at each move
        double enb[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) enb[k] = elev[static_cast<size_t>(static_cast<long>(c) + ox[k] * P + oy[k])];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const size_t nb = static_cast<size_t>(static_cast<long>(c) + ox[k] * P + oy[k]);
          const double en = enb[k];
          const double drop = (ox[k] != 0 && oy[k] != 0) ? (e - en) / 1.41 : (e - en);
          double d = 0.0;
          if (drop > wet && en > -9999) {
            d = drop - wet;
            if ((e - d) < bpa) d = e - bpa;
            if (d > EF) d = EF;
            e -= d;
            elev[nb] = en + d;
            if (d > 0) pos = 1;
            else neg = 1;
          }
          moves[k * plane + c] = d;
        }
        elev[c] = e;
      }
      __syncthreads();  // this wavefront's moves are seen by the next
    }
    if (pos) atomicOr(&s_pos, 1);
    if (neg) atomicOr(&s_neg, 1);
    __syncthreads();
    if (tid == 0) {
      double total = s_pos ? 1.0 : 0.0;
      if (s_neg) {  // the reference's sum, in its order
        total = 0;
        for (int x = 2; x < imax; ++x)
          for (int y = 2; y < jmax; ++y)
            for (int k = 0; k < 8; ++k) total += moves[k * plane + static_cast<size_t>(x) * P + y];
      }
      s_again = (total > 0 && inc < 200) ? 1 : 0;
      if (!s_again) *sweeps = inc;
    }
    __syncthreads();
    if (!s_again) break;
  }
}

// creep (LCM.cpp:5293-5427).  Every interior cell above its bedrock moves
// ((z - z_nb) / D) * rate * time / D to each lower neighbour (D = DX on the
// sides, the reference's constant root = 7.07 on the diagonals, LCM.hpp:437),
// cut at its bedrock and at 0; the moves collect in tempcreep from the
// unchanged elevations and are added at the end, and a cell with grain
// slides the moved grain with each move, in the sweep's (row-major) order.
// Neighbour order of the reference's moves:
__device__ constexpr int kCreepOx[8] = {0, 1, 1, 1, 0, -1, -1, -1};
__device__ constexpr int kCreepOy[8] = {-1, -1, 0, 1, 1, 1, 0, -1};
// the same offsets for a direction known only at run time, from arithmetic
// rather than a table load: oy(k) = -ox(k + 2 mod 8)
__device__ __forceinline__ int creep_ox(int k) { return (k & 3) == 0 ? 0 : (k < 4 ? 1 : -1); }
__device__ __forceinline__ int creep_oy(int k) { return -creep_ox((k + 2) & 7); }

// The grain slides: the reference's row-major sweep as wavefronts t = 3x + y
// (a slide touches the cell and one neighbour, so sources interact within
// Chebyshev distance 2), a grid-wide barrier per wavefront.  Runs before
// the elevations change (the slides' grain allocation reads them); no other
// cell of a wavefront touches a cell's 3x3 block.  A cell's row, its
// neighbours' elevations and all eight neighbours' rows are fetched in one
// batch (the neighbours' rows staged in LDS, in the thread's own column);
// the moves then run from registers and LDS and store each neighbour's row
// once.  A wavefront's cells are one long instruction stream each, so they
// are spread one per lane over blocks of two wavefronts (one block per
// compute unit: the staging takes most of its LDS), as many as the device
// holds at once, so that every block is resident for the barriers.
constexpr int kCreepThreads = 128;

// A barrier over a grid whose blocks are all resident at once (the launch is
// sized to what the device holds): each block adds one arrival per pass to
// *arrive and waits for target = passes * blocks.  Every wavefront releases
// its stores at device scope before it and acquires after it, so blocks on
// different XCDs see each other's rows.  A wait that outlasts
// kGridBarrierPolls polls (seconds) sets *err, and no barrier waits once it
// is set, so a grid that could not be made resident ends with an error
// rather than a hang.
constexpr unsigned kGridBarrierPolls = 1u << 22;
__device__ __forceinline__ void grid_barrier(unsigned long long* arrive, unsigned long long target,
                                             unsigned long long* err) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(arrive, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned polls = 0;
    while (__hip_atomic_load(arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target &&
           __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
      if (++polls > kGridBarrierPolls) {
        __hip_atomic_fetch_or(err, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

template <typename E>
__global__ __launch_bounds__(kCreepThreads) void k_creep_grain(ErodeArgs a, int imax, int jmax, int P,
                                                               double DX, double rate, double time,
                                                               const double* __restrict__ elev,
                                                               const double* __restrict__ bedrock,
                                                               GrainT<E> gr, uint32_t ep,
                                                               unsigned long long* __restrict__ bar) {
  __shared__ E s_row[8 * (kG + 1)][kCreepThreads];
  __shared__ int32_t s_has[8][kCreepThreads];
  __shared__ double s_en[8][kCreepThreads];
  unsigned long long target = 0;
  const int tl = static_cast<int>(threadIdx.x);
  const int gtid = static_cast<int>(blockIdx.x) * kCreepThreads + tl;
  const int nthreads = static_cast<int>(gridDim.x) * kCreepThreads;
  const int t0 = 3 * 2 + 2, t1 = 3 * (imax - 1) + (jmax - 1);
  for (int t = t0; t <= t1; ++t) {
    const int xlo = max(2, (t - (jmax - 1) + 2) / 3);
    const int xhi = min(imax - 1, (t - 2) / 3);
    for (int x = xlo + gtid; x <= xhi; x += nthreads) {
      const int y = t - 3 * x;
      const size_t c = static_cast<size_t>(x) * P + y;
      auto nb = [&](int k) {
        return static_cast<size_t>(static_cast<long>(c) + static_cast<long>(kCreepOx[k]) * P + kCreepOy[k]);
      };
      // one batch of loads (x in 2..imax-1, y in 2..jmax-1: every neighbour
      // is inside the padded grid); a cell that does not slide discards it
      const double es = elev[c];
      const double bc = bedrock[c];
      RegRow<E> own;
      reg_row_load(own, gr, c);
      double en[8];
      int32_t nh[8];
      E nv[8][kG + 1];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        en[k] = elev[nb(k)];
        nh[k] = gr.has[nb(k)];
#pragma unroll
        for (int n = 0; n <= kG; ++n) nv[k][n] = gr.G(nb(k), n);
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        s_en[k][tl] = en[k];
        s_has[k][tl] = nh[k];
#pragma unroll
        for (int n = 0; n <= kG; ++n) s_row[k * (kG + 1) + n][tl] = nv[k][n];
      }
      // slide_GS is called only where the cell has a grain row (LCM.cpp:5315-5415)
      if (!(es > bc) || !own.h) continue;
      unsigned moves = 0;  // bit k: the slide towards neighbour k happens
#pragma unroll
      for (int k = 0; k < 8; ++k) moves |= (en[k] < es && en[k] > -9999) ? (1u << k) : 0u;
#pragma unroll 1
      while (moves) {
        const int k = __builtin_ctz(moves);
        moves &= moves - 1;
        const size_t d = static_cast<size_t>(static_cast<long>(c) + static_cast<long>(creep_ox(k)) * P + creep_oy(k));
        RegRow<E> cur;
        cur.hasp = gr.has;
        cur.c = d;
        cur.h = s_has[k][tl];
#pragma unroll
        for (int n = 0; n <= kG; ++n) cur.v[n] = s_row[k * (kG + 1) + n][tl];
        // rate * time / D as the reference groups it: ((dz / D) * rate) * time / D
        const double D = (k & 1) ? 7.07 : DX;  // odd directions are the diagonals
        double temp = ((es - s_en[k][tl]) / D) * rate * time / D;
        if ((es - temp) < bc) temp = es - bc;
        if (temp < 0) temp = 0;
        // a row made here is numbered at (x, y), direction k
        slide_gs_r(gr, c, own, d, cur, temp, elev, bedrock, a, hc::row_key(ep, x, y, k));
#pragma unroll
        for (int n = 0; n <= kG; ++n) gr.G(d, n) = static_cast<E>(cur.v[n]);
      }
#pragma unroll
      for (int n = 0; n <= kG; ++n) gr.G(c, n) = static_cast<E>(own.v[n]);
    }
    target += gridDim.x;
    grid_barrier(bar, target, bar + 1);
  }
}

// tempcreep of every cell x = 1..imax, y = 1..jmax, summed in the order the
// reference's sweep adds to it: its 3x3 block's interior cells in row-major
// order, each contributing its move into this cell, or, for the cell itself,
// its 8 moves out.
__global__ __launch_bounds__(256) void k_creep_sum(int imax, int jmax, int P, double DX,
                                                   double rate, double time,
                                                   const double* __restrict__ elev,
                                                   const double* __restrict__ bedrock,
                                                   double* __restrict__ tempcreep) {
  const int y = 1 + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int x = 1 + static_cast<int>(blockIdx.y);
  if (y > jmax || x > imax) return;
  const size_t c = static_cast<size_t>(x) * P + y;
  double tc = 0;
  for (int dx = -1; dx <= 1; ++dx) {
    for (int dy = -1; dy <= 1; ++dy) {
      const int sx = x + dx, sy = y + dy;
      if (sx < 2 || sx > imax - 1 || sy < 2 || sy > jmax - 1) continue;
      const size_t s = static_cast<size_t>(sx) * P + sy;
      if (!(elev[s] > bedrock[s])) continue;
      for (int k = 0; k < 8; ++k) {
        const bool self = (dx == 0 && dy == 0);
        // the source's move k lands on this cell when its offset is -(dx, dy)
        if (!self && !(kCreepOx[k] == -dx && kCreepOy[k] == -dy)) continue;
        const long off = static_cast<long>(kCreepOx[k]) * P + kCreepOy[k];
        const double es = elev[s];
        const double en = elev[static_cast<size_t>(static_cast<long>(s) + off)];
        if (en < es && en > -9999) {
          const double D = (kCreepOx[k] != 0 && kCreepOy[k] != 0) ? 7.07 : DX;
          double temp = ((es - en) / D) * rate * time / D;
          if ((es - temp) < bedrock[s]) temp = es - bedrock[s];
          if (temp < 0) temp = 0;
          if (self) tc -= temp;
          else tc += temp;
        }
      }
    }
  }
  tempcreep[c] = tc;
}

__global__ __launch_bounds__(256) void k_creep_add(int imax, int jmax, int P,
                                                   double* __restrict__ elev,
                                                   const double* __restrict__ tempcreep) {
  const int y = 1 + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int x = 1 + static_cast<int>(blockIdx.y);
  if (y > jmax || x > imax) return;
  const size_t c = static_cast<size_t>(x) * P + y;
  elev[c] += tempcreep[c];
}

// in_channel_landsliding (LCM.cpp:5429-5528).  The reference walks the
// columns y = 2..jmax-1 in order and, in each, the rows of its active list
// (down_scan: the mask) ascending, row 1 taken as row 2 and row imax as
// imax-1; a cell moves material in place to each neighbour, in the order
// below, whose drop (over 1.41 on the diagonals) exceeds tan(failure angle)
// DX (10000 where the cell is at or below bedrock + active), cut so the cell
// stays at bedrock + active, sliding the moved grain with each move.  As for
// global_landsliding, cells interact only within Chebyshev distance 2, so
// the column-major order runs as wavefronts t = 3y + x, one block, a barrier
// per wavefront.  A cell runs once for each list entry taken as its row (its
// own, and row 1's for row 2, row imax's for row imax-1), consecutively as
// in the reference.
template <typename E>
__global__ __launch_bounds__(1024) void k_in_channel_landslide(ErodeArgs a, int imax, int jmax,
                                                               int P, int W, double factor,
                                                               const uint64_t* __restrict__ mask,
                                                               double* __restrict__ elev,
                                                               const double* __restrict__ bedrock,
                                                               GrainT<E> gr,
                                                               const unsigned long long* __restrict__ any,
                                                               uint32_t ep) {
  if (*any == 0) return;  // no cell can move
  constexpr int ox[8] = {1, 0, -1, -1, -1, 0, 1, 1};
  constexpr int oy[8] = {1, 1, 1, 0, -1, -1, -1, 0};
  const int xmax = max(imax - 1, 2);  // the rows list entries are taken as
  const int t0 = 3 * 2 + 2, t1 = 3 * (jmax - 1) + xmax;
  for (int t = t0; t <= t1; ++t) {
    const int ylo = max(2, (t - xmax + 2) / 3);  // x = t - 3y <= xmax
    const int yhi = min(jmax - 1, (t - 2) / 3);  // x >= 2
    for (int y = ylo + static_cast<int>(threadIdx.x); y <= yhi; y += blockDim.x) {
      const int x = t - 3 * y;
      // Everything the cell reads is loaded at once: no other cell of this
      // wavefront touches its 3x3 block, so the values stay in registers;
      // memory is written at each move, before the grain slide that reads it.
      const size_t c = static_cast<size_t>(x) * P + y;
      const bool m0 = mask_bit(mask, W, x, y);
      const bool m1 = x == 2 && mask_bit(mask, W, 1, y);
      const bool m2 = imax >= 3 && x == imax - 1 && mask_bit(mask, W, imax, y);
      double e = elev[c];
      const double bpa = bedrock[c] + a.active;
      double en[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) en[k] = elev[static_cast<size_t>(static_cast<long>(c) + ox[k] * P + oy[k])];
      // entries taken as row x: x itself, row 1 (taken as 2), row imax (taken as imax-1)
      const int runs = (m0 ? 1 : 0) + (m1 ? 1 : 0) + (m2 ? 1 : 0);
      // the reference visits the entries in row order (1, 2, ..., imax):
      // the first of this cell's consecutive entries numbers its new rows
      const int first = m1 ? 1 : (m0 ? x : imax);
      for (int run = 0; run < runs; ++run) {
        const double wet = (e <= bpa) ? 10000.0 : factor;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const double drop = (ox[k] != 0 && oy[k] != 0) ? (e - en[k]) / 1.41 : (e - en[k]);
          if (drop > wet && en[k] > -9999) {
            double diff = drop - wet;
            if ((e - diff) < bpa) diff = (e - bpa);
            e -= diff;
            en[k] += diff;
            const size_t nb = static_cast<size_t>(static_cast<long>(c) + ox[k] * P + oy[k]);
            elev[c] = e;
            elev[nb] = en[k];
            slide_gs(gr, c, nb, diff, elev, bedrock, a, hc::row_key(ep, y, first, run * 8 + k));
          }
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace

namespace {

// The epoch of the next launch that can make grain rows.
int next_row_epoch(hc_ctx* c, uint32_t* ep) {
  if (c->row_epoch >= (1u << (64 - hc::kRowEpochShift)) - 1)
    return hc_fail(c, HC_ERR_STATE, "grain rows: creation epochs exhausted");
  *ep = ++c->row_epoch;
  return HC_OK;
}

// k_row_stamp over the padded grid with a fresh epoch.
int stamp_rows(hc_ctx* c) {
  uint32_t ep;
  if (int rc = next_row_epoch(c, &ep)) return rc;
  const dim3 grd(static_cast<unsigned>((c->g.jmax + 2 + 255) / 256), static_cast<unsigned>(c->NR));
  hipLaunchKernelGGL(k_row_stamp, grd, dim3(256), 0, c->stream, c->g.imax, c->g.jmax, c->P,
                     c->hasgrain, c->rowkey, ep);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

template <typename E>
GrainT<E> grain_of(hc_ctx* c) {
  return GrainT<E>{static_cast<E*>(c->grain), static_cast<E*>(c->strata), c->hasgrain, c->ncell,
                   c->rowkey};
}

template <typename E>
int erode_t(hc_ctx* c, double mult, double* time_step, double* tempbmax, float* gtot_out) {
  const ErodeArgs a = erode_args(c);
  const GrainT<E> gr = grain_of<E>(c);
  const E* h = static_cast<const E*>(c->h[c->cur]);
  double* elev = static_cast<double*>(c->z);
  const E* nman = static_cast<const E*>(c->nman);
  E* V = static_cast<E*>(c->V);
  E* Tau = static_cast<E*>(c->Tau);
  E* ss = static_cast<E*>(c->ss);
  E* sdir = static_cast<E*>(c->sdir);
  E* etot = static_cast<E*>(c->etot);
  E* etot3 = static_cast<E*>(c->etot3);
  const dim3 blk(64, 4);
  const dim3 cells(static_cast<unsigned>(c->P / 64), static_cast<unsigned>((c->NR + 3) / 4));
  double ts = *time_step * 1.5;  // LCM.cpp:4219
  double tb = 0;
  for (int pass = 0;; ++pass) {
    if (pass > 10000) return hc_fail(c, HC_ERR_UNSTABLE, "erode: timestep loop did not converge");
    HC_HIP(c, hipMemsetAsync(c->d_red + 1, 0, sizeof(unsigned long long), c->stream));
    uint32_t ep;
    if (int rc = next_row_epoch(c, &ep)) return rc;
    hipLaunchKernelGGL(k_entrain<E>, cells, blk, 0, c->stream, a, ts, mult, h, elev, c->bedrock,
                       nman, c->mask, static_cast<const E*>(V), static_cast<const E*>(c->veldir),
                       static_cast<E*>(c->Vel), Tau, ss, sdir, gr, c->needgrain, c->d_red + 1, ep);
    HC_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_alloc<E>, dim3(1024), dim3(256), 0, c->stream, a, elev, c->bedrock, gr,
                       c->needgrain, c->ncell);
    HC_HIP(c, hipGetLastError());
    HC_HIP(c, hipMemcpyAsync(c->h_pinned, c->d_red + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HC_HIP(c, hipStreamSynchronize(c->stream));
    std::memcpy(&tb, c->h_pinned, sizeof(double));
    if (tb > a.ERODEFACTOR) ts *= (a.ERODEFACTOR / tb) * 0.5;  // LCM.cpp:4550-4553
    if (!(tb > a.ERODEFACTOR)) break;
  }
  HC_HIP(c, hipMemsetAsync(etot, 0, c->ncell * sizeof(E), c->stream));
  HC_HIP(c, hipMemsetAsync(etot3, 0, c->ncell * sizeof(E), c->stream));
  if (!c->d_dep)
    HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_dep), kDepPlanes * c->ncell * sizeof(double)));
  hipLaunchKernelGGL(k_deposit_pre<E>, cells, blk, 0, c->stream, a, ts, h, c->mask, V,
                     static_cast<const E*>(ss), static_cast<const E*>(sdir), c->ncell, etot, etot3,
                     c->d_dep);
  HC_HIP(c, hipGetLastError());
  const int ncol = std::max(0, c->g.jmax - 2);
  // The deposition sweep runs each column top to bottom and the y-slide each
  // row left to right (the reference's in-place order): one wavefront per
  // column / row, its lanes sharing the per-cell loads and strata moves.
  if (ncol > 0) {
    hipLaunchKernelGGL(k_deposit<E>, dim3(ncol), dim3(64), 0, c->stream, a, ts, mult, h, elev,
                       c->bedrock, c->mask, static_cast<const E*>(Tau),
                       static_cast<const E*>(ss), gr, static_cast<const double*>(c->d_dep));
    HC_HIP(c, hipGetLastError());
    if (int rc = stamp_rows(c)) return rc;
  }
  hipLaunchKernelGGL(k_lateral_y<E>, dim3(c->g.imax), dim3(64), 0, c->stream, a, ts, mult, h, elev,
                     c->bedrock, c->mask, static_cast<const E*>(Tau), gr,
                     static_cast<const E*>(etot3));
  HC_HIP(c, hipGetLastError());
  if (int rc = stamp_rows(c)) return rc;
  hipLaunchKernelGGL(k_edges<E>, dim3(1), dim3(512), 0, c->stream, a, h, V,
                     static_cast<const E*>(sdir), c->ncell, c->gtot);
  HC_HIP(c, hipGetLastError());
  float g[20];
  HC_HIP(c, hipMemcpyAsync(g, c->gtot, sizeof(float) * 10, hipMemcpyDeviceToHost, c->stream));
  HC_HIP(c, hipStreamSynchronize(c->stream));
  *time_step = ts;
  if (tempbmax) *tempbmax = tb;
  if (gtot_out)
    for (int n = 0; n < 10; ++n) gtot_out[n] = g[n];
  return HC_OK;
}

}  // namespace

int hc_erode_impl(hc_ctx* c, double mult, double* time_step, double* tempbmax, float* gtot_out) {
  return c->sesz == 4 ? erode_t<float>(c, mult, time_step, tempbmax, gtot_out)
                      : erode_t<double>(c, mult, time_step, tempbmax, gtot_out);
}

namespace {

// d_red[4] := whether any cell of rows 2..xmax, columns 2..jmax-1 passes
// the landsliding move test (k_slide_candidates); needs xmax >= 2, jmax >= 3.
int slide_candidates(hc_ctx* c, double factor, double wet_low, int xmax) {
  HC_HIP(c, hipMemsetAsync(c->d_red + 4, 0, sizeof(unsigned long long), c->stream));
  const int jmax = c->g.jmax;
  const dim3 blk(256), grd(static_cast<unsigned>((jmax - 2 + 255) / 256), static_cast<unsigned>(xmax - 1));
  hipLaunchKernelGGL(k_slide_candidates, grd, blk, 0, c->stream, xmax, jmax, c->P, factor, wet_low,
                     c->prm.active, static_cast<const double*>(c->z), c->bedrock, c->d_red + 4);
  return HC_OK;
}

}  // namespace

int hc_global_landsliding_impl(hc_ctx* c, double failureangle, int* sweeps) {
  if (c->g.imax < 3 || c->g.jmax < 3) {  // no interior cell
    if (sweeps) *sweeps = 1;
    return HC_OK;
  }
  // the reference's factor, from the host's libm as in the reference
  const double factor = std::tan((failureangle * (3.141592654 / 180))) * c->g.DX;
  const size_t plane = c->ncell;
  if (!c->d_slide_moves) {
    HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_slide_moves), 8 * plane * sizeof(double)));
  }
  if (int rc = slide_candidates(c, factor, 10 * c->g.DX, c->g.imax - 1)) return rc;
  hipLaunchKernelGGL(k_global_landslide, dim3(1), dim3(1024), 0, c->stream, c->g.imax,
                     c->g.jmax, c->P, c->g.DX, factor, c->prm.active, c->prm.ERODEFACTOR,
                     static_cast<double*>(c->z), c->bedrock, c->d_slide_moves, plane,
                     reinterpret_cast<int*>(c->d_red + 3), c->d_red + 4);
  HC_HIP(c, hipGetLastError());
  int n = 0;
  HC_HIP(c, hipMemcpyAsync(c->h_pinned, c->d_red + 3, sizeof(int), hipMemcpyDeviceToHost,
                           c->stream));
  HC_HIP(c, hipStreamSynchronize(c->stream));
  std::memcpy(&n, c->h_pinned, sizeof(int));
  if (sweeps) *sweeps = n;
  return HC_OK;
}

int hc_creep_impl(hc_ctx* c, double time, double creep_rate) {
  const size_t plane = c->ncell;
  if (!c->d_slide_moves) {
    HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_slide_moves), 8 * plane * sizeof(double)));
  }
  const ErodeArgs a = erode_args(c);
  double* elev = static_cast<double*>(c->z);
  const int imax = c->g.imax, jmax = c->g.jmax;
  if (imax >= 3 && jmax >= 3) {
    uint32_t ep;
    if (int rc = next_row_epoch(c, &ep)) return rc;
    // one lane per cell of the longest wavefront, within what the device
    // holds at once
    const int cells = std::min(imax - 2, (jmax - 2 + 2) / 3);
    int nb = std::max(1, (cells + kCreepThreads - 1) / kCreepThreads);
    int dev = 0, ncu = 0, per_cu = 0;
    HC_HIP(c, hipGetDevice(&dev));
    HC_HIP(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const void* fn = c->sesz == 4 ? reinterpret_cast<const void*>(&k_creep_grain<float>)
                                  : reinterpret_cast<const void*>(&k_creep_grain<double>);
    HC_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kCreepThreads, 0));
    nb = std::min(nb, std::max(1, per_cu) * ncu);
    double DX = c->g.DX;
    const double* bedrock = c->bedrock;
    GrainT<float> g4 = grain_of<float>(c);
    GrainT<double> g8 = grain_of<double>(c);
    // d_red[5]: the barrier's arrivals, d_red[6]: its time-out flag
    unsigned long long* bar = c->d_red + 5;
    HC_HIP(c, hipMemsetAsync(bar, 0, 2 * sizeof(unsigned long long), c->stream));
    void* args[] = {const_cast<ErodeArgs*>(&a), const_cast<int*>(&imax), const_cast<int*>(&jmax),
                    &c->P, &DX, &creep_rate, &time, &elev, const_cast<const double**>(&bedrock),
                    c->sesz == 4 ? static_cast<void*>(&g4) : static_cast<void*>(&g8), &ep, &bar};
    HC_HIP(c, hipLaunchKernel(fn, dim3(nb), dim3(kCreepThreads), args, 0, c->stream));
    HC_HIP(c, hipMemcpyAsync(c->h_pinned, bar + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             c->stream));
    HC_HIP(c, hipStreamSynchronize(c->stream));
    unsigned long long timed_out = 0;
    std::memcpy(&timed_out, c->h_pinned, sizeof(timed_out));
    if (timed_out != 0)
      return hc_fail(c, HC_ERR_HIP, "creep: the blocks of the grain sweep were not all resident");
  }
  const dim3 blk(256), grd(static_cast<unsigned>((jmax + 255) / 256), static_cast<unsigned>(imax));
  hipLaunchKernelGGL(k_creep_sum, grd, blk, 0, c->stream, imax, jmax, c->P, c->g.DX, creep_rate,
                     time, elev, c->bedrock, c->d_slide_moves);
  hipLaunchKernelGGL(k_creep_add, grd, blk, 0, c->stream, imax, jmax, c->P, elev,
                     c->d_slide_moves);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_in_channel_landsliding_impl(hc_ctx* c, double failureangle) {
  if (c->g.imax < 2 || c->g.jmax < 3) return HC_OK;  // no column 2..jmax-1 or row 2
  // the reference's factor, from the host's libm as in the reference
  const double factor = std::tan((failureangle * (3.141592654 / 180))) * c->g.DX;
  const ErodeArgs a = erode_args(c);
  if (int rc = slide_candidates(c, factor, 10000.0, std::max(c->g.imax - 1, 2))) return rc;
  uint32_t ep;
  if (int rc = next_row_epoch(c, &ep)) return rc;
  if (c->sesz == 4)
    hipLaunchKernelGGL(k_in_channel_landslide<float>, dim3(1), dim3(1024), 0, c->stream, a,
                       c->g.imax, c->g.jmax, c->P, c->W, factor, c->mask,
                       static_cast<double*>(c->z), c->bedrock, grain_of<float>(c), c->d_red + 4,
                       ep);
  else
    hipLaunchKernelGGL(k_in_channel_landslide<double>, dim3(1), dim3(1024), 0, c->stream, a,
                       c->g.imax, c->g.jmax, c->P, c->W, factor, c->mask,
                       static_cast<double*>(c->z), c->bedrock, grain_of<double>(c), c->d_red + 4,
                       ep);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_reach_inputs_impl(hc_ctx* c, int n, const int* x, const int* y, const double* sediment,
                         const double* water, double threshold) {
  if (n < 0 || n > 16 || (n > 0 && (!x || !y || !water))) return hc_fail(c, HC_ERR_ARG, "reach_inputs: bad arguments");
  if (n == 0) return HC_OK;
  if (sediment && !c->g.sediment)
    return hc_fail(c, HC_ERR_STATE, "reach_inputs: sediment needs a sediment context");
  int idx[16];
  double sed[16 * 9], wat[16];
  for (int k = 0; k < n; ++k) {
    if (x[k] < 0 || x[k] >= c->NR || y[k] < 0 || y[k] >= c->NC)
      return hc_fail(c, HC_ERR_ARG, "reach_inputs: point outside the padded grid");
    idx[k] = x[k] * c->P + y[k];
    wat[k] = water[k];
    for (int f = 0; f < 9 && sediment; ++f) sed[k * 9 + f] = sediment[k * 9 + f];
  }
  // the small arrays travel in the point-input buffers
  if (!c->d_pts || c->pts_cap < 16 * 9 + 32) {
    HC_HIP(c, hipStreamSynchronize(c->stream));
    if (c->d_pts) (void)hipFree(c->d_pts);
    if (c->d_amt) (void)hipFree(c->d_amt);
    c->d_pts = nullptr;
    c->d_amt = nullptr;
    const int cap = 1024;
    HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_pts), cap * sizeof(int)));
    HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_amt), cap * sizeof(double)));
    c->pts_cap = cap;
  }
  HC_HIP(c, hipMemcpyAsync(c->d_pts, idx, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HC_HIP(c, hipMemcpyAsync(c->d_amt, wat, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (sediment)
    HC_HIP(c, hipMemcpyAsync(c->d_amt + 16, sed, n * 9 * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
  // the host arrays above are pageable: the copies have completed on return
  const ErodeArgs a = c->g.sediment ? erode_args(c) : ErodeArgs{};
  const double* sedp = (c->g.sediment && sediment) ? c->d_amt + 16 : nullptr;
  double* zp = c->g.sediment ? static_cast<double*>(c->z) : nullptr;
  uint32_t ep = 0;
  if (sedp)
    if (int rc = next_row_epoch(c, &ep)) return rc;
  if (c->prec == HC_FP64)
    hipLaunchKernelGGL(k_reach<double>, dim3(1), dim3(64), 0, c->stream, a,
                       c->g.sediment ? grain_of<double>(c) : GrainT<double>{}, n, c->d_pts, sedp,
                       c->d_amt, threshold, static_cast<double*>(c->h[c->cur]), zp, c->bedrock,
                       static_cast<double*>(c->V), ep);
  else
    hipLaunchKernelGGL(k_reach<float>, dim3(1), dim3(64), 0, c->stream, a,
                       c->g.sediment ? grain_of<float>(c) : GrainT<float>{}, n, c->d_pts, sedp,
                       c->d_amt, threshold, static_cast<float*>(c->h[c->cur]), zp, c->bedrock,
                       static_cast<float*>(c->V), ep);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}
