// Hydro kernels of the hot path: flow_route, depth_update (separate and
// fused), scan_area, water_flux_out, input points, health check.
//
// FP64 kernels follow the reference's operation order exactly
// (LSDCatchmentModel.cpp:3063-3299) and are compiled without FP contraction;
// the one deliberate difference is hflow^3, which the reference takes from
// libm pow(hflow, 3) (integer 10/3, LSDCatchmentModel.cpp:3115) and which is
// formed here as a correctly rounded cube from error-free products.
//
// FP32 kernels store h, qx, qy, z (and n) in single precision and use the
// algebraically identical form of the face update that never builds the
// water-surface elevation eta = z + h (which cancels catastrophically in
// float): with dz = z_nb - z_cell,
//    hflow = dz >= 0 ? max(h_cell - dz, h_nb) : max(h_cell, h_nb + dz)
//    slope = (dz + (h_nb - h_cell)) / DX
// and faces with |dz| < flat_dz take dz from the FP64 DEM.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include <hip/hip_ext.h>

#include "hc_device.hpp"
#include "hc_internal.hpp"

using hc::DrainArgs;
using hc::EnsMember;
using hc::fused_edge_tile;
using hc::HydroArgs;
using hc::kGravity;

using namespace hcdev;

namespace {

// One face of the FP64 plain form (LSDCatchmentModel.cpp:3094-3191 for the
// x-face, 3195-3242 for the y-face).  Cell values (hc, zc), neighbour across
// the face (hn, zn).  `edge` is -1 for x<=2 / y<=2, +1 for x==imax /
// y==jmax, else 0 (the `<= 2` override is applied last in the reference).
// Leaves q (and qs) unchanged unless the face condition holds.
template <bool SED, typename VT = double>
__device__ __forceinline__ void face64(double& q, double& qs, double hc, double hn, double zc,
                                       double zn, double n, double dt, const HydroArgs& a,
                                       int edge, bool xface, double Vc, double Vn,
                                       VT* vel_pos, VT* vel_neg) {
  if (!((hc > 0 || hn > 0) && zn > -9999)) return;
  const double eta_c = zc + hc;
  const double eta_n = zn + hn;
  const double etamax = dmax_ref(eta_c, eta_n);
  const double zmax = xface ? dmax_ref(zn, zc) : dmax_ref(zc, zn);
  const double hflow = etamax - zmax;
  if (hflow > a.hthr) {
    double slope = (eta_n - eta_c) / a.DX;
    if (edge > 0) slope = a.edgeslope;
    if (edge < 0) slope = 0 - a.edgeslope;
    const double h3 = cube_rn(hflow);
    double qv = q;
    qv = ((qv - (kGravity * hflow * dt * slope)) /
          (1 + kGravity * hflow * dt * (n * n) * fabs(qv) / h3));
    const double sg = sqrt(kGravity * hflow);
    if (qv > 0 && (qv / hflow) / sg > a.froude) qv = hflow * (sg * a.froude);
    if (qv < 0 && fabs(qv / hflow) / sg > a.froude) qv = 0 - (hflow * (sg * a.froude));
    if (qv > 0 && (qv * dt / a.DX) > (hc / 4)) qv = ((hc * a.DX) / 5) / dt;
    if (qv < 0 && fabs(qv * dt / a.DX) > (hn / 4)) qv = 0 - ((hn * a.DX) / 5) / dt;
    q = qv;
    if (SED) {
      double s = qs;
      if (qv > 0) s = qv * (Vc / hc);
      if (qv < 0) s = qv * (Vn / hn);
      if (s > 0 && s * dt > (Vc * a.DX) / 4) s = ((Vc * a.DX) / 5) / dt;
      if (s < 0 && fabs(s * dt) > (Vn * a.DX) / 4) s = 0 - ((Vn * a.DX) / 5) / dt;
      qs = s;
      if (qv > 0) *vel_pos = qv / hflow;
      if (qv < 0) *vel_neg = (0 - qv) / hflow;
    }
  } else {
    q = 0;
    if (SED) qs = 0;
  }
}

// One face of the FP32 form, branch-free: every lane evaluates it and the
// caller selects.  dz = z_nb - z_cell.  The friction denominator is cleared
// of 1/h^3,
//   (q - g h dt S) / (1 + g h dt n^2 |q| / h^3)
//     = (q - g h dt S) h^3 / (h^3 + g h dt n^2 |q|),
// so one reciprocal serves; the Froude cap is a clamp to [-qF, qF]
// (qF = h sqrt(g h) Fr = sqrt(g Fr^2 h^3) > 0 wherever the result is kept,
// from the h^3 already formed) and the depth limiter
// tests are products instead of quotients.  0 where hflow <= the flow
// threshold (LSDCatchmentModel.cpp:3094-3191).  The slope is taken as
// slope * smul + sadd: (1, 0) inside, (0, +-edgeslope) on the edge faces.
struct Face32K {
  float gdt, gfr2, invDX, hthr, k4, k5;  // g dt, g Fr^2, 1/DX, flow threshold, limiters
};

__device__ __forceinline__ float face32(float q, float hc, float hn, float dz, float n2,
                                        float smul, float sadd, const Face32K& K) {
  // dz >= 0 ? max(hc - dz, hn) : max(hc, hn + dz)
  const float hflow = fmaxf(hc - fmaxf(dz, 0.0f), hn + fminf(dz, 0.0f));
  const float slope = __builtin_fmaf((dz + (hn - hc)) * K.invDX, smul, sadd);
  const float ghdt = K.gdt * hflow;
  const float h3 = (hflow * hflow) * hflow;
  const float num = __builtin_fmaf(-ghdt, slope, q);
  float qv = (num * h3) * __builtin_amdgcn_rcpf(__builtin_fmaf(ghdt * n2, fabsf(q), h3));
  const float qF = __builtin_amdgcn_sqrtf(K.gfr2 * h3);
  qv = __builtin_amdgcn_fmed3f(qv, -qF, qF);
  qv = (qv > 0.0f && qv > hc * K.k4) ? hc * K.k5 : qv;
  qv = (qv < 0.0f && -qv > hn * K.k4) ? -(hn * K.k5) : qv;
  return hflow > K.hthr ? qv : 0.0f;
}

// a && b && c && d evaluated without branches (the selects it feeds stay
// selects instead of exec-masked branches)
__device__ __forceinline__ bool and4(bool a, bool b, bool c, bool d) {
  return (static_cast<int>(a) & static_cast<int>(b) & static_cast<int>(c) & static_cast<int>(d)) != 0;
}

__device__ __forceinline__ int edge_code(int i, int imaxlike) {
  return (i <= 2) ? -1 : ((i == imaxlike) ? 1 : 0);
}

// ------------------------------------------------------------ flow_route
// In-place per-cell map (each face is written by exactly one cell).
// T: storage of depths, fluxes, n and the sediment planes (double, or float
// in the FP32 storage form of erosion runs); arithmetic in FP64 throughout,
// rounded once at each store.
template <bool SED, typename T>
__global__ __launch_bounds__(256) void k_flow_route64(HydroArgs a, double dt,
                                                      const T* __restrict__ h,
                                                      const double* __restrict__ z,
                                                      T* __restrict__ qx,
                                                      T* __restrict__ qy,
                                                      const T* __restrict__ nman,
                                                      const uint64_t* __restrict__ mask,
                                                      T* __restrict__ qxs,
                                                      T* __restrict__ qys,
                                                      const T* __restrict__ V,
                                                      T* __restrict__ vd, size_t plane) {
  const int y = blockIdx.x * 64 + threadIdx.x;
  const int x = blockIdx.y * blockDim.y + threadIdx.y;
  if (x < 1 || x > a.imax || y < 1 || y > a.jmax) return;
  if (!mask_bit(mask, a.mwords, x, y)) return;
  const size_t P = a.pitch;
  const size_t c = static_cast<size_t>(x) * P + y;
  const double zc = z[c];
  if (!(zc > -9999)) return;
  const double n = a.spatial_n ? static_cast<double>(nman[c]) : a.mannings;
  const double hc = h[c];
  double qv, qsv = 0;
  // x face (neighbour x-1)
  {
    const size_t cn = c - P;
    qv = qx[c];
    if (SED) qsv = qxs[c];
    face64<SED, T>(qv, qsv, hc, h[cn], zc, z[cn], n, dt, a, edge_code(x, a.imax), true,
                   SED ? static_cast<double>(V[c]) : 0.0, SED ? static_cast<double>(V[cn]) : 0.0,
                   SED ? &vd[3 * plane + c] : nullptr, SED ? &vd[1 * plane + cn] : nullptr);
    qx[c] = static_cast<T>(qv);
    if (SED) qxs[c] = static_cast<T>(qsv);
  }
  // y face (neighbour y-1)
  {
    const size_t cn = c - 1;
    qv = qy[c];
    if (SED) qsv = qys[c];
    face64<SED, T>(qv, qsv, hc, h[cn], zc, z[cn], n, dt, a, edge_code(y, a.jmax), false,
                   SED ? static_cast<double>(V[c]) : 0.0, SED ? static_cast<double>(V[cn]) : 0.0,
                   SED ? &vd[0 * plane + c] : nullptr, SED ? &vd[2 * plane + cn] : nullptr);
    qy[c] = static_cast<T>(qv);
    if (SED) qys[c] = static_cast<T>(qsv);
  }
}

// ---------------------------------------------------------- depth_update
template <bool SED, typename T>
__global__ __launch_bounds__(256) void k_depth64(HydroArgs a, double dt, T* __restrict__ h,
                                                 const double* __restrict__ z,
                                                 const T* __restrict__ qx,
                                                 const T* __restrict__ qy,
                                                 const uint64_t* __restrict__ mask,
                                                 T* __restrict__ V,
                                                 const T* __restrict__ qxs,
                                                 const T* __restrict__ qys,
                                                 unsigned long long* red) {
  const int y = blockIdx.x * 64 + threadIdx.x;
  const int x = blockIdx.y * blockDim.y + threadIdx.y;
  double md = 0;
  if (x >= 1 && x <= a.imax && y >= 1 && y <= a.jmax && mask_bit(mask, a.mwords, x, y)) {
    const size_t P = a.pitch;
    const size_t c = static_cast<size_t>(x) * P + y;
    const double qxc = qx[c], qxp = qx[c + P], qyc = qy[c], qyp = qy[c + 1];
    double hv = static_cast<double>(h[c]) + dt * (qxp - qxc + qyp - qyc) / a.DX;
    if (SED) {
      const double sxc = qxs[c], sxp = qxs[c + P], syc = qys[c], syp = qys[c + 1];
      V[c] = static_cast<T>(static_cast<double>(V[c]) + dt * (sxp - sxc + syp - syc) / a.DX);
    }
    hv = static_cast<double>(static_cast<T>(hv));  // the stored depth
    if (hv > 0) {
      if (z[c] == -9999) hv = 0;
      if (hv > md) md = hv;
    }
    if (!(hv >= hc::kUnstableDepth)) md = INFINITY;  // NaN, -inf or a large negative
    h[c] = static_cast<T>(hv);
  }
  block_max_pos(red, md);
}

// ------------------------------------------------------------ fused kernel
// flow_route followed by depth_update in one pass, reading generation-0
// buffers and writing generation 1 for every cell (unmasked cells are copied,
// which keeps the reference's stale-flux semantics).
//
// Tiling: a wavefront owns 64 consecutive columns starting at 64*w; the
// y-faces on its two sides take their neighbour values from the adjacent
// wavefronts through LDS, or from one extra column load on a block's sides.
// Each wavefront takes R output rows and first loads its whole tile (R+2
// rows of h and z, R+1 rows of qx, qy, n and the mask) so all of a tile's
// loads are in flight together; then it computes the rows in order.
// Blocks are four wavefronts side by side; consecutive block ids go to the
// eight XCDs in turn, so the block id is remapped to give each XCD a
// contiguous band of tiles (whose halo rows then hit the same L2).
//
// MODE 0: FP64 reference form.  MODE 1: FP32 state with FP32 elevations (and
// the FP64 fallback for near-flat faces when flat_dz > 0).  MODE 2: FP32
// state with FP64 elevations (exact face differences).  MODE 3: FP32 state
// with 32-bit fixed-point elevations (differences exact to one quantum).
template <int MODE>
struct FusedTypes {
  using T = typename std::conditional<MODE == 0, double, float>::type;
  using Z = typename std::conditional<
      MODE == 3, int32_t, typename std::conditional<MODE == 1, float, double>::type>::type;
};

// Elevation predicates of the reference (`elev > -9999`, `elev == -9999`),
// for every elevation type of the fused kernels.
template <typename Z>
__device__ __forceinline__ bool z_ok(Z z) {
  return z > Z(-9999);
}
template <>
__device__ __forceinline__ bool z_ok<int32_t>(int32_t z) {
  return z != hc::kZqNodata;
}
template <typename Z>
__device__ __forceinline__ bool z_nodata(Z z) {
  return z == Z(-9999);
}
template <>
__device__ __forceinline__ bool z_nodata<int32_t>(int32_t z) {
  return z == hc::kZqNodata;
}

// Device pointers of one fused launch (one generation in, the other out).
template <int MODE>
struct FusedPtrs {
  using T = typename FusedTypes<MODE>::T;
  using Z = typename FusedTypes<MODE>::Z;
  const T* h0;
  const Z* zz;
  const double* z64;
  const T* qx0;
  const T* qy0;
  const T* nman;
  const uint64_t* mask;
  T* h1;
  T* qx1;
  T* qy1;
  unsigned long long* red;
  // device loop: one maximum depth per wavefront (slot 4 * block + wavefront),
  // reduced into red by the scan kernel that follows; null: a block maximum
  // and one conditional atomic on red per block
  double* wmax;
  hc::InjectArgs inj;  // rain input at the input points (device loop only)
  // device loop, aligned tiles, the kernel variant for a body that rebuilds
  // the active mask: every stored depth also leaves its wet bit (h > 0) in
  // wet, one 64-column word per row and wavefront, for the scan kernel's
  // 3x3 dilation
  uint64_t* wet = nullptr;
};

// One tile: a strip of nsub x R output rows by four wavefronts side by side.
//
// FP32 forms (aligned): wavefront w owns the 64 columns 64w .. 64w+63 (its
// rows are whole aligned 256-byte segments).  The y-faces need one column on
// either side: column 64w-1's depth and elevation for lane 0's west face,
// and the new flux of the face east of lane 63, which belongs to column
// 64w+64.  Inside the block the neighbouring wavefronts supply them through
// LDS (one barrier per sub-tile); on the block's own west and east sides
// wavefronts 0 and 3 fetch that column with one row per lane.  The east
// faces of all the sub-tile's rows are computed together, lane k taking row
// k, by the same y-face code as the row's own faces, and enter the row's
// lane shift as the value shifted into lane 63 (DPP with the old value kept).
//
// Overlapping (stride 62): wavefront w covers the 64 columns from 62w and
// owns the 62 in the middle (lane 0 of wavefront 0 also owns column 0); its
// first and last lanes are halo columns, so the y-neighbours come from lane
// shifts alone, with no exchange or barrier.  The FP64 form is bound by its
// arithmetic, which the aligned form's recomputed east faces would add to
// (1/(2R) more y-faces), and the ensemble's tiles are small and mostly on a
// member's edge, where the aligned form's per-tile exchange does not pay;
// both use this form.  The FP32 single grid is bound by its loads, which the
// aligned rows make cheaper.
//
// The strip is walked in sub-tiles of R rows: the first loads rows r0-1 ..
// r0+R of every field into registers up front, each later one loads only
// its R new rows and carries the two rows above it and the new fluxes of its
// first row from the sub-tile before (no halo rows are re-read and no face
// is computed twice).  Rows past the grid's first and last rows load the
// nearest grid row instead: their mask bits are 0, so no face or depth of
// theirs is used or stored, and the loads need no branch.  Loads and stores
// address a row by a 32-bit byte offset from the field's base.  Every thread
// of the block calls this function (it has block barriers).
template <int MODE, int R, bool AL, bool WET = false>
__device__ __forceinline__ void fused_tile(const HydroArgs& a, int bx, int by, int nbx, int nsub,
                                           double dt_arg, const double* dtp, const int* haltp,
                                           const FusedPtrs<MODE>& p, const DrainArgs& drain) {
  using T = typename FusedTypes<MODE>::T;
  using Z = typename FusedTypes<MODE>::Z;
  const T* __restrict__ h0 = p.h0;
  const Z* __restrict__ zz = p.zz;
  const double* __restrict__ z64 = p.z64;
  const T* __restrict__ qx0 = p.qx0;
  const T* __restrict__ qy0 = p.qy0;
  const T* __restrict__ nman = p.nman;
  const uint64_t* __restrict__ mask = p.mask;
  T* __restrict__ h1 = p.h1;
  T* __restrict__ qx1 = p.qx1;
  T* __restrict__ qy1 = p.qy1;
  unsigned long long* red = p.red;
  static_assert(R + 2 <= 32, "rows of a sub-tile are indexed by lanes 0..31");
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wavefront in the block
  const int kk = lane & 31;         // row of the lane-indexed vectors
  const int wu = __builtin_amdgcn_readfirstlane(bx * 4 + wv);
  const int NR = a.imax + 2, NC = a.jmax + 2;
  const int H = R * nsub;  // rows per block
  constexpr int SW = AL ? hc::kStrideAligned : hc::kStrideOverlap;  // columns per wavefront
  const bool active_wave = AL ? (64 * wu <= NC - 1) : (62 * wu + 1 <= NC - 1);
  const int c = SW * wu + lane;
  const bool owned = (AL || (lane >= 1 && lane <= 62) || (wu == 0 && lane == 0)) && c <= NC - 1;
  const uint32_t P = static_cast<uint32_t>(a.pitch);
  const uint32_t cl = static_cast<uint32_t>(min(c, a.pitch - 1));
  // the block's outer halo columns (column 0 stands in west of column 0,
  // whose faces are never used)
  const uint32_t cw = static_cast<uint32_t>(max(64 * wu - 1, 0));
  const uint32_t ce = static_cast<uint32_t>(min(64 * wu + 64, a.pitch - 1));

  T hh[R + 2];
  Z zt[R + 2];
  T qxo[R + 2], qyo[R + 2], nn[R + 2];
  // rain input at input points: one amount for the whole grid when every
  // cell has the same rain cell (added with a select, no branch)
  const bool inject = p.inj.pmask != nullptr;
  double amt_u = 0.0;  // read with the loop state below
  auto rain = [&](T v, bool pt, uint32_t o) -> T {  // LCM.cpp:3463-3477
    if (p.inj.rf_uniform >= 0) {
      const T hp = static_cast<T>(static_cast<double>(v) + amt_u);
      return pt ? hp : v;
    }
    return pt ? static_cast<T>(static_cast<double>(v) + p.inj.amt[p.inj.rfarea[o]]) : v;
  };
  // Bit masks (active cells, input points).  The wavefront's 64 columns are
  // one 64-bit word of a row (aligned) or span two (stride 62).  One load per
  // sub-tile: lane k (< R+2) fetches the active-mask words of row r0-1+k,
  // lane 32+k those of the input-point mask, and funnels out the 64 bits
  // under the wavefront (bit l = column SW*w+l; columns past the last word
  // read 0).  A row's 64 bits, read back with v_readlane, are then the lane
  // mask of its predicate as they stand (inverse ballot): no per-lane bit
  // work.  Aligned, lane k also keeps bit 0 of the next word: the east
  // column's active bit.
  const int w0 = AL ? wu : (62 * wu) >> 6;
  const int wsh = AL ? 0 : 62 * wu - 64 * w0;  // bit of column SW*w in word w0
  uint32_t east_m = 0;
  auto load_win = [&](int r0) -> uint64_t {
    const bool pm = lane >= 32;
    const int x = r0 - 1 + kk;
    uint64_t m0 = 0, m1 = 0;
    if (kk < R + 2 && x >= 0 && x < NR && (!pm || inject)) {
      const uint64_t* base = pm ? p.inj.pmask : mask;
      const size_t ro = static_cast<size_t>(x) * a.mwords + w0;
      m0 = base[ro];
      if (w0 + 1 < a.mwords) m1 = base[ro + 1];
    }
    if constexpr (AL) {
      east_m = static_cast<uint32_t>(m1) & 1u;
      return m0;
    } else {
      return wsh ? ((m0 >> wsh) | (m1 << (64 - wsh))) : m0;
    }
  };
  // whether any of the sub-tile's rows has an input point under the
  // wavefront (lanes 32.. hold the point-mask words)
  auto any_point = [&](uint64_t win) -> bool {
    return __ballot(lane >= 32 && lane < 32 + R + 2 && win != 0) != 0;
  };
  auto row_pred = [&](uint64_t win, int src) -> bool {
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(win), src);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(win >> 32), src);
    return __builtin_amdgcn_inverse_ballot_w64((static_cast<uint64_t>(hi) << 32) | lo);
  };
  const T nu = static_cast<T>(a.mannings);
  T md = 0;
  bool sick = false;  // an updated depth is non-finite or below kUnstableDepth
  // water_flux_out (LCM.cpp:3005-3047) on this tile's edge cells, after
  // depth_update and before the store; maxdepth is taken before the drain
  const bool drains = drain.ts != nullptr && fused_edge_tile(bx, by, H, a.imax, a.jmax, SW);
  double ts = 1.0;
  double dacc = 0;
  // In the device-resident loop dt, the halt flag, the time step and the
  // rain amount are read from the loop state.  They are read after the first
  // sub-tile's loads are issued: scalar loads complete out of order, so a
  // wait for any of them is a wait for all, and the vector loads would
  // otherwise queue behind that chain.
  double dtd = dt_arg;
  constexpr bool wetb = WET && AL;  // this launch leaves wet-cell bits
  uint32_t wet_lo = 0, wet_hi = 0;  // lane k: the wet bits of the sub-tile's row r0 + k
  // the sub-tile's R words of wet bits in one store (word wu of each row)
  uint64_t* const wetp = p.wet;
  auto flush_wet = [&](int r0) {
    const int xr = r0 + lane;
    const uint32_t wo = static_cast<uint32_t>(xr * a.mwords + wu) * 8u;
    if (lane < R && xr < NR)  // an ordinary store: the words of a row meet in L2
      *reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(wetp) + wo) =
          (static_cast<uint64_t>(wet_hi) << 32) | wet_lo;
  };
  float dtDX = 0.0f;  // FP32 form's depth-update factor
  const float fthr = static_cast<float>(a.flat_dz);
  Face32K K;
  auto dt_setup = [&]() {
    if constexpr (MODE != 0) {
      const float DX = static_cast<float>(a.DX);
      const float invdt = 1.0f / static_cast<float>(dtd);
      dtDX = static_cast<float>(dtd / a.DX);
      const float g = static_cast<float>(kGravity);
      const float fr = static_cast<float>(a.froude);
      K.gdt = g * static_cast<float>(dtd);
      K.gfr2 = (g * fr) * fr;
      K.invDX = 1.0f / DX;
      K.hthr = static_cast<float>(a.hthr);
      K.k4 = 0.25f * DX * invdt;  // q*dt/DX > h/4  <=>  q > h*DX/(4 dt)
      K.k5 = 0.2f * DX * invdt;
      // computed once here: left to itself the compiler repeats the
      // divisions at every use
      K.gdt = pin_value(K.gdt);
      K.invDX = pin_value(K.invDX);
      K.k4 = pin_value(K.k4);
      K.k5 = pin_value(K.k5);
      dtDX = pin_value(dtDX);
    }
  };
  const float es = static_cast<float>(a.edgeslope);
  const int ycode = edge_code(c, a.jmax);  // y-faces: per column
  // y-face slope = slope * ymul + yadd: the edge override without a lane mask
  const float ymul = ycode != 0 ? 0.0f : 1.0f;
  const float yadd = ycode < 0 ? -es : (ycode > 0 ? es : 0.0f);
  const int ycode_e = edge_code(64 * wu + 64, a.jmax);  // the east column's
  const float ymul_e = ycode_e != 0 ? 0.0f : 1.0f;
  const float yadd_e = ycode_e < 0 ? -es : (ycode_e > 0 ? es : 0.0f);

  // New flux of the y-face between a cell (qo, h, z, n, active bit m, offset
  // o, slope terms) and its west neighbour (hl, zl): the face of row() below,
  // also used for the lane-batched east faces, so both give the same bits.
  auto yface = [&](T qo, T hcell, T hl, Z zc, Z zl, T ncell, bool m, int code, float smul,
                   float sadd, uint32_t o) -> T {
    T qN = qo;
    if constexpr (MODE == 0) {
      if (m && zc > Z(-9999)) {
        double dummy = 0;
        face64<false, double>(qN, dummy, hcell, hl, zc, zl, ncell, dtd, a, code, false, 0, 0, nullptr,
                      nullptr);
      }
    } else {
      float dz;
      if constexpr (MODE == 2) {
        dz = static_cast<float>(zl - zc);
      } else if constexpr (MODE == 3) {
        dz = static_cast<float>(static_cast<int32_t>(static_cast<uint32_t>(zl) -
                                                     static_cast<uint32_t>(zc))) *
             a.zq_inv;
      } else {
        dz = zl - zc;
        if (fthr > 0.0f && m && fabsf(dz) < fthr) dz = static_cast<float>(z64[o - 1] - z64[o]);
      }
      float fy = face32(qo, hcell, hl, dz, ncell * ncell, smul, sadd, K);
      asm volatile("" : "+v"(fy));  // on every lane, not under the select
      qN = and4(m, z_ok(zc), (hcell > 0.0f) || (hl > 0.0f), z_ok(zl)) ? fy : qo;
    }
    return qN;
  };

  // Element offset of the lane's cell in row k (x = r0 - 1 + k; rows past
  // the grid: the nearest grid row).  The row part is uniform and made
  // afresh in a scalar register at each use, so no per-row offsets stay live
  // in vector registers.
  auto off = [&](int k, int r0) -> uint32_t {
    uint32_t rb = __builtin_amdgcn_readfirstlane(min(max(r0 - 1 + k, 0), NR - 1) * P);
    asm volatile("" : "+s"(rb));
    return rb + cl;
  };
  // load row k of every field
  auto load_row = [&](int k, int r0) {
    const uint32_t o = off(k, r0);
    hh[k] = ld_elem(h0, o);
    zt[k] = ld_elem(zz, o);
    if (k >= 1) {  // the fluxes' only re-read is the tile below's first row
      qxo[k] = ld_nt(qx0, o);
      qyo[k] = ld_nt(qy0, o);
    }
  };
  // (Manning's n, like the elevations, is read every step and never written:
  // ordinary loads leave it in the caches between steps, where the fields
  // that stream through -- the fluxes in, every field out -- go nontemporal;
  // 106.5 against 108.5 us per FP32 launch at 4096^2)
  auto load_n = [&](int k0, int r0) {
    if (a.spatial_n) {
#pragma unroll
      for (int k = k0; k < R + 2; ++k) nn[k] = ld_elem(nman, off(k, r0));
    } else {
#pragma unroll
      for (int k = k0; k < R + 2; ++k) nn[k] = nu;
    }
  };

  // Halo columns as lane-indexed vectors (lane k = row r0-1+k): west column
  // (depth, elevation), east column (depth, elevation, old y-flux, n, offset)
  // and this wavefront's own column 63 (depth, elevation).
  T hw = 0, he = 0, qe = 0, ne = 0, h63 = 0;
  Z zw = 0, ze = 0, z63 = 0;
  uint32_t ork = 0;  // lane k: offset of row r0-1+k (rows past the grid: the nearest)
  // the block's outer columns, one row per lane (wavefronts 0 and 3).  The
  // halo variables are assigned on every path by value selects: a store to
  // one of two variables in the arms of a branch becomes a store through a
  // selected address, which puts both variables in scratch memory.
  auto load_outer = [&](int r0) {
    ork = static_cast<uint32_t>(min(max(r0 - 1 + min(kk, R + 1), 0), NR - 1)) * P;
    if (wv == 0 || wv == 3) {
      const bool west = wv == 0;
      const uint32_t o = ork + (west ? cw : ce);
      const T hv = ld_elem(h0, o);
      const Z zv = ld_elem(zz, o);
      hw = west ? hv : hw;
      zw = west ? zv : zw;
      he = west ? he : hv;
      ze = west ? ze : zv;
    }
    if (wv == 3) {
      qe = ld_elem(qy0, ork + ce);
      ne = nu;
      if (a.spatial_n) ne = ld_elem(nman, ork + ce);
    }
  };
  auto rain_outer = [&](int r0) {
    if (wv != 0 && wv != 3) return;
    const bool west = wv == 0;
    const uint32_t col = west ? cw : ce;
    const int x = r0 - 1 + kk;
    bool pt = false;
    if (kk < R + 2 && x >= 0 && x < NR)
      pt = (p.inj.pmask[static_cast<size_t>(x) * a.mwords + (col >> 6)] >> (col & 63)) & 1u;
    const T hv = rain(west ? hw : he, pt, ork + col);
    hw = west ? hv : hw;
    he = west ? he : hv;
  };
  // Edge values through LDS, double-buffered by sub-tile parity so one
  // barrier per sub-tile separates each write from the reads of it and the
  // reads from the next writes to the same buffer.
  constexpr int LB = AL ? 2 : 1, LW = AL ? 4 : 1, LR = AL ? 32 : 1;  // none for stride 62
  __shared__ T s_eh[LB][LW][LR], s_wh[LB][LW][LR], s_wq[LB][LW][LR], s_wn[LB][LW][LR];
  __shared__ Z s_ez[LB][LW][LR], s_wz[LB][LW][LR];
  auto exchange = [&](auto k0c, int par) {
    constexpr int k0 = decltype(k0c)::value;  // first row written: a constant
    if (active_wave && (lane == 0 || lane == 63)) {
      T* dh = lane == 0 ? s_wh[par][wv] : s_eh[par][wv];
      Z* dz = lane == 0 ? s_wz[par][wv] : s_ez[par][wv];
#pragma unroll
      for (int k = k0; k < R + 2; ++k) {
        dh[k] = hh[k];
        dz[k] = zt[k];
      }
      if (lane == 0) {
#pragma unroll
        for (int k = k0; k < R + 2; ++k) {
          s_wq[par][wv][k] = qyo[k];
          s_wn[par][wv][k] = nn[k];
        }
      }
    }
    __syncthreads();
    if (active_wave) {
      h63 = s_eh[par][wv][kk];
      z63 = s_ez[par][wv][kk];
      if (wv > 0) {
        hw = s_eh[par][wv - 1][kk];
        zw = s_ez[par][wv - 1][kk];
      }
      if (wv < 3) {
        he = s_wh[par][wv + 1][kk];
        ze = s_wz[par][wv + 1][kk];
        qe = s_wq[par][wv + 1][kk];
        ne = s_wn[par][wv + 1][kk];
      }
    }
  };
  // new fluxes of the faces east of lane 63, lane k = row k
  T qye = 0;
  auto east_faces = [&]() {
    qye = yface(qe, he, h63, ze, z63, ne, east_m != 0, ycode_e, ymul_e, yadd_e, ork + ce);
  };

  T qxNp = 0, qyNp = 0, qyRp = 0;  // new fluxes of the row above (carried)
  bool mprev = false;                // active-mask predicate of the row above
  // faces of row k, and depth_update + store of row k-1 (k >= 2)
  auto row = [&](uint64_t win, int k, int r0) {
    const int xr = r0 - 1 + k;
    T hl;
    Z zl;
    if constexpr (AL) {
      hl = lane_up1(hh[k], lane_value(hw, k));
      zl = lane_up1(zt[k], lane_value(zw, k));
    } else {
      hl = lane_up1(hh[k]);
      zl = lane_up1(zt[k]);
    }
    const bool m = row_pred(win, k);
    T qxN = qxo[k];
    T qyN = qyo[k];
    if constexpr (MODE == 0) {
      if (m && zt[k] > Z(-9999)) {  // both faces in one branch
        double dummy = 0;
        face64<false, double>(qxN, dummy, hh[k], hh[k - 1], zt[k], zt[k - 1], nn[k], dtd, a,
                      edge_code(xr, a.imax), true, 0, 0, nullptr, nullptr);
        face64<false, double>(qyN, dummy, hh[k], hl, zt[k], zl, nn[k], dtd, a, ycode, false, 0, 0,
                      nullptr, nullptr);
      }
    } else {
      // both faces on every lane, then a select: in a wet grid nearly every
      // lane takes them, and branches would only add exec-mask work
      float dzx;
      if constexpr (MODE == 2) {
        dzx = static_cast<float>(zt[k - 1] - zt[k]);
      } else if constexpr (MODE == 3) {
        // exact integer differences (wrapping only where a no-data cell is
        // involved, whose result is discarded), scaled by a power of two
        dzx = static_cast<float>(static_cast<int32_t>(static_cast<uint32_t>(zt[k - 1]) -
                                                      static_cast<uint32_t>(zt[k]))) *
              a.zq_inv;
      } else {
        dzx = zt[k - 1] - zt[k];
        if (fthr > 0.0f && m && fabsf(dzx) < fthr)  // near-flat: dz from the FP64 DEM
          dzx = static_cast<float>(z64[off(k - 1, r0)] - z64[off(k, r0)]);
      }
      const int xcode = edge_code(xr, a.imax);  // uniform
      float fx = face32(qxo[k], hh[k], hh[k - 1], dzx, nn[k] * nn[k], xcode != 0 ? 0.0f : 1.0f,
                        xcode < 0 ? -es : (xcode > 0 ? es : 0.0f), K);
      asm volatile("" : "+v"(fx));  // on every lane, not under the select
      qxN = and4(m, z_ok(zt[k]), (hh[k] > 0.0f) || (hh[k - 1] > 0.0f), z_ok(zt[k - 1])) ? fx
                                                                                       : qxo[k];
    }
    if constexpr (MODE != 0)
      qyN = yface(qyo[k], hh[k], hl, zt[k], zl, nn[k], m, ycode, ymul, yadd, off(k, r0));
    T qyR;
    if constexpr (AL) qyR = lane_down1(qyN, lane_value(qye, k));
    else qyR = lane_down1(qyN);
    if (k >= 2) {  // row k-1 has both of its x-faces and y-faces now
      const int xo = xr - 1;
      const bool mu = mprev;
      T hn = hh[k - 1];
      if constexpr (MODE == 0) {
        if (mu) {
          hn = hh[k - 1] + dtd * (qxN - qxNp + qyRp - qyNp) / a.DX;
          if (hn > T(0)) {
            if (zt[k - 1] == Z(-9999)) hn = T(0);
            if (hn > md) md = hn;
          }
          sick |= !(hn >= T(hc::kUnstableDepth));
        }
      } else {
        const float t = __builtin_fmaf(dtDX, (qxN - qxNp) + (qyRp - qyNp), hh[k - 1]);
        const bool pos = t > 0.0f;
        const bool zbad = z_nodata(zt[k - 1]);
        hn = mu ? ((pos && zbad) ? 0.0f : t) : hh[k - 1];
        md = fmaxf(md, (mu && pos && !zbad) ? t : 0.0f);
        sick |= mu && !(t >= static_cast<float>(hc::kUnstableDepth));
      }
      if (drains && owned && xo >= 1 && xo <= a.imax && c >= 1 && c <= a.jmax &&
          (xo == 1 || xo == a.imax || c == 1 || c == a.jmax)) {
        const double hv = static_cast<double>(hn);
        if (hv > drain.thr) {  // counted out down to the stored threshold
          dacc += (hv - static_cast<double>(static_cast<T>(drain.thr))) * a.DX * a.DX / ts;
          hn = static_cast<T>(drain.thr);
        }
      }
      if (owned && xo < NR) {
        const uint32_t o = off(k - 1, r0);
        st_nt(h1, o, hn);
        st_nt(qx1, o, qxNp);
        st_nt(qy1, o, qyNp);
      }
      if constexpr (wetb) {
        {  // row xo's word (the wavefront's 64 columns) waits in lane k-2
          const uint64_t bits = __ballot(owned && hn > T(0));
          const bool mine = lane == k - 2;
          wet_lo = mine ? static_cast<uint32_t>(bits) : wet_lo;
          wet_hi = mine ? static_cast<uint32_t>(bits >> 32) : wet_hi;
        }
      }
    }
    qxNp = qxN;
    qyNp = qyN;
    qyRp = qyR;
    mprev = m;
  };

  // first sub-tile: rows r0-1 .. r0+R loaded, faces of rows r0 .. r0+R
  int r0 = by * H;
  uint64_t win = 0;
  if (active_wave) {
    win = load_win(r0);
#pragma unroll
    for (int k = 0; k < R + 2; ++k) load_row(k, r0);
    load_n(1, r0);
    if constexpr (AL) load_outer(r0);
  }
  asm volatile("" ::: "memory");  // the loads above stay ahead of the loop-state reads
  if (haltp && *haltp) return;  // the loop has finished: uniform over the grid
  if (dtp) dtd = *dtp;
  if (drains) ts = *drain.ts;
  if (inject && p.inj.rf_uniform >= 0) amt_u = p.inj.amt[p.inj.rf_uniform];
  dt_setup();
  if (active_wave && inject) {
    if (any_point(win)) {  // uniform: input points are sparse, most sub-tiles have none
#pragma unroll
      for (int k = 0; k < R + 2; ++k) hh[k] = rain(hh[k], row_pred(win, 32 + k), off(k, r0));
    }
    if constexpr (AL) rain_outer(r0);
  }
  if constexpr (AL) exchange(std::integral_constant<int, 1>{}, 0);
  if (active_wave) {
    if constexpr (AL) east_faces();
#pragma unroll
    for (int k = 1; k < R + 2; ++k) row(win, k, r0);
    if constexpr (wetb) flush_wet(r0);
  }
  for (int t = 1; t < nsub; ++t) {
    r0 += R;
    if (r0 > NR) break;  // uniform: no row of this sub-tile is in the grid
    if (active_wave) {
      // rows r0-1 and r0 were the last two rows of the sub-tile before (their
      // rain input included), and the faces of row r0 are already computed
      // (qxNp, qyNp, qyRp)
      hh[0] = hh[R];
      hh[1] = hh[R + 1];
      zt[0] = zt[R];
      zt[1] = zt[R + 1];
      win = load_win(r0);
#pragma unroll
      for (int k = 2; k < R + 2; ++k) load_row(k, r0);
      load_n(2, r0);
      if constexpr (AL) load_outer(r0);
      if (inject) {
        if (any_point(win)) {
#pragma unroll
          for (int k = 2; k < R + 2; ++k) hh[k] = rain(hh[k], row_pred(win, 32 + k), off(k, r0));
        }
        if constexpr (AL) rain_outer(r0);
      }
    }
    if constexpr (AL) exchange(std::integral_constant<int, 2>{}, t & 1);
    if (active_wave) {
      if constexpr (AL) east_faces();
#pragma unroll
      for (int k = 2; k < R + 2; ++k) row(win, k, r0);
      if constexpr (wetb) flush_wet(r0);
    }
  }
  const double mine = owned && active_wave ? (sick ? INFINITY : static_cast<double>(md)) : 0.0;
  if (p.wmax) {  // uniform
    const double wm = wave_max(mine);
    if (lane == 0) p.wmax[4 * blockIdx.x + wv] = wm;
  } else {
    block_max_pos(red, mine);
  }
  if (drains) {  // uniform per block
    __shared__ double s_dw[4];
    dacc = wave_sum(dacc);
    if (lane == 0) s_dw[threadIdx.x >> 6] = dacc;
    __syncthreads();
    if (threadIdx.x == 0)
      drain.partials[hc::EdgeSlots(nbx, H, a.imax, a.jmax, SW).slot(bx, by)] =
          (s_dw[0] + s_dw[1]) + (s_dw[2] + s_dw[3]);
  }
}

// One grid: one block per tile, block ids remapped so that each XCD (block
// b runs on XCD b mod 8) takes a contiguous band of tiles in row-major order.
// (Orders that walk groups of tile columns, so that vertically adjacent
// tiles run close together, measured slower: the row-major stream keeps the
// DRAM accesses sequential.)
template <int MODE, int R, bool WET = false>
__global__ __launch_bounds__(256) void k_fused(HydroArgs a, double dt_arg, int nbx, int nblocks,
                                               int nsub, FusedPtrs<MODE> p,
                                               const double* __restrict__ dtp,
                                               const int* __restrict__ haltp,
                                               const DrainArgs drain) {
  const int b = blockIdx.x;
  const int q8 = nblocks / 8, r8 = nblocks % 8;
  const int xcd = b % 8;
  const int bb = xcd * q8 + min(xcd, r8) + b / 8;
  fused_tile<MODE, R, MODE != 0, WET>(a, bb % nbx, bb / nbx, nbx, nsub, dt_arg, dtp, haltp, p,
                                      drain);
}

// Ensemble: a persistent grid walks the ensemble's tile list (member-major).
// Each XCD takes a contiguous eighth of the list and each block a contiguous
// run of it; a member's own dt, time step and halt flag come from its loop
// state, and its generation for this step from its parity at run start.
template <int MODE, int R>
__global__ __launch_bounds__(256) void k_ens_fused(const EnsMember* __restrict__ mem,
                                                   const int2* __restrict__ tiles, int ntiles,
                                                   int step_parity) {
  using T = typename FusedTypes<MODE>::T;
  using Z = typename FusedTypes<MODE>::Z;
  const int nb = gridDim.x, b = blockIdx.x;
  const int q8 = nb / 8, r8 = nb % 8;
  const int xcd = b % 8;
  const int bb = xcd * q8 + min(xcd, r8) + b / 8;  // rank of this block in list order
  const long t0 = static_cast<long>(ntiles) * bb / nb;
  const long t1 = static_cast<long>(ntiles) * (bb + 1) / nb;
  for (long t = t0; t < t1; ++t) {
    const int2 tile = tiles[t];
    const EnsMember& M = mem[tile.x];
    const hc::LoopCtl* ctl = table_ptr<const hc::LoopCtl>(M.ctl);
    // a finished member's tiles are skipped before any of their loads
    if (ctl->halt) continue;  // uniform per block
    const int g = M.cur0 ^ step_parity;
    FusedPtrs<MODE> p;
    p.h0 = table_ptr<const T>(M.h[g]);
    p.zz = table_ptr<const Z>(M.z);
    p.z64 = table_ptr<const double>(M.z64);
    p.qx0 = table_ptr<const T>(M.qx[g]);
    p.qy0 = table_ptr<const T>(M.qy[g]);
    p.nman = table_ptr<const T>(M.nman);
    p.mask = table_ptr<const uint64_t>(M.mask);
    p.h1 = table_ptr<T>(M.h[g ^ 1]);
    p.qx1 = table_ptr<T>(M.qx[g ^ 1]);
    p.qy1 = table_ptr<T>(M.qy[g ^ 1]);
    p.red = table_ptr<unsigned long long>(M.red);
    p.wmax = nullptr;
    p.inj.pmask = table_ptr<const uint64_t>(M.pmask);
    p.inj.rf_uniform = M.rf_uniform;
    p.inj.rfarea = table_ptr<const int>(M.rfarea);
    p.inj.amt = ctl->amt;
    DrainArgs d;
    d.ts = &ctl->time_step;
    d.thr = M.thr;
    d.partials = table_ptr<double>(M.dpart);
    __syncthreads();  // the block's shared scratch is reused tile to tile
    fused_tile<MODE, R, false>(M.ha, tile.y % M.nbx, tile.y / M.nbx, M.nbx, 1, 0.0,
                               &ctl->dt_flow, nullptr, p, d);
  }
}

// ---------------------------------------------------------------- scan
template <typename T>
__global__ __launch_bounds__(256) void k_scan(int imax, int jmax, int P, int W,
                                              const T* __restrict__ h,
                                              uint64_t* __restrict__ mask) {
  const int items = scan_items(imax, W, static_cast<int>(sizeof(T)));
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform
  for (int it = blockIdx.x * 4 + wv; it < items; it += gridDim.x * 4)
    scan_item<T>(it, imax, jmax, P, W, h, mask);
}

// --------------------------------------------------------- water_flux_out
template <typename T>
__global__ __launch_bounds__(256) void k_flux_out(int imax, int jmax, int P, double DX,
                                                  double thr, double ts, T* __restrict__ h,
                                                  double* __restrict__ partials,
                                                  unsigned int* __restrict__ count,
                                                  double* __restrict__ out) {
  const double bs = flux_block<T>(imax, jmax, P, DX, thr, ts, h);
  double tot;
  if (grid_sum_last(bs, partials, count, &tot) && threadIdx.x == 0) out[0] = tot;
}

// ------------------------------------------------------------- points
template <typename T>
__global__ void k_points(int n, const int* __restrict__ idx, const double* __restrict__ amt,
                         T* __restrict__ h) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    T* p = &h[idx[k]];
    *p = static_cast<T>(static_cast<double>(*p) + amt[k]);
  }
}

// Point inputs of the persistent list: the amount by the point's rain cell.
template <typename T>
__global__ void k_points_rc(int n, const int* __restrict__ idx, const int* __restrict__ rc,
                            const double* __restrict__ amt, T* __restrict__ h) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    T* p = &h[idx[k]];
    *p = static_cast<T>(static_cast<double>(*p) + amt[rc[k]]);
  }
}

// Point inputs where a cell may appear more than once: the reference adds
// the amounts one after another in list order (LCM.cpp:3463-3477).  Every
// point scans the whole list through LDS; the first occurrence of a cell adds
// its own amount and then those of the later points on that cell, in list
// order, and writes the cell once; later occurrences write nothing.  Cells
// come from idx (x*P+y) or, when idx is null, from the x and y lists.  Costs
// n/256 LDS passes per block: for short lists and lists with repeats.
template <typename T>
__global__ __launch_bounds__(256) void k_points_ordered(int n, int P, const int* __restrict__ idx,
                                                        const int* __restrict__ xs,
                                                        const int* __restrict__ ys,
                                                        const double* __restrict__ amt,
                                                        T* __restrict__ h) {
  __shared__ int s_cell[256];
  const int k = blockIdx.x * 256 + threadIdx.x;
  const bool live = k < n;
  const int my = live ? (idx ? idx[k] : xs[k] * P + ys[k]) : -1;
  bool first = live;
  T v = T(0);
  if (live) v = static_cast<T>(static_cast<double>(h[my]) + amt[k]);
  for (int base = 0; base < n; base += 256) {
    const int j = base + static_cast<int>(threadIdx.x);
    __syncthreads();
    s_cell[threadIdx.x] = (j < n) ? (idx ? idx[j] : xs[j] * P + ys[j]) : -2;
    __syncthreads();
    const int m = min(256, n - base);
    for (int t = 0; t < m; ++t) {
      if (s_cell[t] != my) continue;  // cells are >= 0: -1 and -2 never match
      const int jj = base + t;
      if (jj < k)
        first = false;
      else if (jj > k && first)
        v = static_cast<T>(static_cast<double>(v) + amt[jj]);
    }
  }
  if (first) h[my] = v;
}

// ------------------------------------------------------------- health
template <typename T>
__global__ __launch_bounds__(256) void k_health(int imax, int jmax, int P, const T* __restrict__ h,
                                                unsigned long long* __restrict__ bad) {
  const int y = blockIdx.x * 64 + threadIdx.x;
  const int x = blockIdx.y * blockDim.y + threadIdx.y;
  bool b = false;
  if (x >= 1 && x <= imax && y >= 1 && y <= jmax) {
    const double v = static_cast<double>(h[static_cast<size_t>(x) * P + y]);
    b = !(v >= 0.0) || isinf(v);
  }
  const uint64_t m = __ballot(b);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(bad, static_cast<unsigned long long>(__popcll(m)));
}

template <typename T>
__global__ void k_fill(T* p, size_t n, T v) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

HydroArgs hydro_args(const hc_ctx* c) {
  HydroArgs a;
  a.imax = c->g.imax;
  a.jmax = c->g.jmax;
  a.pitch = c->P;
  a.mwords = c->W;
  a.DX = c->g.DX;
  a.hthr = c->prm.hflow_threshold;
  a.edgeslope = c->prm.edgeslope;
  a.froude = c->prm.froude_limit;
  a.mannings = c->prm.mannings;
  a.spatial_n = c->prm.spatially_var_mannings;
  a.flat_dz = c->prm.flat_dz;
  a.zq_inv = std::ldexp(1.0f, -c->zq_shift);
  return a;
}

dim3 cell_grid(const hc_ctx* c) {
  return dim3(static_cast<unsigned>(c->P / 64), static_cast<unsigned>((c->NR + 3) / 4));
}

}  // namespace


// The separate flow_route and depth_update: FP64 contexts, and the FP32
// storage form of erosion runs (FP64 elevations; float depths, fluxes and
// sediment planes, FP64 arithmetic).
template <typename T>
void launch_flow_route(hc_ctx* c, const HydroArgs& a, double dt) {
  const int k = c->cur;
  const dim3 blk(64, 4);
  if (c->g.sediment)
    hipLaunchKernelGGL((k_flow_route64<true, T>), cell_grid(c), blk, 0, c->stream, a, dt,
                       (const T*)c->h[k], (const double*)c->z, (T*)c->qx[k], (T*)c->qy[k],
                       (const T*)c->nman, c->mask, (T*)c->qxs, (T*)c->qys, (const T*)c->V,
                       (T*)c->veldir, c->ncell);
  else
    hipLaunchKernelGGL((k_flow_route64<false, T>), cell_grid(c), blk, 0, c->stream, a, dt,
                       (const T*)c->h[k], (const double*)c->z, (T*)c->qx[k], (T*)c->qy[k],
                       (const T*)c->nman, c->mask, nullptr, nullptr, nullptr, nullptr, c->ncell);
}

template <typename T>
void launch_depth_update(hc_ctx* c, const HydroArgs& a, double dt) {
  const int k = c->cur;
  const dim3 blk(64, 4);
  if (c->g.sediment)
    hipLaunchKernelGGL((k_depth64<true, T>), cell_grid(c), blk, 0, c->stream, a, dt, (T*)c->h[k],
                       (const double*)c->z, (const T*)c->qx[k], (const T*)c->qy[k], c->mask,
                       (T*)c->V, (const T*)c->qxs, (const T*)c->qys, c->d_red);
  else
    hipLaunchKernelGGL((k_depth64<false, T>), cell_grid(c), blk, 0, c->stream, a, dt, (T*)c->h[k],
                       (const double*)c->z, (const T*)c->qx[k], (const T*)c->qy[k], c->mask,
                       nullptr, nullptr, nullptr, c->d_red);
}

int hc_launch_flow_route(hc_ctx* c, double dt) {
  if (c->prec != HC_FP64 && !c->g.sediment)
    return hc_fail(c, HC_ERR_STATE, "flow_route: FP32 hydro-only contexts step with hc_flow_depth");
  const HydroArgs a = hydro_args(c);
  if (c->prec == HC_FP64) launch_flow_route<double>(c, a, dt);
  else launch_flow_route<float>(c, a, dt);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_launch_depth_update(hc_ctx* c, double dt) {
  if (c->prec != HC_FP64 && !c->g.sediment)
    return hc_fail(c, HC_ERR_STATE, "depth_update: FP32 hydro-only contexts step with hc_flow_depth");
  const HydroArgs a = hydro_args(c);
  HC_HIP(c, hipMemsetAsync(c->d_red, 0, sizeof(unsigned long long), c->stream));
  if (c->prec == HC_FP64) launch_depth_update<double>(c, a, dt);
  else launch_depth_update<float>(c, a, dt);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

hc::HydroArgs hc_hydro_args(const hc_ctx* c) { return hydro_args(c); }

const void* hc_fused_z(const hc_ctx* c) {
  if (c->prec == HC_FP32 && c->zq) return c->zq;
  return (c->prec == HC_FP32 && c->prm.flat_dz < 0) ? static_cast<const void*>(c->z64) : c->z;
}

int hc_launch_ens_fused(hipStream_t st, int mode, int R, const EnsMember* mem, const int2* tiles,
                        int ntiles, int nblocks, int parity) {
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nblocks), dim3(256), 0, st, mem, tiles, ntiles, parity);
  };
  if (mode == 0) {
    if (R == 4) go(k_ens_fused<0, 4>);
    else if (R == 16) go(k_ens_fused<0, 16>);
    else go(k_ens_fused<0, 8>);
  } else if (mode == 2) {
    if (R == 4) go(k_ens_fused<2, 4>);
    else if (R == 16) go(k_ens_fused<2, 16>);
    else go(k_ens_fused<2, 8>);
  } else if (mode == 3) {
    if (R == 4) go(k_ens_fused<3, 4>);
    else if (R == 16) go(k_ens_fused<3, 16>);
    else go(k_ens_fused<3, 8>);
  } else {
    if (R == 4) go(k_ens_fused<1, 4>);
    else if (R == 16) go(k_ens_fused<1, 16>);
    else go(k_ens_fused<1, 8>);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int hc_fused_stride(const hc_ctx* c) {
  return c->prec == HC_FP64 ? hc::kStrideOverlap : hc::kStrideAligned;
}

void hc_fused_geometry(const hc_ctx* c, int R, int sw, int* nbx, int* nby) {
  const int nwaves = sw == hc::kStrideAligned ? (c->NC + 63) / 64 : (c->NC - 1 + 61) / 62;
  *nbx = (nwaves + 3) / 4;
  *nby = (c->NR + R - 1) / R;
}

int hc_launch_flow_depth(hc_ctx* c, double dt, const double* dtp, const int* haltp, bool swap,
                         const DrainArgs* drain, const hc::InjectArgs* inject, double* wmax,
                         hipEvent_t ev_start, hipEvent_t ev_stop, uint64_t* wet) {
  if (c->g.sediment)
    return hc_fail(c, HC_ERR_STATE, "flow_depth: fused kernel is for hydro-only contexts");
  if (!hc::fused_offsets_fit(c->NR, c->P))
    return hc_fail(c, HC_ERR_ARG, "flow_depth: grid too large for the fused kernel's 32-bit offsets");
  const HydroArgs a = hydro_args(c);
  const int R = c->fused_rows;
  const int ns = c->fused_strip;
  int nbx, nby;
  hc_fused_geometry(c, R * ns, hc_fused_stride(c), &nbx, &nby);
  const int nblocks = nbx * nby;
  if (wmax && static_cast<size_t>(4) * nblocks > c->wmax_cap)
    return hc_fail(c, HC_ERR_STATE, "flow_depth: wavefront maxima buffer too small");
  const DrainArgs dr = drain ? *drain : DrainArgs{};
  if (dr.ts && static_cast<size_t>(hc::EdgeSlots(nbx, R * ns, c->g.imax, c->g.jmax,
                                                  hc_fused_stride(c)).count(nby)) >
                   c->dpart_cap)
    return hc_fail(c, HC_ERR_STATE, "flow_depth: drain partials buffer too small");
  const int k = c->cur, o = c->cur ^ 1;
  // in the device-resident loop the control kernel clears the reduction slot
  if (!dtp) HC_HIP(c, hipMemsetAsync(c->d_red, 0, sizeof(unsigned long long), c->stream));
  auto go = [&](auto kern, auto* tag_t, auto* tag_z) {
    using T = std::remove_pointer_t<decltype(tag_t)>;
    using Z = std::remove_pointer_t<decltype(tag_z)>;
    const Z* zp = reinterpret_cast<const Z*>(hc_fused_z(c));
    constexpr int M = std::is_same<T, double>::value
                          ? 0
                          : (std::is_same<Z, double>::value ? 2 : (std::is_same<Z, float>::value ? 1 : 3));
    FusedPtrs<M> p;
    p.h0 = (const T*)c->h[k];
    p.zz = zp;
    p.z64 = (const double*)c->z64;
    p.qx0 = (const T*)c->qx[k];
    p.qy0 = (const T*)c->qy[k];
    p.nman = (const T*)c->nman;
    p.mask = (const uint64_t*)c->mask;
    p.h1 = (T*)c->h[o];
    p.qx1 = (T*)c->qx[o];
    p.qy1 = (T*)c->qy[o];
    p.red = c->d_red;
    p.wmax = wmax;
    p.inj = inject ? *inject : hc::InjectArgs{};
    p.wet = wet;
    if (ev_start)  // the events time the kernel's own execution (dispatch packet timestamps)
      hipExtLaunchKernelGGL(kern, dim3(nblocks), dim3(256), 0, c->stream, ev_start, ev_stop, 0, a,
                            dt, nbx, nblocks, ns, p, dtp, haltp, dr);
    else
      hipLaunchKernelGGL(kern, dim3(nblocks), dim3(256), 0, c->stream, a, dt, nbx, nblocks, ns, p,
                         dtp, haltp, dr);
  };
  double* td = nullptr;
  float* tf = nullptr;
  int32_t* tq = nullptr;
  if (c->prec == HC_FP64) {
    if (R == 4) go(k_fused<0, 4>, td, td);
    else if (R == 16) go(k_fused<0, 16>, td, td);
    else go(k_fused<0, 8>, td, td);
  } else if (wet) {  // FP32 forms on a body that rebuilds the active mask
    if (c->zq) {
      if (R == 4) go(k_fused<3, 4, true>, tf, tq);
      else if (R == 16) go(k_fused<3, 16, true>, tf, tq);
      else go(k_fused<3, 8, true>, tf, tq);
    } else if (c->prm.flat_dz < 0) {
      if (R == 4) go(k_fused<2, 4, true>, tf, td);
      else if (R == 16) go(k_fused<2, 16, true>, tf, td);
      else go(k_fused<2, 8, true>, tf, td);
    } else {
      if (R == 4) go(k_fused<1, 4, true>, tf, tf);
      else if (R == 16) go(k_fused<1, 16, true>, tf, tf);
      else go(k_fused<1, 8, true>, tf, tf);
    }
  } else if (c->zq) {
    if (R == 4) go(k_fused<3, 4>, tf, tq);
    else if (R == 16) go(k_fused<3, 16>, tf, tq);
    else go(k_fused<3, 8>, tf, tq);
  } else if (c->prm.flat_dz < 0) {
    if (R == 4) go(k_fused<2, 4>, tf, td);
    else if (R == 16) go(k_fused<2, 16>, tf, td);
    else go(k_fused<2, 8>, tf, td);
  } else {
    if (R == 4) go(k_fused<1, 4>, tf, tf);
    else if (R == 16) go(k_fused<1, 16>, tf, tf);
    else go(k_fused<1, 8>, tf, tf);
  }
  HC_HIP(c, hipGetLastError());
  if (swap) c->cur = o;
  return HC_OK;
}

// Tile tuner: time the fused kernel for each candidate row count on the
// context's current state (the kernel reads one generation and writes the
// other, so repeated launches without swapping leave the state unchanged).
extern "C" int hc_tune_tiles(hc_ctx* c, double dt, int* best_rows, int* best_strip,
                             double* best_us) {
  if (!c) return HC_ERR_ARG;
  if (c->g.sediment) return hc_fail(c, HC_ERR_STATE, "tune: hydro-only contexts");
  // the reduction slot carries the last maxdepth into a device-resident loop
  unsigned long long saved = 0;
  HC_HIP(c, hipStreamSynchronize(c->stream));
  HC_HIP(c, hipMemcpy(&saved, c->d_red, sizeof(saved), hipMemcpyDeviceToHost));
  hipEvent_t e0, e1;
  HC_HIP(c, hipEventCreate(&e0));
  HC_HIP(c, hipEventCreate(&e1));
  double best = 1e30;
  int bestR = c->fused_rows, bestS = c->fused_strip;
  for (int R : {4, 8, 16}) {
    for (int S : {1, 2, 4, 8}) {
      c->fused_rows = R;
      c->fused_strip = S;
      int rc = hc_launch_flow_depth(c, dt, nullptr, nullptr, false);  // warm
      if (rc) return rc;
      HC_HIP(c, hipEventRecord(e0, c->stream));
      const int reps = 10;
      for (int k = 0; k < reps; ++k) {
        rc = hc_launch_flow_depth(c, dt, nullptr, nullptr, false);
        if (rc) return rc;
      }
      HC_HIP(c, hipEventRecord(e1, c->stream));
      HC_HIP(c, hipEventSynchronize(e1));
      float ms = 0;
      HC_HIP(c, hipEventElapsedTime(&ms, e0, e1));
      const double us = 1000.0 * ms / reps;
      if (us < best) {
        best = us;
        bestR = R;
        bestS = S;
      }
    }
  }
  HC_HIP(c, hipMemcpy(c->d_red, &saved, sizeof(saved), hipMemcpyHostToDevice));
  hc_drop_loop_graphs(c);  // recapture with the chosen tiles
  c->fused_rows = bestR;
  c->fused_strip = bestS;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best_rows) *best_rows = bestR;
  if (best_strip) *best_strip = bestS;
  if (best_us) *best_us = best;
  return HC_OK;
}

extern "C" int hc_set_tiles(hc_ctx* c, int rows, int strip) {
  if (!c || (rows != 4 && rows != 8 && rows != 16) || strip < 1 || strip > 64) return HC_ERR_ARG;
  if (rows != c->fused_rows || strip != c->fused_strip) hc_drop_loop_graphs(c);
  c->fused_rows = rows;
  c->fused_strip = strip;
  return HC_OK;
}

extern "C" int hc_bench_fused(hc_ctx* c, double dt, int reps, double* us) {
  if (!c || reps < 1) return HC_ERR_ARG;
  unsigned long long saved = 0;
  HC_HIP(c, hipStreamSynchronize(c->stream));
  HC_HIP(c, hipMemcpy(&saved, c->d_red, sizeof(saved), hipMemcpyDeviceToHost));
  hipEvent_t e0, e1;
  HC_HIP(c, hipEventCreate(&e0));
  HC_HIP(c, hipEventCreate(&e1));
  int rc = hc_launch_flow_depth(c, dt, nullptr, nullptr, false);
  if (rc) return rc;
  HC_HIP(c, hipEventRecord(e0, c->stream));
  for (int k = 0; k < reps; ++k) {
    rc = hc_launch_flow_depth(c, dt, nullptr, nullptr, false);
    if (rc) return rc;
  }
  HC_HIP(c, hipEventRecord(e1, c->stream));
  HC_HIP(c, hipEventSynchronize(e1));
  float ms = 0;
  HC_HIP(c, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HC_HIP(c, hipMemcpy(c->d_red, &saved, sizeof(saved), hipMemcpyHostToDevice));
  if (us) *us = 1000.0 * ms / reps;
  return HC_OK;
}


int hc_launch_fused64_ptrs(hc_ctx* c, double dt, const double* h0, const double* qx0,
                           const double* qy0, double* h1, double* qx1, double* qy1,
                           const uint64_t* mask) {
  if (c->prec != HC_FP64 || c->g.sediment)
    return hc_fail(c, HC_ERR_STATE, "fused64: FP64 hydro-only contexts");
  if (!hc::fused_offsets_fit(c->NR, c->P))
    return hc_fail(c, HC_ERR_ARG, "fused64: grid too large for the fused kernel's 32-bit offsets");
  const HydroArgs a = hydro_args(c);
  const int R = c->fused_rows;
  const int ns = c->fused_strip;
  int nbx, nby;
  hc_fused_geometry(c, R * ns, hc_fused_stride(c), &nbx, &nby);
  const int nblocks = nbx * nby;
  FusedPtrs<0> p;
  p.h0 = h0;
  p.zz = static_cast<const double*>(c->z);
  p.z64 = nullptr;
  p.qx0 = qx0;
  p.qy0 = qy0;
  p.nman = static_cast<const double*>(c->nman);
  p.mask = mask;
  p.h1 = h1;
  p.qx1 = qx1;
  p.qy1 = qy1;
  p.red = c->d_red;
  p.wmax = nullptr;
  p.inj = hc::InjectArgs{};
  const DrainArgs dr{};
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nblocks), dim3(256), 0, c->stream, a, dt, nbx, nblocks, ns, p,
                       nullptr, nullptr, dr);
  };
  if (R == 4) go(k_fused<0, 4>);
  else if (R == 16) go(k_fused<0, 16>);
  else go(k_fused<0, 8>);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_launch_scan_ptr(hc_ctx* c, const void* h, uint64_t* mask) {
  const int items = scan_items(c->g.imax, c->W, static_cast<int>(c->esz));
  const int nb = std::max(1, std::min((items + 3) / 4, 2048));
  if (c->prec == HC_FP64)
    hipLaunchKernelGGL(k_scan<double>, dim3(nb), dim3(256), 0, c->stream, c->g.imax, c->g.jmax,
                       c->P, c->W, static_cast<const double*>(h), mask);
  else
    hipLaunchKernelGGL(k_scan<float>, dim3(nb), dim3(256), 0, c->stream, c->g.imax, c->g.jmax,
                       c->P, c->W, static_cast<const float*>(h), mask);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_launch_scan(hc_ctx* c) {
  const int k = c->cur;
  const int items = scan_items(c->g.imax, c->W, static_cast<int>(c->esz));
  const int nb = std::max(1, std::min((items + 3) / 4, 2048));
  if (c->prec == HC_FP64)
    hipLaunchKernelGGL(k_scan<double>, dim3(nb), dim3(256), 0, c->stream, c->g.imax, c->g.jmax,
                       c->P, c->W, (const double*)c->h[k], c->mask);
  else
    hipLaunchKernelGGL(k_scan<float>, dim3(nb), dim3(256), 0, c->stream, c->g.imax, c->g.jmax,
                       c->P, c->W, (const float*)c->h[k], c->mask);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_launch_flux_out(hc_ctx* c, double ts) {
  const int imax = c->g.imax, jmax = c->g.jmax;
  const long nedge = jmax + (imax > 1 ? jmax : 0) + (imax > 2 ? (long)(imax - 2) * (jmax > 1 ? 2 : 1) : 0);
  int nb = static_cast<int>((nedge + 255) / 256);
  if (nb > 256) nb = 256;
  if (nb < 1) nb = 1;
  const double thr = c->prm.water_depth_erosion_threshold;
  if (c->prec == HC_FP64)
    hipLaunchKernelGGL(k_flux_out<double>, dim3(nb), dim3(256), 0, c->stream, imax, jmax, c->P,
                       c->g.DX, thr, ts, (double*)c->h[c->cur], c->d_partials, c->d_count,
                       c->d_out);
  else
    hipLaunchKernelGGL(k_flux_out<float>, dim3(nb), dim3(256), 0, c->stream, imax, jmax, c->P,
                       c->g.DX, thr, ts, (float*)c->h[c->cur], c->d_partials, c->d_count,
                       c->d_out);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_launch_points(hc_ctx* c, int n) {
  if (n <= 0) return HC_OK;
  const int nb = (n + 255) / 256;
  if (c->prec == HC_FP64)
    hipLaunchKernelGGL(k_points<double>, dim3(nb), dim3(256), 0, c->stream, n, c->d_pts,
                       c->d_amt, (double*)c->h[c->cur]);
  else
    hipLaunchKernelGGL(k_points<float>, dim3(nb), dim3(256), 0, c->stream, n, c->d_pts,
                       c->d_amt, (float*)c->h[c->cur]);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_launch_points_ordered(hc_ctx* c, int n, const int* idx, const int* xs, const int* ys,
                             const double* amt) {
  if (n <= 0) return HC_OK;
  const int nb = (n + 255) / 256;
  if (c->prec == HC_FP64)
    hipLaunchKernelGGL(k_points_ordered<double>, dim3(nb), dim3(256), 0, c->stream, n, c->P, idx,
                       xs, ys, amt, (double*)c->h[c->cur]);
  else
    hipLaunchKernelGGL(k_points_ordered<float>, dim3(nb), dim3(256), 0, c->stream, n, c->P, idx,
                       xs, ys, amt, (float*)c->h[c->cur]);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_launch_points_rc(hc_ctx* c, int n, const int* idx, const int* rc, const double* amt) {
  if (n <= 0) return HC_OK;
  const int nb = (n + 255) / 256;
  if (c->prec == HC_FP64)
    hipLaunchKernelGGL(k_points_rc<double>, dim3(nb), dim3(256), 0, c->stream, n, idx, rc, amt,
                       (double*)c->h[c->cur]);
  else
    hipLaunchKernelGGL(k_points_rc<float>, dim3(nb), dim3(256), 0, c->stream, n, idx, rc, amt,
                       (float*)c->h[c->cur]);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_launch_health(hc_ctx* c) {
  const dim3 blk(64, 4);
  HC_HIP(c, hipMemsetAsync(c->d_red + 2, 0, sizeof(unsigned long long), c->stream));
  if (c->prec == HC_FP64)
    hipLaunchKernelGGL(k_health<double>, cell_grid(c), blk, 0, c->stream, c->g.imax, c->g.jmax,
                       c->P, (const double*)c->h[c->cur], c->d_red + 2);
  else
    hipLaunchKernelGGL(k_health<float>, cell_grid(c), blk, 0, c->stream, c->g.imax, c->g.jmax,
                       c->P, (const float*)c->h[c->cur], c->d_red + 2);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_launch_fill(hc_ctx* c, void* p, size_t n, double v, bool is_float) {
  if (n == 0) return HC_OK;
  const unsigned nb = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 4096));
  if (is_float)
    hipLaunchKernelGGL(k_fill<float>, dim3(nb), dim3(256), 0, c->stream, (float*)p, n,
                       static_cast<float>(v));
  else
    hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(256), 0, c->stream, (double*)p, n, v);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}
