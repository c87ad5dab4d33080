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

#include "hc_internal.hpp"

using hc::HydroArgs;
using hc::kGravity;

namespace {

// ---------------------------------------------------------------- helpers
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

// One face of the FP64 plain form (LSDCatchmentModel.cpp:3094-3191 for the
// x-face, 3195-3242 for the y-face).  Cell values (hc, zc), neighbour across
// the face (hn, zn).  `edge` is -1 for x<=2 / y<=2, +1 for x==imax /
// y==jmax, else 0 (the `<= 2` override is applied last in the reference).
// Leaves q (and qs) unchanged unless the face condition holds.
template <bool SED>
__device__ __forceinline__ void face64(double& q, double& qs, double hc, double hn, double zc,
                                       double zn, double n, double dt, const HydroArgs& a,
                                       int edge, bool xface, double Vc, double Vn,
                                       double* vel_pos, double* vel_neg) {
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

// One face of the FP32 form.  dz = z_nb - z_cell.  Same branch structure as
// face64; limiter tests are written as products instead of quotients.
__device__ __forceinline__ float face32(float q, float hc, float hn, float dz, float n, float dt,
                                        float invdt, const HydroArgs& a, int edge) {
  const float hflow = (dz >= 0.0f) ? fmaxf(hc - dz, hn) : fmaxf(hc, hn + dz);
  if (!(hflow > static_cast<float>(a.hthr))) return 0.0f;
  const float DX = static_cast<float>(a.DX);
  const float invDX = 1.0f / DX;
  float slope = (dz + (hn - hc)) * invDX;
  if (edge > 0) slope = static_cast<float>(a.edgeslope);
  if (edge < 0) slope = -static_cast<float>(a.edgeslope);
  const float gh = static_cast<float>(kGravity) * hflow;
  const float ghdt = gh * dt;
  const float h3 = hflow * hflow * hflow;
  const float num = q - ghdt * slope;
  const float den = 1.0f + ghdt * (n * n) * fabsf(q) * __builtin_amdgcn_rcpf(h3);
  float qv = num * __builtin_amdgcn_rcpf(den);
  const float qF = hflow * (__builtin_amdgcn_sqrtf(gh) * static_cast<float>(a.froude));
  if (qv > qF) qv = qF;
  if (qv < -qF) qv = -qF;
  const float k4 = 0.25f * DX * invdt;   // q*dt/DX > h/4  <=>  q > h*DX/(4 dt)
  const float k5 = 0.2f * DX * invdt;
  if (qv > 0.0f && qv > hc * k4) qv = hc * k5;
  if (qv < 0.0f && -qv > hn * k4) qv = -(hn * k5);
  return qv;
}

__device__ __forceinline__ int edge_code(int i, int imaxlike) {
  return (i <= 2) ? -1 : ((i == imaxlike) ? 1 : 0);
}

// ------------------------------------------------------------ flow_route
// In-place per-cell map (each face is written by exactly one cell).
template <bool SED>
__global__ __launch_bounds__(256) void k_flow_route64(HydroArgs a, double dt,
                                                      const double* __restrict__ h,
                                                      const double* __restrict__ z,
                                                      double* __restrict__ qx,
                                                      double* __restrict__ qy,
                                                      const double* __restrict__ nman,
                                                      const uint64_t* __restrict__ mask,
                                                      double* __restrict__ qxs,
                                                      double* __restrict__ qys,
                                                      const double* __restrict__ V,
                                                      double* __restrict__ vd, size_t plane) {
  const int y = blockIdx.x * 64 + threadIdx.x;
  const int x = blockIdx.y * blockDim.y + threadIdx.y;
  if (x < 1 || x > a.imax || y < 1 || y > a.jmax) return;
  if (!mask_bit(mask, a.mwords, x, y)) return;
  const size_t P = a.pitch;
  const size_t c = static_cast<size_t>(x) * P + y;
  const double zc = z[c];
  if (!(zc > -9999)) return;
  const double n = a.spatial_n ? nman[c] : a.mannings;
  const double hc = h[c];
  double qv, qsv = 0;
  // x face (neighbour x-1)
  {
    const size_t cn = c - P;
    qv = qx[c];
    if (SED) qsv = qxs[c];
    face64<SED>(qv, qsv, hc, h[cn], zc, z[cn], n, dt, a, edge_code(x, a.imax), true,
                SED ? V[c] : 0.0, SED ? V[cn] : 0.0, SED ? &vd[3 * plane + c] : nullptr,
                SED ? &vd[1 * plane + cn] : nullptr);
    qx[c] = qv;
    if (SED) qxs[c] = qsv;
  }
  // y face (neighbour y-1)
  {
    const size_t cn = c - 1;
    qv = qy[c];
    if (SED) qsv = qys[c];
    face64<SED>(qv, qsv, hc, h[cn], zc, z[cn], n, dt, a, edge_code(y, a.jmax), false,
                SED ? V[c] : 0.0, SED ? V[cn] : 0.0, SED ? &vd[0 * plane + c] : nullptr,
                SED ? &vd[2 * plane + cn] : nullptr);
    qy[c] = qv;
    if (SED) qys[c] = qsv;
  }
}

// ---------------------------------------------------------- depth_update
template <bool SED>
__global__ __launch_bounds__(256) void k_depth64(HydroArgs a, double dt, double* __restrict__ h,
                                                 const double* __restrict__ z,
                                                 const double* __restrict__ qx,
                                                 const double* __restrict__ qy,
                                                 const uint64_t* __restrict__ mask,
                                                 double* __restrict__ V,
                                                 const double* __restrict__ qxs,
                                                 const double* __restrict__ qys,
                                                 unsigned long long* red) {
  const int y = blockIdx.x * 64 + threadIdx.x;
  const int x = blockIdx.y * blockDim.y + threadIdx.y;
  double md = 0;
  if (x >= 1 && x <= a.imax && y >= 1 && y <= a.jmax && mask_bit(mask, a.mwords, x, y)) {
    const size_t P = a.pitch;
    const size_t c = static_cast<size_t>(x) * P + y;
    double hv = h[c] + dt * (qx[c + P] - qx[c] + qy[c + 1] - qy[c]) / a.DX;
    if (SED) V[c] = V[c] + dt * (qxs[c + P] - qxs[c] + qys[c + 1] - qys[c]) / a.DX;
    if (hv > 0) {
      if (z[c] == -9999) hv = 0;
      if (hv > md) md = hv;
    }
    h[c] = hv;
  }
  md = wave_max(md);
  if ((threadIdx.x & 63) == 0) atomic_max_pos(red, md);
}

// ----------------------------------------------------- fused, FP64 form
// Each wavefront owns a 64-column strip and marches down R output rows,
// keeping the previous row in registers.  Reads generation `0` buffers and
// writes generation `1` buffers for every cell of its rows (unmasked cells
// are copied through, which keeps the reference's stale-flux semantics).
// Column neighbours come from lane shuffles; lane 0 loads the cell left of
// the strip and lane 63 recomputes the y-face right of the strip, which the
// depth update of its own cell needs.
template <int R>
__global__ __launch_bounds__(256) void k_flow_depth64(HydroArgs a, double dt,
                                                      const double* __restrict__ h0,
                                                      const double* __restrict__ z,
                                                      const double* __restrict__ qx0,
                                                      const double* __restrict__ qy0,
                                                      const double* __restrict__ nman,
                                                      const uint64_t* __restrict__ mask,
                                                      double* __restrict__ h1,
                                                      double* __restrict__ qx1,
                                                      double* __restrict__ qy1,
                                                      unsigned long long* red) {
  const int lane = threadIdx.x & 63;
  const int c0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
  if (c0 >= a.pitch) return;  // uniform per wavefront
  const int c = c0 + lane;
  const int NR = a.imax + 2;
  const int r0 = blockIdx.y * R;
  const int rend = min(r0 + R, NR);
  const int xlast = (rend < NR) ? rend : NR - 1;
  const size_t P = a.pitch;
  const int W = a.mwords;
  const int wcol = c0 >> 6;
  const bool has_right = (c0 + 64) < a.pitch;

  double hP = 0, zP = -9999;
  if (r0 >= 1) {
    hP = h0[static_cast<size_t>(r0 - 1) * P + c];
    zP = z[static_cast<size_t>(r0 - 1) * P + c];
  }
  double qxNp = 0, qyNp = 0, qyRp = 0;
  bool mP = false;
  double md = 0;

  for (int x = r0; x <= xlast; ++x) {
    const size_t row = static_cast<size_t>(x) * P;
    const double hx = h0[row + c];
    const double zx = z[row + c];
    const double qxo = qx0[row + c];
    const double qyo = qy0[row + c];
    const uint64_t mw = mask[static_cast<size_t>(x) * W + wcol];
    const bool mx = (mw >> lane) & 1ull;
    double hl = __shfl_up(hx, 1, 64);
    double zl = __shfl_up(zx, 1, 64);
    if (lane == 0) {
      if (c0 > 0) {
        hl = h0[row + c - 1];
        zl = z[row + c - 1];
      } else {
        hl = 0;
        zl = -9999;
      }
    }
    double qxN = qxo, qyN = qyo;
    if (mx && zx > -9999) {
      const double n = a.spatial_n ? nman[row + c] : a.mannings;
      double dummy = 0;
      face64<false>(qxN, dummy, hx, hP, zx, zP, n, dt, a, edge_code(x, a.imax), true, 0, 0,
                    nullptr, nullptr);
      face64<false>(qyN, dummy, hx, hl, zx, zl, n, dt, a, edge_code(c, a.jmax), false, 0, 0,
                    nullptr, nullptr);
    }
    double qyR = __shfl_down(qyN, 1, 64);
    if (lane == 63) {
      qyR = 0;
      if (has_right) {
        const size_t cr = row + c + 1;
        const uint64_t mwr = mask[static_cast<size_t>(x) * W + wcol + 1];
        qyR = qy0[cr];
        const double zr = z[cr];
        if ((mwr & 1ull) && zr > -9999) {
          const double n = a.spatial_n ? nman[cr] : a.mannings;
          double dummy = 0;
          face64<false>(qyR, dummy, h0[cr], hx, zr, zx, n, dt, a, edge_code(c + 1, a.jmax), false,
                        0, 0, nullptr, nullptr);
        }
      }
    }
    if (x > r0) {  // finalise row x-1
      const size_t o = row - P + c;
      double hn = hP;
      if (mP) {
        hn = hP + dt * (qxN - qxNp + qyRp - qyNp) / a.DX;
        if (hn > 0) {
          if (zP == -9999) hn = 0;
          if (hn > md) md = hn;
        }
      }
      h1[o] = hn;
      qx1[o] = qxNp;
      qy1[o] = qyNp;
    }
    hP = hx;
    zP = zx;
    qxNp = qxN;
    qyNp = qyN;
    qyRp = qyR;
    mP = mx;
  }
  if (rend == NR) {  // last padding row: never on the mask, copy through
    const size_t o = static_cast<size_t>(NR - 1) * P + c;
    h1[o] = hP;
    qx1[o] = qxNp;
    qy1[o] = qyNp;
  }
  md = wave_max(md);
  if (lane == 0) atomic_max_pos(red, md);
}

// ----------------------------------------------------- fused, FP32 form
__device__ __forceinline__ float flat_dz(float dz, const double* __restrict__ z64, size_t cn,
                                         size_t cc, float thr) {
  if (fabsf(dz) < thr) dz = static_cast<float>(z64[cn] - z64[cc]);
  return dz;
}

template <int R>
__global__ __launch_bounds__(256) void k_flow_depth32(HydroArgs a, float dt,
                                                      const float* __restrict__ h0,
                                                      const float* __restrict__ z,
                                                      const double* __restrict__ z64,
                                                      const float* __restrict__ qx0,
                                                      const float* __restrict__ qy0,
                                                      const float* __restrict__ nman,
                                                      const uint64_t* __restrict__ mask,
                                                      float* __restrict__ h1,
                                                      float* __restrict__ qx1,
                                                      float* __restrict__ qy1,
                                                      unsigned long long* red) {
  const int lane = threadIdx.x & 63;
  const int c0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
  if (c0 >= a.pitch) return;
  const int c = c0 + lane;
  const int NR = a.imax + 2;
  const int r0 = blockIdx.y * R;
  const int rend = min(r0 + R, NR);
  const int xlast = (rend < NR) ? rend : NR - 1;
  const size_t P = a.pitch;
  const int W = a.mwords;
  const int wcol = c0 >> 6;
  const bool has_right = (c0 + 64) < a.pitch;
  const float invdt = 1.0f / dt;
  const float fthr = static_cast<float>(a.flat_dz);
  const float dtDX = dt / static_cast<float>(a.DX);
  const float nu = static_cast<float>(a.mannings);

  float hP = 0.0f, zP = -9999.0f;
  if (r0 >= 1) {
    hP = h0[static_cast<size_t>(r0 - 1) * P + c];
    zP = z[static_cast<size_t>(r0 - 1) * P + c];
  }
  float qxNp = 0, qyNp = 0, qyRp = 0;
  bool mP = false;
  float md = 0;

  for (int x = r0; x <= xlast; ++x) {
    const size_t row = static_cast<size_t>(x) * P;
    const float hx = h0[row + c];
    const float zx = z[row + c];
    const float qxo = qx0[row + c];
    const float qyo = qy0[row + c];
    const float nx = a.spatial_n ? nman[row + c] : nu;
    const uint64_t mw = mask[static_cast<size_t>(x) * W + wcol];
    const bool mx = (mw >> lane) & 1ull;
    float hl = __shfl_up(hx, 1, 64);
    float zl = __shfl_up(zx, 1, 64);
    if (lane == 0) {
      if (c0 > 0) {
        hl = h0[row + c - 1];
        zl = z[row + c - 1];
      } else {
        hl = 0.0f;
        zl = -9999.0f;
      }
    }
    float qxN = qxo, qyN = qyo;
    if (mx && zx > -9999.0f) {
      if ((hx > 0.0f || hP > 0.0f) && zP > -9999.0f) {
        const float dz = flat_dz(zP - zx, z64, row - P + c, row + c, fthr);
        qxN = face32(qxo, hx, hP, dz, nx, dt, invdt, a, edge_code(x, a.imax));
      }
      if ((hx > 0.0f || hl > 0.0f) && zl > -9999.0f) {
        const float dz = flat_dz(zl - zx, z64, row + c - 1, row + c, fthr);
        qyN = face32(qyo, hx, hl, dz, nx, dt, invdt, a, edge_code(c, a.jmax));
      }
    }
    float qyR = __shfl_down(qyN, 1, 64);
    if (lane == 63) {
      qyR = 0.0f;
      if (has_right) {
        const size_t cr = row + c + 1;
        const uint64_t mwr = mask[static_cast<size_t>(x) * W + wcol + 1];
        qyR = qy0[cr];
        const float zr = z[cr];
        const float hr = h0[cr];
        if ((mwr & 1ull) && zr > -9999.0f && (hr > 0.0f || hx > 0.0f) && zx > -9999.0f) {
          const float nr = a.spatial_n ? nman[cr] : nu;
          const float dz = flat_dz(zx - zr, z64, row + c, cr, fthr);
          qyR = face32(qyR, hr, hx, dz, nr, dt, invdt, a, edge_code(c + 1, a.jmax));
        }
      }
    }
    if (x > r0) {
      const size_t o = row - P + c;
      float hn = hP;
      if (mP) {
        hn = hP + dtDX * ((qxN - qxNp) + (qyRp - qyNp));
        if (hn > 0.0f) {
          if (zP == -9999.0f) hn = 0.0f;
          md = fmaxf(md, hn);
        }
      }
      h1[o] = hn;
      qx1[o] = qxNp;
      qy1[o] = qyNp;
    }
    hP = hx;
    zP = zx;
    qxNp = qxN;
    qyNp = qyN;
    qyRp = qyR;
    mP = mx;
  }
  if (rend == NR) {
    const size_t o = static_cast<size_t>(NR - 1) * P + c;
    h1[o] = hP;
    qx1[o] = qxNp;
    qy1[o] = qyNp;
  }
  double mdd = wave_max(static_cast<double>(md));
  if (lane == 0) atomic_max_pos(red, mdd);
}

// ---------------------------------------------------------------- scan
template <typename T>
__global__ __launch_bounds__(256) void k_scan(int imax, int jmax, int P, int W,
                                              const T* __restrict__ h,
                                              uint64_t* __restrict__ mask) {
  const int lane = threadIdx.x;
  const int c = blockIdx.x * 64 + lane;
  const int x = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= imax + 2) return;  // uniform per wavefront
  bool wet = false;
  if (x >= 1 && x <= imax && c >= 1 && c <= jmax) {
    const size_t b = static_cast<size_t>(x) * P + c;
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy)
        wet = wet || (h[b + static_cast<long>(dx) * P + dy] > T(0));
  }
  const uint64_t bits = __ballot(wet);
  if (lane == 0) mask[static_cast<size_t>(x) * W + blockIdx.x] = bits;
}

// --------------------------------------------------------- water_flux_out
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

template <typename T>
__global__ __launch_bounds__(256) void k_flux_out(int imax, int jmax, int P, double DX,
                                                  double thr, double ts, T* __restrict__ h,
                                                  double* __restrict__ partials,
                                                  unsigned int* __restrict__ count,
                                                  double* __restrict__ out) {
  const int nedge = jmax + (imax > 1 ? jmax : 0) + (imax > 2 ? (imax - 2) * (jmax > 1 ? 2 : 1) : 0);
  double acc = 0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nedge; k += gridDim.x * blockDim.x) {
    int x, y;
    edge_cell(k, imax, jmax, x, y);
    T* hp = &h[static_cast<size_t>(x) * P + y];
    const double hv = static_cast<double>(*hp);
    if (hv > thr) {
      acc += (hv - thr) * DX * DX / ts;
      *hp = static_cast<T>(thr);
    }
  }
  __shared__ double s[256];
  __shared__ bool last;
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = s[0];
    __threadfence();
    const unsigned prev = atomicAdd(count, 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    double tot = 0;
    for (unsigned b = 0; b < gridDim.x; ++b)
      tot += __longlong_as_double(__atomic_load_n(reinterpret_cast<const long long*>(&partials[b]), __ATOMIC_RELAXED));
    out[0] = tot;
    *count = 0;
  }
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
  return a;
}

dim3 cell_grid(const hc_ctx* c) {
  return dim3(static_cast<unsigned>(c->P / 64), static_cast<unsigned>((c->NR + 3) / 4));
}

}  // namespace

// Rows per wavefront of the fused kernels (set by the tile tuner).
int g_fused_rows = 32;

int hc_launch_flow_route(hc_ctx* c, double dt) {
  if (c->prec != HC_FP64)
    return hc_fail(c, HC_ERR_STATE, "flow_route: the separate kernels are FP64-only; use hc_flow_depth");
  const HydroArgs a = hydro_args(c);
  const dim3 blk(64, 4);
  const int k = c->cur;
  if (c->g.sediment) {
    hipLaunchKernelGGL(k_flow_route64<true>, cell_grid(c), blk, 0, c->stream, a, dt,
                       (const double*)c->h[k], (const double*)c->z, (double*)c->qx[k],
                       (double*)c->qy[k], (const double*)c->nman, c->mask, c->qxs, c->qys,
                       (const double*)c->V, c->veldir, c->ncell);
  } else {
    hipLaunchKernelGGL(k_flow_route64<false>, cell_grid(c), blk, 0, c->stream, a, dt,
                       (const double*)c->h[k], (const double*)c->z, (double*)c->qx[k],
                       (double*)c->qy[k], (const double*)c->nman, c->mask, nullptr, nullptr,
                       nullptr, nullptr, c->ncell);
  }
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_launch_depth_update(hc_ctx* c, double dt) {
  if (c->prec != HC_FP64)
    return hc_fail(c, HC_ERR_STATE, "depth_update: the separate kernels are FP64-only; use hc_flow_depth");
  const HydroArgs a = hydro_args(c);
  const dim3 blk(64, 4);
  const int k = c->cur;
  HC_HIP(c, hipMemsetAsync(c->d_red, 0, sizeof(unsigned long long), c->stream));
  if (c->g.sediment) {
    hipLaunchKernelGGL(k_depth64<true>, cell_grid(c), blk, 0, c->stream, a, dt, (double*)c->h[k],
                       (const double*)c->z, (const double*)c->qx[k], (const double*)c->qy[k],
                       c->mask, c->V, (const double*)c->qxs, (const double*)c->qys, c->d_red);
  } else {
    hipLaunchKernelGGL(k_depth64<false>, cell_grid(c), blk, 0, c->stream, a, dt, (double*)c->h[k],
                       (const double*)c->z, (const double*)c->qx[k], (const double*)c->qy[k],
                       c->mask, nullptr, nullptr, nullptr, c->d_red);
  }
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_launch_flow_depth(hc_ctx* c, double dt) {
  if (c->g.sediment)
    return hc_fail(c, HC_ERR_STATE, "flow_depth: fused kernel is for hydro-only contexts");
  const HydroArgs a = hydro_args(c);
  const int R = g_fused_rows;
  const dim3 grid(static_cast<unsigned>((c->P / 64 + 3) / 4),
                  static_cast<unsigned>((c->NR + R - 1) / R));
  const dim3 blk(256);
  const int k = c->cur, o = c->cur ^ 1;
  HC_HIP(c, hipMemsetAsync(c->d_red, 0, sizeof(unsigned long long), c->stream));
  if (c->prec == HC_FP64) {
    auto args = [&](auto kern) {
      hipLaunchKernelGGL(kern, grid, blk, 0, c->stream, a, dt, (const double*)c->h[k],
                         (const double*)c->z, (const double*)c->qx[k], (const double*)c->qy[k],
                         (const double*)c->nman, (const uint64_t*)c->mask, (double*)c->h[o],
                         (double*)c->qx[o], (double*)c->qy[o], c->d_red);
    };
    if (R == 16) args(k_flow_depth64<16>);
    else if (R == 64) args(k_flow_depth64<64>);
    else if (R == 8) args(k_flow_depth64<8>);
    else args(k_flow_depth64<32>);
  } else {
    auto args = [&](auto kern) {
      hipLaunchKernelGGL(kern, grid, blk, 0, c->stream, a, static_cast<float>(dt),
                         (const float*)c->h[k], (const float*)c->z, (const double*)c->z64,
                         (const float*)c->qx[k], (const float*)c->qy[k], (const float*)c->nman,
                         (const uint64_t*)c->mask, (float*)c->h[o], (float*)c->qx[o],
                         (float*)c->qy[o], c->d_red);
    };
    if (R == 16) args(k_flow_depth32<16>);
    else if (R == 64) args(k_flow_depth32<64>);
    else if (R == 8) args(k_flow_depth32<8>);
    else args(k_flow_depth32<32>);
  }
  HC_HIP(c, hipGetLastError());
  c->cur = o;
  return HC_OK;
}

int hc_launch_scan(hc_ctx* c) {
  const dim3 blk(64, 4);
  const int k = c->cur;
  if (c->prec == HC_FP64)
    hipLaunchKernelGGL(k_scan<double>, cell_grid(c), blk, 0, c->stream, c->g.imax, c->g.jmax,
                       c->P, c->W, (const double*)c->h[k], c->mask);
  else
    hipLaunchKernelGGL(k_scan<float>, cell_grid(c), blk, 0, c->stream, c->g.imax, c->g.jmax,
                       c->P, c->W, (const float*)c->h[k], c->mask);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_launch_flux_out(hc_ctx* c, double ts) {
  const int imax = c->g.imax, jmax = c->g.jmax;
  const long nedge = jmax + (imax > 1 ? jmax : 0) + (imax > 2 ? (long)(imax - 2) * (jmax > 1 ? 2 : 1) : 0);
  int nb = static_cast<int>((nedge + 255) / 256);
  if (nb > 64) nb = 64;
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
