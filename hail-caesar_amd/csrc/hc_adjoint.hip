// Hand-written adjoint of flow_route + depth_update (with the edge drain and a
// per-block rain source) at a fixed time step, for gradients of an outlet
// hydrograph misfit with respect to per-cell Manning n and rain-block
// multipliers.  No counterpart in the reference, which has no derivatives;
// the forward it differentiates is the FP64 plain form of
// LSDCatchmentModel::flow_route / depth_update / water_flux_out
// (LCM.cpp:3063-3299, 3005-3047) with dt held fixed (set_maximum_timestep
// bypassed) and rain added directly to every cell of its block.
//
// Forward step s (state h, qx, qy and the active mask at its start):
//   h1 = h + dt * rain[s] * mult[block]                     (cells with block >= 0)
//   (qx', qy', h2) = fused flow_route + depth_update (h1, qx, qy; mask)
//   Q[s] = sum over edge cells of (h2 - thr) DX^2 / dt where h2 > thr; h3 = thr there
//   scan_area on h3 when (s + 1) % 5 == 0 (and once before step 0)
// Loss L = sum_s (Q[s] - Qtarget[s])^2.
//
// Reverse step s, from the adjoints of (h3, qx', qy'):
//   lam2 = dL/dh2: g_s DX^2 / dt on drained edge cells, dL/dh3 elsewhere,
//          0 on NODATA cells (depth_update zeroes them)
//   face adjoint  lq = dL/dq' + dt/DX (dep(upstream cell) - dep(cell)),
//                 dep = lam2 on cells whose depth was updated (mask), else 0
//   k_adj_face: per cell, its two faces' derivatives (the branch the forward
//          took: plain update, Froude cap, depth cap, below-threshold zero,
//          untouched) -> dL/dq, the contributions to dL/dh1 of the cell and
//          of its two upstream neighbours, and dL/dn accumulated
//   k_adj_cell: dL/dh1 = lam2 + own + neighbours' contributions; dL/dh = dL/dh1
//          (the rain is additive) and dL/dmult[b] += dt rain[s] sum_b dL/dh1.
//
// Checkpointing, two levels: the forward keeps the state every C steps; the
// reverse sweep re-runs each segment once from its checkpoint, writing every
// step's input state into its own slot (C+1 slots: no copies), with the
// active mask and the drained edge depths of each step beside them.  The
// extra cost is one forward; device memory is (C+1) states, small next to
// HBM.  All reductions are in fixed order: repeated runs are bit-identical.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "hc_device.hpp"
#include "hc_internal.hpp"

using hc::HydroArgs;
using hc::kGravity;
using namespace hcdev;

struct hc_adj {
  hc_ctx* c = nullptr;
  int N = 0, C = 0, nb = 0, nedge = 0, nck = 0;
  double dt = 0, thr = 0;
  std::vector<double> rain, target, mult, Q;
  size_t ncell = 0, nwords = 0;
  int* d_block = nullptr;         // rain block per cell (pitched), -1: none
  double* d_mult = nullptr;       // nb
  double* d_Q = nullptr;          // N
  double* d_g = nullptr;          // N: dL/dQ
  double* d_ck = nullptr;         // nck x 3 x ncell (h, qx, qy at checkpoint starts)
  uint64_t* d_ckmask = nullptr;   // nck x nwords
  double* d_slot = nullptr;       // (C+1) x 3 x ncell (segment states)
  uint64_t* d_smask = nullptr;    // C x nwords
  double* d_h2e = nullptr;        // C x nedge: edge depths before the drain
  double* d_init = nullptr;       // 3 x ncell: the initial state
  uint64_t* d_initmask = nullptr;
  uint64_t* d_mask = nullptr;     // working mask
  double* d_l = nullptr;          // 7 x ncell: lh, lqx, lqy, own, nbx, nby, gn
  double* d_gm = nullptr;         // gm_blocks x nb partial rain-multiplier gradients
  int gm_blocks = 0;
  bool have_forward = false;
  std::string err;
};

namespace {

int adj_fail(hc_adj* a, int code, const std::string& msg) {
  a->err = msg;
  return code;
}

#define ADJ_HIP(a, call)                                                         \
  do {                                                                           \
    hipError_t _r = (call);                                                      \
    if (_r != hipSuccess)                                                        \
      return adj_fail((a), HC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_r)); \
  } while (0)

__device__ __forceinline__ int edge_code_d(int i, int imaxlike) {
  return (i <= 2) ? -1 : ((i == imaxlike) ? 1 : 0);
}

// Index of edge cell (x, y) in the enumeration of edge_cell().
__device__ __forceinline__ int edge_index(int x, int y, int imax, int jmax) {
  if (x == 1) return y - 1;
  if (imax > 1 && x == imax) return jmax + y - 1;
  const int nside = (jmax > 1) ? 2 : 1;
  return jmax + (imax > 1 ? jmax : 0) + (x - 2) * nside + ((y == 1) ? 0 : 1);
}

__device__ __forceinline__ bool on_edge(int x, int y, int imax, int jmax) {
  return x == 1 || x == imax || y == 1 || y == jmax;
}

struct FaceD {
  double dq, dhc, dhn, dn;
};

// Derivatives of one FP64 face update (the forward is face64 in hc_hydro.hip,
// LCM.cpp:3094-3191 / 3195-3242): q' as a function of the old flux q, the
// cell depth hc, the neighbour depth hn and the cell's Manning n, on the
// branch the forward takes for these inputs.
__device__ FaceD face64_d(double q, double hc, double hn, double zc, double zn, double n,
                          double dt, const HydroArgs& a, int edge, bool xface) {
  FaceD r{1.0, 0.0, 0.0, 0.0};
  if (!((hc > 0 || hn > 0) && zn > -9999)) return r;  // flux left as it was
  const double eta_c = zc + hc;
  const double eta_n = zn + hn;
  const double etamax = dmax_ref(eta_c, eta_n);
  const double zmax = xface ? dmax_ref(zn, zc) : dmax_ref(zc, zn);
  const double hflow = etamax - zmax;
  // d hflow / d(hc, hn): the argument std::max returned
  const double fc = (eta_c < eta_n) ? 0.0 : 1.0;
  const double fn = 1.0 - fc;
  if (!(hflow > a.hthr)) return FaceD{0.0, 0.0, 0.0, 0.0};  // q' = 0
  double slope = (eta_n - eta_c) / a.DX;
  double sc = -1.0 / a.DX, sn = 1.0 / a.DX;
  if (edge > 0) {
    slope = a.edgeslope;
    sc = sn = 0.0;
  }
  if (edge < 0) {
    slope = 0 - a.edgeslope;
    sc = sn = 0.0;
  }
  const double h3 = cube_rn(hflow);
  const double A = q - (kGravity * hflow * dt * slope);
  const double D = 1 + kGravity * hflow * dt * (n * n) * fabs(q) / h3;
  double qv = A / D;
  // D = 1 + K |q| / hflow^2 with K = g dt n^2
  const double K = kGravity * dt * (n * n);
  const double sq = (q > 0) ? 1.0 : ((q < 0) ? -1.0 : 0.0);
  const double hf2 = hflow * hflow;
  const double dD_dq = K * sq / hf2;
  const double dD_dh = -2.0 * K * fabs(q) / (hf2 * hflow);
  const double dD_dn = 2.0 * kGravity * dt * n * fabs(q) / hf2;
  const double D2 = D * D;
  const double dq = (D - A * dD_dq) / D2;
  const double dhf = ((-kGravity * dt * slope) * D - A * dD_dh) / D2;
  const double dsl = (-kGravity * hflow * dt) / D;
  FaceD d{dq, dhf * fc + dsl * sc, dhf * fn + dsl * sn, -A * dD_dn / D2};
  const double sg = sqrt(kGravity * hflow);
  if (qv > 0 && (qv / hflow) / sg > a.froude) {  // q' = Fr sqrt(g) hflow^1.5
    qv = hflow * (sg * a.froude);
    const double k = 1.5 * sg * a.froude;
    d = FaceD{0.0, k * fc, k * fn, 0.0};
  }
  if (qv < 0 && fabs(qv / hflow) / sg > a.froude) {
    qv = 0 - (hflow * (sg * a.froude));
    const double k = -1.5 * sg * a.froude;
    d = FaceD{0.0, k * fc, k * fn, 0.0};
  }
  if (qv > 0 && (qv * dt / a.DX) > (hc / 4)) {  // q' = hc DX / 5 / dt
    qv = ((hc * a.DX) / 5) / dt;
    d = FaceD{0.0, a.DX / (5.0 * dt), 0.0, 0.0};
  }
  if (qv < 0 && fabs(qv * dt / a.DX) > (hn / 4)) {
    qv = 0 - ((hn * a.DX) / 5) / dt;
    d = FaceD{0.0, 0.0, -a.DX / (5.0 * dt), 0.0};
  }
  return d;
}

struct AdjArgs {
  HydroArgs a;
  int NR, P, W, nedge;
  double dt, thr, gs;          // gs = dL/dQ of this step
  const double* h1;            // step input depth (after rain)
  const double* qx0;
  const double* qy0;
  const uint64_t* mask;        // active mask during the step
  const double* h2e;           // edge depths before the drain
  const double* z;
  const double* n;
  const int* block;
  double* lh;
  double* lqx;
  double* lqy;
  double* own;
  double* nbx;
  double* nby;
  double* gn;
};

__device__ __forceinline__ bool cell_mask(const AdjArgs& A, int x, int y) {
  return (A.mask[static_cast<size_t>(x) * A.W + (y >> 6)] >> (y & 63)) & 1ull;
}

// dL/dh2 of an interior cell.
__device__ __forceinline__ double lam2(const AdjArgs& A, int x, int y) {
  const size_t o = static_cast<size_t>(x) * A.P + y;
  if (!(A.z[o] > -9999)) return 0.0;
  if (on_edge(x, y, A.a.imax, A.a.jmax) &&
      A.h2e[edge_index(x, y, A.a.imax, A.a.jmax)] > A.thr)
    return A.gs * A.a.DX * A.a.DX / A.dt;
  return A.lh[o];
}

// the part of dL/dh2 that reaches the fluxes through depth_update
__device__ __forceinline__ double dep(const AdjArgs& A, int x, int y) {
  if (x < 1 || x > A.a.imax || y < 1 || y > A.a.jmax) return 0.0;
  return cell_mask(A, x, y) ? lam2(A, x, y) : 0.0;
}

__global__ __launch_bounds__(256) void k_adj_face(AdjArgs A) {
  const int y = blockIdx.x * 64 + (threadIdx.x & 63);
  const int x = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x < 1 || x > A.a.imax || y < 1 || y > A.a.jmax) return;
  const size_t o = static_cast<size_t>(x) * A.P + y;
  const double dc = dep(A, x, y);
  const double lqxT = A.lqx[o] + A.dt / A.a.DX * (dep(A, x - 1, y) - dc);
  const double lqyT = A.lqy[o] + A.dt / A.a.DX * (dep(A, x, y - 1) - dc);
  FaceD fx{1.0, 0.0, 0.0, 0.0}, fy{1.0, 0.0, 0.0, 0.0};
  if (cell_mask(A, x, y) && A.z[o] > -9999) {  // the forward updated this cell's faces
    const double nn = A.a.spatial_n ? A.n[o] : A.a.mannings;
    fx = face64_d(A.qx0[o], A.h1[o], A.h1[o - A.P], A.z[o], A.z[o - A.P], nn, A.dt, A.a,
                  edge_code_d(x, A.a.imax), true);
    fy = face64_d(A.qy0[o], A.h1[o], A.h1[o - 1], A.z[o], A.z[o - 1], nn, A.dt, A.a,
                  edge_code_d(y, A.a.jmax), false);
  }
  A.lqx[o] = lqxT * fx.dq;
  A.lqy[o] = lqyT * fy.dq;
  A.own[o] = lqxT * fx.dhc + lqyT * fy.dhc;
  A.nbx[o] = lqxT * fx.dhn;
  A.nby[o] = lqyT * fy.dhn;
  A.gn[o] += lqxT * fx.dn + lqyT * fy.dn;
}

constexpr int kMaxRainBlocks = 16;

// dL/dh1 per cell and the rain-multiplier gradient; gm holds one row of nb
// partial sums per block of this launch, each accumulated across steps by
// the same block in the same order.
__global__ __launch_bounds__(256) void k_adj_cell(AdjArgs A, double rain_dt, int nb,
                                                  double* __restrict__ gm) {
  __shared__ double s_part[4][kMaxRainBlocks];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long ncell = static_cast<long>(A.a.imax) * A.a.jmax;
  double acc[kMaxRainBlocks];
#pragma unroll
  for (int b = 0; b < kMaxRainBlocks; ++b) acc[b] = 0.0;
  for (long k = blockIdx.x * 256L + threadIdx.x; k < ncell; k += gridDim.x * 256L) {
    const int x = 1 + static_cast<int>(k / A.a.jmax);
    const int y = 1 + static_cast<int>(k % A.a.jmax);
    const size_t o = static_cast<size_t>(x) * A.P + y;
    double l1 = 0.0;
    if (A.z[o] > -9999)
      l1 = lam2(A, x, y) + A.own[o] + A.nbx[o + A.P] + A.nby[o + 1];
    A.lh[o] = l1;
    const int b = A.block[o];
#pragma unroll
    for (int j = 0; j < kMaxRainBlocks; ++j)
      if (j < nb && b == j) acc[j] += rain_dt * l1;
  }
  // fixed-order block reduction per rain block
#pragma unroll
  for (int j = 0; j < kMaxRainBlocks; ++j) {
    if (j < nb) {  // uniform
      const double v = wave_sum(acc[j]);
      if (lane == 0) s_part[wv][j] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < nb) {
    const int j = threadIdx.x;
    gm[static_cast<size_t>(blockIdx.x) * nb + j] +=
        (s_part[0][j] + s_part[1][j]) + (s_part[2][j] + s_part[3][j]);
  }
}

__global__ __launch_bounds__(256) void k_adj_rain(int imax, int jmax, int P,
                                                  double* __restrict__ h,
                                                  const double* __restrict__ z,
                                                  const int* __restrict__ block,
                                                  const double* __restrict__ mult,
                                                  double rain_dt) {
  const int y = blockIdx.x * 64 + (threadIdx.x & 63);
  const int x = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x < 1 || x > imax || y < 1 || y > jmax) return;
  const size_t o = static_cast<size_t>(x) * P + y;
  const int b = block[o];
  if (b >= 0 && z[o] > -9999) h[o] = h[o] + rain_dt * mult[b];
}

// Edge drain of one step (one block, fixed-order sum): Q = sum (h - thr)
// DX^2 / dt over edge cells above the threshold, which are set to it; the
// depths before the drain are kept for the reverse sweep.
__global__ __launch_bounds__(256) void k_adj_drain(int imax, int jmax, int P, double DX,
                                                   double thr, double dt,
                                                   double* __restrict__ h,
                                                   double* __restrict__ h2e,
                                                   double* __restrict__ Q) {
  __shared__ double s[256];
  const int nedge = jmax + (imax > 1 ? jmax : 0) + (imax > 2 ? (imax - 2) * (jmax > 1 ? 2 : 1) : 0);
  double acc = 0;
  for (int k = threadIdx.x; k < nedge; k += 256) {
    int x, y;
    edge_cell(k, imax, jmax, x, y);
    double* hp = &h[static_cast<size_t>(x) * P + y];
    const double hv = *hp;
    if (h2e) h2e[k] = hv;
    if (hv > thr) {
      acc += (hv - thr) * DX * DX / dt;
      *hp = thr;
    }
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *Q = s[0];
}

dim3 cell_grid2(const hc_ctx* c) {
  return dim3(static_cast<unsigned>(c->P / 64), static_cast<unsigned>((c->NR + 3) / 4));
}

double* fld(double* base, size_t ncell, int k) { return base + static_cast<size_t>(k) * ncell; }

// One forward step from (h, qx, qy) (rain added to h in place) to the output
// buffers; records the mask and the pre-drain edge depths when asked.
int fwd_step(hc_adj* A, int s, double* h, double* qx, double* qy, double* ho, double* qxo,
             double* qyo, uint64_t* mask_rec, double* h2e_rec) {
  hc_ctx* c = A->c;
  const double rdt = A->rain[s] * A->dt;
  hipLaunchKernelGGL(k_adj_rain, cell_grid2(c), dim3(256), 0, c->stream, c->g.imax, c->g.jmax,
                     c->P, h, static_cast<const double*>(c->z), A->d_block, A->d_mult, rdt);
  ADJ_HIP(A, hipGetLastError());
  if (mask_rec)
    ADJ_HIP(A, hipMemcpyAsync(mask_rec, A->d_mask, A->nwords * sizeof(uint64_t),
                              hipMemcpyDeviceToDevice, c->stream));
  if (hc_launch_fused64_ptrs(c, A->dt, h, qx, qy, ho, qxo, qyo, A->d_mask))
    return adj_fail(A, HC_ERR_HIP, "adjoint: fused launch failed");
  hipLaunchKernelGGL(k_adj_drain, dim3(1), dim3(256), 0, c->stream, c->g.imax, c->g.jmax, c->P,
                     c->g.DX, A->thr, A->dt, ho, h2e_rec, A->d_Q + s);
  ADJ_HIP(A, hipGetLastError());
  if ((s + 1) % 5 == 0 && hc_launch_scan_ptr(c, ho, A->d_mask))
    return adj_fail(A, HC_ERR_HIP, "adjoint: scan launch failed");
  return HC_OK;
}

}  // namespace

extern "C" {

int hc_adj_destroy(hc_adj* A);

int hc_adj_create(hc_adj** out, hc_ctx* c, int nsteps, double dt, int checkpoint_every,
                  const int* block, int nblocks, const double* rain) {
  if (!out || !c || nsteps < 1 || !(dt > 0) || checkpoint_every < 1 || !block || !rain ||
      nblocks < 1 || nblocks > kMaxRainBlocks)
    return HC_ERR_ARG;
  *out = nullptr;
  if (c->prec != HC_FP64 || c->g.sediment || !c->prm.spatially_var_mannings || !c->nman)
    return HC_ERR_STATE;
  hc_adj* A = new (std::nothrow) hc_adj;
  if (!A) return HC_ERR_ARG;
  A->c = c;
  A->N = nsteps;
  A->C = std::min(checkpoint_every, nsteps);
  A->nck = (nsteps + A->C - 1) / A->C;
  A->nb = nblocks;
  A->dt = dt;
  A->thr = c->prm.water_depth_erosion_threshold;
  A->rain.assign(rain, rain + nsteps);
  A->target.assign(nsteps, 0.0);
  A->mult.assign(nblocks, 1.0);
  A->Q.assign(nsteps, 0.0);
  A->ncell = c->ncell;
  A->nwords = static_cast<size_t>(c->NR) * c->W;
  const int imax = c->g.imax, jmax = c->g.jmax;
  A->nedge = jmax + (imax > 1 ? jmax : 0) + (imax > 2 ? (imax - 2) * (jmax > 1 ? 2 : 1) : 0);
  // rain blocks, pitched
  std::vector<int> bl(A->ncell, -1);
  for (int x = 0; x < c->NR; ++x)
    for (int y = 0; y < c->NC; ++y) {
      const int v = block[static_cast<size_t>(x) * c->NC + y];
      if (v >= nblocks) {
        hc_adj_destroy(A);
        return HC_ERR_ARG;
      }
      bl[static_cast<size_t>(x) * c->P + y] = (x >= 1 && x <= imax && y >= 1 && y <= jmax) ? v : -1;
    }
  const size_t cs = A->ncell * sizeof(double);
  const size_t ms = A->nwords * sizeof(uint64_t);
  A->gm_blocks = 1024;
  int rc = HC_OK;
  auto al = [&](void** p, size_t bytes) {
    if (rc == HC_OK && hipMalloc(p, bytes) != hipSuccess) rc = HC_ERR_HIP;
  };
  al(reinterpret_cast<void**>(&A->d_block), A->ncell * sizeof(int));
  al(reinterpret_cast<void**>(&A->d_mult), nblocks * sizeof(double));
  al(reinterpret_cast<void**>(&A->d_Q), nsteps * sizeof(double));
  al(reinterpret_cast<void**>(&A->d_g), nsteps * sizeof(double));
  al(reinterpret_cast<void**>(&A->d_ck), A->nck * 3 * cs);
  al(reinterpret_cast<void**>(&A->d_ckmask), A->nck * ms);
  al(reinterpret_cast<void**>(&A->d_slot), (A->C + 1) * 3 * cs);
  al(reinterpret_cast<void**>(&A->d_smask), A->C * ms);
  al(reinterpret_cast<void**>(&A->d_h2e), static_cast<size_t>(A->C) * A->nedge * sizeof(double));
  al(reinterpret_cast<void**>(&A->d_init), 3 * cs);
  al(reinterpret_cast<void**>(&A->d_initmask), ms);
  al(reinterpret_cast<void**>(&A->d_mask), ms);
  al(reinterpret_cast<void**>(&A->d_l), 7 * cs);
  al(reinterpret_cast<void**>(&A->d_gm), static_cast<size_t>(A->gm_blocks) * nblocks * sizeof(double));
  if (rc != HC_OK) {
    hc_adj_destroy(A);
    return rc;
  }
  // the context's current state is the initial state of every forward run;
  // its active mask is the scan of the initial depths
  const int k = c->cur;
  hipStream_t st = c->stream;
  if (hipMemcpy(A->d_block, bl.data(), A->ncell * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpyAsync(fld(A->d_init, A->ncell, 0), c->h[k], cs, hipMemcpyDeviceToDevice, st) != hipSuccess ||
      hipMemcpyAsync(fld(A->d_init, A->ncell, 1), c->qx[k], cs, hipMemcpyDeviceToDevice, st) != hipSuccess ||
      hipMemcpyAsync(fld(A->d_init, A->ncell, 2), c->qy[k], cs, hipMemcpyDeviceToDevice, st) != hipSuccess ||
      hc_launch_scan_ptr(c, c->h[k], A->d_initmask) != HC_OK ||
      hipStreamSynchronize(st) != hipSuccess) {
    hc_adj_destroy(A);
    return HC_ERR_HIP;
  }
  *out = A;
  return HC_OK;
}

int hc_adj_destroy(hc_adj* A) {
  if (!A) return HC_OK;
  if (A->c) (void)hipStreamSynchronize(A->c->stream);
  void* ptrs[] = {A->d_block, A->d_mult, A->d_Q, A->d_g, A->d_ck, A->d_ckmask, A->d_slot,
                  A->d_smask, A->d_h2e, A->d_init, A->d_initmask, A->d_mask, A->d_l, A->d_gm};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete A;
  return HC_OK;
}

const char* hc_adj_last_error(const hc_adj* A) { return A ? A->err.c_str() : "null adjoint"; }

int hc_adj_set_target(hc_adj* A, const double* target) {
  if (!A || !target) return HC_ERR_ARG;
  A->target.assign(target, target + A->N);
  return HC_OK;
}

int hc_adj_forward(hc_adj* A, const double* mult, double* loss, double* Q) {
  if (!A || !mult) return HC_ERR_ARG;
  hc_ctx* c = A->c;
  hipStream_t st = c->stream;
  const size_t cs = A->ncell * sizeof(double);
  const size_t ms = A->nwords * sizeof(uint64_t);
  A->mult.assign(mult, mult + A->nb);
  ADJ_HIP(A, hipMemcpyAsync(A->d_mult, A->mult.data(), A->nb * sizeof(double),
                            hipMemcpyHostToDevice, st));
  // start from the initial state in slots 0 (ping) and 1 (pong)
  double* S = A->d_slot;
  for (int f = 0; f < 3; ++f)
    ADJ_HIP(A, hipMemcpyAsync(fld(S, A->ncell, f), fld(A->d_init, A->ncell, f), cs,
                              hipMemcpyDeviceToDevice, st));
  ADJ_HIP(A, hipMemcpyAsync(A->d_mask, A->d_initmask, ms, hipMemcpyDeviceToDevice, st));
  for (int s = 0; s < A->N; ++s) {
    const int in = s & 1, o = in ^ 1;
    double* hi = fld(S, A->ncell, 3 * in);
    if (s % A->C == 0) {  // checkpoint: the state at the start of step s
      const int j = s / A->C;
      for (int f = 0; f < 3; ++f)
        ADJ_HIP(A, hipMemcpyAsync(fld(A->d_ck, A->ncell, 3 * j + f), fld(S, A->ncell, 3 * in + f),
                                  cs, hipMemcpyDeviceToDevice, st));
      ADJ_HIP(A, hipMemcpyAsync(A->d_ckmask + j * A->nwords, A->d_mask, ms,
                                hipMemcpyDeviceToDevice, st));
    }
    const int rc = fwd_step(A, s, hi, fld(S, A->ncell, 3 * in + 1), fld(S, A->ncell, 3 * in + 2),
                            fld(S, A->ncell, 3 * o), fld(S, A->ncell, 3 * o + 1),
                            fld(S, A->ncell, 3 * o + 2), nullptr, nullptr);
    if (rc) return rc;
  }
  ADJ_HIP(A, hipMemcpyAsync(A->Q.data(), A->d_Q, A->N * sizeof(double), hipMemcpyDeviceToHost, st));
  ADJ_HIP(A, hipStreamSynchronize(st));
  double L = 0;
  for (int s = 0; s < A->N; ++s) {
    const double d = A->Q[s] - A->target[s];
    L += d * d;
  }
  if (loss) *loss = L;
  if (Q) std::memcpy(Q, A->Q.data(), A->N * sizeof(double));
  A->have_forward = true;
  return HC_OK;
}

int hc_adj_backward(hc_adj* A, double* grad_n, double* grad_mult) {
  if (!A) return HC_ERR_ARG;
  if (!A->have_forward) return adj_fail(A, HC_ERR_STATE, "adjoint: run hc_adj_forward first");
  hc_ctx* c = A->c;
  hipStream_t st = c->stream;
  const size_t cs = A->ncell * sizeof(double);
  const size_t ms = A->nwords * sizeof(uint64_t);
  std::vector<double> g(A->N);
  for (int s = 0; s < A->N; ++s) g[s] = 2.0 * (A->Q[s] - A->target[s]);
  ADJ_HIP(A, hipMemcpyAsync(A->d_g, g.data(), A->N * sizeof(double), hipMemcpyHostToDevice, st));
  ADJ_HIP(A, hipMemsetAsync(A->d_l, 0, 7 * cs, st));
  ADJ_HIP(A, hipMemsetAsync(A->d_gm, 0, static_cast<size_t>(A->gm_blocks) * A->nb * sizeof(double), st));
  hc::HydroArgs ha = hc_hydro_args(c);
  AdjArgs R;
  R.a = ha;
  R.NR = c->NR;
  R.P = c->P;
  R.W = c->W;
  R.nedge = A->nedge;
  R.dt = A->dt;
  R.thr = A->thr;
  R.z = static_cast<const double*>(c->z);
  R.n = static_cast<const double*>(c->nman);
  R.block = A->d_block;
  R.lh = fld(A->d_l, A->ncell, 0);
  R.lqx = fld(A->d_l, A->ncell, 1);
  R.lqy = fld(A->d_l, A->ncell, 2);
  R.own = fld(A->d_l, A->ncell, 3);
  R.nbx = fld(A->d_l, A->ncell, 4);
  R.nby = fld(A->d_l, A->ncell, 5);
  R.gn = fld(A->d_l, A->ncell, 6);
  const long ncell_int = static_cast<long>(c->g.imax) * c->g.jmax;
  const int cell_blocks = static_cast<int>(std::min<long>(A->gm_blocks, (ncell_int + 255) / 256));
  for (int j = A->nck - 1; j >= 0; --j) {
    const int s0 = j * A->C, s1 = std::min(A->N, s0 + A->C);
    // re-run the segment from its checkpoint, one slot per step
    for (int f = 0; f < 3; ++f)
      ADJ_HIP(A, hipMemcpyAsync(fld(A->d_slot, A->ncell, f), fld(A->d_ck, A->ncell, 3 * j + f), cs,
                                hipMemcpyDeviceToDevice, st));
    ADJ_HIP(A, hipMemcpyAsync(A->d_mask, A->d_ckmask + j * A->nwords, ms,
                              hipMemcpyDeviceToDevice, st));
    for (int s = s0; s < s1; ++s) {
      const int i = s - s0;
      const int rc = fwd_step(A, s, fld(A->d_slot, A->ncell, 3 * i),
                              fld(A->d_slot, A->ncell, 3 * i + 1), fld(A->d_slot, A->ncell, 3 * i + 2),
                              fld(A->d_slot, A->ncell, 3 * i + 3), fld(A->d_slot, A->ncell, 3 * i + 4),
                              fld(A->d_slot, A->ncell, 3 * i + 5), A->d_smask + i * A->nwords,
                              A->d_h2e + static_cast<size_t>(i) * A->nedge);
      if (rc) return rc;
    }
    // reverse sweep over the segment
    for (int s = s1 - 1; s >= s0; --s) {
      const int i = s - s0;
      R.gs = g[s];
      R.h1 = fld(A->d_slot, A->ncell, 3 * i);
      R.qx0 = fld(A->d_slot, A->ncell, 3 * i + 1);
      R.qy0 = fld(A->d_slot, A->ncell, 3 * i + 2);
      R.mask = A->d_smask + i * A->nwords;
      R.h2e = A->d_h2e + static_cast<size_t>(i) * A->nedge;
      hipLaunchKernelGGL(k_adj_face, cell_grid2(c), dim3(256), 0, st, R);
      ADJ_HIP(A, hipGetLastError());
      hipLaunchKernelGGL(k_adj_cell, dim3(cell_blocks), dim3(256), 0, st, R, A->rain[s] * A->dt,
                         A->nb, A->d_gm);
      ADJ_HIP(A, hipGetLastError());
    }
  }
  // gradients out: dL/dn on the reference layout, dL/dmult summed in block order
  std::vector<double> gm(static_cast<size_t>(cell_blocks) * A->nb);
  ADJ_HIP(A, hipMemcpyAsync(gm.data(), A->d_gm, gm.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  std::vector<double> gn(A->ncell);
  ADJ_HIP(A, hipMemcpyAsync(gn.data(), R.gn, cs, hipMemcpyDeviceToHost, st));
  ADJ_HIP(A, hipStreamSynchronize(st));
  if (grad_mult)
    for (int b = 0; b < A->nb; ++b) {
      double t = 0;
      for (int k = 0; k < cell_blocks; ++k) t += gm[static_cast<size_t>(k) * A->nb + b];
      grad_mult[b] = t;
    }
  if (grad_n)
    for (int x = 0; x < c->NR; ++x)
      for (int y = 0; y < c->NC; ++y)
        grad_n[static_cast<size_t>(x) * c->NC + y] = gn[static_cast<size_t>(x) * c->P + y];
  return HC_OK;
}

}  // extern "C"
