// Device-resident stepping: the loop body of src/main.cpp:139-204 for
// hydro-only runs with every scalar kept on the device, so that n iterations
// are enqueued (or replayed from a captured HIP graph) without a host
// round trip per step.
//
// Per iteration, three launches:
//   k_loop_control  (1 block)  set_loop_cycle, set_inputoutput_diff,
//                              set_maximum_timestep, increment_counters,
//                              catchment_water_input_and_hydrology (input
//                              point injection, TOPMODEL catch-up from the
//                              host-tabulated sequence, calchydrograph,
//                              baseflow triggers), get_flow_timestep, and
//                              the end-of-run test (main.cpp:204), which
//                              takes effect before the next body
//   fused flow_route + depth_update + water_flux_out (dt and time_step read
//                              from the control block; the edge tiles drain
//                              and store one partial sum each)
//   k_loop_scan                temptot from the edge partials (block 0, in
//                              tile order), scan_area when counter % 5 == 0
//                              (FP32 forms: from the wet-cell bits the fused
//                              kernel stored on that body, not the depths),
//                              and the input-point rebuild after a baseflow
//                              trigger (get_catchment_input_points,
//                              LCM.cpp:3741-3767)
// Draining the edges before the scan leaves the mask unchanged: a drained
// cell keeps the threshold depth, which is still wet.  The rebuilt point list
// is first used by the next iteration's injection, as in the reference.
// The scalar arithmetic follows the reference's operation order in FP64.
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "hc_device.hpp"
#include "hc_internal.hpp"

using hc::kGravity;
using hc::kRfMax;
using hc::EnsMember;
using hc::LoopArgs;
using hc::LoopCtl;
using hc::LoopRecord;
using namespace hcdev;

namespace {

__device__ __forceinline__ double cfl(const LoopArgs& a, double maxdepth) {
  return a.courant * (a.DX / sqrt(kGravity * (maxdepth)));  // LCM.cpp:1884-1887
}

// The scalar part of one loop body for one catchment (thread 0 of the block
// calling it).  The rain input itself is added by the fused kernel as it
// loads the input-point cells (amounts in ctl->amt).
__device__ __forceinline__ void loop_control(const LoopArgs& a, LoopCtl* __restrict__ ctl,
                                             const double* __restrict__ top,
                                             uint64_t* __restrict__ pmask,
                                             unsigned long long* __restrict__ red,
                                             LoopRecord* __restrict__ rec) {
  if (threadIdx.x == 0) {
    LoopCtl& C = *ctl;
    bool stop = C.halt || C.err;
    if (!stop && C.halt_pending) {  // the previous body ended the run
      C.halt = 1;
      stop = true;
    }
    if (!stop) {
      // health of the previous body, tested before anything of this one
      // moves: its depth kernel reports +inf for a non-finite or large
      // negative depth.  This body's flow timestep is min(time_step, the
      // Courant limit at that maximum); set_maximum_timestep keeps time_step
      // at or above the limit and the TOPMODEL catch-up lowers it to no less
      // than max_time_step, which hc_loop_begin holds at or above the floor,
      // so the timestep falls below the floor exactly when the limit does.
      const double md = __longlong_as_double(static_cast<long long>(red[0]));
      if (!(md < INFINITY))
        C.err = hc::kLoopErrDepth;
      else if (a.dt_floor > 0 && cfl(a, md <= 0.1 ? 0.1 : md) < a.dt_floor)
        C.err = hc::kLoopErrDtFloor;
      if (C.err) {
        C.halt = 1;
        stop = true;
      }
    }
    if (!stop) {
      // a list rebuilt by the previous iteration: an empty one keeps a single
      // stale point, in_x[1]/in_y[1] of the last non-empty list (LCM.cpp:3764)
      if (C.pts_dirty) {
        if (C.npts == 0) {
          const int px = C.last_first / a.pitch, py = C.last_first % a.pitch;
          pmask[static_cast<size_t>(px) * a.mwords + (py >> 6)] |= 1ull << (py & 63);
          C.npts = 1;
        } else {
          C.last_first = C.first_new;
        }
        C.pts_dirty = 0;
      }
      C.maxdepth = __longlong_as_double(static_cast<long long>(red[0]));
      // set_loop_cycle, set_inputoutput_diff (LCM.cpp:1873-1882)
      C.previous = C.cycle;
      C.old_cycle = fmod(C.cycle, a.osi);
      C.iod = fabs(C.waterinput - C.waterOut);
      // set_maximum_timestep (LCM.cpp:1890-1910)
      if (C.maxdepth <= 0.1) C.maxdepth = 0.1;
      if (C.time_step < cfl(a, C.maxdepth)) C.time_step = cfl(a, C.maxdepth);
      if (C.iod > a.in_out_allowed && C.time_step > cfl(a, C.maxdepth))
        C.time_step = cfl(a, C.maxdepth);
      // increment_counters (LCM.cpp:1925-1933)
      C.counter++;
      C.cycle += C.time_step / 60;
      C.new_cycle = fmod(C.cycle, a.osi);
      // catchment_waterinputs (LCM.cpp:2027-2039, 3455-3541)
      C.waterinput = 0;
      double fts = C.time_step;
      if (fts > cfl(a, C.maxdepth)) fts = cfl(a, C.maxdepth);
      for (int i = 1; i <= a.rfnum; ++i) C.waterinput += C.j_mean[i] * C.nActual[i] * a.DX * a.DX;
      for (int rf = 0; rf <= a.rfnum; ++rf) {
        double amt = (C.j_mean[rf] * C.nActual[rf]) / (C.input_counter[rf]) * fts;
        if (amt > a.ERODEFACTOR) amt = a.ERODEFACTOR;
        C.amt[rf] = amt;
      }
      if (C.cycle >= C.time_1) {
        do {
          C.time_1++;
          const int t = static_cast<int>(C.time_1);
          if (t < 0 || t >= a.ntop) {
            C.err = hc::kLoopErrTable;
            C.halt = 1;
            break;
          }
          for (int n = 1; n <= a.rfnum; ++n) {
            C.old_j_mean[n] = C.new_j_mean[n];
            C.new_j_mean[n] = top[t * kRfMax + n];
          }
          if (C.time_step > a.max_time_step) {
            double jmt = 0;
            for (int n = 1; n <= a.rfnum; ++n)
              if (C.new_j_mean[n] > jmt) jmt = C.new_j_mean[n];
            const unsigned cells = static_cast<unsigned>(a.imax) * static_cast<unsigned>(a.jmax);
            if (jmt > (0.2 / (cells * a.DX * a.DX))) {
              C.cycle = C.time_1 + (a.max_time_step / 60);
              C.time_step = a.max_time_step;
            }
          }
        } while (C.time_1 < C.cycle);
      }

      const double tt = C.time_1 - C.cycle;  // calchydrograph (LCM.cpp:3717-3724)
      for (int n = 1; n <= a.rfnum; ++n)
        C.j_mean[n] = C.old_j_mean[n] + (((C.new_j_mean[n] - C.old_j_mean[n]) / 2) * (2 - tt));
      double jmeanmax = 0;
      for (int n = 1; n <= a.rfnum; ++n)
        if (C.j_mean[n] > jmeanmax) jmeanmax = C.j_mean[n];
      bool rebuild = false;
      if (jmeanmax >= C.baseflow) {
        C.baseflow = C.baseflow * 3;
        rebuild = true;
      }
      if (C.baseflow > (jmeanmax * 3) && C.baseflow > 0.000001) {
        C.baseflow = jmeanmax * 1.25;
        rebuild = true;
      }
      if (rebuild) {
        C.pts_dirty = 1;
        C.npts = 0;
        for (int rf = 0; rf < kRfMax; ++rf) C.input_counter[rf] = 0;
        C.first_new = INT_MAX;
      }
      double f = C.time_step;  // get_flow_timestep as flow_route sees it
      if (f > cfl(a, C.maxdepth)) f = cfl(a, C.maxdepth);
      C.dt_flow = f;
      red[0] = 0ull;
      LoopRecord& r = rec[C.iter];
      r.cycle = C.cycle;
      r.previous = C.previous;
      r.old_cycle = C.old_cycle;
      r.new_cycle = C.new_cycle;
      r.temptot = 0;
      r.dt_flow = C.dt_flow;
      r.time_step = C.time_step;
      r.waterinput = C.waterinput;
      for (int n = 0; n < kRfMax; ++n) r.j_mean[n] = C.j_mean[n];
      // main.cpp:204 on this body's clock (nothing later in the body moves it)
      if (!(C.cycle / 60 < a.maxcycle)) C.halt_pending = 1;
    }
  }
}

__global__ __launch_bounds__(64) void k_loop_control(LoopArgs a, LoopCtl* __restrict__ ctl,
                                                     const double* __restrict__ top,
                                                     uint64_t* __restrict__ pmask,
                                                     unsigned long long* __restrict__ red,
                                                     LoopRecord* __restrict__ rec) {
  loop_control(a, ctl, top, pmask, red, rec);
}

// get_catchment_input_points (LCM.cpp:3741-3767) after a baseflow trigger:
// rewrites the whole input-point mask (one 64-column word per wavefront,
// grid-stride) and accumulates the point count, the per-rain-cell counters
// and the first point in list order (row-major) through the block.  Every
// thread of the block calls it.
__device__ __forceinline__ void rebuild_points(const LoopArgs& a, LoopCtl* __restrict__ ctl,
                                               const double* __restrict__ area,
                                               const int* __restrict__ rfarea,
                                               uint64_t* __restrict__ pmask) {
  __shared__ int s_cnt[kRfMax];
  __shared__ int s_n, s_first;
  if (threadIdx.x < kRfMax) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    s_n = 0;
    s_first = INT_MAX;
  }
  __syncthreads();
  const double b = ctl->baseflow;
  const int lane = threadIdx.x & 63;
  const long nwords = static_cast<long>(a.imax + 2) * a.mwords;
  const long nwaves = static_cast<long>(gridDim.x) * (blockDim.x >> 6);
  int my_n = 0, my_first = INT_MAX;
  for (long wi = static_cast<long>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
       wi < nwords; wi += nwaves) {
    const int x = static_cast<int>(wi / a.mwords);
    const int y = static_cast<int>(wi % a.mwords) * 64 + lane;
    bool pt = false;
    int idx = 0;
    if (x >= 1 && x <= a.imax && y >= 1 && y <= a.jmax) {
      idx = x * a.pitch + y;
      const double v = area[idx] * b * 3 * a.DX * a.DX;
      pt = v > a.MIN_Q && v < a.MIN_Q_MAXVAL;
    }
    const uint64_t bits = __ballot(pt);
    if (lane == 0) pmask[wi] = bits;
    if (pt) {
      atomicAdd(&s_cnt[rfarea[idx]], 1);
      ++my_n;
      my_first = min(my_first, idx);
    }
  }
  if (my_n) {
    atomicAdd(&s_n, my_n);
    atomicMin(&s_first, my_first);
  }
  __syncthreads();
  if (threadIdx.x < kRfMax && s_cnt[threadIdx.x]) atomicAdd(&ctl->input_counter[threadIdx.x], s_cnt[threadIdx.x]);
  if (threadIdx.x == 0 && s_n) {
    atomicAdd(&ctl->npts, s_n);
    atomicMin(&ctl->first_new, s_first);
  }
}

// scan_area from the wet-cell bits the fused kernel left (bit = stored depth
// > 0, a NaN counting as dry, rows 0..imax+1): one thread per mask word ORs
// the words of the rows above and below, dilates the result one column either
// way (the side words supply the bits past the word's ends) and keeps
// columns 1..jmax; rows 0 and imax+1 are cleared.  The same mask as
// scan_item's from the depths themselves, from 8 bytes per 64 cells.
__device__ __forceinline__ void scan_from_wet(int imax, int jmax, int W,
                                              const uint64_t* __restrict__ wet,
                                              uint64_t* __restrict__ mask) {
  const long n = static_cast<long>(imax + 2) * W;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const int x = static_cast<int>(i / W), w = static_cast<int>(i % W);
    uint64_t m = 0;
    if (x >= 1 && x <= imax) {
      const uint64_t* r = wet + static_cast<size_t>(x - 1) * W + w;  // rows x-1 .. x+1
      const uint64_t v = r[0] | r[W] | r[2 * W];
      const uint64_t west = w > 0 ? (r[-1] | r[W - 1] | r[2 * W - 1]) >> 63 : 0ull;
      const uint64_t east = w + 1 < W ? (r[1] | r[W + 1] | r[2 * W + 1]) << 63 : 0ull;
      m = v | (v << 1) | west | (v >> 1) | east;
      const int lo = w * 64;
      uint64_t keep = ~0ull;
      if (lo == 0) keep &= ~1ull;
      const int hi = jmax + 1 - lo;  // first column past jmax, relative to the word
      if (hi <= 0) keep = 0;
      else if (hi < 64) keep &= (1ull << hi) - 1;
      m &= keep;
    }
    mask[i] = m;
  }
}

// Partial sums of the edge tiles in slot order, then a fixed tree: the
// water_flux_out total for the iteration.
__device__ void finish_drain(LoopCtl* __restrict__ ctl, LoopRecord* __restrict__ rec,
                             const double* __restrict__ partials, int nby, int nbx, int R,
                             int imax, int jmax, int sw) {
  __shared__ double s_sum[256];
  const int t = threadIdx.x;
  const hc::EdgeSlots es(nbx, R, imax, jmax, sw);
  const int n = es.count(nby);
  double acc = 0;
  for (int k = t; k < n; k += 256) {
    const double v = partials[k];  // holes hold stale values: selected out
    acc += es.valid(k) ? v : 0.0;
  }
  s_sum[t] = acc;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if (t < o) s_sum[t] += s_sum[t + o];
    __syncthreads();
  }
  if (t == 0) {
    LoopCtl& C = *ctl;
    const double tot = s_sum[0];
    C.temptot = tot;
    C.waterOut = tot;
    rec[C.iter].temptot = tot;
    C.iter++;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_loop_scan(LoopArgs a, LoopCtl* __restrict__ ctl, int W,
                                                   const T* __restrict__ h,
                                                   uint64_t* __restrict__ mask,
                                                   const double* __restrict__ area,
                                                   const int* __restrict__ rfarea,
                                                   uint64_t* __restrict__ pmask,
                                                   const double* __restrict__ partials,
                                                   int nby, int nbx, int R, int sw,
                                                   LoopRecord* __restrict__ rec,
                                                   const double* __restrict__ wmax, int nwmax,
                                                   unsigned long long* __restrict__ red,
                                                   const uint64_t* __restrict__ wet) {
  // the control kernel alone sets halt, so the fused kernel of this
  // iteration ran exactly when it is still clear here
  const bool live = !ctl->halt && !ctl->err;
  const bool scan = live && (ctl->counter % a.scan_interval) == 0;
  const bool rebuild = live && ctl->pts_dirty;
  if (live) {  // uniform: the fused kernel's per-wavefront maxima into maxdepth
    double m = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nwmax; i += gridDim.x * 256)
      m = fmax(m, wmax[i]);
    block_max_pos(red, m);
  }
  if (live && blockIdx.x == 0)
    finish_drain(ctl, rec, partials, nby, nbx, R, a.imax, a.jmax, sw);
  if (!scan && !rebuild) return;  // uniform over the grid
  if (scan && wet) {
    scan_from_wet(a.imax, a.jmax, W, wet, mask);
  } else if (scan) {
    const int items = scan_items(a.imax, W, static_cast<int>(sizeof(T)));
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform
    for (int it = blockIdx.x * 4 + wv; it < items; it += gridDim.x * 4)
      scan_item<T>(it, a.imax, a.jmax, a.pitch, W, h, mask);
  }
  if (rebuild) rebuild_points(a, ctl, area, rfarea, pmask);
}

__global__ void k_loop_reset_iter(LoopCtl* ctl) { ctl->iter = 0; }

// ---------------------------------------------------------------- ensemble
// One block per member: that member's loop_control on its current depth.
// It also leaves the member's two per-step flags for the scan kernel:
// flags[m] = this step scans the active area (counter % 5 == 0, as
// k_loop_scan tests it), flags[n + m] = this step rebuilds the input points.
__global__ __launch_bounds__(64) void k_ens_control(const EnsMember* __restrict__ mem, int n,
                                                    int* __restrict__ flags) {
  const int m = blockIdx.x;
  const EnsMember& M = mem[m];
  LoopCtl* ctl = table_ptr<LoopCtl>(M.ctl);
  loop_control(M.la, ctl, table_ptr<const double>(M.top), table_ptr<uint64_t>(M.pmask),
               table_ptr<unsigned long long>(M.red), table_ptr<LoopRecord>(M.rec));
  if (threadIdx.x == 0) {
    const bool live = !ctl->halt && !ctl->err;
    flags[m] = (live && (ctl->counter % M.la.scan_interval) == 0) ? 1 : 0;
    flags[n + m] = (live && ctl->pts_dirty) ? 1 : 0;
  }
}

// Per member: temptot from its edge partials (one block each), scan_area on
// every 5th of its own steps, and its input-point rebuild.  Scan work items
// of all members form one list (item_base holds each member's first item,
// item_member each item's member); the control kernel's flags say which
// members scan or rebuild in this step, so an item of a member that does not
// scan costs two loads.
template <typename T>
__global__ __launch_bounds__(256) void k_ens_scan(const EnsMember* __restrict__ mem, int n,
                                                  const int* __restrict__ item_base,
                                                  const int* __restrict__ item_member,
                                                  const int* __restrict__ flags, int R,
                                                  int parity) {
  for (int m = blockIdx.x; m < n; m += gridDim.x) {
    const EnsMember& M = mem[m];
    LoopCtl* ctl = table_ptr<LoopCtl>(M.ctl);
    if (!ctl->halt && !ctl->err)
      finish_drain(ctl, table_ptr<LoopRecord>(M.rec), table_ptr<const double>(M.dpart), M.nby,
                   M.nbx, R, M.la.imax, M.la.jmax, hc::kStrideOverlap);
  }
  const int total = item_base[n];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform
  for (int it = blockIdx.x * 4 + wv; it < total; it += gridDim.x * 4) {
    const int m = item_member[it];
    if (!flags[m]) continue;  // per wave
    const EnsMember& M = mem[m];
    scan_item<T>(it - item_base[m], M.la.imax, M.la.jmax, M.la.pitch, M.W,
                 table_ptr<const T>(M.h[M.cur0 ^ parity ^ 1]), table_ptr<uint64_t>(M.mask));
  }
  // input-point rebuilds: the flags of a chunk of members are read at once
  // and the chunk skipped when none is set (the usual case)
  __shared__ int s_dirty[256];
  for (int base = 0; base < n; base += 256) {
    const int m = base + static_cast<int>(threadIdx.x);
    const int d = (m < n) ? flags[n + m] : 0;
    s_dirty[threadIdx.x] = d;
    if (!__syncthreads_or(d)) continue;
    for (int k = 0; k < 256 && base + k < n; ++k) {
      if (!s_dirty[k]) continue;  // uniform
      const EnsMember& M = mem[base + k];
      rebuild_points(M.la, table_ptr<LoopCtl>(M.ctl), table_ptr<const double>(M.area),
                     table_ptr<const int>(M.rfarea), table_ptr<uint64_t>(M.pmask));
    }
    __syncthreads();  // s_dirty is rewritten by the next chunk
  }
}

__global__ void k_ens_reset(const EnsMember* __restrict__ mem, int n) {
  for (int m = threadIdx.x; m < n; m += blockDim.x) table_ptr<LoopCtl>(mem[m].ctl)->iter = 0;
}

// One iteration's launches on the context stream.  On a body that the
// host's counter forecast says rebuilds the active mask (FP32 forms), the
// fused kernel's variant that also stores the wet-cell bits runs, and the
// scan kernel dilates those; the scan kernel still tests the counter itself
// (and scans the depths if no bits were stored), so the forecast only
// chooses the faster path.
int a_scan_interval(const hc_ctx* c) { return c->largs.scan_interval; }

int enqueue_iteration(hc_ctx* c, int k_in_run = -1) {
  const LoopArgs& a = c->largs;
  const bool scan_body = (c->loop_counter + 1) % a.scan_interval == 0;
  ++c->loop_counter;
  uint64_t* wet = (c->prec != HC_FP64 && scan_body) ? c->d_wet : nullptr;
  hipLaunchKernelGGL(k_loop_control, dim3(1), dim3(64), 0, c->stream, a, c->d_ctl, c->d_top,
                     c->d_pmask, c->d_red, c->d_rec);
  HC_HIP(c, hipGetLastError());
  const bool prof = c->prof_on && k_in_run >= 0 && 2 * k_in_run + 1 < static_cast<int>(c->prof_ev.size());
  hc::DrainArgs dr;
  dr.ts = &c->d_ctl->time_step;
  dr.thr = c->prm.water_depth_erosion_threshold;
  dr.partials = c->d_dpart;
  hc::InjectArgs inj;
  inj.pmask = c->d_pmask;
  inj.rfarea = c->d_rfarea;
  inj.amt = c->d_ctl->amt;
  inj.rf_uniform = c->rf_uniform;
  int rc = hc_launch_flow_depth(c, 0.0, &c->d_ctl->dt_flow, &c->d_ctl->halt, true, &dr, &inj,
                                c->d_wmax, prof ? c->prof_ev[2 * k_in_run] : nullptr,
                                prof ? c->prof_ev[2 * k_in_run + 1] : nullptr, wet);
  if (rc) return rc;
  if (prof) c->prof_n = k_in_run + 1;
  const int o = c->cur;  // flow_depth flipped the generation
  const int items = hcdev::scan_items(c->g.imax, c->W, static_cast<int>(c->esz));
  const int nb = std::max(1, std::min((items + 3) / 4, 1024));
  const int R = c->fused_rows * c->fused_strip;  // rows per fused block
  int nbx, nby;
  const int sw = hc_fused_stride(c);
  hc_fused_geometry(c, R, sw, &nbx, &nby);
  const int nwmax = 4 * nbx * nby;
  if (c->prec == HC_FP64)
    hipLaunchKernelGGL(k_loop_scan<double>, dim3(nb), dim3(256), 0, c->stream, a, c->d_ctl, c->W,
                       (const double*)c->h[o], c->mask, c->d_area, c->d_rfarea, c->d_pmask,
                       c->d_dpart, nby, nbx, R, sw, c->d_rec, c->d_wmax, nwmax, c->d_red,
                       nullptr);
  else
    hipLaunchKernelGGL(k_loop_scan<float>, dim3(nb), dim3(256), 0, c->stream, a, c->d_ctl, c->W,
                       (const float*)c->h[o], c->mask, c->d_area, c->d_rfarea, c->d_pmask,
                       c->d_dpart, nby, nbx, R, sw, c->d_rec, c->d_wmax, nwmax, c->d_red,
                       wet);
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

template <typename T>
int upload_pitched(hc_ctx* c, T** dst, const T* host) {
  if (!*dst) HC_HIP(c, hipMalloc(reinterpret_cast<void**>(dst), c->ncell * sizeof(T)));
  std::vector<T> tmp(c->ncell, T(0));
  for (int x = 0; x < c->NR; ++x)
    for (int y = 0; y < c->NC; ++y)
      tmp[static_cast<size_t>(x) * c->P + y] = host[static_cast<size_t>(x) * c->NC + y];
  HC_HIP(c, hipMemcpy(*dst, tmp.data(), c->ncell * sizeof(T), hipMemcpyHostToDevice));
  return HC_OK;
}

}  // namespace

extern "C" {

int hc_loop_begin(hc_ctx* c, const hc_loop_init* in) {
  if (!c || !in || !in->topmodel || !in->area || !in->rfarea) return HC_ERR_ARG;
  if (c->g.sediment)
    return hc_fail(c, HC_ERR_STATE, "loop: device-resident stepping is for hydro-only contexts");
  if (in->rfnum < 1 || in->rfnum >= kRfMax)
    return hc_fail(c, HC_ERR_ARG, "loop: rfnum out of range (1..16)");
  if (in->npoints < 0 || in->npoints > c->g.imax * c->g.jmax + 1)
    return hc_fail(c, HC_ERR_ARG, "loop: bad input point count");
  if (in->first_x < 0 || in->first_x >= c->NR || in->first_y < 0 || in->first_y >= c->NC)
    return hc_fail(c, HC_ERR_ARG, "loop: first input point outside the padded grid");
  if (!(in->dt_floor >= 0) || in->dt_floor > in->max_time_step)
    return hc_fail(c, HC_ERR_ARG, "loop: dt_floor must lie in [0, max_time_step]");
  HC_HIP(c, hipStreamSynchronize(c->stream));
  LoopArgs& a = c->largs;
  a.courant = in->courant_number;
  a.DX = c->g.DX;
  a.in_out_allowed = in->in_out_difference_allowed;
  a.osi = in->output_file_save_interval;
  a.ERODEFACTOR = in->ERODEFACTOR;
  a.MIN_Q = in->MIN_Q;
  a.MIN_Q_MAXVAL = in->MIN_Q_MAXVAL;
  a.dt_floor = in->dt_floor;
  a.max_time_step = in->max_time_step;
  a.maxcycle = in->maxcycle;
  a.rfnum = in->rfnum;
  a.imax = c->g.imax;
  a.jmax = c->g.jmax;
  a.pitch = c->P;
  a.ntop = in->ntop;
  a.scan_interval = 5;
  a.mwords = c->W;
  // tables and inputs
  if (!c->d_ctl) HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_ctl), sizeof(LoopCtl)));
  {
    // edge-drain partials: the largest slot count over the tile heights and
    // both wavefront strides (single grid and ensemble)
    size_t need = 1;
    for (int sw : {hc::kStrideAligned, hc::kStrideOverlap}) {
      for (int R : {4, 8, 16}) {
        int nbx, nby;
        hc_fused_geometry(c, R, sw, &nbx, &nby);
        need = std::max(need, static_cast<size_t>(
                                  hc::EdgeSlots(nbx, R, c->g.imax, c->g.jmax, sw).count(nby)));
      }
    }
    if (need > c->dpart_cap) {
      if (c->d_dpart) (void)hipFree(c->d_dpart);
      c->d_dpart = nullptr;
      HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_dpart), need * sizeof(double)));
      c->dpart_cap = need;
    }
    // per-wavefront maxima of the single-grid fused kernel: 4 per block at
    // the most blocks any tiling makes (4 rows per block)
    int nbx, nby;
    hc_fused_geometry(c, 4, hc_fused_stride(c), &nbx, &nby);
    const size_t wneed = static_cast<size_t>(4) * nbx * nby;
    if (wneed > c->wmax_cap) {
      if (c->d_wmax) (void)hipFree(c->d_wmax);
      c->d_wmax = nullptr;
      hc_drop_loop_graphs(c);
      HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_wmax), wneed * sizeof(double)));
      c->wmax_cap = wneed;
    }
  }
  if (c->d_top) (void)hipFree(c->d_top);
  HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_top), sizeof(double) * kRfMax * std::max(1, in->ntop)));
  HC_HIP(c, hipMemcpy(c->d_top, in->topmodel, sizeof(double) * kRfMax * in->ntop, hipMemcpyHostToDevice));
  int rc = upload_pitched(c, &c->d_area, in->area);
  if (rc) return rc;
  rc = upload_pitched(c, &c->d_rfarea, in->rfarea);
  if (rc) return rc;
  // every rain cell index must address the amounts table (any cell can
  // become an input point after a rebuild)
  c->rf_uniform = in->rfarea[0];
  for (size_t k = 0; k < static_cast<size_t>(c->NR) * c->NC; ++k) {
    if (in->rfarea[k] < 0 || in->rfarea[k] >= kRfMax)
      return hc_fail(c, HC_ERR_ARG, "loop: rain cell index out of range (0..16)");
    if (in->rfarea[k] != c->rf_uniform) c->rf_uniform = -1;
  }
  const size_t nwords = static_cast<size_t>(c->NR) * c->W;
  if (!c->d_pmask) HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_pmask), nwords * sizeof(uint64_t)));
  // FP32 forms (aligned tiles): the fused kernel's wet bits for the scan
  if (c->prec != HC_FP64 && !c->d_wet)
    HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_wet), nwords * sizeof(uint64_t)));
  std::vector<uint64_t> pm(nwords, 0ull);
  for (int k = 0; k < in->npoints; ++k) {
    const int x = in->in_x[k], y = in->in_y[k];
    if (x < 0 || x >= c->NR || y < 0 || y >= c->NC)
      return hc_fail(c, HC_ERR_ARG, "loop: input point outside the padded grid");
    pm[static_cast<size_t>(x) * c->W + (y >> 6)] |= 1ull << (y & 63);
  }
  HC_HIP(c, hipMemcpy(c->d_pmask, pm.data(), nwords * sizeof(uint64_t), hipMemcpyHostToDevice));
  LoopCtl C;
  std::memset(&C, 0, sizeof(C));
  C.time_step = in->time_step;
  C.cycle = in->cycle;
  C.maxdepth = in->maxdepth;
  C.waterinput = in->waterinput;
  C.waterOut = in->waterOut;
  C.time_1 = in->time_1;
  C.baseflow = in->baseflow;
  C.counter = in->counter;
  c->loop_counter = in->counter;
  for (int n = 0; n < kRfMax; ++n) {
    C.j_mean[n] = in->j_mean[n];
    C.old_j_mean[n] = in->old_j_mean[n];
    C.new_j_mean[n] = in->new_j_mean[n];
    C.nActual[n] = in->nActual[n];
    C.input_counter[n] = in->input_counter[n];
  }
  C.npts = in->npoints;
  C.last_first = in->first_x * c->P + in->first_y;
  C.first_new = INT_MAX;
  HC_HIP(c, hipMemcpy(c->d_ctl, &C, sizeof(C), hipMemcpyHostToDevice));
  // the control kernel reads the previous maxdepth from the reduction slot
  unsigned long long bits;
  std::memcpy(&bits, &in->maxdepth, sizeof(bits));
  HC_HIP(c, hipMemcpy(c->d_red, &bits, sizeof(bits), hipMemcpyHostToDevice));
  hc_drop_loop_graphs(c);
  c->loop_ready = true;
  return HC_OK;
}

int hc_loop_run(hc_ctx* c, int n, int use_graph) {
  if (!c || n < 0) return HC_ERR_ARG;
  if (!c->loop_ready) return hc_fail(c, HC_ERR_STATE, "loop: hc_loop_begin first");
  if (n > c->rec_cap) {
    // the captured graphs write records through the old address
    HC_HIP(c, hipStreamSynchronize(c->stream));
    hc_drop_loop_graphs(c);
    if (c->d_rec) (void)hipFree(c->d_rec);
    c->d_rec = nullptr;
    const int cap = std::max(n, std::max(1024, 2 * c->rec_cap));
    HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_rec), sizeof(LoopRecord) * (cap + 1)));
    c->rec_cap = cap;
  }
  c->loop_cur0 = c->cur;
  c->loop_enqueued = n;
  hipLaunchKernelGGL(k_loop_reset_iter, dim3(1), dim3(1), 0, c->stream, c->d_ctl);
  HC_HIP(c, hipGetLastError());
  constexpr int kBlock = 10;  // even (ping-pong) and a multiple of the scan interval
  int left = n;
  if (use_graph && n >= kBlock) {
    // a graph per starting generation and counter phase (which of its
    // bodies rebuild the active mask); kBlock keeps both fixed from one
    // launch of the graph to the next
    const int parity = c->cur;
    const int phase = static_cast<int>(c->loop_counter % a_scan_interval(c));
    hipGraphExec_t& gx = c->loop_graph[parity][phase];
    if (!gx) {
      hipGraph_t graph;
      const long counter = c->loop_counter;
      HC_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rc = HC_OK;
      for (int k = 0; k < kBlock && rc == HC_OK; ++k) rc = enqueue_iteration(c);
      hipError_t e = hipStreamEndCapture(c->stream, &graph);
      c->cur = parity;  // capture does not execute; restore the generation and forecast
      c->loop_counter = counter;
      if (rc) return rc;
      if (e != hipSuccess) return hc_fail(c, HC_ERR_HIP, "loop: graph capture failed");
      e = hipGraphInstantiate(&gx, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return hc_fail(c, HC_ERR_HIP, "loop: graph instantiate failed");
    }
    while (left >= kBlock) {
      HC_HIP(c, hipGraphLaunch(gx, c->stream));
      c->loop_counter += kBlock;
      left -= kBlock;
    }
  }
  c->prof_n = 0;
  if (c->prof_on && !use_graph) {
    const size_t need = 2 * static_cast<size_t>(left);
    while (c->prof_ev.size() < need) {
      hipEvent_t e;
      HC_HIP(c, hipEventCreate(&e));
      c->prof_ev.push_back(e);
    }
  }
  for (int k = 0; left > 0; --left, ++k) {
    const int rc = enqueue_iteration(c, k);
    if (rc) return rc;
  }
  return HC_OK;
}

int hc_loop_profile(hc_ctx* c, int on) {
  if (!c) return HC_ERR_ARG;
  c->prof_on = on != 0;
  return HC_OK;
}

int hc_loop_kernel_time(hc_ctx* c, double* total_ms, int* launches) {
  if (!c) return HC_ERR_ARG;
  HC_HIP(c, hipStreamSynchronize(c->stream));
  double tot = 0;
  int n = 0;
  int iters = 0;
  {
    LoopCtl C;
    HC_HIP(c, hipMemcpy(&C, c->d_ctl, sizeof(C), hipMemcpyDeviceToHost));
    iters = C.iter;
  }
  for (int k = 0; k < c->prof_n && k < iters; ++k) {
    float ms = 0;
    HC_HIP(c, hipEventElapsedTime(&ms, c->prof_ev[2 * k], c->prof_ev[2 * k + 1]));
    tot += ms;
    ++n;
  }
  if (total_ms) *total_ms = tot;
  if (launches) *launches = n;
  return HC_OK;
}

int hc_loop_sync(hc_ctx* c, hc_loop_state* st, hc_loop_record* rec, int max_records, int* nrec) {
  if (!c) return HC_ERR_ARG;
  if (!c->loop_ready) return hc_fail(c, HC_ERR_STATE, "loop: hc_loop_begin first");
  HC_HIP(c, hipStreamSynchronize(c->stream));
  LoopCtl C;
  HC_HIP(c, hipMemcpy(&C, c->d_ctl, sizeof(C), hipMemcpyDeviceToHost));
  // iterations that ran flipped the ping-pong generation once each
  c->cur = c->loop_cur0 ^ (C.iter & 1);
  c->loop_counter = C.counter;
  c->loop_enqueued = 0;
  if (st) {
    st->time_step = C.time_step;
    st->cycle = C.cycle;
    unsigned long long bits;
    HC_HIP(c, hipMemcpy(&bits, c->d_red, sizeof(bits), hipMemcpyDeviceToHost));
    std::memcpy(&st->maxdepth, &bits, sizeof(bits));
    st->temptot = C.temptot;
    st->waterinput = C.waterinput;
    st->waterOut = C.waterOut;
    st->time_1 = C.time_1;
    st->baseflow = C.baseflow;
    st->counter = C.counter;
    st->finished = C.halt || C.halt_pending;
    st->npoints = C.npts;
    st->error = C.err;
    for (int n = 0; n < kRfMax; ++n) {
      st->j_mean[n] = C.j_mean[n];
      st->old_j_mean[n] = C.old_j_mean[n];
      st->new_j_mean[n] = C.new_j_mean[n];
    }
  }
  const int nr = std::min(C.iter, std::max(0, max_records));
  if (rec && nr > 0)
    HC_HIP(c, hipMemcpy(rec, c->d_rec, sizeof(LoopRecord) * nr, hipMemcpyDeviceToHost));
  if (nrec) *nrec = C.iter;
  if (C.err == hc::kLoopErrTable) return hc_fail(c, HC_ERR_HALTED, "loop: TOPMODEL table exhausted");
  if (C.err == hc::kLoopErrDepth)
    return hc_fail(c, HC_ERR_UNSTABLE, "loop: non-finite or negative water depth after iteration " +
                                           std::to_string(C.iter) + " of the call");
  if (C.err == hc::kLoopErrDtFloor)
    return hc_fail(c, HC_ERR_UNSTABLE, "loop: flow timestep below the floor (" +
                                           std::to_string(c->largs.dt_floor) + " s) after iteration " +
                                           std::to_string(C.iter) + " of the call");
  if (C.err) return hc_fail(c, HC_ERR_HALTED, "loop: halted");
  return HC_OK;
}

}  // extern "C"

// -------------------------------------------------------------- ensemble API
struct hc_ens {
  std::vector<hc_ctx*> m;
  int mode = 0;           // 0 FP64, 1 FP32 with FP32 elevations, 2 FP32 with FP64 elevations,
                          // 3 FP32 with fixed-point elevations
  int R = 8;              // fused tile rows
  int nblocks = 1024;     // persistent grid of the fused step
  int scan_blocks = 1024;
  hipStream_t stream = nullptr;
  std::vector<EnsMember> mem;
  EnsMember* d_mem = nullptr;
  int2* d_tiles = nullptr;
  size_t tiles_cap = 0;
  int ntiles = 0;
  int* d_items = nullptr;  // n+1 prefix sums of scan items
  int* d_item_member = nullptr;  // member of each scan item
  int* d_flags = nullptr;        // 2n per-step flags from the control kernel
  hipGraphExec_t graph = nullptr;
  std::string err;
};

namespace {

int ens_fail(hc_ens* e, int code, const std::string& msg) {
  e->err = msg;
  return code;
}

#define ENS_HIP(e, call)                                                         \
  do {                                                                           \
    hipError_t _r = (call);                                                      \
    if (_r != hipSuccess)                                                        \
      return ens_fail((e), HC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_r)); \
  } while (0)

int ens_mode(const hc_ctx* c) {
  if (c->prec == HC_FP64) return 0;
  if (c->zq) return 3;
  return c->prm.flat_dz < 0 ? 2 : 1;
}

void ens_drop_graph(hc_ens* e) {
  if (e->graph) (void)hipGraphExecDestroy(e->graph);
  e->graph = nullptr;
}

// Tile list (member-major) and per-member tile geometry for the current R.
int ens_layout(hc_ens* e) {
  std::vector<int2> tiles;
  for (size_t i = 0; i < e->m.size(); ++i) {
    int nbx, nby;
    hc_fused_geometry(e->m[i], e->R, hc::kStrideOverlap, &nbx, &nby);
    e->mem[i].nbx = nbx;
    e->mem[i].nby = nby;
    for (int t = 0; t < nbx * nby; ++t) tiles.push_back(make_int2(static_cast<int>(i), t));
  }
  e->ntiles = static_cast<int>(tiles.size());
  if (tiles.size() > e->tiles_cap) {
    if (e->d_tiles) (void)hipFree(e->d_tiles);
    e->d_tiles = nullptr;
    ENS_HIP(e, hipMalloc(reinterpret_cast<void**>(&e->d_tiles), tiles.size() * sizeof(int2)));
    e->tiles_cap = tiles.size();
  }
  ENS_HIP(e, hipMemcpy(e->d_tiles, tiles.data(), tiles.size() * sizeof(int2),
                       hipMemcpyHostToDevice));
  ens_drop_graph(e);
  return HC_OK;
}

// Refresh the member table (generations, record buffers).  A run (run =
// true) also marks each member as stepped by it, so that hc_loop_sync on the
// member afterwards returns the run's records and restores its generation.
int ens_prepare(hc_ens* e, int steps, bool run) {
  for (size_t i = 0; i < e->m.size(); ++i) {
    hc_ctx* c = e->m[i];
    ENS_HIP(e, hipStreamSynchronize(c->stream));
    if (run && steps > c->rec_cap) {
      hc_drop_loop_graphs(c);
      if (c->d_rec) (void)hipFree(c->d_rec);
      c->d_rec = nullptr;
      const int cap = std::max(steps, std::max(1024, 2 * c->rec_cap));
      ENS_HIP(e, hipMalloc(reinterpret_cast<void**>(&c->d_rec), sizeof(LoopRecord) * (cap + 1)));
      c->rec_cap = cap;
      ens_drop_graph(e);  // the table changes, not the graph's own arguments; kept simple
    }
    if (run) {
      c->loop_cur0 = c->cur;
      c->loop_enqueued = steps;
    }
    e->mem[i].rec = c->d_rec;
    e->mem[i].cur0 = c->cur;
  }
  ENS_HIP(e, hipMemcpyAsync(e->d_mem, e->mem.data(), e->mem.size() * sizeof(EnsMember),
                            hipMemcpyHostToDevice, e->stream));
  return HC_OK;
}

int ens_step(hc_ens* e, int parity) {
  const int n = static_cast<int>(e->m.size());
  hipLaunchKernelGGL(k_ens_control, dim3(n), dim3(64), 0, e->stream, e->d_mem, n, e->d_flags);
  ENS_HIP(e, hipGetLastError());
  if (hc_launch_ens_fused(e->stream, e->mode, e->R, e->d_mem, e->d_tiles, e->ntiles, e->nblocks,
                          parity))
    return ens_fail(e, HC_ERR_HIP, "ensemble: fused launch failed");
  if (e->mode == 0)
    hipLaunchKernelGGL(k_ens_scan<double>, dim3(e->scan_blocks), dim3(256), 0, e->stream,
                       e->d_mem, n, e->d_items, e->d_item_member, e->d_flags, e->R, parity);
  else
    hipLaunchKernelGGL(k_ens_scan<float>, dim3(e->scan_blocks), dim3(256), 0, e->stream,
                       e->d_mem, n, e->d_items, e->d_item_member, e->d_flags, e->R, parity);
  ENS_HIP(e, hipGetLastError());
  return HC_OK;
}

}  // namespace

extern "C" {

int hc_ens_create(hc_ens** out, hc_ctx* const* members, int n) {
  if (!out || !members || n < 1) return HC_ERR_ARG;
  *out = nullptr;
  hc_ens* e = new (std::nothrow) hc_ens;
  if (!e) return HC_ERR_ARG;
  for (int i = 0; i < n; ++i) {
    hc_ctx* c = members[i];
    if (!c || !c->loop_ready || c->g.sediment) {
      delete e;
      return HC_ERR_STATE;
    }
    if ((i > 0 && ens_mode(c) != ens_mode(members[0])) || !hc::fused_offsets_fit(c->NR, c->P)) {
      delete e;
      return HC_ERR_ARG;
    }
    for (int j = 0; j < i; ++j)
      if (members[j] == c) {
        delete e;
        return HC_ERR_ARG;
      }
    e->m.push_back(c);
  }
  e->mode = ens_mode(members[0]);
  e->mem.resize(n);
  std::vector<int> items(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    hc_ctx* c = members[i];
    EnsMember& M = e->mem[i];
    std::memset(&M, 0, sizeof(M));
    M.la = c->largs;
    M.ha = hc_hydro_args(c);
    M.ctl = c->d_ctl;
    M.rec = c->d_rec;
    M.top = c->d_top;
    M.pmask = c->d_pmask;
    M.rf_uniform = c->rf_uniform;
    M.rfarea = c->d_rfarea;
    M.area = c->d_area;
    for (int k = 0; k < 2; ++k) {
      M.h[k] = c->h[k];
      M.qx[k] = c->qx[k];
      M.qy[k] = c->qy[k];
    }
    M.z = hc_fused_z(c);
    M.z64 = c->z64;
    M.nman = c->nman;
    M.mask = c->mask;
    M.red = c->d_red;
    M.dpart = c->d_dpart;
    M.thr = c->prm.water_depth_erosion_threshold;
    M.W = c->W;
    items[i + 1] = items[i] + hcdev::scan_items(c->g.imax, c->W, static_cast<int>(c->esz));
  }
  std::vector<int> item_member(std::max(items[n], 1), 0);
  for (int i = 0; i < n; ++i)
    for (int t = items[i]; t < items[i + 1]; ++t) item_member[t] = i;
  int rc = HC_OK;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&e->d_mem), n * sizeof(EnsMember)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&e->d_items), (n + 1) * sizeof(int)) != hipSuccess ||
      hipMemcpy(e->d_items, items.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&e->d_item_member), item_member.size() * sizeof(int)) !=
          hipSuccess ||
      hipMemcpy(e->d_item_member, item_member.data(), item_member.size() * sizeof(int),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&e->d_flags), 2 * n * sizeof(int)) != hipSuccess ||
      hipMemset(e->d_flags, 0, 2 * n * sizeof(int)) != hipSuccess)
    rc = HC_ERR_HIP;
  if (rc == HC_OK) rc = ens_layout(e);
  if (rc != HC_OK) {
    hc_ens_destroy(e);
    return rc;
  }
  e->scan_blocks = std::max(1, std::min((items[n] + 3) / 4, 1024));
  *out = e;
  return HC_OK;
}

int hc_ens_destroy(hc_ens* e) {
  if (!e) return HC_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  ens_drop_graph(e);
  if (e->d_mem) (void)hipFree(e->d_mem);
  if (e->d_tiles) (void)hipFree(e->d_tiles);
  if (e->d_items) (void)hipFree(e->d_items);
  if (e->d_item_member) (void)hipFree(e->d_item_member);
  if (e->d_flags) (void)hipFree(e->d_flags);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return HC_OK;
}

const char* hc_ens_last_error(const hc_ens* e) { return e ? e->err.c_str() : "null ensemble"; }

int hc_ens_set_tiles(hc_ens* e, int rows, int blocks) {
  if (!e || (rows != 4 && rows != 8 && rows != 16) || blocks < 1 || blocks > 65536)
    return HC_ERR_ARG;
  ENS_HIP(e, hipStreamSynchronize(e->stream));
  e->nblocks = blocks;
  ens_drop_graph(e);
  if (rows != e->R) {
    e->R = rows;
    return ens_layout(e);
  }
  return HC_OK;
}

int hc_ens_run(hc_ens* e, int steps, int use_graph) {
  if (!e || steps < 0) return HC_ERR_ARG;
  int rc = ens_prepare(e, steps, true);
  if (rc) return rc;
  const int n = static_cast<int>(e->m.size());
  hipLaunchKernelGGL(k_ens_reset, dim3(1), dim3(256), 0, e->stream, e->d_mem, n);
  ENS_HIP(e, hipGetLastError());
  constexpr int kBlock = 10;  // even: every block starts on parity 0
  int s = 0;
  if (use_graph && steps >= kBlock) {
    if (!e->graph) {
      hipGraph_t g;
      ENS_HIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      int crc = HC_OK;
      for (int k = 0; k < kBlock && crc == HC_OK; ++k) crc = ens_step(e, k & 1);
      const hipError_t ce = hipStreamEndCapture(e->stream, &g);
      if (crc) return crc;
      if (ce != hipSuccess) return ens_fail(e, HC_ERR_HIP, "ensemble: graph capture failed");
      const hipError_t ie = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return ens_fail(e, HC_ERR_HIP, "ensemble: graph instantiate failed");
    }
    for (; s + kBlock <= steps; s += kBlock) ENS_HIP(e, hipGraphLaunch(e->graph, e->stream));
  }
  for (; s < steps; ++s) {
    rc = ens_step(e, s & 1);
    if (rc) return rc;
  }
  return HC_OK;
}

int hc_ens_sync(hc_ens* e) {
  if (!e) return HC_ERR_ARG;
  ENS_HIP(e, hipStreamSynchronize(e->stream));
  return HC_OK;
}

int hc_ens_tune(hc_ens* e, int* best_rows, int* best_blocks, double* best_us) {
  if (!e) return HC_ERR_ARG;
  int rc = ens_prepare(e, 0, false);
  if (rc) return rc;
  ENS_HIP(e, hipStreamSynchronize(e->stream));
  // the reduction slots carry each member's last maxdepth into its next step
  std::vector<unsigned long long> saved(e->m.size());
  for (size_t i = 0; i < e->m.size(); ++i)
    ENS_HIP(e, hipMemcpy(&saved[i], e->m[i]->d_red, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost));
  hipEvent_t e0, e1;
  ENS_HIP(e, hipEventCreate(&e0));
  ENS_HIP(e, hipEventCreate(&e1));
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  double best = 1e30;
  int bR = e->R, bB = e->nblocks;
  for (int R : {4, 8, 16}) {
    e->R = R;
    rc = ens_layout(e);
    if (rc) return rc;
    ENS_HIP(e, hipMemcpyAsync(e->d_mem, e->mem.data(), e->mem.size() * sizeof(EnsMember),
                              hipMemcpyHostToDevice, e->stream));
    for (int per_cu : {2, 4, 8, 0}) {  // 0: one block per tile (no persistent loop)
      const int nb = per_cu ? std::max(8, std::min(ncu * per_cu, e->ntiles)) : e->ntiles;
      // parity 0 reads each member's current generation and writes the other:
      // repeated launches leave the state unchanged
      if (hc_launch_ens_fused(e->stream, e->mode, R, e->d_mem, e->d_tiles, e->ntiles, nb, 0))
        return ens_fail(e, HC_ERR_HIP, "ensemble tune: launch failed");
      ENS_HIP(e, hipEventRecord(e0, e->stream));
      const int reps = 5;
      for (int k = 0; k < reps; ++k)
        (void)hc_launch_ens_fused(e->stream, e->mode, R, e->d_mem, e->d_tiles, e->ntiles, nb, 0);
      ENS_HIP(e, hipEventRecord(e1, e->stream));
      ENS_HIP(e, hipEventSynchronize(e1));
      float ms = 0;
      ENS_HIP(e, hipEventElapsedTime(&ms, e0, e1));
      const double us = 1000.0 * ms / reps;
      if (us < best) {
        best = us;
        bR = R;
        bB = nb;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (size_t i = 0; i < e->m.size(); ++i)
    ENS_HIP(e, hipMemcpy(e->m[i]->d_red, &saved[i], sizeof(unsigned long long),
                         hipMemcpyHostToDevice));
  e->R = bR;
  e->nblocks = bB;
  rc = ens_layout(e);
  if (rc) return rc;
  if (best_rows) *best_rows = bR;
  if (best_blocks) *best_blocks = bB;
  if (best_us) *best_us = best;
  return HC_OK;
}

int hc_ens_bench_fused(hc_ens* e, int reps, double* us) {
  if (!e || reps < 1 || !us) return HC_ERR_ARG;
  int rc = ens_prepare(e, 0, false);
  if (rc) return rc;
  std::vector<unsigned long long> saved(e->m.size());
  ENS_HIP(e, hipStreamSynchronize(e->stream));
  for (size_t i = 0; i < e->m.size(); ++i)
    ENS_HIP(e, hipMemcpy(&saved[i], e->m[i]->d_red, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost));
  hipEvent_t e0, e1;
  ENS_HIP(e, hipEventCreate(&e0));
  ENS_HIP(e, hipEventCreate(&e1));
  (void)hc_launch_ens_fused(e->stream, e->mode, e->R, e->d_mem, e->d_tiles, e->ntiles, e->nblocks, 0);
  ENS_HIP(e, hipEventRecord(e0, e->stream));
  for (int k = 0; k < reps; ++k)
    (void)hc_launch_ens_fused(e->stream, e->mode, e->R, e->d_mem, e->d_tiles, e->ntiles, e->nblocks, 0);
  ENS_HIP(e, hipEventRecord(e1, e->stream));
  ENS_HIP(e, hipEventSynchronize(e1));
  float ms = 0;
  ENS_HIP(e, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (size_t i = 0; i < e->m.size(); ++i)
    ENS_HIP(e, hipMemcpy(e->m[i]->d_red, &saved[i], sizeof(unsigned long long),
                         hipMemcpyDeviceToHost));
  *us = 1000.0 * ms / reps;
  return HC_OK;
}

}  // extern "C"
