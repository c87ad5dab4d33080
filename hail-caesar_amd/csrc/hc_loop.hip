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
//                              tile order), scan_area when counter % 5 == 0,
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
using hc::LoopArgs;
using hc::LoopCtl;
using hc::LoopRecord;
using namespace hcdev;

namespace {

__device__ __forceinline__ double cfl(const LoopArgs& a, double maxdepth) {
  return a.courant * (a.DX / sqrt(kGravity * (maxdepth)));  // LCM.cpp:1884-1887
}

template <typename T>
__global__ __launch_bounds__(256) void k_loop_control(LoopArgs a, LoopCtl* __restrict__ ctl,
                                                      const double* __restrict__ top,
                                                      int* __restrict__ pts,
                                                      int* __restrict__ prf,
                                                      const int* __restrict__ rfarea,
                                                      T* __restrict__ h,
                                                      unsigned long long* __restrict__ red,
                                                      LoopRecord* __restrict__ rec) {
  __shared__ int s_n, s_stop;
  __shared__ double s_amt[kRfMax];
  if (threadIdx.x == 0) {
    LoopCtl& C = *ctl;
    s_stop = (C.halt || C.err) ? 1 : 0;
    if (!s_stop && C.halt_pending) {  // the previous body ended the run
      C.halt = 1;
      s_stop = 1;
    }
    if (!s_stop) {
      // a list rebuilt by the previous iteration: an empty one keeps a single
      // stale point, in_x[1]/in_y[1] of the last non-empty list (LCM.cpp:3764)
      if (C.pts_dirty) {
        if (C.npts == 0) {
          pts[0] = C.last_first;
          prf[0] = rfarea[C.last_first];
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
        s_amt[rf] = amt;
      }
      s_n = C.npts;
      if (C.cycle >= C.time_1) {
        do {
          C.time_1++;
          const int t = static_cast<int>(C.time_1);
          if (t < 0 || t >= a.ntop) {
            C.err = 1;
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
      if (C.err) {
        s_stop = 1;
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
  __syncthreads();
  if (s_stop) return;
  // water_depth[i][j] += water_add_amt at every input point (LCM.cpp:3463-3477),
  // with the list as it was before any rebuild of this iteration
  for (int k = threadIdx.x; k < s_n; k += blockDim.x) {
    T* p = &h[pts[k]];
    *p = static_cast<T>(static_cast<double>(*p) + s_amt[prf[k]]);
  }
}

// Partial sums of the edge tiles in slot order, then a fixed tree: the
// water_flux_out total for the iteration.
__device__ void finish_drain(LoopCtl* __restrict__ ctl, LoopRecord* __restrict__ rec,
                             const double* __restrict__ partials, int nby, int nbx, int R,
                             int imax, int jmax) {
  __shared__ double s_sum[256];
  const int t = threadIdx.x;
  const hc::EdgeSlots es(nbx, R, imax, jmax);
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
                                                   int* __restrict__ pts, int* __restrict__ prf,
                                                   const double* __restrict__ partials,
                                                   int nby, int nbx, int R,
                                                   LoopRecord* __restrict__ rec) {
  // the control kernel alone sets halt, so the fused kernel of this
  // iteration ran exactly when it is still clear here
  const bool live = !ctl->halt && !ctl->err;
  const bool scan = live && (ctl->counter % a.scan_interval) == 0;
  const bool rebuild = live && ctl->pts_dirty;
  if (live && blockIdx.x == 0) finish_drain(ctl, rec, partials, nby, nbx, R, a.imax, a.jmax);
  if (!scan && !rebuild) return;  // uniform over the grid
  if (scan) {
    const int items = scan_items(a.imax, W);
    for (int it = blockIdx.x * 4 + (threadIdx.x >> 6); it < items; it += gridDim.x * 4)
      scan_item<T>(it, a.imax, a.jmax, a.pitch, W, h, mask);
  }
  if (rebuild) {
    // the slot order of the rebuilt list is arbitrary: every cell appears at
    // most once and injection adds to distinct cells
    const double b = ctl->baseflow;
    const long ncell = static_cast<long>(a.imax) * a.jmax;
    for (long k = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; k < ncell;
         k += static_cast<long>(gridDim.x) * blockDim.x) {
      const int x = 1 + static_cast<int>(k / a.jmax);
      const int y = 1 + static_cast<int>(k % a.jmax);
      const int idx = x * a.pitch + y;
      const double v = area[idx] * b * 3 * a.DX * a.DX;
      if (v > a.MIN_Q && v < a.MIN_Q_MAXVAL) {
        const int slot = atomicAdd(&ctl->npts, 1);
        const int rf = rfarea[idx];
        pts[slot] = idx;
        prf[slot] = rf;
        atomicAdd(&ctl->input_counter[rf], 1);
        atomicMin(&ctl->first_new, idx);
      }
    }
  }
}

__global__ void k_loop_reset_iter(LoopCtl* ctl) { ctl->iter = 0; }

// One iteration's launches on the context stream.
int enqueue_iteration(hc_ctx* c, int k_in_run = -1) {
  const LoopArgs& a = c->largs;
  const int k = c->cur;
  if (c->prec == HC_FP64)
    hipLaunchKernelGGL(k_loop_control<double>, dim3(1), dim3(256), 0, c->stream, a, c->d_ctl,
                       c->d_top, c->d_lpts, c->d_lrf, c->d_rfarea, (double*)c->h[k], c->d_red,
                       c->d_rec);
  else
    hipLaunchKernelGGL(k_loop_control<float>, dim3(1), dim3(256), 0, c->stream, a, c->d_ctl,
                       c->d_top, c->d_lpts, c->d_lrf, c->d_rfarea, (float*)c->h[k], c->d_red,
                       c->d_rec);
  HC_HIP(c, hipGetLastError());
  const bool prof = c->prof_on && k_in_run >= 0 && 2 * k_in_run + 1 < static_cast<int>(c->prof_ev.size());
  if (prof) HC_HIP(c, hipEventRecord(c->prof_ev[2 * k_in_run], c->stream));
  hc::DrainArgs dr;
  dr.ts = &c->d_ctl->time_step;
  dr.thr = c->prm.water_depth_erosion_threshold;
  dr.partials = c->d_dpart;
  int rc = hc_launch_flow_depth(c, 0.0, &c->d_ctl->dt_flow, &c->d_ctl->halt, true, &dr);
  if (rc) return rc;
  if (prof) {
    HC_HIP(c, hipEventRecord(c->prof_ev[2 * k_in_run + 1], c->stream));
    c->prof_n = k_in_run + 1;
  }
  const int o = c->cur;  // flow_depth flipped the generation
  const int nb = std::max(1, std::min((hcdev::scan_items(c->g.imax, c->W) + 3) / 4, 1024));
  const int R = c->fused_rows;
  int nbx, nby;
  hc_fused_geometry(c, R, &nbx, &nby);
  if (c->prec == HC_FP64)
    hipLaunchKernelGGL(k_loop_scan<double>, dim3(nb), dim3(256), 0, c->stream, a, c->d_ctl, c->W,
                       (const double*)c->h[o], c->mask, c->d_area, c->d_rfarea, c->d_lpts,
                       c->d_lrf, c->d_dpart, nby, nbx, R, c->d_rec);
  else
    hipLaunchKernelGGL(k_loop_scan<float>, dim3(nb), dim3(256), 0, c->stream, a, c->d_ctl, c->W,
                       (const float*)c->h[o], c->mask, c->d_area, c->d_rfarea, c->d_lpts,
                       c->d_lrf, c->d_dpart, nby, nbx, R, c->d_rec);
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
  HC_HIP(c, hipStreamSynchronize(c->stream));
  LoopArgs& a = c->largs;
  a.courant = in->courant_number;
  a.DX = c->g.DX;
  a.in_out_allowed = in->in_out_difference_allowed;
  a.osi = in->output_file_save_interval;
  a.ERODEFACTOR = in->ERODEFACTOR;
  a.MIN_Q = in->MIN_Q;
  a.MIN_Q_MAXVAL = in->MIN_Q_MAXVAL;
  a.max_time_step = in->max_time_step;
  a.maxcycle = in->maxcycle;
  a.rfnum = in->rfnum;
  a.imax = c->g.imax;
  a.jmax = c->g.jmax;
  a.pitch = c->P;
  a.ntop = in->ntop;
  a.scan_interval = 5;
  // tables and inputs
  if (!c->d_ctl) HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_ctl), sizeof(LoopCtl)));
  {
    // edge-drain partials: the largest slot count over the tile heights
    size_t need = 1;
    for (int R : {4, 8, 16}) {
      int nbx, nby;
      hc_fused_geometry(c, R, &nbx, &nby);
      need = std::max(need, static_cast<size_t>(
                                hc::EdgeSlots(nbx, R, c->g.imax, c->g.jmax).count(nby)));
    }
    if (need > c->dpart_cap) {
      if (c->d_dpart) (void)hipFree(c->d_dpart);
      c->d_dpart = nullptr;
      HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_dpart), need * sizeof(double)));
      c->dpart_cap = need;
    }
  }
  if (c->d_top) (void)hipFree(c->d_top);
  HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_top), sizeof(double) * kRfMax * std::max(1, in->ntop)));
  HC_HIP(c, hipMemcpy(c->d_top, in->topmodel, sizeof(double) * kRfMax * in->ntop, hipMemcpyHostToDevice));
  int rc = upload_pitched(c, &c->d_area, in->area);
  if (rc) return rc;
  rc = upload_pitched(c, &c->d_rfarea, in->rfarea);
  if (rc) return rc;
  const size_t cap = static_cast<size_t>(c->g.imax) * c->g.jmax + 2;
  if (!c->d_lpts) {
    HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_lpts), cap * sizeof(int)));
    HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_lrf), cap * sizeof(int)));
  }
  std::vector<int> pts(std::max(1, in->npoints)), prf(std::max(1, in->npoints));
  for (int k = 0; k < in->npoints; ++k) {
    const int x = in->in_x[k], y = in->in_y[k];
    if (x < 0 || x >= c->NR || y < 0 || y >= c->NC)
      return hc_fail(c, HC_ERR_ARG, "loop: input point outside the padded grid");
    pts[k] = x * c->P + y;
    prf[k] = in->rfarea[static_cast<size_t>(x) * c->NC + y];
    if (prf[k] < 0 || prf[k] >= kRfMax) return hc_fail(c, HC_ERR_ARG, "loop: rain cell index out of range");
  }
  if (in->npoints > 0) {
    HC_HIP(c, hipMemcpy(c->d_lpts, pts.data(), in->npoints * sizeof(int), hipMemcpyHostToDevice));
    HC_HIP(c, hipMemcpy(c->d_lrf, prf.data(), in->npoints * sizeof(int), hipMemcpyHostToDevice));
  }
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
    const int parity = c->cur;
    if (!c->loop_graph[parity]) {
      hipGraph_t graph;
      HC_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rc = HC_OK;
      for (int k = 0; k < kBlock && rc == HC_OK; ++k) rc = enqueue_iteration(c);
      hipError_t e = hipStreamEndCapture(c->stream, &graph);
      c->cur = parity;  // capture does not execute; restore the generation
      if (rc) return rc;
      if (e != hipSuccess) return hc_fail(c, HC_ERR_HIP, "loop: graph capture failed");
      e = hipGraphInstantiate(&c->loop_graph[parity], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return hc_fail(c, HC_ERR_HIP, "loop: graph instantiate failed");
    }
    while (left >= kBlock) {
      HC_HIP(c, hipGraphLaunch(c->loop_graph[parity], c->stream));
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
  if (C.err) return hc_fail(c, HC_ERR_HALTED, "loop: TOPMODEL table exhausted");
  return HC_OK;
}

}  // extern "C"
