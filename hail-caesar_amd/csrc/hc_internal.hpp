// Internal declarations shared by the libhc translation units.
//
// Device layout: every 2-D field is stored row-major with the reference's
// padding ring, (imax+2) rows by `pitch` columns, pitch = jmax+2 rounded up
// to a multiple of 64 so each 64-lane wavefront reads one aligned run of a
// row.  Columns jmax+2..pitch-1 are inert (elev = -9999, everything else 0).
// The active mask is one bit per cell, one 64-bit word per wavefront-wide
// run: word (x, w) holds columns 64w..64w+63 of row x.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hc.h"

#define HC_WAVE 64

namespace hc {

constexpr double kGravity = 9.81;
constexpr double kNoData = -9999.0;
// A depth update that leaves a cell non-finite or below this depth (m) marks
// the step unstable: the reported maximum depth becomes +inf.  The outflow
// limiter (no face drains more than a quarter of its cell) keeps a sound
// update at or above zero up to rounding, a few ulp of the depth, which stays
// far above -1e-4 m for any depth below 100 m even in FP32.
constexpr double kUnstableDepth = -1e-4;

// Hydro kernel arguments (all passed by value).
struct HydroArgs {
  int imax, jmax;
  int pitch;      // elements per row
  int mwords;     // mask words per row
  double DX;
  double hthr;    // hflow_threshold
  double edgeslope;
  double froude;
  double mannings;
  int spatial_n;
  double flat_dz; // FP32 form only
  float zq_inv;   // FP32 form, fixed-point elevations: metres per quantum (a power of two)
};

// Erosion arguments.
struct ErodeArgs {
  int imax, jmax, pitch, mwords;
  double DX, wdet, ERODEFACTOR, active, Csuspmax, max_vel, mannings;
  int spatial_n, wilcock, einstein;
  double d[10], logd[10], dprop[11], fall[11];
  double bedrock_thr, bedrock_rate, lateral_constant, chann_lateral, veg_lat;
};

// Device-resident loop state (one per context, in device memory).  Mirrors
// the reference's scalar members that the loop body of src/main.cpp:139-204
// reads and writes.
constexpr int kRfMax = 17;  // rain cells 1..16
struct LoopCtl {
  double time_step, cycle, maxdepth, temptot, waterinput, waterOut, iod;
  double previous, old_cycle, new_cycle, time_1, baseflow, dt_flow;
  double j_mean[kRfMax], old_j_mean[kRfMax], new_j_mean[kRfMax], amt[kRfMax];
  int input_counter[kRfMax], nActual[kRfMax];
  int counter, halt, err, npts, pts_dirty, first_new, last_first, iter;  // err: kLoopErr*
  int halt_pending;  // the end-of-run test passed: halt before the next body
};

static_assert(kRfMax == HC_RF_MAX, "rain cell capacity");
using LoopRecord = hc_loop_record;

// LoopCtl::err values: the loop halts and hc_loop_sync reports the cause.
constexpr int kLoopErrTable = 1;     // the TOPMODEL table ran out (HC_ERR_HALTED)
constexpr int kLoopErrDepth = 2;     // a non-finite or negative depth (HC_ERR_UNSTABLE)
constexpr int kLoopErrDtFloor = 3;   // the flow timestep fell below the floor (HC_ERR_UNSTABLE)

struct LoopArgs {
  double courant, DX, in_out_allowed, osi, ERODEFACTOR, MIN_Q, MIN_Q_MAXVAL;
  double dt_floor;  // stop (unstable) before a body whose flow timestep is below this
  int max_time_step, maxcycle, rfnum, imax, jmax, pitch, ntop, scan_interval;
  int mwords;  // words per row of the active and input-point bit masks
};

// Rain input folded into the fused kernel of the device-resident loop: the
// input points are a bit mask over the padded grid, and a point cell's depth
// gets amt[rfarea] as the kernel loads it (LCM.cpp:3463-3477 adds it in place
// before flow_route; the points are distinct cells, so the order is free).
struct InjectArgs {
  const uint64_t* pmask = nullptr;  // null: no injection in this launch
  const int* rfarea = nullptr;      // rain cell per cell (pitched)
  const double* amt = nullptr;      // per rain cell amount of this step
  int rf_uniform = -1;              // >= 0: every cell has this rain cell
};

// water_flux_out folded into the fused kernel of the device-resident loop:
// the tiles holding edge cells drain them and store one partial sum each;
// the next kernel on the stream adds the partials in tile order.
struct DrainArgs {
  const double* ts = nullptr;       // time_step on the device; null: no drain
  double thr = 0;                   // water_depth_erosion_threshold
  double* partials = nullptr;       // one slot per tile of the launch
};

// One member of an ensemble: the device state of an independent catchment
// context, read by the batched kernels (hc_ens_*).
struct EnsMember {
  LoopArgs la;
  HydroArgs ha;
  LoopCtl* ctl;
  LoopRecord* rec;
  const double* top;
  uint64_t* pmask;       // input points
  int rf_uniform;        // >= 0: the one rain cell of every cell
  const int* rfarea;
  const double* area;
  void* h[2];
  void* qx[2];
  void* qy[2];
  const void* z;        // the elevations the fused kernel reads (FP64, or FP32 copy)
  const double* z64;    // FP64 elevations (FP32 form)
  const void* nman;
  uint64_t* mask;
  unsigned long long* red;
  double* dpart;
  double thr;           // water_depth_erosion_threshold
  int nbx, nby;         // fused tiles at the ensemble's tile height
  int cur0;             // generation holding the state when the run starts
  int W;                // mask words per row
};

// Creation order of grain rows.  The reference numbers a cell's grain row
// (index[x][y]) by when it was made: rows of the grain data file by their
// line (LCM.cpp:611-683), every later row by the next grain_array_tot in
// addGS (LCM.cpp:3949-3955), which its single-threaded loops call in a fixed
// order.  Each cell here keeps a key that sorts its row into that order:
// epoch 0 = a row of the initial data (the low bits are its line number);
// epoch e > 0 = made by the e-th kernel launch that can make rows, ordered
// within the launch by the position in the reference's loop (major, minor
// loop indices) and a sub-step (neighbour direction, repeat).  Rows made in
// one launch that the reference makes in a different order get distinct
// keys in that order.  Grids up to 32767 cells a side.
constexpr uint64_t kRowNone = ~0ull;
constexpr int kRowSubBits = 5, kRowMinorBits = 15, kRowMajorBits = 15;
constexpr int kRowEpochShift = kRowSubBits + kRowMinorBits + kRowMajorBits;  // 35
__host__ __device__ constexpr uint64_t row_key(uint32_t epoch, uint32_t major, uint32_t minor,
                                               uint32_t sub) {
  return (static_cast<uint64_t>(epoch) << kRowEpochShift) |
         (static_cast<uint64_t>(major) << (kRowSubBits + kRowMinorBits)) |
         (static_cast<uint64_t>(minor) << kRowSubBits) | sub;
}
__host__ __device__ constexpr uint64_t row_epoch_base(uint32_t epoch) {
  return static_cast<uint64_t>(epoch) << kRowEpochShift;
}

// No-data value of the fixed-point elevations.
constexpr int32_t kZqNodata = INT32_MIN;

// The fused kernels address a field by 32-bit byte offsets (8-byte elements
// at most).
inline bool fused_offsets_fit(int NR, int P) {
  return static_cast<uint64_t>(NR) * static_cast<uint64_t>(P) * 8ull <= 0xffffffffull;
}

// Fused-kernel tile (bx, by) holds an edge cell: row 1 or imax, column 1 or
// jmax.  Tile rows by*R .. by*R+R-1; four wavefronts per tile column bx,
// wavefront w owning columns 64w .. 64w+63 (sw = 64, FP32 forms) or
// 62w+1 .. 62w+62 and column 0 for w = 0 (sw = 62, FP64 form).
// Columns per wavefront of the fused tiles: aligned (FP32 single grid) or
// overlapping (FP64 form and the ensemble).
constexpr int kStrideAligned = 64, kStrideOverlap = 62;
__host__ __device__ inline int fused_last_tile_col(int jmax, int sw) {
  return sw == 64 ? jmax / 256 : ((jmax - 1) / 62) / 4;
}
__host__ __device__ inline bool fused_edge_tile(int bx, int by, int R, int imax, int jmax, int sw) {
  return by == 0 || by == imax / R || bx == 0 || bx == fused_last_tile_col(jmax, sw);
}

// Compact slots for the edge tiles' partial sums: the tile rows 0 and imax/R
// first (nbx slots each), then two slots (one if they coincide) per tile row
// for tile columns 0 and the last one holding column jmax.  The latter slots
// of rows 0 and imax/R are holes that no tile writes and the sum skips.
struct EdgeSlots {
  int nbx, r1, c1, nr, nc;
  __host__ __device__ EdgeSlots(int nbx_, int R, int imax, int jmax, int sw)
      : nbx(nbx_), r1(imax / R), c1(fused_last_tile_col(jmax, sw)), nr(imax / R == 0 ? 1 : 2),
        nc(fused_last_tile_col(jmax, sw) == 0 ? 1 : 2) {}
  __host__ __device__ int count(int nby) const { return nr * nbx + nby * nc; }
  // slot of tile (bx, by), -1 when it holds no edge cell
  __host__ __device__ int slot(int bx, int by) const {
    if (by == 0) return bx;
    if (by == r1) return nbx + bx;
    const int ci = (bx == 0) ? 0 : (bx == c1 ? 1 : -1);
    return ci < 0 ? -1 : nr * nbx + by * nc + ci;
  }
  __host__ __device__ bool valid(int s) const {
    if (s < nr * nbx) return true;
    const int by = (s - nr * nbx) / nc;
    return by != 0 && by != r1;
  }
};

}  // namespace hc

struct hc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hc_grid g{};
  hc_params prm{};
  int NR = 0, NC = 0;   // padded rows/cols of the reference layout
  int P = 0;            // device pitch (elements)
  int W = 0;            // mask words per row
  size_t ncell = 0;     // NR * P
  int prec = HC_FP64;
  size_t esz = 8;       // bytes per state element

  // hydro state (element type double or float); h/qx/qy are ping-pong pairs
  void* h[2] = {nullptr, nullptr};
  void* qx[2] = {nullptr, nullptr};
  void* qy[2] = {nullptr, nullptr};
  int cur = 0;
  int fused_rows = 8;         // output rows per sub-tile of the fused kernel (tuned)
  int fused_strip = 1;        // sub-tiles per block: a block covers rows * strip rows
  void* z = nullptr;          // elevation (state type)
  double* z64 = nullptr;      // FP32 form: FP64 elevation for near-flat faces
  int32_t* zq = nullptr;      // FP32 form, elev_fixed: fixed-point elevations
  double zq_base = 0.0;       // elevation of quantum 0
  int zq_shift = 0;           // quanta per metre = 2^zq_shift
  void* nman = nullptr;       // Manning n (state type), spatial only
  uint64_t* mask = nullptr;

  // sediment state.  The planes below hold sesz-byte elements: double in an
  // FP64 context, float in an FP32 one (the FP32 storage form; every kernel
  // computes in FP64 and rounds once where it stores).  Elevations (z) and
  // bedrock stay FP64 in both: erosion moves a cell by as little as 1e-9 m
  // per call, below FP32's spacing of elevations (7.6e-6 m at 100 m).
  size_t sesz = 8;
  void* V = nullptr;          // Vsusptot
  void* qxs = nullptr;
  void* qys = nullptr;
  void* veldir = nullptr;     // 4 planes: directions 1,3,5,7
  void* grain = nullptr;      // 11 planes
  void* strata = nullptr;     // 90 planes (layer-major: z*9+n)
  int32_t* hasgrain = nullptr;
  int32_t* needgrain = nullptr;
  // Grain-row creation order, for the reference's row numbering (see
  // hc::row_key): a key per cell (kRowNone without a row), the last epoch
  // handed to a kernel that can create rows, and the rows the initial grain
  // data held (grain_array_tot after ingest_graindata_from_file).
  uint64_t* rowkey = nullptr;
  uint32_t row_epoch = 0;
  long long rows_base = 0;
  double* bedrock = nullptr;
  void* Tau = nullptr;
  void* Vel = nullptr;
  void* ss = nullptr;
  void* sdir = nullptr;       // 4 directions x 10 planes: su, sr, sd, sl
  void* etot = nullptr;       // erodetot
  void* etot3 = nullptr;      // erodetot3
  double* d_dep = nullptr;    // deposition pre-pass terms (11 FP64 planes; scratch)
  float* gtot = nullptr;      // 20 floats

  // reductions and small device scalars
  unsigned long long* d_red = nullptr;  // [0] maxdepth bits, [1] tempbmax bits, [2] health count,
                                        // [3] global_landsliding sweep count,
                                        // [4] landsliding: some cell can move,
                                        // [5] [6] creep's grid barrier: arrivals, time-out
  double* d_slide_moves = nullptr;      // hillslope scratch: 8 planes (landslide moves, tempcreep)
  double* d_runoff = nullptr;           // spatially complex rainfall: old, new j_mean planes
  double* d_partials = nullptr;         // per-block partial sums
  unsigned int* d_count = nullptr;      // last-block counters
  double* d_dpart = nullptr;            // fused-kernel edge drain: one slot per tile
  size_t dpart_cap = 0;
  double* d_wmax = nullptr;             // device loop: fused kernel's per-wavefront maxdepth
  uint64_t* d_wet = nullptr;            // device loop (FP32 forms): wet-cell bits, NR x W words
  size_t wmax_cap = 0;
  double* d_out = nullptr;              // [0] temptot
  double* h_pinned = nullptr;           // pinned host mirror (8 doubles)
  // device-resident loop
  hc::LoopCtl* d_ctl = nullptr;
  hc::LoopRecord* d_rec = nullptr;
  int rec_cap = 0;
  double* d_top = nullptr;              // TOPMODEL new_j_mean table [ntop][kRfMax]
  double* d_area = nullptr;             // drainage area (pitched)
  int* d_rfarea = nullptr;              // rain cell index (pitched)
  uint64_t* d_pmask = nullptr;          // loop input points: one bit per padded cell
  int rf_uniform = -1;                  // >= 0: every cell has this rain cell index
  hc::LoopArgs largs{};
  bool loop_ready = false;
  int loop_cur0 = 0;                    // `cur` when the last run was enqueued
  int loop_enqueued = 0;                // iterations enqueued by the last run
  // the loop's counter as the next enqueued body starts (the host's forecast:
  // each body adds one unless the loop halts; hc_loop_sync resets it), which
  // tells which bodies rebuild the active mask
  long loop_counter = 0;
  hipGraphExec_t loop_graph[2][5] = {};  // per starting parity and counter % 5
  bool prof_on = false;
  std::vector<hipEvent_t> prof_ev;      // start/stop pairs around fused launches
  int prof_n = 0;
  unsigned char* scratch = nullptr;     // this device's static staging array (hc_capi.hip)
  int* d_pts = nullptr;                 // point input cells (x*P+y)
  double* d_amt = nullptr;              // point input amounts
  int pts_cap = 0;
  std::vector<uint64_t> pt_bits;        // host: repeated-point test, one bit per padded cell
  int* d_plist = nullptr;               // persistent input points: cells (x*P+y), then rain cells
  double* d_pamt = nullptr;             // their per-rain-cell amounts (HC_RF_MAX)
  int plist_n = -1, plist_cap = 0;      // -1: no list set

  std::string err;
};
// Captured loop graphs hold the device addresses of every buffer they touch;
// any reallocation of such a buffer, or a change of launch shapes, must drop
// them so the next run captures again.
inline void hc_drop_loop_graphs(hc_ctx* c) {
  for (auto& row : c->loop_graph)
    for (auto& gexec : row) {
      if (gexec) (void)hipGraphExecDestroy(gexec);
      gexec = nullptr;
    }
}


// helpers implemented in hc_capi.hip
int hc_fail(hc_ctx* c, int code, const std::string& msg);
#define HC_HIP(ctx, call)                                                     \
  do {                                                                        \
    hipError_t _e = (call);                                                   \
    if (_e != hipSuccess)                                                     \
      return hc_fail((ctx), HC_ERR_HIP,                                       \
                     std::string(#call) + ": " + hipGetErrorString(_e));      \
  } while (0)

// launchers (hc_hydro.hip)
int hc_launch_flow_route(hc_ctx* c, double dt);
int hc_launch_depth_update(hc_ctx* c, double dt);
int hc_launch_flow_depth(hc_ctx* c, double dt, const double* dtp = nullptr,
                         const int* haltp = nullptr, bool swap = true,
                         const hc::DrainArgs* drain = nullptr,
                         const hc::InjectArgs* inject = nullptr, double* wmax = nullptr,
                         hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr,
                         uint64_t* wet = nullptr);
// Geometry of the fused kernel's launch for the context's tile rows.
// Columns per wavefront of this context's single-grid fused kernel.
int hc_fused_stride(const hc_ctx* c);
// Tile grid of the fused kernel at R rows per block and sw columns per wavefront.
void hc_fused_geometry(const hc_ctx* c, int R, int sw, int* nbx, int* nby);
hc::HydroArgs hc_hydro_args(const hc_ctx* c);
// Elevation plane the fused kernel reads for this context's mode.
const void* hc_fused_z(const hc_ctx* c);
// Ensemble fused step over a tile list (mode: 0 FP64, 1 FP32 z32, 2 FP32 z64,
// 3 FP32 fixed-point z).
int hc_launch_ens_fused(hipStream_t s, int mode, int R, const hc::EnsMember* mem,
                        const int2* tiles, int ntiles, int nblocks, int parity);
int hc_launch_scan(hc_ctx* c);
// FP64 fused flow_route + depth_update between explicit state buffers (fixed
// dt, no drain or injection), and scan_area of an explicit depth buffer.
int hc_launch_fused64_ptrs(hc_ctx* c, double dt, const double* h0, const double* qx0,
                           const double* qy0, double* h1, double* qx1, double* qy1,
                           const uint64_t* mask);
int hc_launch_scan_ptr(hc_ctx* c, const void* h, uint64_t* mask);
int hc_launch_flux_out(hc_ctx* c, double ts);
int hc_launch_points(hc_ctx* c, int n);
// Point inputs applied in list order where cells repeat; cells from idx
// (x*P+y), or from xs/ys when idx is null.
int hc_launch_points_ordered(hc_ctx* c, int n, const int* idx, const int* xs, const int* ys,
                             const double* amt);
// h[idx[k]] += amt[rc[k]] for the persistent point list (distinct cells).
int hc_launch_points_rc(hc_ctx* c, int n, const int* idx, const int* rc, const double* amt);
int hc_launch_health(hc_ctx* c);
int hc_launch_fill(hc_ctx* c, void* p, size_t n, double v, bool is_float);
// erosion (hc_erode.hip)
int hc_erode_impl(hc_ctx* c, double mult, double* time_step, double* tempbmax, float* gtot);
int hc_global_landsliding_impl(hc_ctx* c, double failureangle, int* sweeps);
int hc_creep_impl(hc_ctx* c, double time, double creep_rate);
int hc_in_channel_landsliding_impl(hc_ctx* c, double failureangle);
int hc_reach_inputs_impl(hc_ctx* c, int n, const int* x, const int* y, const double* sediment,
                         const double* water, double threshold);
