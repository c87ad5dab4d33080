/* hc.h - C-ABI of the MI355X-native HAIL-CAESAR hot path (libhc.so).
 *
 * The reference has no plugin or FFI layer: the hot path is the private
 * member functions of LSDCatchmentModel acting on TNT arrays
 * (include/catchmentmodel/LSDCatchmentModel.hpp:175-396), called in a fixed
 * order by main() (src/main.cpp:139-204).  Each entry point below replaces one
 * of those methods; the comment on each names the reference method (file:line,
 * paths relative to the reference root, LCM.cpp =
 * src/catchmentmodel/LSDCatchmentModel.cpp).
 *
 * Conventions
 *   - Every function returns 0 (HC_OK) or a negative hc_status; the message
 *     is available from hc_last_error(ctx).  Nothing throws across the ABI.
 *   - Host arrays use the reference's padded layout: (imax+2) x (jmax+2)
 *     row-major, x = row 1..imax (first index), y = column 1..jmax
 *     (contiguous), a one-cell padding ring (LCM.cpp:1622-1705).  Multi-plane
 *     fields are cell-major with planes last, exactly like the reference's
 *     TNT::Array3D (vel_dir: 9 planes, grain: 11, strata: 10x9, sr/sl/su/sd: 10).
 *   - Host pointers are borrowed for the duration of the call only.
 *   - A context owns its device buffers and one HIP stream and is used by one
 *     host thread at a time.  Contexts are independent of each other.
 *   - Entry points that return a host scalar (maxdepth, temptot, tempbmax)
 *     synchronise the context's stream; the others only enqueue work.
 */
#ifndef HC_H
#define HC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hc_ctx hc_ctx;

typedef enum hc_status {
  HC_OK = 0,
  HC_ERR_ARG = -1,      /* bad argument (size, field id, null pointer) */
  HC_ERR_HIP = -2,      /* a HIP runtime call failed */
  HC_ERR_STATE = -3,    /* call not valid in this context's configuration */
  HC_ERR_UNSTABLE = -4, /* health check tripped (NaN/Inf/negative depth) */
  HC_ERR_HALTED = -5    /* device-resident stepping stopped early */
} hc_status;

typedef enum hc_precision {
  HC_FP64 = 0, /* plain form: FP64 state, reference operation order */
  HC_FP32 = 1  /* FP32 state, FP64 mass-balance accumulators */
} hc_precision;

typedef enum hc_transport_law { HC_WILCOCK = 1, HC_EINSTEIN = 2 } hc_transport_law;

/* Grid shape and storage options (LCM.cpp:92-138 reads imax/jmax/DX from the
 * DEM header). */
typedef struct hc_grid {
  int imax;       /* rows (nrows) */
  int jmax;       /* columns (ncols) */
  double DX;      /* cell size, metres */
  int precision;  /* hc_precision */
  int sediment;   /* 1: allocate erosion state (hydro_model_only: no) */
} hc_grid;

/* Scalar model parameters.  Names and meanings follow the reference's
 * members (LCM.hpp:437-568) as set by initialise_variables (LCM.cpp:842-1610). */
typedef struct hc_params {
  double courant_number;        /* courant_number */
  double froude_limit;          /* froude_limit */
  double hflow_threshold;       /* hflow_threshold */
  double edgeslope;             /* edgeslope (slope_on_edge_cell) */
  double mannings;              /* uniform Manning n (mannings_n) */
  int spatially_var_mannings;   /* 1: per-cell n uploaded as HC_FIELD_MANNINGS */
  double water_depth_erosion_threshold; /* 0.01 in the reference (LCM.hpp:448) */
  double ERODEFACTOR;           /* erosion_limit */
  double active;                /* active_layer_thickness */
  double Csuspmax;              /* 0.05 (LCM.hpp:548) */
  double max_vel;               /* max_tau_velocity */
  int transport_law;            /* hc_transport_law */
  double d[10];                 /* grain sizes d1..d9 in d[1..9] (LCM.hpp:557-565) */
  double dprop[11];             /* default proportions (LCM.hpp:568) */
  double fallVelocity[11];      /* LCM.cpp:1848-1860 */
  double bedrock_erosion_threshold;
  double bedrock_erosion_rate;
  double lateral_constant;
  double chann_lateral_erosion;
  double veg_lat_restriction;
  /* FP32 form only.  0: single-precision elevations; > 0: faces whose
   * single-precision elevation difference is smaller than this (metres) take
   * the difference from the FP64 DEM; < 0: every face takes its elevation
   * difference from the FP64 DEM (4 more bytes per cell and step). */
  double flat_dz;
  /* FP32 form only.  1: elevations held as 32-bit fixed point, a power-of-two
   * scale about the DEM's mid-range chosen on upload so that the relief fits
   * in 2^31 quanta; every elevation difference is then exact to one quantum
   * (relief / 2^31, about 2e-8 m for 40 m of relief) at 4 bytes per cell.
   * flat_dz is ignored.  Cells at -9999 or below stay no-data. */
  int elev_fixed;
} hc_params;

/* Field identifiers for hc_upload / hc_download. */
typedef enum hc_field {
  HC_FIELD_WATER_DEPTH = 0, /* water_depth           double */
  HC_FIELD_ELEV = 1,        /* elev                  double */
  HC_FIELD_QX = 2,          /* qx                    double */
  HC_FIELD_QY = 3,          /* qy                    double */
  HC_FIELD_MANNINGS = 4,    /* spat_var_mannings     double */
  HC_FIELD_VSUSP = 5,       /* Vsusptot              double */
  HC_FIELD_QXS = 6,         /* qxs                   double */
  HC_FIELD_QYS = 7,         /* qys                   double */
  HC_FIELD_VEL_DIR = 8,     /* vel_dir (x,y,9)       double */
  HC_FIELD_MASK = 9,        /* down_scan membership  int32 0/1 */
  HC_FIELD_GRAIN = 10,      /* grain (x,y,11)        double, dense per cell */
  HC_FIELD_STRATA = 11,     /* strata (x,y,10,9)     double, dense per cell */
  HC_FIELD_HASGRAIN = 12,   /* index[x][y] != -9999  int32 0/1 */
  HC_FIELD_BEDROCK = 13,    /* bedrock               double */
  HC_FIELD_TAU = 14,        /* Tau                   double */
  HC_FIELD_VEL = 15,        /* Vel                   double */
  HC_FIELD_SS = 16,         /* ss                    double */
  HC_FIELD_SR = 17,         /* sr (x,y,10)           double */
  HC_FIELD_SL = 18,         /* sl (x,y,10)           double */
  HC_FIELD_SU = 19,         /* su (x,y,10)           double */
  HC_FIELD_SD = 20          /* sd (x,y,10)           double */
} hc_field;

/* ---- lifetime -------------------------------------------------------- */

/* Replaces LSDCatchmentModel construction + initialise_arrays/zero_values
 * (LCM.cpp:1617-1843, 2891-2982): allocates zeroed state with elev = -9999
 * on device 0 (or the device current on the calling thread). */
int hc_create(hc_ctx** ctx, const hc_grid* grid, const hc_params* params);
int hc_destroy(hc_ctx* ctx);
const char* hc_last_error(const hc_ctx* ctx);
/* Element count of one host array of `field` for this grid. */
size_t hc_field_count(const hc_ctx* ctx, int field);

/* Copy a whole field between a host array in the reference layout and the
 * device (load_data's array fills, LCM.cpp:140-606).  `count` must equal
 * hc_field_count(ctx, field).  FP32 contexts convert on the way. */
int hc_upload(hc_ctx* ctx, int field, const void* host, size_t count);
int hc_download(hc_ctx* ctx, int field, void* host, size_t count);

/* ---- the per-timestep hot path ---------------------------------------- */

/* LSDCatchmentModel::flow_route (LCM.cpp:3063-3247): local-inertial qx/qy
 * update on the cells of the active mask; with sediment state also
 * qxs/qys and vel_dir.  `dt_flow` is get_flow_timestep() (LCM.cpp:1913). */
int hc_flow_route(hc_ctx* ctx, double dt_flow);

/* LSDCatchmentModel::depth_update (LCM.cpp:3259-3299): h (and Vsusptot) from
 * the flux divergence; returns maxdepth (max h over updated cells, h > 0). */
int hc_depth_update(hc_ctx* ctx, double dt_flow, double* maxdepth);

/* flow_route followed by depth_update in one fused launch (hydro-only
 * contexts).  Results are identical to the two calls above. */
int hc_flow_depth(hc_ctx* ctx, double dt_flow, double* maxdepth);

/* LSDCatchmentModel::scan_area (LCM.cpp:3821-3849): rebuild the active mask
 * (3x3 wet dilation).  check_wetted_area (LCM.cpp:2019) calls it every 5th
 * iteration. */
int hc_scan_area(hc_ctx* ctx);

/* LSDCatchmentModel::water_flux_out (LCM.cpp:3005-3047): drain the four
 * edges to water_depth_erosion_threshold; returns temptot (m^3/s), summed
 * in FP64. */
int hc_water_flux_out(hc_ctx* ctx, double time_step, double* temptot);

/* Input-point injection of catchment_water_input_and_hydrology
 * (LCM.cpp:3463-3477): water_depth[x[k]][y[k]] += amt[k] for k = 0..n-1.  A
 * cell listed more than once takes each of its amounts in turn, in list
 * order, as the reference's loop does.  Points must lie in the padded grid
 * (0 <= x <= imax+1, 0 <= y <= jmax+1). */
int hc_add_point_inputs(hc_ctx* ctx, const int* x, const int* y,
                        const double* amt, int n);

/* The same injection with the point list kept on the device, for loops
 * that inject every step: the reference's amount at a point depends only on
 * the point's rain cell, (j_mean[rf] * nActualGridCells[rf]) /
 * catchment_input_counter[rf] * flow_timestep capped at ERODEFACTOR
 * (LCM.cpp:3463-3477).  hc_points_set uploads the list (cells x[k], y[k]
 * in the padded grid, which must be distinct, and each point's rain cell
 * rain_cell[k] in 0..HC_RF_MAX-1) when it changes (get_catchment_input_points,
 * LCM.cpp:3741-3767); hc_points_add then adds amt[rain_cell[k]] to every
 * point's depth.  hc_points_add before any hc_points_set returns
 * HC_ERR_STATE; n = 0 sets an empty list. */
int hc_points_set(hc_ctx* ctx, const int* x, const int* y, const int* rain_cell, int n);
int hc_points_add(hc_ctx* ctx, const double* amt, int nrf);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): hc_add_point_inputs without any allocation.  Where
 * hc_add_point_inputs stages the points in context buffers that it grows on
 * demand, this stages them in a static device array of HC_NOALLOC_MAX_POINTS
 * entries (256 KiB, one per device, held until the process ends); more points
 * return HC_ERR_ARG.  Results are those of hc_add_point_inputs, repeated
 * cells included.  Every context on one device shares that device's array
 * and the copy into it is asynchronous: calls on one context are ordered by
 * its stream, but calls on two contexts of one device race unless the caller
 * orders them (hc_synchronize on the first context before calling on the
 * second). */
#define HC_NOALLOC_MAX_POINTS 16384
int hc_add_point_inputs_noalloc(hc_ctx* ctx, const int* x, const int* y,
                                const double* amt, int n);

/* LSDCatchmentModel::erode (LCM.cpp:4211-4845) with its grain helpers
 * (LCM.cpp:3874-4171).  *time_step is read and updated exactly as the
 * reference updates its member; returns tempbmax and the per-fraction edge
 * outputs gtot2[0..9] (accumulated in float, as the reference's float
 * gtot2[20]). */
int hc_erode(hc_ctx* ctx, double mult_factor, double* time_step,
             double* tempbmax, float gtot[10]);

/* LSDCatchmentModel::global_landsliding (LCM.cpp:5529-5654), which the
 * reference's loop calls through call_global_landsliding (LCM.cpp:1986-1996)
 * once per simulated day in erosion runs: repeated sweeps moving material
 * down every slope steeper than `failureangle` degrees (the parameter-file
 * key slope_failure_thresh), each move capped at ERODEFACTOR and kept above
 * bedrock + active, until a sweep's total is not positive or 200 sweeps.
 * Elevations are updated in place, bit-identical to the reference's sweep
 * order.  FP64 sediment contexts; *sweeps (may be NULL) receives the number
 * of sweeps.  Synchronous. */
int hc_global_landsliding(hc_ctx* ctx, double failureangle, int* sweeps);

/* LSDCatchmentModel::creep(time) (LCM.cpp:5293-5427), which the reference's
 * loop calls through slope_creep (LCM.cpp:1976-1984) with time = 0.028 every
 * 14400 simulated minutes in erosion runs: each cell above its bedrock moves
 * material to its lower neighbours in proportion to the slope and
 * `creep_rate` (the parameter-file key creep_rate), sliding the moved grain
 * fractions in the reference's sweep order.  Elevations, grain rows and
 * strata bit-identical to the reference.  Sediment contexts (grain stored in
 * FP64 or FP32).  Synchronous.  The grain sweep runs on a grid of blocks
 * that must all be resident on the device at once (it is sized to what the
 * device holds); should its grid-wide barrier wait for seconds, as when
 * other work holds the device's compute units, it returns HC_ERR_HIP before
 * the elevations move, with the grain rows part-way through the sweep
 * (restore a saved state before going on). */
int hc_creep(hc_ctx* ctx, double time, double creep_rate);

/* LSDCatchmentModel::in_channel_landsliding (LCM.cpp:5429-5528), which the
 * reference's loop calls every 10 iterations in erosion runs
 * (call_channel_landsliding, LCM.cpp:1968-1974): over the active mask as the
 * scan left it (the reference's down_scan), column by column, each listed
 * cell moves material down every slope steeper than `failureangle` degrees,
 * kept above bedrock + active, sliding the moved grain fractions.
 * Elevations, grain rows and strata bit-identical to the reference.  FP64
 * sediment contexts. */
int hc_in_channel_landsliding(hc_ctx* ctx, double failureangle);

/* The reference's grain-row numbers, index[x][y] (LCM.hpp; -9999 = no row),
 * over the padded grid ((imax+2) x (jmax+2) int32).  The reference numbers
 * the rows of its grain data file by line (ingest_graindata_from_file,
 * LCM.cpp:611-683) and every later row by the next grain_array_tot when
 * addGS makes it (LCM.cpp:3949-3955); its Grainz output (LGM.cpp:59-118)
 * prints that number on every line.  Grain rows here are per cell; each
 * keeps the position in the reference's single-threaded loops where it was
 * made, so the numbers come out as the reference's.
 *   set: the initial rows (also sets the has-grain flags) and rows_total,
 *        the reference's grain_array_tot after ingest (the line count, which
 *        duplicate lines make larger than the number of rows).  Uploading
 *        HC_FIELD_HASGRAIN numbers the flagged cells 1..n in row-major order.
 *   get: every row's number; rows made since are numbered rows_total+1,...
 *        in creation order; *rows_total (may be NULL) receives the new
 *        grain_array_tot.
 * Sediment contexts only. */
int hc_grain_index_set(hc_ctx* ctx, const int32_t* index, size_t count, long long rows_total);
int hc_grain_index_get(hc_ctx* ctx, int32_t* index, size_t count, long long* rows_total);

/* Spatially complex rainfall (spatially_complex_rainfall_on).  The
 * reference keeps a per-cell TOPMODEL (runoffGrid, LSDRainfallRunoff.cpp:
 * 175-241) and each step adds every cell's runoff to its water depth
 * (catchment_water_input_and_hydrology with the runoff grid, LCM.cpp:
 * 3559-3573).  hc_runoff_set uploads the grid's old and new j_mean planes
 * (reference layout, (imax+2) x (jmax+2)) after a TOPMODEL update;
 * hc_runoff_add adds to every cell 1..imax-1, 1..jmax-1 the runoff
 * old + ((new - old)/2)(2 - time) (calchydrograph, LCM.cpp:3728-3739) times
 * flow_timestep, capped at `cap` (ERODEFACTOR), formed in the reference's
 * order.  hc_runoff_add before any hc_runoff_set returns HC_ERR_STATE. */
/* reach_water_and_sediment_input (LCM.cpp:3301-3452): n (at most 16) reach
 * points at (x[k], y[k]), in order.  With `sediment` (n x 9 amounts, m, for
 * fractions 1..9 as the reference forms them, FP64 sediment contexts), a
 * point whose water depth exceeds `threshold` gets a grain row if it has
 * none, fraction 1 into Vsusptot and fractions 2..9 into its grain row and
 * elevation; then every point's depth takes water[k] (m).  NULL sediment:
 * water only (hydro-only runs). */
int hc_reach_inputs(hc_ctx* ctx, int n, const int* x, const int* y, const double* sediment,
                    const double* water, double threshold);
int hc_runoff_set(hc_ctx* ctx, const double* old_j_mean, const double* new_j_mean);
int hc_runoff_add(hc_ctx* ctx, double time, double flow_timestep, double cap);

/* Block until every call enqueued on the context has finished. */
int hc_synchronize(hc_ctx* ctx);

/* Health check: counts non-finite or negative water depths on interior
 * cells.  Returns HC_ERR_UNSTABLE when any are found (count in *bad). */
int hc_check_health(hc_ctx* ctx, long long* bad);

/* Exact binary state image of a host-stepped context.  Replaces the
 * reference's lossy hot start: water depth re-read from a raster written at
 * 6 significant figures (water_init_from_raster_on, LCM.cpp:331-373,
 * 1258-1262; save_raster_data, LR.cpp:1216) and grain state from a Grainz
 * file (read_in_graindata_from_file, LCM.cpp:590-683), with qx, qy, vel_dir
 * and the clocks not restored.  The image holds every persistent device
 * buffer in its stored layout and the values that decide how the next step
 * computes, so a context of the same grid, precision and configuration that
 * loads it continues bit for bit.  hc_state_bytes gives the image size;
 * save and load take exactly that many bytes.  Both return HC_ERR_STATE on
 * a context whose device-resident loop (hc_loop_begin) has started, and
 * load returns HC_ERR_STATE for an image of another grid or configuration
 * and HC_ERR_ARG for a foreign or truncated one. */
int hc_state_bytes(hc_ctx* ctx, size_t* bytes);
int hc_state_save(hc_ctx* ctx, void* host, size_t bytes);
int hc_state_load(hc_ctx* ctx, const void* host, size_t bytes);

/* ---- device-resident stepping ----------------------------------------- */

/* Rain cells 1..HC_RF_MAX-1 (index 0 unused, as in the reference's vectors). */
#define HC_RF_MAX 17

/* Everything the loop body of src/main.cpp:139-204 needs besides the raster
 * state already uploaded: the reference's scalar members at the start of the
 * run, the rainfall-runoff tables and the input points (hydro-only runs).
 * `topmodel[t*HC_RF_MAX + n]` is new_j_mean[n] after the TOPMODEL call made
 * when time_1 reaches t (LCM.cpp:3484-3505; the call sequence does not depend
 * on the flow state, so the host tabulates it with libm once). */
typedef struct hc_loop_init {
  double time_step, cycle, maxdepth, waterinput, waterOut, time_1, baseflow;
  int counter;
  double courant_number, in_out_difference_allowed, output_file_save_interval;
  double ERODEFACTOR, MIN_Q, MIN_Q_MAXVAL;
  int max_time_step, maxcycle, rfnum;
  double j_mean[HC_RF_MAX], old_j_mean[HC_RF_MAX], new_j_mean[HC_RF_MAX];
  int nActual[HC_RF_MAX], input_counter[HC_RF_MAX];
  const double* topmodel; /* [ntop][HC_RF_MAX] */
  int ntop;
  const double* area;     /* drainage area, reference layout */
  const int* rfarea;      /* rain cell per cell, reference layout */
  const int* in_x;        /* current input points (LCM.cpp:3741-3767) */
  const int* in_y;
  int npoints;            /* totalinputpoints (an empty list is 1 stale point) */
  int first_x, first_y;   /* in_x[1], in_y[1] kept for that case */
  /* Failure detection (the reference has none): a body whose flow timestep
   * would fall below dt_floor seconds is not run and the loop halts with
   * HC_ERR_UNSTABLE.  0 disables the test; at most max_time_step. */
  double dt_floor;
} hc_loop_init;

/* One record per iteration for output_data's interval series. */
typedef struct hc_loop_record {
  double cycle, previous, old_cycle, new_cycle, temptot, dt_flow, time_step, waterinput;
  double j_mean[HC_RF_MAX];
} hc_loop_record;

typedef struct hc_loop_state {
  double time_step, cycle, maxdepth, temptot, waterinput, waterOut, time_1, baseflow;
  /* error: 0, or why the loop halted: 1 the TOPMODEL table ran out, 2 a
   * depth went non-finite or below -1e-4 m (maxdepth is then +inf), 3 the
   * flow timestep fell below dt_floor */
  int counter, finished, npoints, error;
  double j_mean[HC_RF_MAX], old_j_mean[HC_RF_MAX], new_j_mean[HC_RF_MAX];
} hc_loop_state;

/* Start device-resident stepping (hydro-only contexts): uploads the scalar
 * state and tables; the flow state is whatever the context holds. */
int hc_loop_begin(hc_ctx* ctx, const hc_loop_init* init);

/* Enqueue n loop bodies with no host synchronisation: Courant timestep,
 * clocks, TOPMODEL, input points, fused flow_route+depth_update, scan_area
 * every 5th iteration and water_flux_out, with every scalar on the device.
 * Iterations after the end of the run (main.cpp:204) are no-ops.  With
 * use_graph, blocks of 10 iterations are captured once into a HIP graph and
 * replayed. */
int hc_loop_run(hc_ctx* ctx, int n, int use_graph);

/* Wait for the enqueued iterations and read the scalar state and up to
 * max_records per-iteration records of the last hc_loop_run.  A loop that
 * halted on a failure returns HC_ERR_UNSTABLE (a non-finite or large
 * negative depth: the state holds the body that produced it; or a flow
 * timestep below dt_floor: the state holds the last body run) or
 * HC_ERR_HALTED (TOPMODEL table exhausted), with the state and the records of
 * the bodies that ran filled in; later calls of hc_loop_run are no-ops until
 * hc_loop_begin starts over. */
int hc_loop_sync(hc_ctx* ctx, hc_loop_state* state, hc_loop_record* records,
                 int max_records, int* nrecords);

/* Tile tuner: times the fused flow_route+depth_update kernel for each
 * supported sub-tile height (4, 8, 16 rows held in registers) and strip
 * length (1, 2, 4, 8 sub-tiles walked by one block) on the context's current
 * state, which it leaves unchanged, and keeps the fastest for this context.
 * Call before hc_loop_run captures a graph.  hc_set_tiles fixes the choice. */
int hc_tune_tiles(hc_ctx* ctx, double dt_flow, int* best_rows, int* best_strip,
                  double* best_us);
int hc_set_tiles(hc_ctx* ctx, int rows, int strip);
/* Average duration in microseconds of `reps` fused launches on the current
 * state (which is left unchanged); for the tile tuner and benchmarks. */
int hc_bench_fused(hc_ctx* ctx, double dt_flow, int reps, double* us);

/* Kernel timing for the benchmark: with profiling on, hc_loop_run (without
 * graphs) launches every fused flow_route+depth_update kernel with a start
 * and a stop HIP event on the context stream, taken from the kernel's own
 * dispatch (hipExtLaunchKernel), so they time its execution alone;
 * hc_loop_kernel_time returns their summed duration (ms) and count for the
 * last run. */
int hc_loop_profile(hc_ctx* ctx, int on);
int hc_loop_kernel_time(hc_ctx* ctx, double* total_ms, int* launches);

/* ---- batched ensemble ----------------------------------------------------
 * No counterpart in the reference, whose ensembles are independent processes
 * (parallel_use.md:67-68).  An ensemble steps several independent catchment
 * contexts together: each step is one control block per member (its own
 * Courant dt, clocks and rain input, as hc_loop_run), one persistent fused
 * flow_route+depth_update+water_flux_out kernel over the tile list of all
 * members, and one scan/drain/rebuild launch.  Each member's results are
 * those of hc_loop_run on that member alone.  Members must be hydro-only, of
 * one precision mode, and have had hc_loop_begin; the ensemble borrows their
 * device state (contexts outlive the ensemble).  A member that reaches the end
 * of its run stops while the others continue. */
typedef struct hc_ens hc_ens;
int hc_ens_create(hc_ens** ens, hc_ctx* const* members, int n);
int hc_ens_destroy(hc_ens* ens);
const char* hc_ens_last_error(const hc_ens* ens);
/* Enqueue `steps` ensemble steps on the ensemble's own stream (with
 * use_graph, blocks of 10 steps replay one captured HIP graph). */
int hc_ens_run(hc_ens* ens, int steps, int use_graph);
/* Wait for the run; then hc_loop_sync on each member returns that member's
 * state and per-step records, exactly as after hc_loop_run. */
int hc_ens_sync(hc_ens* ens);
/* Autotuner: times the fused ensemble kernel for tile rows {4, 8, 16} and
 * grids of {2, 4, 8} blocks per CU (each block walking a run of tiles) or one
 * block per tile, on the current state (left unchanged), and keeps the
 * fastest.  Call with members synced. */
int hc_ens_tune(hc_ens* ens, int* best_rows, int* best_blocks, double* best_us);
int hc_ens_set_tiles(hc_ens* ens, int rows, int blocks);
/* Average duration (us) of `reps` fused ensemble launches; state unchanged. */
int hc_ens_bench_fused(hc_ens* ens, int reps, double* us);

/* ---- adjoint -------------------------------------------------------------
 * No counterpart in the reference (it has no derivatives).  Gradient of an
 * outlet-hydrograph misfit L = sum_s (Q_s - Qtarget_s)^2, Q_s the edge outflow
 * of water_flux_out (LCM.cpp:3005-3047) at step s, through nsteps of the FP64
 * flow_route + depth_update (LCM.cpp:3063-3299) at a fixed dt (the reference's
 * set_maximum_timestep bypassed), with the active mask rebuilt every 5th
 * step and rain added at the start of each step to every cell of a rain block:
 * h += dt * rain[s] * mult[block].  Gradients are with respect to the
 * per-cell Manning n (the context's HC_FIELD_MANNINGS) and the block
 * multipliers.  The context must be FP64, hydro-only, with spatially variable
 * n; its state at hc_adj_create is the initial state of every forward run.
 * Two-level checkpointing: the forward keeps every checkpoint_every-th state;
 * the reverse sweep re-runs each segment once into per-step device slots. */
typedef struct hc_adj hc_adj;
/* block: rain block per cell, reference layout ((imax+2) x (jmax+2)), values
 * 0..nblocks-1 or -1 for no rain (nblocks <= 16); rain: nsteps rates (m/s). */
int hc_adj_create(hc_adj** adj, hc_ctx* ctx, int nsteps, double dt, int checkpoint_every,
                  const int* block, int nblocks, const double* rain);
int hc_adj_destroy(hc_adj* adj);
const char* hc_adj_last_error(const hc_adj* adj);
/* target discharge series (nsteps values); zero until set */
int hc_adj_set_target(hc_adj* adj, const double* target);
/* forward run with the given multipliers and the context's current n:
 * loss, and the discharge series when Q is not null */
int hc_adj_forward(hc_adj* adj, const double* mult, double* loss, double* Q);
/* gradients of the last forward's loss: grad_n on the reference layout,
 * grad_mult[nblocks] (either may be null) */
int hc_adj_backward(hc_adj* adj, double* grad_n, double* grad_mult);

#ifdef __cplusplus
}
#endif
#endif /* HC_H */
