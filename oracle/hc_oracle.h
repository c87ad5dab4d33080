/* Test infrastructure only: the oracle.
 *
 * A plain FP64 CPU restatement of the HAIL-CAESAR per-timestep hot path,
 * written from the reference's semantics (file:line cited per function in
 * hc_oracle.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it, and only as the checker.
 *
 * Layout: every 2-D field is the reference's padded row-major
 * (imax+2) x (jmax+2) array, x = row (1..imax), y = column (1..jmax),
 * y contiguous (LSDCatchmentModel.cpp:1622-1705).
 */
#ifndef HC_ORACLE_H
#define HC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  int imax, jmax;
  double DX;
  double courant_number;
  double froude_limit;
  double hflow_threshold;
  double edgeslope;
  double mannings;
  int spatially_var_mannings;
  double water_depth_erosion_threshold; /* fixed 0.01 in the reference */
  double ERODEFACTOR;
  double active;
  double Csuspmax;
  double max_vel;
  int wilcock;
  int einstein;
  double d[10];         /* d[1..9] grain sizes */
  double dprop[11];
  double fallVelocity[11];
  int rfnum;
  double M;
  double rain_data_time_step; /* minutes */
  int max_time_step;          /* seconds (already capped as initialise_arrays does) */
  double in_out_difference_allowed;
  double MIN_Q, MIN_Q_MAXVAL;
  int maxcycle;               /* hours */
  double output_file_save_interval; /* minutes */
  int hydro_only;
  double bedrock_erosion_threshold;
  double bedrock_erosion_rate;
  double lateral_constant;
  double chann_lateral_erosion;
  double veg_lat_restriction;
  double vegTauCrit;
  int scan_interval;          /* 5 (main.cpp:129) */
  int nthreads;               /* OpenMP threads for the flow/depth loops */
  int spatially_complex_rainfall; /* spatially_complex_rainfall_on (LCM.cpp:1355-1359) */
  double no_data_value;       /* the DEM header's NODATA value (LCM.cpp:125) */
  int reach_mode;             /* reach_mode (LCM.cpp:1456-1459) */
  int div_inputs;             /* divide_inputs_by, held as int (LCM.hpp:685) */
  int reach_input_data_timestep; /* minutes, int (LCM.hpp:449) */
  double creep_rate;          /* creep_rate, CREEP_RATE (LCM.hpp:499, LCM.cpp:1547-1551) */
  double failureangle;        /* slope_failure_thresh (LCM.hpp:470, LCM.cpp:1553-1558) */
} orc_params;

typedef struct orc_model orc_model;

orc_model* orc_create(const orc_params* p);
void orc_destroy(orc_model* m);

/* Field access by name.  2-D double fields: elev, water_depth, qx, qy, qxs,
 * qys, Vsusptot, mannings, area, bedrock, Tau, Vel, ss.  Multi-plane fields
 * (planes last, cell-major like the reference's Array3D): vel_dir (9),
 * grain (11), strata (10*9), sr, sl, su, sd (10 each).
 * Returns NULL for an unknown name; *len receives the element count. */
double* orc_field(orc_model* m, const char* name, long* len);
/* int fields: mask (1 = cell on the active list, the reference's down_scan
 * membership), hasgrain (index != -9999), rfarea. */
int* orc_field_int(orc_model* m, const char* name, long* len);
/* reach_water_and_sediment_input (LCM.cpp:3301-3452), one call. */
void orc_reach_input(orc_model* m);
/* reach input point k (0..2) at (x, y) with its file's rows (rows x cols,
 * row-major, held as float as read_reachfile does, LCM.cpp:718-760). */
void orc_set_reach(orc_model* m, int k, int x, int y, const double* rows, int nrows, int ncols);
/* rainfall table: rows x rfnum, row-major (mm/h). */
void orc_set_rain(orc_model* m, const double* rain, int rows, int cols);
/* Scalars by name (time_step, cycle, counter, maxdepth, temptot, waterinput,
 * waterOut, baseflow, time_1, erode_mult, erode_call, ...). */
int orc_scalar(orc_model* m, const char* name, double* value, int set);

/* Set-up that load_data/initialise_* perform after the arrays are filled:
 * nActualGridCells, TOPMODEL first call, drainage area, input points. */
void orc_initialise(orc_model* m);

/* Hot-path methods. */
void orc_flow_route(orc_model* m);
double orc_depth_update(orc_model* m);
void orc_scan_area(orc_model* m);
double orc_water_flux_out(orc_model* m);
double orc_get_flow_timestep(orc_model* m);
void orc_set_maximum_timestep(orc_model* m);
double orc_erode(orc_model* m, double mult_factor);
void orc_call_erosion(orc_model* m);
/* global_landsliding with the given slope failure angle (degrees); returns
 * the number of sweeps. */
int orc_global_landsliding(orc_model* m, double failureangle);
/* creep(time) with the given creep rate (the parameter key creep_rate). */
void orc_creep(orc_model* m, double time, double creep_rate);
/* in_channel_landsliding over the current mask (the active list). */
void orc_in_channel_landsliding(orc_model* m, double failureangle);
void orc_drainage_area(orc_model* m);
int orc_input_points(orc_model* m);
double orc_grain_call(orc_model* m, int which, int x, int y, double amount, int x2, int y2);

/* Whole loop bodies (main.cpp:139-204, hydro processes; erosion via
 * call_erosion when !hydro_only).  Records per-step temptot and cycle.
 * Returns the number of iterations run (stops at the end of the run). */
int orc_iterate(orc_model* m, int n, double* temptot_out, double* cycle_out,
                double* dtflow_out);

/* Time-series rows written by output_data so far: 14 columns per row
 * (hours, Qw, Jw, sand_out, Qs, Qg[1..9]).  Returns the row count; copies at
 * most max_rows rows. */
int orc_series(orc_model* m, double* out, int max_rows);

#ifdef __cplusplus
}
#endif
#endif
