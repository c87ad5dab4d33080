/* hc_model.h - C-ABI of the C++ run-loop driver (libhc.so).
 *
 * The driver (hail-caesar_amd/csrc/driver/hc_model.cpp) is the host side of
 * the drop-in: a C++ class with the method surface of LSDCatchmentModel
 * (include/catchmentmodel/LSDCatchmentModel.hpp:73-423) whose loop is the one
 * of src/main.cpp:139-204, with every hot-path method delegated to the
 * device entry points of hc.h.  It reads the reference's parameter files,
 * ASCII DEMs, rainfall and grain-data files and writes the reference's
 * time-series and ASCII raster outputs.  These functions expose it to
 * non-C++ callers (the Python tests and benchmark).
 */
#ifndef HC_MODEL_H
#define HC_MODEL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hcm_model hcm_model;

typedef struct hcm_options {
  int precision;  /* hc_precision: HC_FP64 (plain form) or HC_FP32 */
  int fused;      /* 1: flow_route+depth_update as one fused launch */
  int quiet;      /* 1: no progress output on stdout */
  int write_files;/* 1: write the time series and rasters like the reference */
  double flat_dz; /* FP32 form: see hc_params.flat_dz */
  int device_loop;/* 1: hydro-only runs step with hc_loop_* (scalars on the device;
                     rasters are not written in this mode) */
  int use_graph;  /* with device_loop: replay captured HIP graphs */
  int elev_fixed; /* FP32 form: see hc_params.elev_fixed */
  /* Failure detection, which the reference does not have: a loop body
   * whose flow timestep would fall below dt_floor seconds is not run and
   * hcm_iterate returns HC_ERR_UNSTABLE, as it does after a body that leaves
   * a depth non-finite or below -1e-4 m.  0 selects 1e-4 s (a Courant
   * limit that small needs a depth above 4e6 m at 1 m cells); < 0 turns the
   * timestep test off (the depth test stays). */
  double dt_floor;
} hcm_options;

/* Constructor + initialise_model_domain_extents + initialise_arrays +
 * load_data + set_time_counters + initialise_rainfall_runoff +
 * initialise_drainage_area (src/main.cpp:103-120).  `path` is prepended to
 * `pfile` exactly as the reference does (pname + pfname). */
int hcm_create(hcm_model** m, const char* path, const char* pfile, const hcm_options* opt);
/* Same set-up with the inputs given in memory instead of files: the text
 * of a parameter file and the rasters (unpadded, row-major nrows x ncols,
 * row 0 = first row of an ASCII raster).  Optional rasters may be NULL;
 * the parameter text decides which are used, as the file names would. */
typedef struct hcm_arrays {
  int nrows, ncols;
  double DX, xll, yll;
  const double* dem;          /* read_fname */
  const double* mannings;     /* spatial_mannings_dem_file */
  const double* water_init;   /* water_init_raster_file */
  const int* hydroindex;      /* hydroindex_file */
  const float* rain;          /* rainfall_data_file: rain_rows x rain_cols, mm/h */
  int rain_rows, rain_cols;
  /* grain_data_file: per cell (nrows x ncols, row-major) the 101 values a
   * line of that file holds after its x y index columns, grain[0..10] then
   * strata[z=0..9][n=0..8] (LCM.cpp:611-683); every cell has a grain row */
  const double* graindata;
} hcm_arrays;
int hcm_create_arrays(hcm_model** m, const char* params_text, const hcm_arrays* arrays,
                      const hcm_options* opt);
int hcm_destroy(hcm_model* m);
const char* hcm_last_error(const hcm_model* m);

/* Run up to n loop bodies (src/main.cpp:139-202), stopping when the loop
 * condition (main.cpp:204) ends the run.  Optional per-iteration records:
 * temptot after water_flux_out, cycle (minutes) and the flow timestep.
 * *done receives the number of iterations performed; *finished 1 when the
 * run is over.  HC_ERR_UNSTABLE: the run stopped on a failure (see
 * hcm_options.dt_floor); *done and the records cover the bodies that ran. */
int hcm_iterate(hcm_model* m, int n, double* temptot, double* cycle, double* dtflow,
                int* done, int* finished);

/* Download a field (hc_field ids of hc.h) in the reference layout. */
int hcm_download(hcm_model* m, int field, void* host, size_t count);
size_t hcm_field_count(const hcm_model* m, int field);

/* Scalars by member name: time_step, cycle, counter, maxdepth, temptot,
 * waterinput, waterOut, baseflow, totalinputpoints, imax, jmax, DX,
 * erode_mult, globalsediq. */
int hcm_scalar(hcm_model* m, const char* name, double* value);

/* Switch graph replay of the device-resident loop on or off between runs. */
int hcm_set_graph(hcm_model* m, int use_graph);

/* The device context the model drives (for hc.h calls such as profiling). */
struct hc_ctx* hcm_context(hcm_model* m);

/* Exact binary checkpoint of a host-stepped run, written to `path`: every
 * scalar and array the loop carries between iterations (clocks, TOPMODEL
 * store, input points, drainage area, erosion and output accumulators, the
 * series so far) and the device state image of hc_state_save.  Replaces the
 * reference's lossy hot start from rasters and a Grainz file
 * (water_init_from_raster_on, LCM.cpp:331-373; read_in_graindata_from_file,
 * LCM.cpp:590-683), which restores neither qx, qy, vel_dir nor the clocks.
 * hcm_restart loads it into a model created from the same inputs, which
 * then continues bit for bit as the run that wrote it.  Both return
 * HC_ERR_STATE with a message (hcm_last_error) for a device-resident loop
 * that has started, an unreadable, foreign or truncated file, or one written
 * by a model of another grid, rain cell count, mode or precision; a restart
 * that fails leaves the model as it was. */
int hcm_checkpoint(hcm_model* m, const char* path);
int hcm_restart(hcm_model* m, const char* path);
/* Rows written by output_data so far (14 columns: hours, Qw, Jw, sand_out,
 * Qs, Qg[1..9]); copies at most max_rows; returns the row count. */
int hcm_series(hcm_model* m, double* out, int max_rows);

/* Device-resident stepping driven from outside the model (an ensemble,
 * hc_ens_*): hcm_loop_begin uploads the loop state (hc_loop_begin) if the
 * model has not started it; after the outside run of n steps,
 * hcm_loop_collect reads the state and replays output_data per step, with
 * hcm_iterate's outputs. */
int hcm_loop_begin(hcm_model* m);
int hcm_loop_collect(hcm_model* m, int n, double* temptot, double* cycle, double* dtflow,
                     int* done, int* finished);

#ifdef __cplusplus
}
#endif
#endif /* HC_MODEL_H */
