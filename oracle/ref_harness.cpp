// Test infrastructure only (never linked into the product).
//
// C-ABI harness around the reference model, compiled by oracle/Makefile
// together with the reference's own sources (src/catchmentmodel/*.cpp,
// src/topotools/*.cpp) into oracle/_ref/libhcref.so.  It lets the Python
// fixture scripts drive the reference one method at a time and read its full
// FP64 state:
//   * ref_create() performs exactly the initialisation sequence of
//     src/main.cpp:103-132 (constructor, domain extents, arrays, load_data,
//     time counters, runoff grid, rainfall-runoff and drainage area init);
//   * ref_iterate() runs whole loop bodies in the order of src/main.cpp:139-204;
//   * ref_call() invokes a single hot-path method;
//   * ref_field*/ref_scalar* expose the private arrays and scalars.
#include "catchmentmodel/LSDCatchmentModel.hpp"
#include "catchmentmodel/LSDRainfallRunoff.hpp"

#include <cstring>
#include <string>

namespace {

struct RefRun {
  LSDCatchmentModel* m = nullptr;
  runoffGrid* runoff = nullptr;
  // main.cpp:127-132
  int in_channel_landslide_iter_interval = 10;
  int creep_time_interval_days = 14400;
  int scan_area_interval_iter = 5;
  int global_landsliding_interval_hours = 1440;
  int vegetation_growth_interval_hours = 1440;
  double creep_coeff = 0.028;
};

template <class T>
T* arr2(TNT::Array2D<T>& a, int* d0, int* d1) {
  *d0 = a.dim1();
  *d1 = a.dim2();
  if (a.dim1() == 0 || a.dim2() == 0) return nullptr;
  return &a[0][0];
}

template <class T>
T* arr3(TNT::Array3D<T>& a, int* d0, int* d1, int* d2) {
  *d0 = a.dim1();
  *d1 = a.dim2();
  *d2 = a.dim3();
  if (a.dim1() == 0 || a.dim2() == 0 || a.dim3() == 0) return nullptr;
  return &a[0][0][0];
}

}  // namespace

extern "C" {

void* ref_create(const char* path, const char* pfile) {
  RefRun* r = new RefRun;
  r->m = new LSDCatchmentModel(std::string(path), std::string(pfile));
  LSDCatchmentModel& s = *r->m;
  s.initialise_model_domain_extents();
  s.initialise_arrays();
  if (s.groundwater_mode()) s.initialise_groundwater();
  s.load_data();
  s.set_time_counters();
  r->runoff = new runoffGrid(s.get_imax(), s.get_jmax());
  s.initialise_rainfall_runoff(*r->runoff);
  s.initialise_drainage_area();
  return r;
}

// OpenMP thread count for the reference's parallel loops.  The shared
// `mannings` member written inside flow_route (LSDCatchmentModel.cpp:3091)
// and the neighbour writes of erode (4684-4747) race with more than one
// thread, so fixtures are made with 1.
void ref_set_threads(int n) { omp_set_num_threads(n); }

void ref_destroy(void* h) {
  RefRun* r = static_cast<RefRun*>(h);
  delete r->runoff;
  delete r->m;
  delete r;
}

// One loop body of main.cpp:139-202.  Returns 1 when the do/while condition
// (main.cpp:204) says the run is over.
static int one_iteration(RefRun* r) {
  LSDCatchmentModel& s = *r->m;
  s.set_loop_cycle();
  s.set_inputoutput_diff();
  s.set_maximum_timestep();
  s.increment_counters();
  s.reach_water_and_sediment_input();
  s.catchment_waterinputs(*r->runoff);
  s.flow_route();
  if (s.groundwater_mode()) s.wpgw_water_input();
  s.depth_update();
  s.check_wetted_area(r->scan_area_interval_iter);
  if (!s.is_hydro_only()) s.call_erosion();
  s.water_flux_out();
  s.call_channel_landsliding(r->in_channel_landslide_iter_interval);
  s.slope_creep(r->creep_time_interval_days, r->creep_coeff);
  s.call_global_landsliding(r->global_landsliding_interval_hours);
  s.grow_vegetation(r->vegetation_growth_interval_hours);
  if (s.groundwater_mode()) s.call_groundwater_routines();
  s.write_output_timeseries(*r->runoff);
  s.print_cycle();
  s.save_raster_output();
  return (s.get_cycle() / 60 < s.get_maxcycle()) ? 0 : 1;
}

// Runs up to n iterations; stops early when the run is over.  Returns the
// number of iterations performed.  If temptot_out is non-null the per-step
// outlet discharge (temptot after water_flux_out) is recorded; if cycle_out
// is non-null the clock (minutes) after each step is recorded.
int ref_iterate(void* h, int n, double* temptot_out, double* cycle_out,
                double* dtflow_out) {
  RefRun* r = static_cast<RefRun*>(h);
  int k = 0;
  for (; k < n; ++k) {
    // flow dt of this iteration is get_flow_timestep() as seen by flow_route;
    // recorded through the same call after the loop body has run would be
    // wrong (time_step may change in erode), so run the body inline.
    LSDCatchmentModel& s = *r->m;
    s.set_loop_cycle();
    s.set_inputoutput_diff();
    s.set_maximum_timestep();
    s.increment_counters();
    s.reach_water_and_sediment_input();
    s.catchment_waterinputs(*r->runoff);
    if (dtflow_out) dtflow_out[k] = s.get_flow_timestep();
    s.flow_route();
    if (s.groundwater_mode()) s.wpgw_water_input();
    s.depth_update();
    s.check_wetted_area(r->scan_area_interval_iter);
    if (!s.is_hydro_only()) s.call_erosion();
    s.water_flux_out();
    s.call_channel_landsliding(r->in_channel_landslide_iter_interval);
    s.slope_creep(r->creep_time_interval_days, r->creep_coeff);
    s.call_global_landsliding(r->global_landsliding_interval_hours);
    s.grow_vegetation(r->vegetation_growth_interval_hours);
    if (s.groundwater_mode()) s.call_groundwater_routines();
    s.write_output_timeseries(*r->runoff);
    s.print_cycle();
    s.save_raster_output();
    if (temptot_out) temptot_out[k] = s.temptot;
    if (cycle_out) cycle_out[k] = s.cycle;
    if (!(s.get_cycle() / 60 < s.get_maxcycle())) {
      ++k;
      break;
    }
  }
  return k;
}

int ref_iterate_plain(void* h, int n) {
  RefRun* r = static_cast<RefRun*>(h);
  int k = 0;
  for (; k < n; ++k) {
    if (one_iteration(r)) {
      ++k;
      break;
    }
  }
  return k;
}

// Single hot-path methods.
//   0 flow_route        1 depth_update       2 scan_area
//   3 water_flux_out    4 set_maximum_timestep
//   5 increment_counters 6 catchment_waterinputs 7 call_erosion
//   8 set_loop_cycle    9 set_inputoutput_diff 10 write_output_timeseries
//  11 zero_and_calc_drainage_area  12 get_catchment_input_points
//  13 erode  14 get_flow_timestep  15 topmodel_runoff  16 global_landsliding
//  17 creep(arg)  18 in_channel_landsliding
double ref_call(void* h, int which, double arg) {
  RefRun* r = static_cast<RefRun*>(h);
  LSDCatchmentModel& s = *r->m;
  switch (which) {
    case 0: s.flow_route(); return 0;
    case 1: s.depth_update(); return s.maxdepth;
    case 2: s.scan_area(); return 0;
    case 3: s.water_flux_out(); return s.temptot;
    case 4: s.set_maximum_timestep(); return s.time_step;
    case 5: s.increment_counters(); return s.cycle;
    case 6: s.catchment_waterinputs(*r->runoff); return s.waterinput;
    case 7: s.call_erosion(); return s.erode_mult;
    case 8: s.set_loop_cycle(); return 0;
    case 9: s.set_inputoutput_diff(); return s.input_output_difference;
    case 10: s.write_output_timeseries(*r->runoff); return 0;
    case 11: s.zero_and_calc_drainage_area(); return 0;
    case 12: s.get_catchment_input_points(); return s.totalinputpoints;
    case 13: return s.erode(arg);
    case 14: return s.get_flow_timestep();
    case 15: s.topmodel_runoff(arg); return s.new_j_mean[1];
    case 16: s.global_landsliding(); return 0;
    case 17: s.creep(arg); return 0;
    case 18: s.in_channel_landsliding(); return 0;
    case 19: s.reach_water_and_sediment_input(); return s.waterinput;
    default: return -1;
  }
}

// Per-cell grain helpers (LSDCatchmentModel.cpp:3874-4171).
double ref_grain_call(void* h, int which, int x, int y, double amount, int x2,
                      int y2) {
  LSDCatchmentModel& s = *static_cast<RefRun*>(h)->m;
  switch (which) {
    case 0: s.sort_active(x, y); return 0;
    case 1: s.addGS(x, y); return s.index[x][y];
    case 2: return s.sand_fraction(x);
    case 3: return s.d50(x);
    case 4: s.slide_GS(x, y, amount, x2, y2); return 0;
    default: return -1;
  }
}

double* ref_field(void* h, const char* name, int* d0, int* d1) {
  LSDCatchmentModel& s = *static_cast<RefRun*>(h)->m;
  std::string n(name);
  if (n == "elev") return arr2(s.elev, d0, d1);
  if (n == "water_depth") return arr2(s.water_depth, d0, d1);
  if (n == "qx") return arr2(s.qx, d0, d1);
  if (n == "qy") return arr2(s.qy, d0, d1);
  if (n == "qxs") return arr2(s.qxs, d0, d1);
  if (n == "qys") return arr2(s.qys, d0, d1);
  if (n == "Vsusptot") return arr2(s.Vsusptot, d0, d1);
  if (n == "spat_var_mannings") return arr2(s.spat_var_mannings, d0, d1);
  if (n == "area") return arr2(s.area, d0, d1);
  if (n == "area_depth") return arr2(s.area_depth, d0, d1);
  if (n == "bedrock") return arr2(s.bedrock, d0, d1);
  if (n == "Tau") return arr2(s.Tau, d0, d1);
  if (n == "Vel") return arr2(s.Vel, d0, d1);
  if (n == "ss") return arr2(s.ss, d0, d1);
  if (n == "grain") return arr2(s.grain, d0, d1);
  if (n == "edge") return arr2(s.edge, d0, d1);
  runoffGrid& ro = *static_cast<RefRun*>(h)->runoff;
  if (n == "runoff_j_mean") return arr2(ro.j_mean_array, d0, d1);
  if (n == "runoff_new_j_mean") return arr2(ro.new_j_mean_array, d0, d1);
  *d0 = *d1 = 0;
  return nullptr;
}

int* ref_field_int(void* h, const char* name, int* d0, int* d1) {
  LSDCatchmentModel& s = *static_cast<RefRun*>(h)->m;
  std::string n(name);
  if (n == "down_scan") return arr2(s.down_scan, d0, d1);
  if (n == "index") return arr2(s.index, d0, d1);
  if (n == "rfarea") return arr2(s.rfarea, d0, d1);
  *d0 = *d1 = 0;
  return nullptr;
}

double* ref_field3(void* h, const char* name, int* d0, int* d1, int* d2) {
  LSDCatchmentModel& s = *static_cast<RefRun*>(h)->m;
  std::string n(name);
  if (n == "vel_dir") return arr3(s.vel_dir, d0, d1, d2);
  if (n == "sr") return arr3(s.sr, d0, d1, d2);
  if (n == "sl") return arr3(s.sl, d0, d1, d2);
  if (n == "su") return arr3(s.su, d0, d1, d2);
  if (n == "sd") return arr3(s.sd, d0, d1, d2);
  if (n == "strata") return arr3(s.strata, d0, d1, d2);
  if (n == "veg") return arr3(s.veg, d0, d1, d2);
  *d0 = *d1 = *d2 = 0;
  return nullptr;
}

// std::vector<double> members.
double* ref_vec(void* h, const char* name, int* len) {
  LSDCatchmentModel& s = *static_cast<RefRun*>(h)->m;
  std::string n(name);
  std::vector<double>* v = nullptr;
  if (n == "j") v = &s.j;
  else if (n == "jo") v = &s.jo;
  else if (n == "j_mean") v = &s.j_mean;
  else if (n == "old_j_mean") v = &s.old_j_mean;
  else if (n == "new_j_mean") v = &s.new_j_mean;
  else if (n == "sum_grain") v = &s.sum_grain;
  else if (n == "fallVelocity") v = &s.fallVelocity;
  if (!v) { *len = 0; return nullptr; }
  *len = static_cast<int>(v->size());
  return v->data();
}

int* ref_vec_int(void* h, const char* name, int* len) {
  LSDCatchmentModel& s = *static_cast<RefRun*>(h)->m;
  std::string n(name);
  std::vector<int>* v = nullptr;
  if (n == "catchment_input_x_coord") v = &s.catchment_input_x_coord;
  else if (n == "catchment_input_y_coord") v = &s.catchment_input_y_coord;
  else if (n == "catchment_input_counter") v = &s.catchment_input_counter;
  else if (n == "nActualGridCells") v = &s.nActualGridCells;
  if (!v) { *len = 0; return nullptr; }
  *len = static_cast<int>(v->size());
  return v->data();
}

// Scalars.  Returns 0 on success, -1 for an unknown name.
int ref_scalar(void* h, const char* name, double* value, int set) {
  LSDCatchmentModel& s = *static_cast<RefRun*>(h)->m;
  std::string n(name);
#define HC_D(field)                                   \
  if (n == #field) {                                  \
    if (set) s.field = *value; else *value = s.field; \
    return 0;                                         \
  }
#define HC_I(field)                                                   \
  if (n == #field) {                                                  \
    if (set) s.field = static_cast<decltype(s.field)>(*value);        \
    else *value = static_cast<double>(s.field);                       \
    return 0;                                                         \
  }
  HC_D(time_step) HC_D(cycle) HC_D(maxdepth) HC_D(temptot) HC_D(waterinput)
  HC_D(waterOut) HC_D(input_output_difference) HC_D(erode_mult)
  HC_D(erode_call) HC_D(baseflow) HC_D(time_1) HC_D(previous) HC_D(DX)
  HC_D(mannings) HC_D(hflow_threshold) HC_D(edgeslope) HC_D(courant_number)
  HC_D(froude_limit) HC_D(ERODEFACTOR) HC_D(active) HC_D(Csuspmax)
  HC_D(max_vel) HC_D(M) HC_D(globalsediq) HC_D(sediQ) HC_D(Qw_newvol)
  HC_D(Qw_hour) HC_D(tx) HC_D(new_cycle) HC_D(old_cycle) HC_D(temptotal)
  HC_D(Jw_newvol) HC_D(Jw_overvol) HC_D(Jw_hour)
  HC_D(d1) HC_D(d2) HC_D(d3) HC_D(d4) HC_D(d5) HC_D(d6) HC_D(d7) HC_D(d8)
  HC_D(d9) HC_D(rain_data_time_step) HC_D(in_out_difference_allowed)
  HC_D(water_depth_erosion_threshold) HC_D(output_file_save_interval)
  HC_D(bedrock_erosion_threshold) HC_D(bedrock_erosion_rate)
  HC_D(lateral_constant) HC_D(chann_lateral_erosion) HC_D(veg_lat_restriction)
  HC_D(vegTauCrit) HC_D(MIN_Q) HC_D(MIN_Q_MAXVAL)
  HC_I(counter) HC_I(imax) HC_I(jmax) HC_I(max_time_step) HC_I(maxcycle)
  HC_I(totalinputpoints) HC_I(rfnum) HC_I(grain_array_tot) HC_I(hours)
  HC_I(hydro_only) HC_I(spatially_var_mannings) HC_I(wilcock) HC_I(einstein)
  HC_I(LIMIT)
#undef HC_D
#undef HC_I
  return -1;
}

}  // extern "C"
