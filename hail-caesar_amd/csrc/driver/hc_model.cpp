// C++ run-loop driver: the host side of the drop-in.
//
// CatchmentModel keeps the method surface of the reference's
// LSDCatchmentModel (include/catchmentmodel/LSDCatchmentModel.hpp:73-423) and
// run() is the loop of src/main.cpp:139-204.  Host-side scalar logic
// (clocks, Courant timestep, TOPMODEL rainfall-runoff, drainage area and
// input points, interval time series) stays here in FP64 exactly as the
// reference computes it; every raster-wide method goes through the device
// entry points of include/hc.h.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <limits>
#include <cctype>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/hc.h"
#include "../../../include/hc_model.h"

namespace hcdrv {

constexpr double kGravity = 9.81;  // LSDCatchmentModel.hpp:437
constexpr int kGMax = 10;          // LSDCatchmentModel.hpp:444
// D8 offsets of LSDCatchmentModel.hpp:445-446 (index 0 unused).
constexpr int kDX8[9] = {0, 0, 1, 1, 1, 0, -1, -1, -1};
constexpr int kDY8[9] = {0, -1, -1, 0, 1, 1, 1, 0, -1};

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
// A run stopped by a failure test (or a device loop halted early): the
// C-ABI code to return and, for a device-resident run, the bodies that ran.
struct Halted : Error {
  int code, done;
  Halted(const std::string& m, int c, int d = -1) : Error(m), code(c), done(d) {}
};
// The timestep floor when hcm_options.dt_floor is 0.
constexpr double kDefaultDtFloor = 1e-4;

// x86-64 behaviour of static_cast<int>(double) on out-of-range values
// (reached at LSDCatchmentModel.cpp:2283 when erode returns 0).
int trunc_int(double v) {
  if (!(v > -2147483649.0 && v < 2147483648.0)) return INT_MIN;
  return static_cast<int>(v);
}

// One `key: value # comment` line, as LSDUtils::parse_line splits it
// (src/catchmentmodel/LSDUtils.cpp:12-77).
bool parse_line(std::istream& in, std::string& key, std::string& value) {
  char c;
  std::string buf;
  int word = 0;
  bool any = false;
  key = "NULL";
  value = "NULL";
  while (in.get(c)) {
    any = true;
    if (c == '#') {
      if (word == 0) { key = "NULL"; value = "NULL"; }
      if (word == 1) value = "NULL";
      word = 2;
    }
    if ((c == ' ' || c == '\t') && buf.empty() && word < 2) continue;
    if ((c == ':' && word == 0) || ((c == ' ' || c == '\n' || c == '\t') && word == 1)) {
      while (!buf.empty() && (buf.back() == ' ' || buf.back() == '\t')) buf.pop_back();
      if (word == 0) key = buf; else value = buf;
      ++word;
      buf.clear();
    } else if (c == '\n' && word == 0) {
      key = "NULL";
      value = buf;
      ++word;
    } else if (word < 2) {
      buf.push_back(c);
    }
    if (c == '\n') break;
  }
  return any;
}

std::string strip_ctrl(std::string s) {
  while (!s.empty() && std::iscntrl(static_cast<unsigned char>(s.back()))) s.pop_back();
  return s;
}

struct Raster {
  int ncols = 0, nrows = 0;
  double xll = 0, yll = 0, cell = 0, nodata = -9999;
  std::vector<double> v;  // nrows x ncols
};

Raster read_asc(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw Error("cannot open raster " + path);
  Raster r;
  std::string s;
  in >> s >> r.ncols >> s >> r.nrows >> s >> r.xll >> s >> r.yll >> s >> r.cell >> s >> r.nodata;
  if (!in || r.ncols <= 0 || r.nrows <= 0) throw Error("bad ASCII raster header in " + path);
  r.v.assign(static_cast<size_t>(r.ncols) * r.nrows, r.nodata);
  for (auto& x : r.v) in >> x;
  return r;
}

class CatchmentModel {
 public:
  CatchmentModel(const std::string& pname, const std::string& pfname, const hcm_options& opt)
      : opt_(opt) {
    std::ifstream in(pname + pfname);
    if (!in) throw Error("cannot open parameter file " + pname + pfname);
    initialise_variables(in);
  }
  // inputs held in memory (hcm_create_arrays)
  CatchmentModel(const std::string& params_text, const hcm_arrays& arr, const hcm_options& opt)
      : opt_(opt), arrays_(&arr) {
    std::istringstream in(params_text);
    initialise_variables(in);
  }
  ~CatchmentModel() {
    if (ctx_) hc_destroy(ctx_);
  }

  // ------------------------------------------------ set-up (main.cpp:103-120)
  void initialise_model_domain_extents() {
    if (arrays_) {
      imax_ = arrays_->nrows;
      jmax_ = arrays_->ncols;
      DX_ = arrays_->DX;
      xll_ = arrays_->xll;
      yll_ = arrays_->yll;
      no_data_value_ = -9999;
      if (imax_ <= 0 || jmax_ <= 0 || !arrays_->dem) throw Error("bad in-memory DEM");
      return;
    }
    const std::string f = read_path_ + "/" + read_fname_ + "." + dem_read_extension_;
    std::ifstream in(f);
    if (!in) throw Error("No terrain DEM found by name of: " + f);
    std::string s;
    in >> s >> jmax_ >> s >> imax_ >> s >> xll_ >> s >> yll_ >> s >> DX_ >> s >> no_data_value_;
    if (!in || imax_ <= 0 || jmax_ <= 0) throw Error("bad DEM header in " + f);
  }

  void initialise_arrays() {
    NR_ = imax_ + 2;
    NC_ = jmax_ + 2;
    const size_t n = static_cast<size_t>(NR_) * NC_;
    elev_.assign(n, -9999);
    rfarea_.assign(n, 1);
    area_.assign(n, 0.0);
    area_depth_.assign(n, 0.0);
    if (rain_data_time_step_ < 1) rain_data_time_step_ = 1;  // LCM.cpp:1680-1683
    if (max_time_step_ / 60 > rain_data_time_step_)
      max_time_step_ = static_cast<int>(rain_data_time_step_) * 60;
    const int rows = static_cast<int>(maxcycle_ * (60 / rain_data_time_step_)) + 100;
    rain_.assign(rows, std::vector<float>(rfnum_ + 1, 0.0f));
    j_.assign(rfnum_ + 1, 0.000000001);
    jo_.assign(rfnum_ + 1, 0.000000001);
    j_mean_.assign(rfnum_ + 1, 0.0);
    old_j_mean_.assign(rfnum_ + 1, 0.0);
    new_j_mean_.assign(rfnum_ + 1, 0.0);
    nActual_.assign(rfnum_ + 1, 0);
    input_counter_.assign(rfnum_ + 1, 0);
    in_x_.assign(static_cast<size_t>(imax_) * jmax_ + 2, 0);
    in_y_.assign(static_cast<size_t>(imax_) * jmax_ + 2, 0);
    sum_grain_.assign(kGMax + 1, 0.0);
    old_sum_grain_.assign(kGMax + 1, 0.0);
    Qg_step_.assign(kGMax + 1, 0.0);
    Qg_hour_.assign(kGMax + 1, 0.0);
    Qg_over_.assign(kGMax + 1, 0.0);
    Qg_last_.assign(kGMax + 1, 0.0);

    hc_grid g{imax_, jmax_, DX_, opt_.precision, hydro_only_ ? 0 : 1};
    hc_params p{};
    p.courant_number = courant_number_;
    p.froude_limit = froude_limit_;
    p.hflow_threshold = hflow_threshold_;
    p.edgeslope = edgeslope_;
    p.mannings = mannings_;
    p.spatially_var_mannings = spatially_var_mannings_ ? 1 : 0;
    p.water_depth_erosion_threshold = water_depth_erosion_threshold_;
    p.ERODEFACTOR = ERODEFACTOR_;
    p.active = active_;
    p.Csuspmax = Csuspmax_;
    p.max_vel = max_vel_;
    p.transport_law = einstein_ ? HC_EINSTEIN : HC_WILCOCK;
    for (int k = 0; k < 10; ++k) p.d[k] = d_[k];
    for (int k = 0; k < 11; ++k) {
      p.dprop[k] = dprop_[k];
      p.fallVelocity[k] = fall_[k];
    }
    p.bedrock_erosion_threshold = bedrock_erosion_threshold_;
    p.bedrock_erosion_rate = bedrock_erosion_rate_;
    p.lateral_constant = lateral_constant_;
    p.chann_lateral_erosion = chann_lateral_erosion_;
    p.veg_lat_restriction = veg_lat_restriction_;
    p.flat_dz = opt_.flat_dz;
    // erosion moves elevations every erode call: they stay FP64 in both
    // precisions, so the fixed-point form applies to flow-only runs
    p.elev_fixed = hydro_only_ ? opt_.elev_fixed : 0;
    check(hc_create(&ctx_, &g, &p), "hc_create");
  }

  // An input raster: from memory when the model was built from arrays.
  Raster input_raster(const std::string& file, const double* mem) const {
    if (!arrays_) return read_asc(read_path_ + "/" + file);
    if (!mem) throw Error("in-memory input missing for " + file);
    Raster r;
    r.nrows = arrays_->nrows;
    r.ncols = arrays_->ncols;
    r.cell = arrays_->DX;
    r.v.assign(mem, mem + static_cast<size_t>(r.nrows) * r.ncols);
    return r;
  }

  void load_data() {
    Raster dem = input_raster(read_fname_ + "." + dem_read_extension_, arrays_ ? arrays_->dem : nullptr);
    if (dem.nrows != imax_ || dem.ncols != jmax_) throw Error("DEM size changed between reads");
    for (int i = 0; i < imax_; ++i)
      for (int j = 0; j < jmax_; ++j) elev_[at(i + 1, j + 1)] = dem.v[static_cast<size_t>(i) * jmax_ + j];
    check_DEM_edge_condition();
    upload(HC_FIELD_ELEV, elev_.data());
    if (spatially_var_rainfall_) {
      if (arrays_) {
        if (!arrays_->hydroindex) throw Error("in-memory hydroindex missing");
        for (int i = 0; i < imax_; ++i)
          for (int j = 0; j < jmax_; ++j)
            rfarea_[at(i + 1, j + 1)] = arrays_->hydroindex[static_cast<size_t>(i) * jmax_ + j];
      } else {
        Raster r = read_asc(read_path_ + "/" + hydroindex_fname_);
        for (int i = 0; i < imax_; ++i)
          for (int j = 0; j < jmax_; ++j)
            rfarea_[at(i + 1, j + 1)] = static_cast<int>(r.v[static_cast<size_t>(i) * jmax_ + j]);
      }
    }
    if (spatially_var_mannings_) {
      Raster r = input_raster(mannings_fname_, arrays_ ? arrays_->mannings : nullptr);
      std::vector<double> nm(elev_.size(), 0.0);
      for (int i = 0; i < imax_; ++i)
        for (int j = 0; j < jmax_; ++j) nm[at(i + 1, j + 1)] = r.v[static_cast<size_t>(i) * jmax_ + j];
      upload(HC_FIELD_MANNINGS, nm.data());
    }
    count_catchment_gridcells();
    if (bedrock_layer_on_) {
      // the bedrock raster is copied without the +1 padding offset
      // (LSDCatchmentModel.cpp:316: bedrock = bedrockR.get_RasterData_dbl())
      Raster r = read_asc(read_path_ + "/" + bedrock_data_file_);
      std::vector<double> b(elev_.size(), -9999.0);
      for (int i = 0; i < std::min(r.nrows, NR_); ++i)
        for (int j = 0; j < std::min(r.ncols, NC_); ++j) b[at(i, j)] = r.v[static_cast<size_t>(i) * r.ncols + j];
      if (!hydro_only_) upload(HC_FIELD_BEDROCK, b.data());
    }
    if (water_init_from_raster_) {
      Raster r = input_raster(water_init_raster_file_, arrays_ ? arrays_->water_init : nullptr);
      std::vector<double> h(elev_.size(), 0.0);
      for (int i = 0; i < imax_; ++i)
        for (int j = 0; j < jmax_; ++j) h[at(i + 1, j + 1)] = r.v[static_cast<size_t>(i) * jmax_ + j];
      upload(HC_FIELD_WATER_DEPTH, h.data());
    }
    if (rainfall_data_on_) {
      if (arrays_) {
        if (!arrays_->rain) throw Error("in-memory rainfall missing");
        rain_.assign(arrays_->rain_rows, std::vector<float>(arrays_->rain_cols));
        for (int r = 0; r < arrays_->rain_rows; ++r)
          for (int c = 0; c < arrays_->rain_cols; ++c)
            rain_[r][c] = arrays_->rain[static_cast<size_t>(r) * arrays_->rain_cols + c];
      } else {
        read_rainfall(read_path_ + "/" + rainfall_data_file_);
      }
    }
    load_reach_files();
    if (graindata_from_file_ && !hydro_only_) {
      if (arrays_) ingest_graindata_arrays();
      else ingest_graindata(read_path_ + "/" + grain_data_file_);
    }
  }

  void set_time_counters() {  // LCM.cpp:1862-1871
    save_time_ = cycle_;
    creep_time_ = cycle_;
    creep_time2_ = cycle_;
    time_1_ = cycle_;
  }
  void initialise_rainfall_runoff() {  // LCM.cpp:2041-2053
    if (complex_rain_) {
      // the runoffGrid (LSDRainfallRunoff.cpp:96-104): j = jo = 1e-9, the means 0
      const size_t n = static_cast<size_t>(NR_) * NC_;
      rj_.assign(n, 0.000000001);
      rjo_.assign(n, 0.000000001);
      rj_mean_.assign(n, 0.0);
      rold_.assign(n, 0.0);
      rnew_.assign(n, 0.0);
      topmodel_runoff_grid(1.0);
      upload_runoff();
      return;
    }
    topmodel_runoff(1.0);
  }
  void initialise_drainage_area() {  // LCM.cpp:2055-2069
    if (complex_rain_) return;       // every cell takes its own runoff: no input points
    zero_and_calc_drainage_area();
    get_catchment_input_points();
  }

  // ------------------------------------------ loop methods (main.cpp:141-202)
  void set_loop_cycle() {  // LCM.cpp:1873-1877
    previous_ = cycle_;
    old_cycle_ = std::fmod(cycle_, output_file_save_interval_);
  }
  void set_inputoutput_diff() { iod_ = std::fabs(waterinput_ - waterOut_); }
  double courant_friedrichs_lewy_condition() const {  // LCM.cpp:1884-1887
    return courant_number_ * (DX_ / std::sqrt(kGravity * (maxdepth_)));
  }
  void set_maximum_timestep() {  // LCM.cpp:1890-1910
    if (maxdepth_ <= 0.1) maxdepth_ = 0.1;
    if (time_step_ < courant_friedrichs_lewy_condition()) time_step_ = courant_friedrichs_lewy_condition();
    if (iod_ > in_out_difference_allowed_ && time_step_ > courant_friedrichs_lewy_condition())
      time_step_ = courant_friedrichs_lewy_condition();
  }
  double get_flow_timestep() const {  // LCM.cpp:1913-1923
    double f = time_step_;
    if (f > courant_friedrichs_lewy_condition()) f = courant_friedrichs_lewy_condition();
    return f;
  }
  void increment_counters() {  // LCM.cpp:1925-1933
    counter_++;
    cycle_ += time_step_ / 60;
    new_cycle_ = std::fmod(cycle_, output_file_save_interval_);
  }
  // reach_water_and_sediment_input (LCM.cpp:3301-3452).  With erosion, a
  // wet point takes each fraction's sediment (file columns 5..13, fraction 1
  // suspended, LCM.cpp:2980), scaled down together when their mean over the
  // points exceeds 0.75 ERODEFACTOR; the total is formed with the float
  // |value| / div_inputs, the amounts with the double product first, as the
  // reference's expressions do.  Every point then takes the discharge
  // (column 1) interpolated to this cycle.  The amounts are formed here and
  // applied on the device in the reference's point order.
  void reach_water_and_sediment_input() {
    const int np = reach_n_;
    if (np == 0) return;
    const double fts = get_flow_timestep();
    const int rdt = reach_rdt_;
    const int row = static_cast<int>(cycle_ / rdt);
    std::vector<double> sed, water(np);
    const bool with_sed = reach_mode_ && !hydro_only_;
    if (with_sed) {
      sed.assign(static_cast<size_t>(np) * 9, 0.0);
      for (int n = 0; n < np; ++n) {
        double adding_factor = 1, added_tot = 0;
        for (int tn = 5; tn <= kGMax + 3; ++tn)
          added_tot += (std::abs(reach_val(n, row, tn))) / div_inputs_ / (DX_ * DX_) / (rdt * 60) * time_step_;
        if (added_tot / np > ERODEFACTOR_ * 0.75) adding_factor = (ERODEFACTOR_ * 0.75) / (added_tot / np);
        for (int tn = 5; tn <= kGMax + 3; ++tn)
          sed[static_cast<size_t>(n) * 9 + (tn - 5)] =
              adding_factor * (std::abs(reach_val(n, row, tn))) / div_inputs_ / (DX_ * DX_) / (rdt * 60) * time_step_;
      }
    }
    for (int n = 0; n < np; ++n) {
      const double i1 = reach_val(n, row, 1), i2 = reach_val(n, row + 1, 1);
      const double prop = (((row + 1) * rdt) - cycle_) / rdt;
      const double input = i1 + ((i2 - i1) * (1 - prop));
      waterinput_ += (input / div_inputs_);
      water[n] = (input / div_inputs_) / (DX_ * DX_) * fts;
    }
    check(hc_reach_inputs(ctx_, np, reach_x_, reach_y_, with_sed ? sed.data() : nullptr, water.data(),
                          water_depth_erosion_threshold_),
          "reach_inputs");
  }
  float reach_val(int k, int row, int col) const {
    if (row < 0 || row >= static_cast<int>(reach_rows_[k].size())) return 0.0f;
    if (col < 0 || col >= static_cast<int>(reach_rows_[k][row].size())) return 0.0f;
    return reach_rows_[k][row][col];
  }
  void catchment_waterinputs() {             // LCM.cpp:2027-2039
    waterinput_ = 0;
    if (complex_rain_) catchment_water_input_grid(get_flow_timestep());
    else catchment_water_input_and_hydrology(get_flow_timestep());
  }
  void flow_route() {  // LCM.cpp:3063-3247
    last_dt_ = get_flow_timestep();
    if (opt_.fused && hydro_only_) return;  // performed together with depth_update
    check(hc_flow_route(ctx_, last_dt_), "flow_route");
  }
  void depth_update() {  // LCM.cpp:3259-3299
    const double dt = get_flow_timestep();
    if (opt_.fused && hydro_only_)
      check(hc_flow_depth(ctx_, dt, &maxdepth_), "flow_depth");
    else
      check(hc_depth_update(ctx_, dt, &maxdepth_), "depth_update");
  }
  void check_wetted_area(int interval) {  // LCM.cpp:2019-2025
    if ((counter_ % interval) == 0) check(hc_scan_area(ctx_), "scan_area");
  }
  void call_erosion() {  // LCM.cpp:2277-2294
    if (counter_ >= erode_call_) {
      erode_mult_ = trunc_int(ERODEFACTOR_ / erode(erode_mult_));
      if (erode_mult_ < 1) erode_mult_ = 1;
      if (erode_mult_ > 5) erode_mult_ = 5;
      erode_call_ = counter_ + erode_mult_;
    }
  }
  double erode(double mult) {  // LCM.cpp:4211-4845 (device), 4829-4842 (host sums)
    double tb = 0;
    float g[10] = {};
    check(hc_erode(ctx_, mult, &time_step_, &tb, g), "erode");
    sediQ_ = 0;
    for (int n = 1; n <= kGMax; ++n) {
      const double gv = (n < 10) ? static_cast<double>(g[n]) : 0.0;
      sediQ_ += gv * DX_ * DX_;
      globalsediq_ += gv * DX_ * DX_;
      sum_grain_[n] += gv * DX_ * DX_;
    }
    return tb;
  }
  void water_flux_out() {  // LCM.cpp:3005-3047
    check(hc_water_flux_out(ctx_, time_step_, &temptot_), "water_flux_out");
    waterOut_ = temptot_;
  }
  void call_channel_landsliding(int interval) {  // LCM.cpp:1968-1974
    if (!hydro_only_ && (counter_ % interval) == 0)
      check(hc_in_channel_landsliding(ctx_, failureangle_), "in_channel_landsliding");
  }
  void slope_creep(int interval_days, double coeff) {  // LCM.cpp:1976-1984
    if (!hydro_only_ && cycle_ > creep_time_) {
      creep_time_ += interval_days;
      check(hc_creep(ctx_, coeff, creep_rate_), "creep");
    }
  }
  void call_global_landsliding(int interval_hours) {  // LCM.cpp:1986-1996
    if (!hydro_only_ && cycle_ > creep_time2_) {
      creep_time2_ += interval_hours;
      check(hc_global_landsliding(ctx_, failureangle_, nullptr), "global_landsliding");
    }
  }
  void write_output_timeseries() {  // LCM.cpp:1952-1966
    temptotal_ = temptot_;
    output_data();
  }
  void print_cycle() {
    if (debug_print_cycle_ && !opt_.quiet) std::cout << "Cycle: " << cycle_ << "                  \r" << std::flush;
  }
  void save_raster_output() {  // LCM.cpp:1935-1942
    if (cycle_ >= save_time_) {
      save_raster_data(std::fabs(cycle_));
      save_time_ += saveinterval_;
    }
  }
  double get_cycle() const { return cycle_; }
  int get_maxcycle() const { return maxcycle_; }
  bool is_hydro_only() const { return hydro_only_; }

  // The timestep floor in force (0: none).
  double dt_floor() const {
    if (opt_.dt_floor < 0) return 0.0;
    return opt_.dt_floor > 0 ? opt_.dt_floor : kDefaultDtFloor;
  }
  // Failure tests before a body, on the previous body's maximum depth: the
  // depth kernels report +inf for a non-finite or large negative depth, and
  // the flow timestep of the next body is the Courant limit at that maximum
  // (set_maximum_timestep holds time_step at or above it; the TOPMODEL
  // catch-up lowers it to max_time_step at the least).  Nothing of the body
  // has run when either throws.  The device loop makes the same tests.
  void check_health() const {
    if (!(maxdepth_ < std::numeric_limits<double>::infinity()))
      throw Halted("unstable: a water depth went non-finite or below -1e-4 m (counter " +
                       std::to_string(counter_) + ")", HC_ERR_UNSTABLE);
    const double floor = std::min(dt_floor(), static_cast<double>(max_time_step_));
    const double md = maxdepth_ <= 0.1 ? 0.1 : maxdepth_;
    if (floor > 0 && courant_number_ * (DX_ / std::sqrt(kGravity * md)) < floor)
      throw Halted("unstable: the flow timestep fell below the floor of " +
                       std::to_string(floor) + " s (maximum depth " + std::to_string(maxdepth_) +
                       " m, counter " + std::to_string(counter_) + ")", HC_ERR_UNSTABLE);
  }

  // One loop body of main.cpp:139-202; returns true when main.cpp:204 ends
  // the run.
  bool iterate() {
    check_health();
    set_loop_cycle();
    set_inputoutput_diff();
    set_maximum_timestep();
    increment_counters();
    reach_water_and_sediment_input();
    catchment_waterinputs();
    flow_route();
    depth_update();
    check_wetted_area(5);
    if (!is_hydro_only()) call_erosion();
    water_flux_out();
    call_channel_landsliding(10);
    slope_creep(14400, 0.028);
    call_global_landsliding(1440);
    write_output_timeseries();
    print_cycle();
    save_raster_output();
    return !(get_cycle() / 60 < get_maxcycle());
  }

  bool device_loop_enabled() const {
    return opt_.device_loop && hydro_only_ && !complex_rain_ && reach_n_ == 0;
  }
  void set_graph(bool g) { opt_.use_graph = g ? 1 : 0; }

  // Device-resident stepping (hc_loop_*): n loop bodies with the scalar
  // state on the device; output_data is replayed here from the per-iteration
  // records so the interval series is the one the host loop would write.
  int iterate_device(int n, double* tt, double* cy, double* dtf, bool* fin) {
    if (!loop_started_) begin_device_loop();
    check(hc_loop_run(ctx_, n, opt_.use_graph), "loop_run");
    return collect_device(n, tt, cy, dtf, fin);
  }

  // After n device-resident steps (hc_loop_run here, or an ensemble run that
  // included this model): read the state and replay output_data per step.
  int collect_device(int n, double* tt, double* cy, double* dtf, bool* fin) {
    std::vector<hc_loop_record> rec(static_cast<size_t>(std::max(n, 1)));
    hc_loop_state st{};
    int nrec = 0;
    const int src = hc_loop_sync(ctx_, &st, rec.data(), n, &nrec);
    if (src != HC_OK && src != HC_ERR_UNSTABLE && src != HC_ERR_HALTED) check(src, "loop_sync");
    const std::string halt_msg = src != HC_OK ? hc_last_error(ctx_) : "";
    const int k = std::min(nrec, n);
    for (int i = 0; i < k; ++i) {
      const hc_loop_record& r = rec[i];
      previous_ = r.previous;
      old_cycle_ = r.old_cycle;
      cycle_ = r.cycle;
      new_cycle_ = r.new_cycle;
      temptot_ = r.temptot;
      for (int m = 0; m <= rfnum_; ++m) j_mean_[m] = r.j_mean[m];
      write_output_timeseries();
      if (tt) tt[i] = r.temptot;
      if (cy) cy[i] = r.cycle;
      if (dtf) dtf[i] = r.dt_flow;
    }
    time_step_ = st.time_step;
    cycle_ = st.cycle;
    counter_ = st.counter;
    maxdepth_ = st.maxdepth;
    temptot_ = st.temptot;
    waterinput_ = st.waterinput;
    waterOut_ = st.waterOut;
    time_1_ = st.time_1;
    baseflow_ = st.baseflow;
    for (int m = 0; m <= rfnum_; ++m) {
      j_mean_[m] = st.j_mean[m];
      old_j_mean_[m] = st.old_j_mean[m];
      new_j_mean_[m] = st.new_j_mean[m];
    }
    *fin = st.finished != 0;
    // a halted loop: the bodies that ran are replayed above, then the cause
    if (src != HC_OK) throw Halted(halt_msg, src, k);
    return k;
  }

  void ensure_device_loop() {
    if (!device_loop_enabled())
      throw std::runtime_error("device loop: model is not hydro-only with device_loop set");
    if (!loop_started_) begin_device_loop();
  }

  void begin_device_loop() {
    if (rfnum_ >= HC_RF_MAX) throw Error("device loop supports at most 16 rain cells");
    hc_loop_init in{};
    in.time_step = time_step_;
    in.cycle = cycle_;
    in.maxdepth = maxdepth_;
    in.waterinput = waterinput_;
    in.waterOut = waterOut_;
    in.time_1 = time_1_;
    in.baseflow = baseflow_;
    in.counter = counter_;
    in.courant_number = courant_number_;
    in.in_out_difference_allowed = in_out_difference_allowed_;
    in.output_file_save_interval = output_file_save_interval_;
    in.ERODEFACTOR = ERODEFACTOR_;
    in.MIN_Q = MIN_Q_;
    in.MIN_Q_MAXVAL = MIN_Q_MAXVAL_;
    in.max_time_step = max_time_step_;
    in.maxcycle = maxcycle_;
    in.rfnum = rfnum_;
    for (int m = 0; m <= rfnum_; ++m) {
      in.j_mean[m] = j_mean_[m];
      in.old_j_mean[m] = old_j_mean_[m];
      in.new_j_mean[m] = new_j_mean_[m];
      in.nActual[m] = nActual_[m];
      in.input_counter[m] = input_counter_[m];
    }
    // TOPMODEL results for every later value of time_1, from the current state
    const int t0 = static_cast<int>(time_1_);
    const int ntop = std::max(t0 + 2, maxcycle_ * 60 + 2 * std::max(1, max_time_step_ / 60) + 64);
    topmodel_table_.assign(static_cast<size_t>(ntop) * HC_RF_MAX, 0.0);
    std::vector<double> j = j_, jo = jo_, oj = old_j_mean_, nj = new_j_mean_;
    for (int m = 0; m <= rfnum_; ++m) topmodel_table_[static_cast<size_t>(t0) * HC_RF_MAX + m] = nj[m];
    for (int t = t0 + 1; t < ntop; ++t) {
      topmodel_step(t, j, jo, oj, nj);
      for (int m = 0; m <= rfnum_; ++m) topmodel_table_[static_cast<size_t>(t) * HC_RF_MAX + m] = nj[m];
    }
    in.topmodel = topmodel_table_.data();
    in.ntop = ntop;
    in.area = area_.data();
    in.rfarea = rfarea_.data();
    std::vector<int> px, py;
    for (int z = 1; z <= totalinputpoints_; ++z) {
      px.push_back(in_x_[z]);
      py.push_back(in_y_[z]);
    }
    in.in_x = px.data();
    in.in_y = py.data();
    in.npoints = totalinputpoints_;
    in.first_x = in_x_[1];
    in.first_y = in_y_[1];
    in.dt_floor = std::min(dt_floor(), static_cast<double>(std::max(0, max_time_step_)));
    check(hc_loop_begin(ctx_, &in), "loop_begin");
    loop_started_ = true;
  }

  double last_dt() const { return last_dt_; }
  double temptot() const { return temptot_; }
  hc_ctx* ctx() { return ctx_; }
  const std::vector<double>& series() const { return series_; }
  int NR() const { return NR_; }
  int NC() const { return NC_; }

  // ------------------------------------------------ checkpoint / restart
  // An exact binary checkpoint: every scalar and array the run loop carries
  // from one iteration to the next (clocks, TOPMODEL store, input points,
  // drainage area, erosion and output accumulators, the series written so
  // far), then the context's device state image (hc_state_save).  The
  // reference restarts only from rasters at 6 significant figures and a
  // Grainz file (LCM.cpp:331-373, 590-683) and does not restore qx, qy,
  // vel_dir or its clocks.  A model created from the same inputs that
  // restarts from the file continues bit for bit.
  void checkpoint(const std::string& path) {
    if (loop_started_)
      throw Error("checkpoint: the device-resident loop holds the run's state on the device; "
                  "checkpoint a host-stepped model");
    // a state with non-finite depths is not saved over a good one (the
    // device check also counts negative depths, which rounding can leave in
    // a sound run: only non-finite ones refuse)
    long long bad = 0;
    if (hc_check_health(ctx_, &bad) != HC_OK) {
      std::vector<double> hv(hc_field_count(ctx_, HC_FIELD_WATER_DEPTH));
      check(hc_download(ctx_, HC_FIELD_WATER_DEPTH, hv.data(), hv.size()), "checkpoint");
      size_t nonfinite = 0;
      for (double v : hv) nonfinite += std::isfinite(v) ? 0 : 1;
      if (nonfinite)
        throw Error("checkpoint: the state has " + std::to_string(nonfinite) +
                    " non-finite water depths; not written");
    }
    size_t nimg = 0;
    check(hc_state_bytes(ctx_, &nimg), "checkpoint");
    std::vector<unsigned char> img(nimg);
    check(hc_state_save(ctx_, img.data(), nimg), "checkpoint");
    // written beside the target and renamed over it once complete and on
    // disk: a write cut short leaves the previous checkpoint as it was
    const std::string tmp = path + ".tmp";
    {
      std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
      if (!f) throw Error("checkpoint: cannot open " + tmp);
      f.write(kCkMagic, sizeof(kCkMagic));
      CkWriter w{f};
      visit_state(w);
      const uint64_t n = nimg;
      f.write(reinterpret_cast<const char*>(&n), sizeof(n));
      f.write(reinterpret_cast<const char*>(img.data()), static_cast<std::streamsize>(nimg));
      f.flush();
      if (!f) {
        std::remove(tmp.c_str());
        throw Error("checkpoint: write to " + tmp + " failed");
      }
    }
    if (!sync_file(tmp)) {
      std::remove(tmp.c_str());
      throw Error("checkpoint: cannot flush " + tmp + " to disk");
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) {
      std::remove(tmp.c_str());
      throw Error("checkpoint: cannot replace " + path);
    }
  }

  static bool sync_file(const std::string& p) {
    const int fd = ::open(p.c_str(), O_RDONLY);
    if (fd < 0) return false;
    const bool ok = ::fsync(fd) == 0;
    ::close(fd);
    return ok;
  }

  void restart(const std::string& path) {
    if (loop_started_)
      throw Error("restart: the device-resident loop has started; restart a freshly created model");
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error("restart: cannot open " + path);
    char magic[sizeof(kCkMagic)];
    f.read(magic, sizeof(magic));
    if (!f || std::memcmp(magic, kCkMagic, sizeof(magic)) != 0)
      throw Error("restart: " + path + " is not a checkpoint of this version");
    // the host state is kept aside: a file that fails its checks leaves the model as it was
    CkReader r{f, path};
    StateCopy saved = take_state();
    try {
      visit_state(r);
      uint64_t n = 0;
      r.raw(&n, sizeof(n));
      size_t nimg = 0;
      check(hc_state_bytes(ctx_, &nimg), "restart");
      if (n != nimg) throw Error("restart: the device state in " + path + " does not fit this model");
      std::vector<unsigned char> img(nimg);
      r.raw(img.data(), nimg);
      if (f.peek() != std::char_traits<char>::eof())
        throw Error("restart: trailing bytes in " + path);
      check(hc_state_load(ctx_, img.data(), nimg), "restart");
    } catch (...) {
      put_state(std::move(saved));
      throw;
    }
    if (complex_rain_) upload_runoff();  // the device planes of the last TOPMODEL update
    points_dirty_ = true;                // the restored list goes to the device again
    if (opt_.write_files) rewrite_series_file();
  }

  bool scalar(const std::string& n, double* v) const {
    if (n == "time_step") *v = time_step_;
    else if (n == "cycle") *v = cycle_;
    else if (n == "counter") *v = counter_;
    else if (n == "maxdepth") *v = maxdepth_;
    else if (n == "temptot") *v = temptot_;
    else if (n == "waterinput") *v = waterinput_;
    else if (n == "waterOut") *v = waterOut_;
    else if (n == "baseflow") *v = baseflow_;
    else if (n == "totalinputpoints") *v = totalinputpoints_;
    else if (n == "imax") *v = imax_;
    else if (n == "jmax") *v = jmax_;
    else if (n == "DX") *v = DX_;
    else if (n == "erode_mult") *v = erode_mult_;
    else if (n == "globalsediq") *v = globalsediq_;
    else return false;
    return true;
  }

 private:
  static constexpr char kCkMagic[8] = {'H', 'C', 'C', 'K', 'P', 'T', '0', '1'};

  struct CkWriter {
    std::ofstream& f;
    void raw(const void* p, size_t n) { f.write(static_cast<const char*>(p), static_cast<std::streamsize>(n)); }
    void same(int v, const char*) { i(v); }
    void i(int& v) { const int64_t x = v; raw(&x, sizeof(x)); }
    void d(double& v) { raw(&v, sizeof(v)); }
    template <class T>
    void vec(std::vector<T>& v, bool) {
      const uint64_t n = v.size();
      raw(&n, sizeof(n));
      if (n) raw(v.data(), n * sizeof(T));
    }
  };

  struct CkReader {
    std::ifstream& f;
    const std::string& path;
    void raw(void* p, size_t n) {
      f.read(static_cast<char*>(p), static_cast<std::streamsize>(n));
      if (!f) throw Error("restart: " + path + " is truncated");
    }
    int64_t get() {
      int64_t x = 0;
      raw(&x, sizeof(x));
      return x;
    }
    void same(int v, const char* what) {
      if (get() != v) throw Error(std::string("restart: the checkpoint's ") + what + " differs from this model's");
    }
    void i(int& v) {
      const int64_t x = get();
      if (x < INT_MIN || x > INT_MAX) throw Error("restart: corrupt checkpoint " + path);
      v = static_cast<int>(x);
    }
    void d(double& v) { raw(&v, sizeof(v)); }
    template <class T>
    void vec(std::vector<T>& v, bool fixed) {
      uint64_t n = 0;
      raw(&n, sizeof(n));
      if (fixed && n != v.size())
        throw Error("restart: an array of the checkpoint does not match this model's grid");
      if (n > (uint64_t{1} << 34) / sizeof(T)) throw Error("restart: corrupt checkpoint " + path);
      v.resize(n);
      if (n) raw(v.data(), n * sizeof(T));
    }
  };

  // Everything the run loop carries between iterations, in file order.  The
  // inputs (parameters, rain series, grid) are the model's own; the first
  // entries check that they match the run that wrote the file.
  template <class V>
  void visit_state(V& v) {
    v.same(imax_, "row count");
    v.same(jmax_, "column count");
    v.same(rfnum_, "rain cell count");
    v.same(hydro_only_ ? 1 : 0, "erosion mode");
    v.same(opt_.precision, "precision");
    v.same(complex_rain_ ? 1 : 0, "rainfall mode");
    v.i(totalinputpoints_);
    v.i(counter_);
    v.i(hours_);
    double* ds[] = {&time_step_, &cycle_, &maxdepth_, &temptot_, &waterinput_, &waterOut_, &iod_,
                    &time_1_, &save_time_, &creep_time_, &creep_time2_, &previous_, &new_cycle_,
                    &old_cycle_, &erode_mult_, &erode_call_, &globalsediq_, &sediQ_, &last_dt_,
                    &baseflow_, &tx_, &tlastcalc_, &Qs_step_, &Qs_hour_, &Qs_over_, &Qs_last_,
                    &Qw_newvol_, &Qw_oldvol_, &Qw_lastvol_, &Qw_stepvol_, &Qw_hourvol_, &Qw_hour_,
                    &Qw_overvol_, &temptotal_, &old_sediq_, &Jw_newvol_, &Jw_oldvol_, &Jw_lastvol_,
                    &Jw_stepvol_, &Jw_hourvol_, &Jw_hour_, &Jw_overvol_, &sand_out_};
    for (double* d : ds) v.d(*d);
    v.vec(elev_, true);
    v.vec(area_, true);
    v.vec(area_depth_, true);
    v.vec(rfarea_, true);
    v.vec(j_, false);
    v.vec(jo_, false);
    v.vec(j_mean_, false);
    v.vec(old_j_mean_, false);
    v.vec(new_j_mean_, false);
    v.vec(nActual_, false);
    v.vec(input_counter_, false);
    v.vec(in_x_, false);
    v.vec(in_y_, false);
    v.vec(pts_x_, false);
    v.vec(pts_y_, false);
    v.vec(pts_a_, false);
    v.vec(sum_grain_, false);
    v.vec(old_sum_grain_, false);
    v.vec(Qg_step_, false);
    v.vec(Qg_hour_, false);
    v.vec(Qg_over_, false);
    v.vec(Qg_last_, false);
    v.vec(series_, false);
    v.d(runoff_t_);
    v.i(runoff_jm_set_);
    v.vec(rj_, false);
    v.vec(rjo_, false);
    v.vec(rj_mean_, false);
    v.vec(rold_, false);
    v.vec(rnew_, false);
  }

  // the host state visit_state covers, kept aside while a restart reads
  struct StateCopy {
    std::vector<unsigned char> bytes;
  };
  struct CkMem {
    std::vector<unsigned char>& b;
    size_t pos = 0;
    bool out = true;
    void raw(void* p, size_t n) {
      if (out) {
        const unsigned char* q = static_cast<const unsigned char*>(p);
        b.insert(b.end(), q, q + n);
      } else {
        std::memcpy(p, b.data() + pos, n);
        pos += n;
      }
    }
    void same(int, const char*) {}
    void i(int& v) { raw(&v, sizeof(v)); }
    void d(double& v) { raw(&v, sizeof(v)); }
    template <class T>
    void vec(std::vector<T>& v, bool) {
      uint64_t n = v.size();
      raw(&n, sizeof(n));
      if (!out) v.resize(n);
      if (n) raw(v.data(), n * sizeof(T));
    }
  };
  StateCopy take_state() {
    StateCopy s;
    CkMem m{s.bytes};
    visit_state(m);
    return s;
  }
  void put_state(StateCopy s) {
    CkMem m{s.bytes, 0, false};
    visit_state(m);
  }

  size_t at(int x, int y) const { return static_cast<size_t>(x) * NC_ + y; }
  void check(int rc, const char* what) {
    if (rc != HC_OK) throw Error(std::string(what) + ": " + hc_last_error(ctx_));
  }
  void upload(int field, const void* host) {
    check(hc_upload(ctx_, field, host, hc_field_count(ctx_, field)), "upload");
  }


  // ---------------------------------------------------------- parameters
  // LSDCatchmentModel::initialise_variables (LCM.cpp:842-1610) for the keys
  // of the hot path and its inputs; options this build does not provide are
  // rejected instead of silently ignored.
  void initialise_variables(std::istream& in) {
    std::string key, value;
    auto yes = [&]() { return value == "yes"; };
    while (in.good()) {
      if (!parse_line(in, key, value)) break;
      if (key == "NULL") continue;
      std::string k = key;
      for (auto& ch : k) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
      value = strip_ctrl(value);
      const char* v = value.c_str();
      if (k == "dem_read_extension") dem_read_extension_ = value;
      else if (k == "dem_write_extension") dem_write_extension_ = value;
      else if (k == "write_path") write_path_ = value;
      else if (k == "write_fname") write_fname_ = value;
      else if (k == "read_path") read_path_ = value;
      else if (k == "read_fname") read_fname_ = value;
      else if (k == "hydroindex_file") hydroindex_fname_ = value;
      else if (k == "rainfall_data_file") rainfall_data_file_ = value;
      else if (k == "reach_mode") reach_mode_ = yes();
      else if (k == "divide_inputs_by") div_inputs_ = static_cast<int>(std::atof(value.c_str()));
      else if (k == "reach_input_data_timestep") reach_rdt_ = std::atoi(value.c_str());
      else if (k.size() > 6 && k.compare(0, 5, "reach") == 0 && k[5] >= '1' && k[5] <= '3' &&
               (k.compare(6, std::string::npos, "_input_on") == 0 ||
                k.compare(6, std::string::npos, "_input_file") == 0 ||
                k.compare(6, std::string::npos, "_x") == 0 ||
                k.compare(6, std::string::npos, "_y") == 0)) {
        const int r = k[5] - '1';
        const std::string f = k.substr(6);
        if (f == "_input_on") reach_on_[r] = yes();
        else if (f == "_input_file") reach_file_[r] = value;
        else if (f == "_x") reach_px_[r] = std::atoi(value.c_str());
        else reach_py_[r] = std::atoi(value.c_str());
      }
      else if (k == "grain_data_file") grain_data_file_ = value;
      else if (k == "bedrock_data_file") bedrock_data_file_ = value;
      else if (k == "water_init_raster_file") water_init_raster_file_ = value;
      else if (k == "spatial_mannings_dem_file") mannings_fname_ = value;
      else if (k == "max_time_step") max_time_step_ = std::atoi(v);
      else if (k == "max_run_duration") maxcycle_ = std::atoi(v);
      else if (k == "timeseries_save_interval") output_file_save_interval_ = std::atoi(v);
      else if (k == "raster_output_interval") saveinterval_ = std::atof(v);
      else if (k == "write_elevation_file") elev_fname_ = value;
      else if (k == "write_elev_file") write_elev_file_ = yes();
      else if (k == "waterdepth_outfile_name") waterdepth_fname_ = value;
      else if (k == "write_waterdepth_file") write_waterd_file_ = yes();
      else if (k == "grainsize_file") grainsize_fname_ = value;
      else if (k == "write_grainsize_file") write_grainsz_file_ = yes();
      else if (k == "read_in_graindata_from_file") graindata_from_file_ = yes();
      else if (k == "bedrock_layer_on") bedrock_layer_on_ = yes();
      else if (k == "transport_law") {
        if (value == "wilcock") { wilcock_ = true; einstein_ = false; }
        else if (value == "einstein") { einstein_ = true; wilcock_ = false; }
        else throw Error("transport_law must be wilcock or einstein");  // LCM.cpp:1134-1142
      }
      else if (k == "max_tau_velocity") max_vel_ = std::atof(v);
      else if (k == "active_layer_thickness") active_ = std::atof(v);
      else if (k == "slope_failure_thresh") failureangle_ = std::atof(v);
      else if (k == "creep_rate") creep_rate_ = std::atof(v);
      else if (k == "chann_lateral_erosion") chann_lateral_erosion_ = std::atof(v);
      else if (k == "erosion_limit") ERODEFACTOR_ = std::atof(v);
      else if (k == "lateral_erosion_const") lateral_constant_ = std::atof(v);
      else if (k == "bedrock_erosion_threshold") bedrock_erosion_threshold_ = std::atof(v);
      else if (k == "hydro_model_only") hydro_only_ = yes();
      else if (k == "rainfall_data_on") rainfall_data_on_ = yes();
      else if (k == "water_init_from_raster_on") water_init_from_raster_ = yes();
      else if (k == "topmodel_m_value") M_ = std::atof(v);
      else if (k == "num_unique_rain_cells") rfnum_ = std::atoi(v);
      else if (k == "rain_data_time_step") rain_data_time_step_ = std::atof(v);
      else if (k == "spatial_var_rain") spatially_var_rainfall_ = yes();
      else if (k == "in_out_difference") in_out_difference_allowed_ = std::atof(v);
      else if (k == "min_q_for_depth_calc") MIN_Q_ = std::atof(v);
      else if (k == "max_q_for_depth_calc") MIN_Q_MAXVAL_ = std::atof(v);
      else if (k == "hflow_threshold") hflow_threshold_ = std::atof(v);
      // the reference's own quirk: this key sets hflow_threshold (LCM.cpp:1318-1323)
      else if (k == "water_depth_erosion_threshold") hflow_threshold_ = std::atof(v);
      else if (k == "slope_on_edge_cell") edgeslope_ = std::atof(v);
      else if (k == "courant_number") courant_number_ = std::atof(v);
      else if (k == "froude_num_limit") froude_limit_ = std::atof(v);
      else if (k == "mannings_n") mannings_ = std::atof(v);
      else if (k == "spatially_variable_mannings_on") spatially_var_mannings_ = yes();
      else if (k == "veg_erosion_prop") veg_lat_restriction_ = std::atof(v);
      else if (k == "debug_print_cycle") debug_print_cycle_ = yes();
      else if (k == "hc_grain_sizes") {  // extension: d1..d9, comma separated
        std::stringstream ss(value);
        std::string tok;
        int idx = 1;
        while (std::getline(ss, tok, ',') && idx <= 9) d_[idx++] = std::atof(tok.c_str());
      }
      else if (k == "spatially_complex_rainfall_on") complex_rain_ = yes();
      else if ((k == "vegetation_on" || k == "groundwater_on" ||
                k == "lateral_erosion_on") && yes())
        throw Error("option '" + k + "' is not provided by this build (hot path only)");
    }
    tx_ = output_file_save_interval_;                // LCM.cpp:1599
    if (!spatially_var_rainfall_) rfnum_ = 1;        // LCM.cpp:1602-1608
  }

  void check_DEM_edge_condition() {  // LCM.cpp:2845-2889
    double temp = -9999;
    for (int n = 1; n <= jmax_; ++n) {
      if (elev_[at(imax_, n)] > -9999) temp = elev_[at(imax_, n)];
      if (elev_[at(1, n)] > -9999) temp = elev_[at(1, n)];
    }
    for (int n = 1; n <= imax_; ++n) {
      if (elev_[at(n, 1)] > -9999) temp = elev_[at(n, 1)];
      if (elev_[at(n, jmax_)] > -9999) temp = elev_[at(n, jmax_)];
    }
    if (temp < -10) throw Error("DEM EDGE CONDITION ERROR: all DEM edges are nodata");
  }

  void count_catchment_gridcells() {  // LCM.cpp:2813-2843
    for (int i = 1; i < imax_; ++i)
      for (int j = 1; j < jmax_; ++j)
        if (elev_[at(i, j)] > -9999) nActual_[rfarea_[at(i, j)]]++;
  }

  // the reach points and files of initialise_arrays (LCM.cpp:1766-1845);
  // read_reachfile (LCM.cpp:718-760) keeps each line's values as float
  void load_reach_files() {
    reach_n_ = 0;
    if (!reach_mode_) return;
    for (int r = 0; r < 3; ++r) {
      if (!reach_on_[r]) continue;
      const std::string f = read_path_ + "/" + reach_file_[r];
      std::ifstream in(f);
      if (!in) throw Error("No reach input data file found by name of: " + f);
      reach_rows_[reach_n_].clear();
      std::string line;
      while (std::getline(in, line)) {
        std::stringstream ss(line);
        std::vector<float> row;
        float val;
        while (ss >> val) row.push_back(val);
        reach_rows_[reach_n_].push_back(row);
      }
      reach_x_[reach_n_] = reach_px_[r];
      reach_y_[reach_n_] = reach_py_[r];
      ++reach_n_;
    }
  }

  void read_rainfall(const std::string& f) {  // LCM.cpp:694-717 (values parsed as float)
    std::ifstream in(f);
    if (!in) throw Error("No rainfall data file found by name of: " + f);
    rain_.clear();
    std::string line;
    while (std::getline(in, line)) {
      std::stringstream ss(line);
      std::vector<float> row;
      float val;
      while (ss >> val) row.push_back(val);
      rain_.push_back(row);
    }
  }

  void ingest_graindata(const std::string& f) {  // LCM.cpp:611-683
    std::ifstream in(f);
    if (!in) throw Error("No grain data file found by name of: " + f);
    const size_t n = elev_.size();
    std::vector<double> grain(n * 11, 0.0), strata(n * 90, 0.0);
    // index[x1][y1] = grain_array_tot, one number per line (LCM.cpp:638-656)
    std::vector<int32_t> index(n, -9999);
    long long tot = 0;
    std::string line;
    while (std::getline(in, line)) {
      std::stringstream ss(line);
      std::vector<double> cols;
      double val;
      while (ss >> val) cols.push_back(val);
      if (cols.size() < 3) continue;
      ++tot;
      int x1 = static_cast<int>(cols[0]), y1 = static_cast<int>(cols[1]);
      if (x1 > imax_) x1 = imax_;
      if (y1 > jmax_) y1 = jmax_;
      if (x1 < 0 || y1 < 0) throw Error("grain data file: negative cell coordinate");
      const size_t c = at(x1, y1);
      index[c] = static_cast<int32_t>(tot);
      for (int k = 0; k <= kGMax && 3 + k < static_cast<int>(cols.size()); ++k) grain[c * 11 + k] = cols[3 + k];
      for (int z = 0; z < 10; ++z)
        for (int k = 0; k < 9; ++k) {
          const size_t col = 14 + z * 9 + k;
          if (col < cols.size()) strata[c * 90 + z * 9 + k] = cols[col];
        }
    }
    upload(HC_FIELD_GRAIN, grain.data());
    upload(HC_FIELD_STRATA, strata.data());
    check(hc_grain_index_set(ctx_, index.data(), n, tot), "grain index");
  }

  // the grain data file's values given in memory, one row per cell
  void ingest_graindata_arrays() {
    if (!arrays_->graindata) throw Error("in-memory grain data missing");
    const size_t n = elev_.size();
    std::vector<double> grain(n * 11, 0.0), strata(n * 90, 0.0);
    std::vector<int32_t> has(n, 0);
    for (int i = 0; i < imax_; ++i)
      for (int j = 0; j < jmax_; ++j) {
        const double* v = arrays_->graindata + (static_cast<size_t>(i) * jmax_ + j) * 101;
        const size_t c = at(i + 1, j + 1);
        has[c] = 1;
        for (int k = 0; k <= kGMax; ++k) grain[c * 11 + k] = v[k];
        for (int k = 0; k < 90; ++k) strata[c * 90 + k] = v[11 + k];
      }
    upload(HC_FIELD_GRAIN, grain.data());
    upload(HC_FIELD_STRATA, strata.data());
    upload(HC_FIELD_HASGRAIN, has.data());
  }

  // ------------------------------------------------ rainfall-runoff (host)
  float rain_at(int row, int col) const {
    if (row < 0 || row >= static_cast<int>(rain_.size())) return 0.0f;
    if (col < 0 || col >= static_cast<int>(rain_[row].size())) return 0.0f;
    return rain_[row][col];
  }
  void topmodel_runoff(double cyc) {
    topmodel_step(cyc, j_, jo_, old_j_mean_, new_j_mean_);
  }
  // LCM.cpp:3649-3715 on explicit state (the device loop tabulates the
  // sequence ahead of time with this same function)
  void topmodel_step(double cyc, std::vector<double>& j, std::vector<double>& jo,
                     std::vector<double>& old_jm, std::vector<double>& new_jm) const {
    const int ts = static_cast<int>(cyc / rain_data_time_step_);
    for (int n = 1; n <= rfnum_; ++n) {
      double local = 0;
      const double ft = 60;
      old_jm[n] = new_jm[n];
      jo[n] = j[n];
      const float r = rain_at(ts, n - 1);
      if (r > 0) local = rain_factor_ * ((r / 1000) / 3600);
      if (local == 0) {
        j[n] = jo[n] / (1 + ((jo[n] * ft) / M_));
        new_jm[n] = M_ / ft * std::log(1 + ((jo[n] * ft) / M_));
      }
      if (local > 0) {
        j[n] = local / (((local - jo[n]) / jo[n]) * std::exp((0 - local) * ft / M_) + 1);
        new_jm[n] = (M_ / ft) * std::log(((local - jo[n]) + jo[n] * std::exp((local * ft) / M_)) / local);
      }
      if (new_jm[n] < 0) new_jm[n] = 0;
    }
  }
  void calchydrograph(double t) {  // LCM.cpp:3717-3724
    for (int n = 1; n <= rfnum_; ++n)
      j_mean_[n] = old_j_mean_[n] + (((new_j_mean_[n] - old_j_mean_[n]) / 2) * (2 - t));
  }
  void catchment_water_input_and_hydrology(double fts) {  // LCM.cpp:3455-3541
    for (int i = 1; i <= rfnum_; ++i) waterinput_ += j_mean_[i] * nActual_[i] * DX_ * DX_;
    // the point list lives on the device and is uploaded when it changes; a
    // point's amount depends only on its rain cell, so each step sends one
    // amount per rain cell (the expression and cap of LCM.cpp:3466-3474)
    if (points_dirty_) {
      pts_x_.assign(in_x_.begin() + 1, in_x_.begin() + 1 + totalinputpoints_);
      pts_y_.assign(in_y_.begin() + 1, in_y_.begin() + 1 + totalinputpoints_);
      pts_rf_.resize(totalinputpoints_);
      for (int z = 0; z < totalinputpoints_; ++z) pts_rf_[z] = rfarea_[at(pts_x_[z], pts_y_[z])];
      check(hc_points_set(ctx_, pts_x_.data(), pts_y_.data(), pts_rf_.data(),
                          static_cast<int>(pts_x_.size())),
            "points_set");
      points_dirty_ = false;
    }
    pts_a_.assign(rfnum_ + 1, 0.0);
    for (int rf = 0; rf <= rfnum_; ++rf) {
      double amt = (j_mean_[rf] * nActual_[rf]) / (input_counter_[rf]) * fts;
      if (amt > ERODEFACTOR_) amt = ERODEFACTOR_;
      pts_a_[rf] = amt;
    }
    check(hc_points_add(ctx_, pts_a_.data(), rfnum_ + 1), "points_add");
    if (cycle_ >= time_1_) {
      do {
        time_1_++;
        topmodel_runoff(time_1_);
        if (time_step_ > max_time_step_) {
          double jmt = 0;
          for (int n = 1; n <= rfnum_; ++n)
            if (new_j_mean_[n] > jmt) jmt = new_j_mean_[n];
          const unsigned cells = static_cast<unsigned>(imax_) * static_cast<unsigned>(jmax_);
          if (jmt > (0.2 / (cells * DX_ * DX_))) {
            cycle_ = time_1_ + (max_time_step_ / 60);
            time_step_ = max_time_step_;
          }
        }
      } while (time_1_ < cycle_);
    }
    calchydrograph(time_1_ - cycle_);
    double jmeanmax = 0;
    for (int n = 1; n <= rfnum_; ++n)
      if (j_mean_[n] > jmeanmax) jmeanmax = j_mean_[n];
    if (jmeanmax >= baseflow_) {
      baseflow_ = baseflow_ * 3;
      zero_and_calc_drainage_area();
      get_catchment_input_points();
    }
    if (baseflow_ > (jmeanmax * 3) && baseflow_ > 0.000001) {
      baseflow_ = jmeanmax * 1.25;
      zero_and_calc_drainage_area();
      get_catchment_input_points();
    }
  }
  // ------------------------------------- spatially complex rainfall ----
  // topmodel_runoff(cycle, runoff) (LCM.cpp:3610-3627): the rain grid of
  // rainGrid::create (LSDRainfallRunoff.cpp:36-63) -- one rain cell: its
  // value everywhere; otherwise cells 1..imax-1, 1..jmax-1 whose hydroindex
  // is 1..rfnum-1 take that column (the last index takes none), all others 0
  // -- held in double, then runoffGrid::calculate_runoff (LSDRainfallRunoff.
  // cpp:175-241) on cells 1..imax-1, 1..jmax-1 above -9999, rain_factor
  // passed as an int, a 60 s step.  Erosion never takes a catchment cell to
  // -9999, so the DEM as loaded gives the same cell set.
  double rain_grid(int ts, size_t c) const {
    if (rfnum_ == 1) return static_cast<double>(rain_at(ts, 0));
    const int rf = rfarea_[c];
    if (rf >= 1 && rf < rfnum_) return static_cast<double>(rain_at(ts, rf - 1));
    return 0.0;
  }
  void topmodel_runoff_grid(double cyc) {
    const int ts = static_cast<int>(cyc / rain_data_time_step_);
    const int rfi = static_cast<int>(rain_factor_);
    for (int m = 1; m < imax_; ++m)
      for (int n = 1; n < jmax_; ++n) {
        const size_t c = at(m, n);
        if (!(elev_[c] > -9999)) continue;
        double local = 0;
        const double lts = 60;
        rold_[c] = rnew_[c];
        rjo_[c] = rj_[c];
        const double r = rain_grid(ts, c);
        if (r > 0) local = rfi * ((r / 1000) / 3600);
        if (local == 0) {
          rj_[c] = rjo_[c] / (1 + ((rjo_[c] * lts) / M_));
          rnew_[c] = M_ / lts * std::log(1 + ((rjo_[c] * lts) / M_));
        }
        if (local > 0) {
          rj_[c] = local / (((local - rjo_[c]) / rjo_[c]) * std::exp((0 - local) * lts / M_) + 1);
          rnew_[c] = (M_ / lts) * std::log(((local - rjo_[c]) + rjo_[c] * std::exp((local * lts) / M_)) / local);
        }
        if (rnew_[c] < 0) rnew_[c] = 0;
      }
  }
  void upload_runoff() { check(hc_runoff_set(ctx_, rold_.data(), rnew_.data()), "runoff_set"); }
  // catchment_water_input_and_hydrology(flow_timestep, runoff)
  // (LCM.cpp:3544-3606).  The water goes in on the device (hc_runoff_add)
  // from the planes of the last TOPMODEL update and the previous call's
  // hydrograph time, which is the j_mean this call's sums read; before the
  // first hydrograph the grid's j_mean is 0 and nothing goes in.
  void catchment_water_input_grid(double fts) {
    if (runoff_jm_set_) check(hc_runoff_add(ctx_, runoff_t_, fts, ERODEFACTOR_), "runoff_add");
    for (int i = 1; i < imax_; ++i)
      for (int j = 1; j < jmax_; ++j) waterinput_ += rj_mean_[at(i, j)] * DX_ * DX_;
    for (int i = 1; i < imax_; ++i)
      for (int j = 1; j < jmax_; ++j) {
        double amt = rj_mean_[at(i, j)] * fts;
        if (amt > ERODEFACTOR_) amt = ERODEFACTOR_;
        waterinput_ += (amt / fts) * DX_ * DX_;
      }
    double njmax = 0;
    for (int m = 1; m <= imax_; ++m)
      for (int n = 1; n <= jmax_; ++n)
        if (rnew_[at(m, n)] > njmax) njmax = rnew_[at(m, n)];
    bool updated = false;
    if (cycle_ >= time_1_) {
      const unsigned cells = static_cast<unsigned>(jmax_) * static_cast<unsigned>(imax_);
      do {
        time_1_++;
        topmodel_runoff_grid(time_1_);
        updated = true;
        if (time_step_ > max_time_step_ && njmax > (0.2 / (cells * DX_ * DX_))) {
          cycle_ = time_1_ + (max_time_step_ / 60);
          time_step_ = max_time_step_;
        }
      } while (time_1_ < cycle_);
    }
    runoff_t_ = time_1_ - cycle_;  // calchydrograph (LCM.cpp:3728-3739)
    for (int m = 1; m <= imax_; ++m)
      for (int n = 1; n <= jmax_; ++n) {
        const size_t c = at(m, n);
        rj_mean_[c] = rold_[c] + (((rnew_[c] - rold_[c]) / 2) * (2 - runoff_t_));
      }
    runoff_jm_set_ = 1;
    if (updated) upload_runoff();
  }
  // output_data(temptotal, runoff)'s runoff volume (LCM.cpp:2487-2501,
  // 2571-2583): each catchment cell's term added in turn
  void add_grid_runoff_volume(double& vol, double minutes) const {
    for (int i = 1; i <= imax_; ++i)
      for (int j = 1; j <= jmax_; ++j)
        if (elev_[at(i, j)] > no_data_value_) vol += (rj_mean_[at(i, j)] * DX_ * DX_) * (minutes * 60);
  }

  void get_catchment_input_points() {  // LCM.cpp:3741-3767
    totalinputpoints_ = 0;
    for (int n = 1; n <= rfnum_; ++n) input_counter_[n] = 0;
    for (int i = 1; i <= imax_; ++i)
      for (int j = 1; j <= jmax_; ++j) {
        const double v = area_[at(i, j)] * baseflow_ * 3 * DX_ * DX_;
        if (v > MIN_Q_ && v < MIN_Q_MAXVAL_) {
          ++totalinputpoints_;
          in_x_[totalinputpoints_] = i;
          in_y_[totalinputpoints_] = j;
          input_counter_[rfarea_[at(i, j)]]++;
        }
      }
    if (totalinputpoints_ == 0) totalinputpoints_ = 1;
    points_dirty_ = true;
  }
  // LCM.cpp:2071-2270.  The elevation sort is std::sort over
  // (imax+2)(jmax+2) pairs exactly as the reference builds them, so ties
  // keep the reference's order.
  void zero_and_calc_drainage_area() {
    if (!hydro_only_) refresh_elev();
    for (int i = 1; i <= imax_; ++i)
      for (int j = 1; j <= jmax_; ++j) {
        area_depth_[at(i, j)] = 1;
        area_[at(i, j)] = 0;
        if (elev_[at(i, j)] == -9999) area_depth_[at(i, j)] = 0.0;
      }
    std::vector<std::pair<size_t, double>> pr(elev_.size(), std::make_pair(size_t(0), 0.0));
    size_t inc = 1;
    for (int i = 1; i <= imax_; ++i)
      for (int j = 1; j <= jmax_; ++j) pr[inc++] = std::make_pair(at(i, j), elev_[at(i, j)]);
    std::sort(pr.begin(), pr.end(), [](const std::pair<size_t, double>& a, const std::pair<size_t, double>& b) {
      return a.second < b.second;
    });
    const size_t ncell = static_cast<size_t>(imax_) * jmax_;
    for (size_t k = ncell; k >= 1; --k) {
      const size_t c = pr[k].first;
      const int i = static_cast<int>(c / NC_), j = static_cast<int>(c % NC_);
      if (!(area_depth_[c] > 0)) continue;
      if (area_depth_[c] > area_[c]) area_[c] = area_depth_[c];
      double difftot = 0;
      for (int dir = 1; dir <= 8; ++dir) {
        int j2 = std::min(std::max(j + kDY8[dir], 1), jmax_);
        int i2 = std::min(std::max(i + kDX8[dir], 1), imax_);
        const double e2 = elev_[at(i2, j2)], e = elev_[c];
        if (e2 < e) difftot += (dir % 2 != 0) ? (e - e2) : (e - e2) / 1.414;
      }
      if (difftot > 0) {
        for (int dir = 1; dir <= 8; ++dir) {
          int i2 = std::min(std::max(i + kDX8[dir], 1), imax_);
          int j2 = std::min(std::max(j + kDY8[dir], 1), jmax_);
          const size_t c2 = at(i2, j2);
          const double e2 = elev_[c2], e = elev_[c];
          if (e2 < e) {
            if (dir % 2 != 0) area_depth_[c2] += area_depth_[c] * ((e - e2) / difftot);
            else area_depth_[c2] += area_depth_[c] * (((e - e2) / 1.414) / difftot);
          }
        }
      }
      area_depth_[c] = 0;
    }
  }
  void refresh_elev() {
    check(hc_download(ctx_, HC_FIELD_ELEV, elev_.data(), elev_.size()), "download elev");
  }

  // --------------------------------------------------- time series output
  void output_data() {  // LCM.cpp:2310-2478
    Qw_newvol_ += temptotal_ * ((cycle_ - previous_) * 60);
    if (complex_rain_) {
      add_grid_runoff_volume(Jw_newvol_, cycle_ - previous_);
    } else {
      for (int nn = 1; nn <= rfnum_; ++nn)
        Jw_newvol_ += (j_mean_[nn] * DX_ * DX_ * nActual_[nn]) * ((cycle_ - previous_) * 60);
    }
    const double osi = output_file_save_interval_;
    if ((new_cycle_ < old_cycle_) || (cycle_ - previous_ >= osi)) {
      while ((tx_ > previous_) && (cycle_ >= tx_)) {
        hours_++;
        Qs_step_ = globalsediq_ - old_sediq_;
        Qs_over_ = Qs_step_ * ((cycle_ - tx_) / (cycle_ - tlastcalc_));
        Qs_hour_ = Qs_step_ - Qs_over_ + Qs_last_;
        if (cycle_ >= tx_ + osi) { Qs_last_ = 0; old_sediq_ = globalsediq_ - Qs_over_; }
        if (cycle_ < tx_ + osi) { Qs_last_ = Qs_over_; old_sediq_ = globalsediq_; }
        for (int n = 1; n <= kGMax - 1; ++n) {
          Qg_step_[n] = sum_grain_[n] - old_sum_grain_[n];
          Qg_over_[n] = Qg_step_[n] * ((cycle_ - tx_) / (cycle_ - tlastcalc_));
          Qg_hour_[n] = Qg_step_[n] - Qg_over_[n] + Qg_last_[n];
          if (cycle_ >= tx_ + osi) { Qg_last_[n] = 0; old_sum_grain_[n] = sum_grain_[n] - Qg_over_[n]; }
          if (cycle_ < tx_ + osi) { Qg_last_[n] = Qg_over_[n]; old_sum_grain_[n] = sum_grain_[n]; }
        }
        Qw_overvol_ = temptotal_ * ((cycle_ - tx_) * 60);
        Qw_stepvol_ = Qw_newvol_ - Qw_oldvol_;
        Qw_hourvol_ = Qw_stepvol_ - Qw_overvol_ + Qw_lastvol_;
        Qw_hour_ = Qw_hourvol_ / (60 * osi);
        if (complex_rain_) {
          add_grid_runoff_volume(Jw_overvol_, cycle_ - tx_);
        } else {
          for (int nn = 1; nn <= rfnum_; ++nn)
            Jw_overvol_ += (j_mean_[nn] * DX_ * DX_ * nActual_[nn]) * ((cycle_ - tx_) * 60);
        }
        Jw_stepvol_ = Jw_newvol_ - Jw_oldvol_;
        Jw_hourvol_ = Jw_stepvol_ - Jw_overvol_ + Jw_lastvol_;
        Jw_hour_ = Jw_hourvol_ / (60 * osi);
        if (cycle_ >= tx_ + osi) {
          Qw_lastvol_ = 0; Qw_oldvol_ = Qw_newvol_ - Qw_overvol_;
          Jw_lastvol_ = 0; Jw_oldvol_ = Jw_newvol_ - Jw_overvol_;
        }
        if (cycle_ < tx_ + osi) {
          Qw_lastvol_ = Qw_overvol_; Qw_oldvol_ = Qw_newvol_;
          Jw_lastvol_ = Jw_overvol_; Jw_oldvol_ = Jw_newvol_;
        }
        const double Tx = tx_;
        tx_ = Tx + osi;
        const double row[5] = {static_cast<double>(hours_), Qw_hour_, Jw_hour_, sand_out_, Qs_hour_};
        series_.insert(series_.end(), row, row + 5);
        for (int n = 1; n <= kGMax - 1; ++n) series_.push_back(Qg_hour_[n]);
        if (opt_.write_files) append_series_row();
        sand_out_ = 0;
      }
      tlastcalc_ = cycle_;
    }
  }
  // one row of the time series in the column formats of LCM.cpp:2427-2474
  static std::string series_line(const double* r) {
    std::ostringstream os;
    os << std::fixed << std::setprecision(0) << r[0];
    os << " " << std::fixed << std::setprecision(6) << r[1];
    os << " " << std::fixed << std::setprecision(6) << r[2];
    os << " " << std::fixed << std::setprecision(6) << r[3];
    os << " " << std::fixed << std::setprecision(10) << r[4];
    for (int n = 0; n < 9; ++n) os << " " << std::fixed << std::setprecision(10) << r[5 + n];
    return os.str();
  }
  void append_series_row() {
    std::ofstream f(write_path_ + "/" + write_fname_, std::ios_base::app | std::ios_base::out);
    f << series_line(&series_[series_.size() - 14]) << std::endl;
  }
  // after a restart: the file holds exactly the rows of the checkpoint (a
  // killed run may have appended rows after its last checkpoint)
  void rewrite_series_file() {
    std::ofstream f(write_path_ + "/" + write_fname_, std::ios_base::trunc | std::ios_base::out);
    if (!f) throw Error("restart: unable to write to " + write_path_ + "/" + write_fname_);
    for (size_t k = 0; k + 14 <= series_.size(); k += 14) f << series_line(&series_[k]) << std::endl;
  }

  // ASCII raster without the padding ring, 6 significant digits
  // (LSDRaster::write_double_asc_raster, src/topotools/LSDRaster.cpp:1194-1221).
  void write_asc(const std::string& fname, const std::vector<double>& a) const {
    std::ofstream out(fname);
    if (!out) throw Error("unable to write to " + fname);
    out << "ncols\t" << jmax_ << "\nnrows\t" << imax_ << "\nxllcorner\t" << std::setprecision(14)
        << xll_ << "\nyllcorner\t" << std::setprecision(14) << yll_ << "\ncellsize\t" << DX_
        << "\nNODATA_value\t" << no_data_value_ << std::endl;
    for (int i = 1; i <= imax_; ++i) {
      for (int j = 1; j <= jmax_; ++j) out << std::setprecision(6) << a[at(i, j)] << " ";
      if (i != imax_) out << std::endl;
    }
  }
  void save_raster_data(double tempcycle) {  // LCM.cpp:2672-2808
    if (!opt_.write_files) return;
    const std::string suffix = std::to_string(static_cast<int>(tempcycle)) + "." + dem_write_extension_;
    if (write_waterd_file_) {
      std::vector<double> h(elev_.size());
      check(hc_download(ctx_, HC_FIELD_WATER_DEPTH, h.data(), h.size()), "download h");
      write_asc(write_path_ + "/" + waterdepth_fname_ + suffix, h);
    }
    if (write_elev_file_) {
      std::vector<double> e(elev_.size());
      check(hc_download(ctx_, HC_FIELD_ELEV, e.data(), e.size()), "download elev");
      write_asc(write_path_ + "/" + elev_fname_ + suffix, e);
    }
    if (write_grainsz_file_) write_grainz(write_path_ + "/" + grainsize_fname_ + suffix);
  }
  // The Grainz file (LSDGrainMatrix::write_grainMatrix_to_ascii_file,
  // LGM.cpp:59-118, called from LCM.cpp:2711-2723): for every cell with a
  // grain row in row-major order, "x y index", the 11 surface fractions and
  // the 10 x 9 strata, each value followed by a space, at the stream's
  // default 6 significant digits.  index is the reference's row number
  // (hc_grain_index_get).  A hydro-only run has no rows: an empty file.
  void write_grainz(const std::string& fname) {
    if (dem_write_extension_ != "asc")
      throw Error("Grainz output: the only extension is asc; dem_write_extension is " +
                  dem_write_extension_);
    std::ofstream out(fname);
    if (!out) throw Error("unable to write to " + fname);
    if (hydro_only_) return;
    const size_t n = elev_.size();
    std::vector<int32_t> idx(n);
    check(hc_grain_index_get(ctx_, idx.data(), n, nullptr), "grain index");
    std::vector<double> g(n * 11), st(n * 90);
    check(hc_download(ctx_, HC_FIELD_GRAIN, g.data(), g.size()), "download grain");
    check(hc_download(ctx_, HC_FIELD_STRATA, st.data(), st.size()), "download strata");
    for (int i = 1; i <= imax_; ++i)
      for (int j = 1; j <= jmax_; ++j) {
        const size_t c = at(i, j);
        if (idx[c] == -9999) continue;
        out << i << " " << j << " " << idx[c] << " ";
        for (int k = 0; k <= kGMax; ++k) out << g[c * 11 + k] << " ";
        for (int k = 0; k < 90; ++k) out << st[c * 90 + k] << " ";
        out << std::endl;
      }
  }

  hcm_options opt_;
  const hcm_arrays* arrays_ = nullptr;
  hc_ctx* ctx_ = nullptr;
  // inputs
  std::string dem_read_extension_ = "asc", dem_write_extension_ = "asc", write_path_, read_path_;
  std::string write_fname_ = "catchment.dat", read_fname_, hydroindex_fname_, rainfall_data_file_,
              grain_data_file_, bedrock_data_file_, water_init_raster_file_, mannings_fname_,
              elev_fname_, waterdepth_fname_, grainsize_fname_;
  int imax_ = 0, jmax_ = 0, NR_ = 0, NC_ = 0;
  double xll_ = 0, yll_ = 0, DX_ = 5.0, no_data_value_ = -9999.9;
  // parameters (LSDCatchmentModel.hpp:437-568 defaults)
  double water_depth_erosion_threshold_ = 0.01, in_out_difference_allowed_ = 0, mannings_ = 0.04;
  int rfnum_ = 1, maxcycle_ = 1000, max_time_step_ = 0;
  double ERODEFACTOR_ = 0.05, MIN_Q_ = 0.01, MIN_Q_MAXVAL_ = 1000.0, active_ = 0.2;
  double M_ = 0.005, baseflow_ = 0.00000005, rain_factor_ = 1;
  double output_file_save_interval_ = 60, max_vel_ = 5, courant_number_ = 0.7;
  double edgeslope_ = 0.001, chann_lateral_erosion_ = 20.0, veg_lat_restriction_ = 0.1;
  double froude_limit_ = 0.8, Csuspmax_ = 0.05, hflow_threshold_ = 0.00001;
  double lateral_constant_ = 0.0000002, bedrock_erosion_threshold_ = 0, bedrock_erosion_rate_ = 0;
  double saveinterval_ = 1000, rain_data_time_step_ = 60;
  double d_[10] = {0.0, 0.000065, 0.001, 0.002, 0.004, 0.008, 0.016, 0.032, 0.064, 0.128};
  double dprop_[11] = {0.0, 0.05, 0.05, 0.15, 0.225, 0.25, 0.1, 0.075, 0.05, 0.05, 0.0};
  double fall_[11] = {0.0, 0.033, 0.109, 0.164, 0.237, 0.338, 0.479, 0.678, 0.959, 1.357, 0.0};
  bool hydro_only_ = false, rainfall_data_on_ = false, water_init_from_raster_ = false,
       spatially_var_rainfall_ = false, spatially_var_mannings_ = false,
       graindata_from_file_ = false, bedrock_layer_on_ = false, wilcock_ = false,
       einstein_ = false, write_elev_file_ = false, write_waterd_file_ = false,
       write_grainsz_file_ = false, debug_print_cycle_ = false;
  // state
  std::vector<double> elev_, area_, area_depth_;
  std::vector<int> rfarea_;
  std::vector<std::vector<float>> rain_;
  std::vector<double> j_, jo_, j_mean_, old_j_mean_, new_j_mean_;
  std::vector<int> nActual_, input_counter_, in_x_, in_y_;
  std::vector<int> pts_x_, pts_y_, pts_rf_;
  std::vector<double> pts_a_;
  bool points_dirty_ = true;  // the input-point list changed since hc_points_set
  int totalinputpoints_ = 0, counter_ = 0, hours_ = 0;
  double time_step_ = 1, cycle_ = 0, maxdepth_ = 10, temptot_ = 0, waterinput_ = 0,
         waterOut_ = 0, iod_ = 0, time_1_ = 1, save_time_ = 0, creep_time_ = 1, creep_time2_ = 1,
         failureangle_ = 45, creep_rate_ = 0.0025, previous_ = 0, new_cycle_ = 0,
         old_cycle_ = 0, erode_mult_ = 1, erode_call_ = 0, globalsediq_ = 0, sediQ_ = 0,
         last_dt_ = 0;
  double tx_ = 60, tlastcalc_ = 0, Qs_step_ = 0, Qs_hour_ = 0, Qs_over_ = 0, Qs_last_ = 0,
         Qw_newvol_ = 0, Qw_oldvol_ = 0, Qw_lastvol_ = 0, Qw_stepvol_ = 0, Qw_hourvol_ = 0,
         Qw_hour_ = 0, Qw_overvol_ = 0, temptotal_ = 0, old_sediq_ = 0, Jw_newvol_ = 0,
         Jw_oldvol_ = 0, Jw_lastvol_ = 0, Jw_stepvol_ = 0, Jw_hourvol_ = 0, Jw_hour_ = 0,
         Jw_overvol_ = 0, sand_out_ = 0;
  std::vector<double> sum_grain_, old_sum_grain_, Qg_step_, Qg_hour_, Qg_over_, Qg_last_;
  std::vector<double> series_;
  std::vector<double> topmodel_table_;
  bool loop_started_ = false;
  // spatially complex rainfall: the runoffGrid (j, jo, j_mean, old, new per
  // cell), the last hydrograph time, and whether j_mean has been formed yet
  bool complex_rain_ = false;
  // reach mode: up to three input points (in file order) and their files
  bool reach_mode_ = false, reach_on_[3] = {false, false, false};
  std::string reach_file_[3];
  int reach_px_[3] = {0, 0, 0}, reach_py_[3] = {0, 0, 0};
  int reach_n_ = 0, reach_x_[3] = {0, 0, 0}, reach_y_[3] = {0, 0, 0};
  int div_inputs_ = 1, reach_rdt_ = 60;
  std::vector<std::vector<float>> reach_rows_[3];
  std::vector<double> rj_, rjo_, rj_mean_, rold_, rnew_;
  double runoff_t_ = 0;
  int runoff_jm_set_ = 0;
};

}  // namespace hcdrv

struct hcm_model {
  hcdrv::CatchmentModel* m = nullptr;
  hcm_arrays arrays{};  // borrowed pointers, valid during hcm_create_arrays only
  std::string err;
};

extern "C" {

static hcm_options default_options(const hcm_options* opt) {
  hcm_options o{};
  o.precision = HC_FP64;
  o.fused = 1;
  o.quiet = 1;
  o.write_files = 1;
  o.flat_dz = 0.0;
  o.device_loop = 0;
  o.use_graph = 0;
  o.elev_fixed = 0;
  o.dt_floor = 0.0;
  if (opt) o = *opt;
  return o;
}

static int finish_create(hcm_model* h, hcm_model** out);

int hcm_create(hcm_model** out, const char* path, const char* pfile, const hcm_options* opt) {
  if (!out || !path || !pfile) return HC_ERR_ARG;
  *out = nullptr;
  hcm_model* h = new hcm_model;
  try {
    h->m = new hcdrv::CatchmentModel(path, pfile, default_options(opt));
  } catch (const std::exception& e) {
    h->err = e.what();
    *out = h;
    return HC_ERR_STATE;
  }
  return finish_create(h, out);
}

int hcm_create_arrays(hcm_model** out, const char* params_text, const hcm_arrays* arrays,
                      const hcm_options* opt) {
  if (!out || !params_text || !arrays) return HC_ERR_ARG;
  *out = nullptr;
  hcm_model* h = new hcm_model;
  try {
    h->arrays = *arrays;
    h->m = new hcdrv::CatchmentModel(params_text, h->arrays, default_options(opt));
  } catch (const std::exception& e) {
    h->err = e.what();
    *out = h;
    return HC_ERR_STATE;
  }
  return finish_create(h, out);
}

static int finish_create(hcm_model* h, hcm_model** out) {
  try {
    h->m->initialise_model_domain_extents();
    h->m->initialise_arrays();
    h->m->load_data();
    h->m->set_time_counters();
    h->m->initialise_rainfall_runoff();
    h->m->initialise_drainage_area();
  } catch (const std::exception& e) {
    h->err = e.what();
    // report through a model that holds only the message
    delete h->m;
    h->m = nullptr;
    *out = h;
    return HC_ERR_STATE;
  }
  *out = h;
  return HC_OK;
}

int hcm_destroy(hcm_model* h) {
  if (!h) return HC_OK;
  delete h->m;
  delete h;
  return HC_OK;
}

const char* hcm_last_error(const hcm_model* h) { return h ? h->err.c_str() : "null model"; }

int hcm_iterate(hcm_model* h, int n, double* tt, double* cy, double* dtf, int* done, int* finished) {
  if (!h || !h->m) return HC_ERR_ARG;
  int k = 0;
  bool fin = false;
  try {
    if (h->m->device_loop_enabled()) {
      k = h->m->iterate_device(n, tt, cy, dtf, &fin);
    } else {
      for (; k < n && !fin; ++k) {
        fin = h->m->iterate();
        if (tt) tt[k] = h->m->temptot();
        if (cy) cy[k] = h->m->get_cycle();
        if (dtf) dtf[k] = h->m->last_dt();
      }
    }
  } catch (const hcdrv::Halted& e) {
    h->err = e.what();
    if (done) *done = e.done >= 0 ? e.done : k;
    if (finished) *finished = 0;
    return e.code;
  } catch (const std::exception& e) {
    h->err = e.what();
    if (done) *done = k;
    return HC_ERR_HIP;
  }
  if (done) *done = k;
  if (finished) *finished = fin ? 1 : 0;
  return HC_OK;
}

int hcm_loop_begin(hcm_model* h) {
  if (!h || !h->m) return HC_ERR_ARG;
  try {
    h->m->ensure_device_loop();
  } catch (const std::exception& e) {
    h->err = e.what();
    return HC_ERR_STATE;
  }
  return HC_OK;
}

int hcm_loop_collect(hcm_model* h, int n, double* tt, double* cy, double* dtf, int* done,
                     int* finished) {
  if (!h || !h->m || n < 0) return HC_ERR_ARG;
  bool fin = false;
  int k = 0;
  try {
    k = h->m->collect_device(n, tt, cy, dtf, &fin);
  } catch (const hcdrv::Halted& e) {
    h->err = e.what();
    if (done) *done = e.done;
    if (finished) *finished = 0;
    return e.code;
  } catch (const std::exception& e) {
    h->err = e.what();
    return HC_ERR_HIP;
  }
  if (done) *done = k;
  if (finished) *finished = fin ? 1 : 0;
  return HC_OK;
}

size_t hcm_field_count(const hcm_model* h, int field) {
  if (!h || !h->m) return 0;
  return hc_field_count(const_cast<hcdrv::CatchmentModel*>(h->m)->ctx(), field);
}

int hcm_download(hcm_model* h, int field, void* host, size_t count) {
  if (!h || !h->m) return HC_ERR_ARG;
  const int rc = hc_download(h->m->ctx(), field, host, count);
  if (rc) h->err = hc_last_error(h->m->ctx());
  return rc;
}

int hcm_scalar(hcm_model* h, const char* name, double* v) {
  if (!h || !h->m || !name || !v) return HC_ERR_ARG;
  return h->m->scalar(name, v) ? HC_OK : HC_ERR_ARG;
}

int hcm_set_graph(hcm_model* h, int use_graph) {
  if (!h || !h->m) return HC_ERR_ARG;
  h->m->set_graph(use_graph != 0);
  return HC_OK;
}

hc_ctx* hcm_context(hcm_model* h) { return (h && h->m) ? h->m->ctx() : nullptr; }

int hcm_checkpoint(hcm_model* h, const char* path) {
  if (!h || !h->m || !path) return HC_ERR_ARG;
  try {
    h->m->checkpoint(path);
  } catch (const std::exception& e) {
    h->err = e.what();
    return HC_ERR_STATE;
  }
  return HC_OK;
}

int hcm_restart(hcm_model* h, const char* path) {
  if (!h || !h->m || !path) return HC_ERR_ARG;
  try {
    h->m->restart(path);
  } catch (const std::exception& e) {
    h->err = e.what();
    return HC_ERR_STATE;
  }
  return HC_OK;
}

int hcm_series(hcm_model* h, double* out, int max_rows) {
  if (!h || !h->m) return 0;
  const auto& s = h->m->series();
  const int rows = static_cast<int>(s.size() / 14);
  const int nc = std::min(rows, max_rows);
  if (out && nc > 0) std::memcpy(out, s.data(), sizeof(double) * 14 * nc);
  return rows;
}

}  // extern "C"
