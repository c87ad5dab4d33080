// Test infrastructure only: the oracle (CPU restatement, FP64).
//
// Written from the reference's semantics; every function cites the reference
// lines it restates (paths relative to the reference root; LCM.cpp =
// src/catchmentmodel/LSDCatchmentModel.cpp, LCM.hpp = its header).  The
// product never links this file: tests, the smoke check and the benchmark's
// CPU baseline load it as the checker.
//
// Differences from the reference that do not change results:
//   * the active list down_scan[y][k] (LCM.cpp:3821-3849) is kept as a 0/1
//     mask; every hot loop visits masked cells in the reference order (column
//     y ascending, row x ascending) where order matters (erode), and as a
//     plain per-cell map where it does not (flow_route, depth_update);
//   * lazily allocated grain rows (index[x][y] -> grain row, LCM.cpp:3949-3983)
//     are a dense per-cell block plus a has-grain flag;
//   * `mannings` is a local, not a shared member written inside the loop
//     (the race at LCM.cpp:3091 cannot change a single-thread result);
//   * veg and edge are identically zero (vegetation and lateral erosion are
//     never enabled, main.cpp:178), so their terms are evaluated with zeros.
#include "hc_oracle.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kGravity = 9.81;   // LCM.hpp:437
constexpr int kGMax = 10;           // LCM.hpp:444
// D8 neighbour offsets, LCM.hpp:445-446 (index 0 unused; odd = cardinal).
constexpr int kDX8[9] = {0, 0, 1, 1, 1, 0, -1, -1, -1};
constexpr int kDY8[9] = {0, -1, -1, 0, 1, 1, 1, 0, -1};
constexpr int kStrataLayers = 10;
constexpr int kStrataFr = 9;

// x86-64 cvttsd2si behaviour of static_cast<int>(double) for out-of-range
// values (LCM.cpp:2283, where ERODEFACTOR/0 reaches it).
int trunc_int(double v) {
  if (!(v > -2147483649.0 && v < 2147483648.0)) return INT_MIN;
  return static_cast<int>(v);
}

}  // namespace

struct orc_model {
  orc_params p{};
  int ni = 0, nj = 0;  // padded dims imax+2, jmax+2
  // 2-D fields
  std::vector<double> elev, h, qx, qy, qxs, qys, V, nman, area, area_depth,
      bedrock, Tau, Vel, ss;
  std::vector<double> vel_dir;  // [cell][9]
  std::vector<double> grain;    // [cell][11]
  std::vector<double> strata;   // [cell][10][9]
  std::vector<double> sr, sl, su, sd;  // [cell][10]
  std::vector<int> mask, hasgrain, rfarea;
  // the reference's row number of each cell's grain row (index[x][y],
  // -9999 = none): file rows by line, later rows by grain_array_tot in addGS
  std::vector<int> gindex;
  // rain-runoff state (LCM.cpp:1779-1788, 3649-3724)
  std::vector<std::vector<float>> rain;
  std::vector<double> j, jo, j_mean, old_j_mean, new_j_mean;
  std::vector<int> nActual, input_counter;
  std::vector<int> in_x, in_y;
  int totalinputpoints = 0;
  // spatially complex rainfall: the runoffGrid's per-cell arrays
  // (LSDRainfallRunoff.cpp:96-104: j = jo = 1e-9, the means 0)
  std::vector<double> rj, rjo, rj_mean, rold, rnew;
  // reach mode: up to three input points and their files (LCM.cpp:1766-1845)
  int reach_n = 0;
  int reach_x[3] = {0, 0, 0}, reach_y[3] = {0, 0, 0};
  std::vector<std::vector<float>> reach_rows[3];
  // scalars (LCM.hpp:456-548 defaults)
  double time_step = 1, cycle = 0, maxdepth = 10, temptot = 0, waterinput = 0,
         waterOut = 0, iod = 0, baseflow = 0.00000005, time_1 = 1, creep_time = 1, creep_time2 = 1,
         previous = 0, new_cycle = 0, old_cycle = 0, erode_mult = 1,
         erode_call = 0, globalsediq = 0, sediQ = 0, rain_factor = 1;
  int counter = 0;
  int grain_array_tot = 0;
  // output_data state (LCM.cpp:2310-2478)
  double tx = 60, Tx = 0, tlastcalc = 0, Qs_step = 0, Qs_hour = 0, Qs_over = 0,
         Qs_last = 0, Qw_newvol = 0, Qw_oldvol = 0, Qw_lastvol = 0,
         Qw_stepvol = 0, Qw_hourvol = 0, Qw_hour = 0, Qw_overvol = 0,
         temptotal = 0, old_sediq = 0, Jw_newvol = 0, Jw_oldvol = 0,
         Jw_lastvol = 0, Jw_stepvol = 0, Jw_hourvol = 0, Jw_hour = 0,
         Jw_overvol = 0, sand_out = 0;
  int hours = 0;
  std::vector<double> sum_grain, old_sum_grain, Qg_step, Qg_hour, Qg_over,
      Qg_last;
  std::vector<double> series;  // 14 columns per row

  long at(int x, int y) const { return static_cast<long>(x) * nj + y; }
  double mannings_at(int x, int y) const {
    return p.spatially_var_mannings ? nman[at(x, y)] : p.mannings;
  }

  // ---------------------------------------------------------------- dt ----
  // LCM.cpp:1884-1887
  double cfl() const {
    return p.courant_number * (p.DX / std::sqrt(kGravity * (maxdepth)));
  }
  // LCM.cpp:1890-1910
  void set_maximum_timestep() {
    if (maxdepth <= 0.1) maxdepth = 0.1;
    if (time_step < cfl()) time_step = cfl();
    if (iod > p.in_out_difference_allowed && time_step > cfl())
      time_step = cfl();
  }
  // LCM.cpp:1913-1923
  double get_flow_timestep() const {
    double f = time_step;
    if (f > cfl()) f = cfl();
    return f;
  }

  // ------------------------------------------------------- flow_route -----
  // LCM.cpp:3063-3247.  One face update; (xn,yn) is the neighbour across
  // the face, `is_x` selects qx/x-edge or qy/y-edge rules.
  inline void face(int x, int y, int xn, int yn, bool is_x, double dt,
                   double n) {
    const long c = at(x, y), cn = at(xn, yn);
    if (!((h[c] > 0 || h[cn] > 0) && elev[cn] > -9999)) return;
    double* q = is_x ? &qx[c] : &qy[c];
    double* qs = is_x ? &qxs[c] : &qys[c];
    const double eta_c = elev[c] + h[c];
    const double eta_n = elev[cn] + h[cn];
    const double zmax = (elev[cn] < elev[c]) ? elev[c] : elev[cn];  // std::max(z_n, z_c)
    const double etamax = (eta_c < eta_n) ? eta_n : eta_c;           // std::max(eta_c, eta_n)
    const double hflow = etamax - zmax;
    if (hflow > p.hflow_threshold) {
      double slope = (eta_n - eta_c) / p.DX;
      if (is_x) {
        if (x == p.imax) slope = p.edgeslope;
        if (x <= 2) slope = 0 - p.edgeslope;
      } else {
        if (y == p.jmax) slope = p.edgeslope;
        if (y <= 2) slope = 0 - p.edgeslope;
      }
      // pow(hflow, 10/3) with integer 10/3 == pow(hflow, 3) (LCM.cpp:3115)
      const double h3 = std::pow(hflow, 3);
      double qv = *q;
      qv = ((qv - (kGravity * hflow * dt * slope)) /
            (1 + kGravity * hflow * dt * (n * n) * std::fabs(qv) / h3));
      const double sg = std::sqrt(kGravity * hflow);
      if (qv > 0 && (qv / hflow) / sg > p.froude_limit)
        qv = hflow * (sg * p.froude_limit);
      if (qv < 0 && std::fabs(qv / hflow) / sg > p.froude_limit)
        qv = 0 - (hflow * (sg * p.froude_limit));
      if (qv > 0 && (qv * dt / p.DX) > (h[c] / 4))
        qv = ((h[c] * p.DX) / 5) / dt;
      if (qv < 0 && std::fabs(qv * dt / p.DX) > (h[cn] / 4))
        qv = 0 - ((h[cn] * p.DX) / 5) / dt;
      *q = qv;
      // suspended fraction (isSuspended[1] is always true, LCM.cpp:2980)
      if (qv > 0) *qs = qv * (V[c] / h[c]);
      if (qv < 0) *qs = qv * (V[cn] / h[cn]);
      if (*qs > 0 && *qs * dt > (V[c] * p.DX) / 4) *qs = ((V[c] * p.DX) / 5) / dt;
      if (*qs < 0 && std::fabs(*qs * dt) > (V[cn] * p.DX) / 4)
        *qs = 0 - ((V[cn] * p.DX) / 5) / dt;
      // directional velocities (LCM.cpp:3177-3184, 3232-3233)
      if (is_x) {
        if (qv > 0) vel_dir[c * 9 + 7] = qv / hflow;
        if (qv < 0) vel_dir[cn * 9 + 3] = (0 - qv) / hflow;
      } else {
        if (qv > 0) vel_dir[c * 9 + 1] = qv / hflow;
        if (qv < 0) vel_dir[cn * 9 + 5] = (0 - qv) / hflow;
      }
    } else {
      *q = 0;
      *qs = 0;
    }
  }

  void flow_route() {
    const double dt = get_flow_timestep();
#pragma omp parallel for schedule(static) num_threads(p.nthreads)
    for (int x = 1; x <= p.imax; ++x) {
      for (int y = 1; y <= p.jmax; ++y) {
        const long c = at(x, y);
        if (!mask[c] || !(elev[c] > -9999)) continue;
        const double n = mannings_at(x, y);
        face(x, y, x - 1, y, true, dt, n);
        face(x, y, x, y - 1, false, dt, n);
      }
    }
  }

  // ----------------------------------------------------- depth_update -----
  // LCM.cpp:3259-3299
  double depth_update() {
    const double dt = get_flow_timestep();
    double md = 0;
#pragma omp parallel for schedule(static) reduction(max : md) num_threads(p.nthreads)
    for (int x = 1; x <= p.imax; ++x) {
      for (int y = 1; y <= p.jmax; ++y) {
        const long c = at(x, y);
        if (!mask[c]) continue;
        h[c] += dt * (qx[at(x + 1, y)] - qx[c] + qy[at(x, y + 1)] - qy[c]) / p.DX;
        V[c] += dt * (qxs[at(x + 1, y)] - qxs[c] + qys[at(x, y + 1)] - qys[c]) / p.DX;
        if (h[c] > 0) {
          if (elev[c] == -9999) h[c] = 0;
          if (h[c] > md) md = h[c];
        }
      }
    }
    maxdepth = md;
    return md;
  }

  // -------------------------------------------------------- scan_area -----
  // LCM.cpp:3821-3849: 3x3 wet dilation over the interior.
  void scan_area() {
#pragma omp parallel for schedule(static) num_threads(p.nthreads)
    for (int x = 1; x <= p.imax; ++x) {
      for (int y = 1; y <= p.jmax; ++y) {
        int wet = 0;
        for (int dx = -1; dx <= 1 && !wet; ++dx)
          for (int dy = -1; dy <= 1; ++dy)
            if (h[at(x + dx, y + dy)] > 0) { wet = 1; break; }
        mask[at(x, y)] = wet;
      }
    }
  }

  // --------------------------------------------------- water_flux_out -----
  // LCM.cpp:3005-3047 (serial, fixed edge order).
  double water_flux_out() {
    const double ts = time_step;
    const double thr = p.water_depth_erosion_threshold;
    double tot = 0;
    auto drain = [&](int x, int y) {
      double& hv = h[at(x, y)];
      if (hv > thr) {
        tot += (hv - thr) * p.DX * p.DX / ts;
        hv = thr;
      }
    };
    for (int i = 1; i <= p.imax; ++i) {
      drain(i, p.jmax);
      drain(i, 1);
    }
    for (int jj = 1; jj <= p.jmax; ++jj) {
      drain(1, jj);
      drain(p.imax, jj);
    }
    temptot = tot;
    waterOut = tot;
    return tot;
  }

  // --------------------------------------------------- rainfall-runoff ----
  // LCM.cpp:3649-3715 (semi-distributed TOPMODEL).  Rain values are float
  // (read_rainfalldata, LCM.cpp:694-717) and the rate is formed in float.
  void topmodel_runoff(double cyc) {
    const int ts = static_cast<int>(cyc / p.rain_data_time_step);
    for (int n = 1; n <= p.rfnum; ++n) {
      double local = 0;
      const double ftime = 60;
      old_j_mean[n] = new_j_mean[n];
      jo[n] = j[n];
      const float r = rain_at(ts, n - 1);
      if (r > 0) local = rain_factor * ((r / 1000) / 3600);
      if (local == 0) {
        j[n] = jo[n] / (1 + ((jo[n] * ftime) / p.M));
        new_j_mean[n] = p.M / ftime * std::log(1 + ((jo[n] * ftime) / p.M));
      }
      if (local > 0) {
        j[n] = local / (((local - jo[n]) / jo[n]) *
                            std::exp((0 - local) * ftime / p.M) + 1);
        new_j_mean[n] =
            (p.M / ftime) *
            std::log(((local - jo[n]) + jo[n] * std::exp((local * ftime) / p.M)) /
                     local);
      }
      if (new_j_mean[n] < 0) new_j_mean[n] = 0;
    }
  }
  float rain_at(int row, int col) const {
    if (row < 0 || row >= static_cast<int>(rain.size())) return 0.0f;
    if (col < 0 || col >= static_cast<int>(rain[row].size())) return 0.0f;
    return rain[row][col];
  }
  // LCM.cpp:3717-3724
  void calchydrograph(double t) {
    for (int n = 1; n <= p.rfnum; ++n)
      j_mean[n] = old_j_mean[n] + (((new_j_mean[n] - old_j_mean[n]) / 2) * (2 - t));
  }

  // ------------------------------------- spatially complex rainfall ----
  // topmodel_runoff(cycle, runoff) (LCM.cpp:3610-3627): the rain grid of
  // rainGrid::create (LSDRainfallRunoff.cpp:36-63) -- with one rain cell the
  // first column's value everywhere; otherwise cells 1..imax-1, 1..jmax-1
  // whose hydroindex is 1..rfnum-1 take column index-1, all others 0 -- held
  // in double; then runoffGrid::calculate_runoff (LSDRainfallRunoff.cpp:
  // 175-241) on cells 1..imax-1, 1..jmax-1 above -9999, with rain_factor
  // passed as an int and a 60 s step.
  double rain_grid(int ts, long c) const {
    if (p.rfnum == 1) return static_cast<double>(rain_at(ts, 0));
    const int rf = rfarea[c];
    if (rf >= 1 && rf < p.rfnum) return static_cast<double>(rain_at(ts, rf - 1));
    return 0.0;
  }
  void topmodel_runoff_grid(double cyc) {
    const int ts = static_cast<int>(cyc / p.rain_data_time_step);
    const int rfi = static_cast<int>(rain_factor);
    for (int m = 1; m < p.imax; ++m) {
      for (int n = 1; n < p.jmax; ++n) {
        const long c = at(m, n);
        if (!(elev[c] > -9999)) continue;
        double local = 0;
        const double lts = 60;
        rold[c] = rnew[c];
        rjo[c] = rj[c];
        const double r = rain_grid(ts, c);
        if (r > 0) local = rfi * ((r / 1000) / 3600);
        if (local == 0) {
          rj[c] = rjo[c] / (1 + ((rjo[c] * lts) / p.M));
          rnew[c] = p.M / lts * std::log(1 + ((rjo[c] * lts) / p.M));
        }
        if (local > 0) {
          rj[c] = local / (((local - rjo[c]) / rjo[c]) * std::exp((0 - local) * lts / p.M) + 1);
          rnew[c] = (p.M / lts) *
                    std::log(((local - rjo[c]) + rjo[c] * std::exp((local * lts) / p.M)) / local);
        }
        if (rnew[c] < 0) rnew[c] = 0;
      }
    }
  }
  // calchydrograph(time, runoff) (LCM.cpp:3728-3739), cells 1..imax, 1..jmax
  void calchydrograph_grid(double t) {
    for (int m = 1; m <= p.imax; ++m)
      for (int n = 1; n <= p.jmax; ++n) {
        const long c = at(m, n);
        rj_mean[c] = rold[c] + (((rnew[c] - rold[c]) / 2) * (2 - t));
      }
  }
  // catchment_water_input_and_hydrology(flow_timestep, runoff)
  // (LCM.cpp:3544-3606): every cell 1..imax-1, 1..jmax-1 takes its own
  // runoff, capped at ERODEFACTOR; the step-size test uses the largest
  // new_j_mean from before the TOPMODEL updates.
  void catchment_water_input_grid(double fts) {
    for (int i = 1; i < p.imax; ++i)
      for (int jj = 1; jj < p.jmax; ++jj) waterinput += rj_mean[at(i, jj)] * p.DX * p.DX;
    for (int i = 1; i < p.imax; ++i)
      for (int jj = 1; jj < p.jmax; ++jj) {
        double amt = rj_mean[at(i, jj)] * fts;
        if (amt > p.ERODEFACTOR) amt = p.ERODEFACTOR;
        waterinput += (amt / fts) * p.DX * p.DX;
        h[at(i, jj)] += amt;
      }
    double njmax = 0;
    for (int m = 1; m <= p.imax; ++m)
      for (int n = 1; n <= p.jmax; ++n)
        if (rnew[at(m, n)] > njmax) njmax = rnew[at(m, n)];
    if (cycle >= time_1) {
      const unsigned cells = static_cast<unsigned>(p.jmax) * static_cast<unsigned>(p.imax);
      do {
        time_1++;
        topmodel_runoff_grid(time_1);
        if (time_step > p.max_time_step && njmax > (0.2 / (cells * p.DX * p.DX))) {
          cycle = time_1 + (p.max_time_step / 60);
          time_step = p.max_time_step;
        }
      } while (time_1 < cycle);
    }
    calchydrograph_grid(time_1 - cycle);
  }
  // output_data(temptotal, runoff)'s runoff volume (LCM.cpp:2487-2501,
  // 2571-2583): each catchment cell's term added in turn
  void add_grid_runoff_volume(double& vol, double minutes) const {
    for (int i = 1; i <= p.imax; ++i)
      for (int jj = 1; jj <= p.jmax; ++jj)
        if (elev[at(i, jj)] > p.no_data_value) vol += (rj_mean[at(i, jj)] * p.DX * p.DX) * (minutes * 60);
  }

  // ------------------------------------------------------- reach mode ----
  float reach_val(int k, int row, int col) const {
    if (row < 0 || row >= static_cast<int>(reach_rows[k].size())) return 0.0f;
    if (col < 0 || col >= static_cast<int>(reach_rows[k][row].size())) return 0.0f;
    return reach_rows[k][row][col];
  }
  // reach_water_and_sediment_input (LCM.cpp:3301-3452).  With erosion, a
  // wet input point takes each fraction's sediment (columns 5..13; fraction
  // 1 is suspended, LCM.cpp:2980), all scaled down together when their
  // mean over the points exceeds 0.75 ERODEFACTOR; the total is formed with
  // the float |value| / div_inputs, the amounts with the double product
  // first, as the reference's expressions do.  Then every point takes the
  // discharge (column 1) interpolated between this data step and the next.
  void reach_water_and_sediment_input() {
    const double fts = get_flow_timestep();
    const int rdt = p.reach_input_data_timestep;
    if (p.reach_mode && !p.hydro_only) {
      for (int n = 0; n <= reach_n - 1; ++n) {
        const int x = reach_x[n], y = reach_y[n];
        const long c = at(x, y);
        const int row = static_cast<int>(cycle / rdt);
        double adding_factor = 1, added_tot = 0;
        if (h[c] > p.water_depth_erosion_threshold) {
          for (int tn = 5; tn <= kGMax + 3; ++tn)
            added_tot += (std::abs(reach_val(n, row, tn))) / p.div_inputs / (p.DX * p.DX) / (rdt * 60) * time_step;
          if (added_tot / reach_n > p.ERODEFACTOR * 0.75)
            adding_factor = (p.ERODEFACTOR * 0.75) / (added_tot / reach_n);
          if (!hasgrain[c]) addGS(x, y);
          for (int tn = 5; tn <= kGMax + 3; ++tn) {
            const double amt = adding_factor * (std::abs(reach_val(n, row, tn))) / p.div_inputs /
                               (p.DX * p.DX) / (rdt * 60) * time_step;
            if (tn - 4 == 1) {
              V[c] += amt;
            } else {
              G(c)[tn - 4] += amt;
              elev[c] += amt;
            }
          }
        }
      }
    }
    for (int n = 0; n <= reach_n - 1; ++n) {
      const long c = at(reach_x[n], reach_y[n]);
      const int row = static_cast<int>(cycle / rdt);
      const double i1 = reach_val(n, row, 1), i2 = reach_val(n, row + 1, 1);
      const double prop = (((row + 1) * rdt) - cycle) / rdt;
      const double input = i1 + ((i2 - i1) * (1 - prop));
      waterinput += (input / p.div_inputs);
      h[c] += (input / p.div_inputs) / (p.DX * p.DX) * fts;
    }
  }

  // LCM.cpp:3741-3767
  int get_catchment_input_points() {
    totalinputpoints = 0;
    for (int n = 1; n <= p.rfnum; ++n) input_counter[n] = 0;
    for (int i = 1; i <= p.imax; ++i) {
      for (int jj = 1; jj <= p.jmax; ++jj) {
        const double v = area[at(i, jj)] * baseflow * 3 * p.DX * p.DX;
        if (v > p.MIN_Q && v < p.MIN_Q_MAXVAL) {
          ++totalinputpoints;
          in_x[totalinputpoints] = i;
          in_y[totalinputpoints] = jj;
          input_counter[rfarea[at(i, jj)]]++;
        }
      }
    }
    if (totalinputpoints == 0) totalinputpoints = 1;
    return totalinputpoints;
  }

  // LCM.cpp:2071-2089 + 2091-2270.  The reference sorts (key, elevation)
  // pairs by elevation with std::sort; a single sort of (cell, elevation)
  // pairs of the same length and order gives the same permutation, because
  // std::sort's control flow depends only on the comparisons.
  void drainage_area() {
    for (int i = 1; i <= p.imax; ++i)
      for (int jj = 1; jj <= p.jmax; ++jj) {
        area_depth[at(i, jj)] = 1;
        area[at(i, jj)] = 0;
        if (elev[at(i, jj)] == -9999) area_depth[at(i, jj)] = 0.0;
      }
    const long total = static_cast<long>(ni) * nj;
    std::vector<std::pair<long, double>> pr(total, std::make_pair(0L, 0.0));
    long inc = 1;
    for (int i = 1; i <= p.imax; ++i)
      for (int jj = 1; jj <= p.jmax; ++jj) {
        pr[inc] = std::make_pair(at(i, jj), elev[at(i, jj)]);
        ++inc;
      }
    std::sort(pr.begin(), pr.end(),
              [](const std::pair<long, double>& a,
                 const std::pair<long, double>& b) { return a.second < b.second; });
    const long ncell = static_cast<long>(p.imax) * p.jmax;
    for (long k = ncell; k >= 1; --k) {
      const long c = pr[k].first;  // fake entries carry cell 0 == (0,0)
      const int i = static_cast<int>(c / nj), jj = static_cast<int>(c % nj);
      if (!(area_depth[c] > 0)) continue;
      if (area_depth[c] > area[c]) area[c] = area_depth[c];
      double difftot = 0;
      for (int dir = 1; dir <= 8; ++dir) {
        int j2 = jj + kDY8[dir], i2 = i + kDX8[dir];
        if (j2 < 1) j2 = 1;
        if (i2 < 1) i2 = 1;
        if (j2 > p.jmax) j2 = p.jmax;
        if (i2 > p.imax) i2 = p.imax;
        const double e2 = elev[at(i2, j2)], e = elev[c];
        if (dir % 2 != 0) {
          if (e2 < e) difftot += e - e2;
        } else {
          if (e2 < e) difftot += (e - e2) / 1.414;
        }
      }
      if (difftot > 0) {
        for (int dir = 1; dir <= 8; ++dir) {
          int i2 = i + kDX8[dir], j2 = jj + kDY8[dir];
          if (j2 < 1) j2 = 1;
          if (i2 < 1) i2 = 1;
          if (j2 > p.jmax) j2 = p.jmax;
          if (i2 > p.imax) i2 = p.imax;
          const long c2 = at(i2, j2);
          const double e2 = elev[c2], e = elev[c];
          if (dir % 2 != 0) {
            if (e2 < e) area_depth[c2] += area_depth[c] * ((e - e2) / difftot);
          } else {
            if (e2 < e)
              area_depth[c2] += area_depth[c] * (((e - e2) / 1.414) / difftot);
          }
        }
      }
      area_depth[c] = 0;
    }
  }

  // LCM.cpp:3455-3541 (uniform / hydroindex rainfall)
  void catchment_water_input_and_hydrology(double fts) {
    for (int i = 1; i <= p.rfnum; ++i)
      waterinput += j_mean[i] * nActual[i] * p.DX * p.DX;
    for (int z = 1; z <= totalinputpoints; ++z) {
      const int i = in_x[z], jj = in_y[z];
      const int rf = rfarea[at(i, jj)];
      double amt = (j_mean[rf] * nActual[rf]) / (input_counter[rf]) * fts;
      if (amt > p.ERODEFACTOR) amt = p.ERODEFACTOR;
      h[at(i, jj)] += amt;
    }
    if (cycle >= time_1) {
      do {
        time_1++;
        topmodel_runoff(time_1);
        if (time_step > p.max_time_step) {
          double jmt = 0;
          for (int n = 1; n <= p.rfnum; ++n)
            if (new_j_mean[n] > jmt) jmt = new_j_mean[n];
          const unsigned cells = static_cast<unsigned>(p.imax) * static_cast<unsigned>(p.jmax);
          if (jmt > (0.2 / (cells * p.DX * p.DX))) {
            cycle = time_1 + (p.max_time_step / 60);
            time_step = p.max_time_step;
          }
        }
      } while (time_1 < cycle);
    }
    calchydrograph(time_1 - cycle);
    double jmeanmax = 0;
    for (int n = 1; n <= p.rfnum; ++n)
      if (j_mean[n] > jmeanmax) jmeanmax = j_mean[n];
    if (jmeanmax >= baseflow) {
      baseflow = baseflow * 3;
      drainage_area();
      get_catchment_input_points();
    }
    if (baseflow > (jmeanmax * 3) && baseflow > 0.000001) {
      baseflow = jmeanmax * 1.25;
      drainage_area();
      get_catchment_input_points();
    }
  }

  // --------------------------------------------------------- output -------
  // LCM.cpp:2310-2478 (rows recorded instead of appended to a text file)
  void output_data() {
    Qw_newvol += temptotal * ((cycle - previous) * 60);
    if (p.spatially_complex_rainfall) {
      add_grid_runoff_volume(Jw_newvol, cycle - previous);
    } else {
      for (int nn = 1; nn <= p.rfnum; ++nn)
        Jw_newvol += (j_mean[nn] * p.DX * p.DX * nActual[nn]) * ((cycle - previous) * 60);
    }
    const double osi = p.output_file_save_interval;
    if ((new_cycle < old_cycle) || (cycle - previous >= osi)) {
      while ((tx > previous) && (cycle >= tx)) {
        hours++;
        Qs_step = globalsediq - old_sediq;
        Qs_over = Qs_step * ((cycle - tx) / (cycle - tlastcalc));
        Qs_hour = Qs_step - Qs_over + Qs_last;
        if (cycle >= tx + osi) {
          Qs_last = 0;
          old_sediq = globalsediq - Qs_over;
        }
        if (cycle < tx + osi) {
          Qs_last = Qs_over;
          old_sediq = globalsediq;
        }
        for (int n = 1; n <= kGMax - 1; ++n) {
          Qg_step[n] = sum_grain[n] - old_sum_grain[n];
          Qg_over[n] = Qg_step[n] * ((cycle - tx) / (cycle - tlastcalc));
          Qg_hour[n] = Qg_step[n] - Qg_over[n] + Qg_last[n];
          if (cycle >= tx + osi) {
            Qg_last[n] = 0;
            old_sum_grain[n] = sum_grain[n] - Qg_over[n];
          }
          if (cycle < tx + osi) {
            Qg_last[n] = Qg_over[n];
            old_sum_grain[n] = sum_grain[n];
          }
        }
        Qw_overvol = temptotal * ((cycle - tx) * 60);
        Qw_stepvol = Qw_newvol - Qw_oldvol;
        Qw_hourvol = Qw_stepvol - Qw_overvol + Qw_lastvol;
        Qw_hour = Qw_hourvol / (60 * osi);
        if (p.spatially_complex_rainfall) {
          add_grid_runoff_volume(Jw_overvol, cycle - tx);
        } else {
          for (int nn = 1; nn <= p.rfnum; ++nn)
            Jw_overvol += (j_mean[nn] * p.DX * p.DX * nActual[nn]) * ((cycle - tx) * 60);
        }
        Jw_stepvol = Jw_newvol - Jw_oldvol;
        Jw_hourvol = Jw_stepvol - Jw_overvol + Jw_lastvol;
        Jw_hour = Jw_hourvol / (60 * osi);
        if (cycle >= tx + osi) {
          Qw_lastvol = 0;
          Qw_oldvol = Qw_newvol - Qw_overvol;
          Jw_lastvol = 0;
          Jw_oldvol = Jw_newvol - Jw_overvol;
        }
        if (cycle < tx + osi) {
          Qw_lastvol = Qw_overvol;
          Qw_oldvol = Qw_newvol;
          Jw_lastvol = Jw_overvol;
          Jw_oldvol = Jw_newvol;
        }
        Tx = tx;
        tx = Tx + osi;
        series.push_back(hours);
        series.push_back(Qw_hour);
        series.push_back(Jw_hour);
        series.push_back(sand_out);
        sand_out = 0;
        series.push_back(Qs_hour);
        for (int n = 1; n <= kGMax - 1; ++n) series.push_back(Qg_hour[n]);
      }
      tlastcalc = cycle;
    }
  }

  // ------------------------------------------------------ grain helpers ---
  double* G(long c) { return &grain[c * 11]; }
  double* S(long c, int z) { return &strata[(c * kStrataLayers + z) * kStrataFr]; }

  // LCM.cpp:3874-3947
  void sort_active(int x, int y) {
    const long c = at(x, y);
    if (!hasgrain[c]) addGS(x, y);
    double* g = G(c);
    double total = 0.0;
    for (int n = 0; n <= kGMax; ++n) total += g[n];
    if (total > (p.active * 1.5)) {
      for (int z = 9; z >= 1; --z)
        for (int n = 0; n <= kGMax - 2; ++n) S(c, z)[n] = S(c, z - 1)[n];
      const double coeff = p.active / total;
      for (int n = 1; n <= kGMax - 1; ++n) {
        if (g[n] > 0.0) {
          const double amount = coeff * (g[n]);
          S(c, 0)[n - 1] = amount;
          g[n] -= amount;
        }
      }
    }
    if (total < (p.active / 4)) {
      for (int n = 1; n <= kGMax - 1; ++n) g[n] += S(c, 0)[n - 1];
      for (int z = 0; z <= 8; ++z)
        for (int n = 0; n <= kGMax - 2; ++n) S(c, z)[n] = S(c, z + 1)[n];
      const double amount = p.active;
      for (int n = 1; n <= kGMax - 1; ++n) S(c, 9)[n - 1] = amount * p.dprop[n];
    }
  }

  // LCM.cpp:3949-3983
  void addGS(int x, int y) {
    const long c = at(x, y);
    grain_array_tot++;
    hasgrain[c] = 1;
    gindex[c] = grain_array_tot;  // LCM.cpp:3954-3955
    double* g = G(c);
    g[0] = 0;
    for (int n = 1; n <= kGMax - 1; ++n) g[n] = p.active * p.dprop[n];
    g[kGMax] = 0;
    for (int n = 0; n <= 9; ++n) {
      for (int n2 = 0; n2 <= kGMax - 2; ++n2) S(c, n)[n2] = (p.active) * p.dprop[n2 + 1];
      if (elev[c] - (p.active * (n + 1)) < (bedrock[c] - p.active))
        for (int q = 0; q <= kGMax - 2; ++q) S(c, n)[q] = 0;
    }
    sort_active(x, y);
  }

  // LCM.cpp:3985-4009
  double sand_fraction(long c) {
    const double* g = G(c);
    double at_ = 0, st = 0;
    for (int n = 1; n <= kGMax; ++n) at_ += (g[n]);
    for (int n = 1; n <= 2; ++n) st += (g[n]);
    if (at_ < 0.0001) return 0.0;
    return st / at_;
  }

  // LCM.cpp:4038-4075
  double d50(long c) {
    const double* g = G(c);
    double at_ = 0, mx = 0, mn = 0;
    double cum[20] = {};
    for (int n = 1; n <= kGMax; ++n) {
      at_ += (g[n]);
      cum[n] += at_;
    }
    int i = 1;
    while (cum[i] < (at_ * 0.5) && i <= 9) i++;
    const double* d = p.d;
    if (i == 1) { mn = std::log(d[1]); mx = std::log(d[1]); }
    if (i == 2) { mn = std::log(d[1]); mx = std::log(d[2]); }
    if (i == 3) { mn = std::log(d[2]); mx = std::log(d[3]); }
    if (i == 4) { mn = std::log(d[3]); mx = std::log(d[4]); }
    if (i == 5) { mn = std::log(d[4]); mx = std::log(d[5]); }
    if (i == 6) { mn = std::log(d[5]); mx = std::log(d[6]); }
    if (i == 7) { mn = std::log(d[6]); mx = std::log(d[7]); }
    if (i == 8) { mn = std::log(d[7]); mx = std::log(d[8]); }
    if (i == 9) { mn = std::log(d[8]); mx = std::log(d[9]); }
    double D = std::exp(mx - ((mx - mn) * ((cum[i] - (at_ * 0.5)) / (cum[i] - cum[i - 1]))));
    if (at_ < 0.0000001) D = 0;
    return D;
  }

  // LCM.cpp:4077-4171
  void slide_GS(int x, int y, double amount, int x2, int y2) {
    const long c = at(x, y), c2 = at(x2, y2);
    double total = 0;
    if (hasgrain[c] && hasgrain[c2]) {
      double* g = G(c);
      double* g2 = G(c2);
      for (int n = 1; n <= kGMax - 1; ++n)
        if (g[n] > 0) total += g[n];
      if (amount > total) {
        for (int n = 1; n <= kGMax - 1; ++n) g2[n] += (amount - total) * p.dprop[n];
        amount = total;
      }
      if (total > 0) {
        for (int n = 1; n <= kGMax - 1; ++n) {
          const double tr = amount * (g[n] / total);
          g2[n] += tr;
          g[n] -= tr;
          if (g[n] < 0) g[n] = 0;
        }
      }
      sort_active(x, y);
      sort_active(x2, y2);
      return;
    }
    if (!hasgrain[c] && hasgrain[c2]) {
      double* g2 = G(c2);
      for (int n = 1; n <= kGMax - 1; ++n) g2[n] += (amount) * p.dprop[n];
      sort_active(x2, y2);
      return;
    }
    if (hasgrain[c] && !hasgrain[c2]) {
      addGS(x2, y2);
      double* g = G(c);
      double* g2 = G(c2);
      if (amount > p.active) {
        for (int n = 1; n <= kGMax - 1; ++n) g2[n] += amount * p.dprop[n];
        amount = p.active;
      }
      for (int n = 1; n <= kGMax - 1; ++n)
        if (g[n] > 0) total += g[n];
      for (int n = 1; n <= kGMax - 1; ++n) {
        if (total > 0) {
          g2[n] += amount * (g[n] / total);
          if (g[n] > 0.0001) g[n] -= amount * (g[n] / total);
          if (g[n] < 0) g[n] = 0;
        }
      }
      sort_active(x, y);
      sort_active(x2, y2);
      return;
    }
  }

  // ------------------------------------------------------------ erode -----
  // LCM.cpp:4211-4845
  double erode(double mult) {
    const double rho = 1000.0;
    const double wdet = p.water_depth_erosion_threshold;
    const double veg0 = 0.0, veg1 = 0.0;  // veg[x][y][0..1] (vegetation off)
    double tempbmax = 0;
    float gtot2[20] = {};
    time_step = time_step * 1.5;
    auto s3 = [&](std::vector<double>& a, long c, int n) -> double& { return a[c * 10 + n]; };
    do {
      tempbmax = 0;
      for (int y = 1; y < p.jmax; ++y) {
        for (int x = 1; x <= p.imax; ++x) {
          const long c = at(x, y);
          if (!mask[c]) continue;
          Vel[c] = 0;
          Tau[c] = 0;
          for (int n = 0; n < kGMax; ++n) {
            s3(sr, c, n) = 0; s3(sl, c, n) = 0; s3(su, c, n) = 0; s3(sd, c, n) = 0;
          }
          ss[c] = 0;
          if (!(h[c] > wdet)) continue;
          double veltot = 0, vel = 0, qtot = 0, tau = 0, slopetot = 0;
          double tempdir[11] = {};
          double temp_dist[11] = {};
          const double mann = mannings_at(x, y);
          if (!hasgrain[c]) addGS(x, y);
          for (int pp = 1; pp <= 8; pp += 2) {
            const int x2 = x + kDX8[pp], y2 = y + kDY8[pp];
            const long c2 = at(x2, y2);
            if (h[c2] > wdet) {
              if (vel_dir[c * 9 + pp] > 0) {
                vel = vel_dir[c * 9 + pp];
                if (vel > p.max_vel) vel = p.max_vel;
                tempdir[pp] = vel * vel;
                veltot += tempdir[pp];
                qtot += (vel * vel);
                slopetot += ((elev[c] - elev[c2]) / p.DX) * vel;
              }
            }
          }
          if (qtot > 0) {
            vel = (std::sqrt(qtot));
            Vel[c] = vel;
            if (vel > p.max_vel) vel = p.max_vel;
            const double ci = kGravity * (mann * mann) * std::pow(h[c], -0.33);
            if (slopetot > 0) slopetot = 0;
            tau = 1000 * ci * vel * vel * (1 + (1 * (slopetot / vel)));
            Tau[c] = tau;
          }
          if (tau > 0) {
            double d_50 = 0, Fs = 0, Di = 0, graintot = 0;
            double* g = G(c);
            if (p.wilcock == 1) {
              d_50 = d50(c);
              if (d_50 < p.d[1]) d_50 = p.d[1];
              Fs = sand_fraction(c);
              for (int n = 1; n <= kGMax; ++n) graintot += (g[n]);
            }
            double temptot1 = 0;
            for (int n = 1; n <= kGMax - 1; ++n) {
              Di = p.d[n];
              if (p.wilcock == 1) {
                const double tau_ri =
                    (0.021 + (0.015 * std::exp(-20 * Fs))) * (rho * kGravity * d_50) *
                    std::pow((Di / d_50), (0.67 / (1 + std::exp(1.5 - (Di / d_50)))));
                const double U_star = std::pow(tau / rho, 0.5);
                const double Fi = g[n] / graintot;
                double Wi_star;
                if ((tau / tau_ri) < 1.35)
                  Wi_star = 0.002 * std::pow(tau / tau_ri, 7.5);
                else
                  Wi_star = 14 * std::pow(1 - (0.894 / std::pow(tau / tau_ri, 0.5)), 4.5);
                temp_dist[n] = mult * time_step *
                               ((Fi * (U_star * U_star * U_star)) / ((2.65 - 1) * kGravity)) *
                               Wi_star / p.DX;
              }
              if (p.einstein == 1) {
                temp_dist[n] =
                    mult * time_step *
                    (40 * std::pow((1 / (((2650 - 1000) * Di) / (tau / kGravity))), 3)) /
                    std::sqrt(1000 / ((2250 - 1000) * kGravity * (Di * Di * Di))) / p.DX;
              }
              if (temp_dist[n] > g[n]) temp_dist[n] = g[n];
              if (n == 1) {
                if ((temp_dist[n] + V[c]) / h[c] > p.Csuspmax)
                  temp_dist[n] = (h[c] * p.Csuspmax) - V[c];
              }
              if (temp_dist[n] < 0) temp_dist[n] = 0;
              temptot1 += temp_dist[n];
            }
            if (elev[c] - temptot1 <= bedrock[c]) {
              const double elevdiff = elev[c] - bedrock[c];
              const double temptot3 = temptot1;
              temptot1 = 0;
              for (int n = 1; n <= kGMax - 1; ++n) {
                if (elev[c] <= bedrock[c]) {
                  temp_dist[n] = 0;
                } else {
                  temp_dist[n] = elevdiff * (temp_dist[n] / temptot3);
                  if (temp_dist[n] < 0) temp_dist[n] = 0;
                }
                temptot1 += temp_dist[n];
              }
              if (tau > p.bedrock_erosion_threshold) {
                const double amount = std::pow(p.bedrock_erosion_rate * tau, 1.5) *
                                      time_step * mult * 0.000000317;
                bedrock[c] -= amount;
                for (int n2 = 1; n2 <= kGMax - 1; ++n2) g[n2] += amount * p.dprop[n2];
              }
            }
            (void)veg0;
            (void)veg1;
            if (temptot1 > tempbmax) tempbmax = temptot1;
            if (temptot1 > 0) {
              for (int pp = 1; pp <= 8; pp += 2) {
                const int x2 = x + kDX8[pp], y2 = y + kDY8[pp];
                const long c2 = at(x2, y2);
                if (h[c2] > wdet) {
                  if (!hasgrain[c2]) addGS(x2, y2);
                  double factor = 0;
                  if (vel_dir[c * 9 + pp] > 0) factor += 0.75 * tempdir[pp] / veltot;
                  for (int n = 1; n <= kGMax - 1; ++n) {
                    if (temp_dist[n] > 0) {
                      if (n == 1) {
                        ss[c] = temp_dist[n];
                      } else {
                        switch (pp) {
                          case 1: s3(su, c, n) = temp_dist[n] * factor; break;
                          case 3: s3(sr, c, n) = temp_dist[n] * factor; break;
                          case 5: s3(sd, c, n) = temp_dist[n] * factor; break;
                          case 7: s3(sl, c, n) = temp_dist[n] * factor; break;
                        }
                      }
                    }
                  }
                }
              }
            }
          }
        }
      }
      if (tempbmax > p.ERODEFACTOR) time_step *= (p.ERODEFACTOR / tempbmax) * 0.5;
    } while (tempbmax > p.ERODEFACTOR);

    std::vector<double> etot(static_cast<size_t>(ni) * nj, 0.0), etot3(etot.size(), 0.0);
    const double EF = p.ERODEFACTOR;
    const double cle = p.chann_lateral_erosion, lc = p.lateral_constant;
    const double vlr = p.veg_lat_restriction;
    // deposition, settling and x-direction lateral slide (LCM.cpp:4570-4682)
    for (int y = 2; y < p.jmax; ++y) {
      for (int x = 1; x <= p.imax; ++x) {
        const long c = at(x, y);
        if (!mask[c]) continue;
        if (!(h[c] > wdet && x < p.imax && x > 1)) continue;
        if (!hasgrain[c]) addGS(x, y);
        for (int n = 1; n <= kGMax - 1; ++n) {
          if (n == 1) {
            V[c] += ss[c];
            G(c)[n] -= ss[c];
            etot[c] -= ss[c];
            double coeff = (p.fallVelocity[n] * time_step) / h[c];
            if (coeff > 1) coeff = 1;
            double Vpdrop = coeff * V[c];
            if (Vpdrop > 0.001) Vpdrop = 0.001;
            G(c)[n] += Vpdrop;
            etot[c] += Vpdrop;
            V[c] -= Vpdrop;
          } else {
            const double val1 = (s3(su, c, n) + s3(sr, c, n) + s3(sd, c, n) + s3(sl, c, n));
            const double val2 = (s3(su, at(x, y + 1), n) + s3(sd, at(x, y - 1), n) +
                                 s3(sl, at(x + 1, y), n) + s3(sr, at(x - 1, y), n));
            G(c)[n] += val2 - val1;
            etot[c] += val2 - val1;
            etot3[c] += val1;
          }
        }
        elev[c] += etot[c];
        if (etot[c] < 0) sort_active(x, y);
        if (etot3[c] > 0) {
          double elev_update = 0;
          const long cu = at(x - 1, y), cd = at(x + 1, y);
          if (elev[cu] > elev[c] && x > 2) {
            double amt = 0;
            if (h[cu] < wdet)
              amt = mult * lc * Tau[c] * 0.0 * time_step / p.DX;
            else
              amt = cle * etot3[c] * (elev[cu] - elev[c]) / p.DX * 0.1;
            if (amt > 0) {
              amt *= 1 - (veg1 * (1 - vlr));
              if ((elev[cu] - amt) < bedrock[cu] || x - 1 == 1) amt = 0;
              if (amt > EF * 0.1) amt = EF * 0.1;
              elev_update += amt;
              elev[cu] -= amt;
              slide_GS(x - 1, y, amt, x, y);
            }
          }
          if (elev[cd] > elev[c] && x < p.imax - 1) {
            double amt = 0;
            if (h[cd] < wdet)
              amt = mult * lc * Tau[c] * 0.0 * time_step / p.DX;
            else
              amt = cle * etot3[c] * (elev[cd] - elev[c]) / p.DX * 0.1;
            if (amt > 0) {
              amt *= 1 - (veg1 * (1 - vlr));
              if ((elev[cd] - amt) < bedrock[cd] || x + 1 == p.imax) amt = 0;
              if (amt > EF * 0.1) amt = EF * 0.1;
              elev_update += amt;
              elev[cd] -= amt;
              slide_GS(x + 1, y, amt, x, y);
            }
          }
          elev[c] += elev_update;
        }
      }
    }
    // y-direction lateral slide (LCM.cpp:4684-4747)
    for (int y = 2; y < p.jmax; ++y) {
      for (int x = 1; x <= p.imax; ++x) {
        const long c = at(x, y);
        if (!mask[c]) continue;
        if (!(etot3[c] > 0)) continue;
        double elev_update = 0;
        const long cl = at(x, y - 1), cr = at(x, y + 1);
        if (elev[cl] > elev[c]) {
          double amt = 0;
          if (h[cl] < wdet)
            amt = mult * lc * Tau[c] * 0.0 * time_step / p.DX;
          else
            amt = cle * etot3[c] * (elev[cl] - elev[c]) / p.DX * 0.1;
          if (amt > 0) {
            amt *= 1 - (veg1 * (1 - vlr));
            if ((elev[cl] - amt) < bedrock[cl] || y - 1 == 1) amt = 0;
            if (amt > EF * 0.1) amt = EF * 0.1;
            elev_update += amt;
            elev[cl] -= amt;
            slide_GS(x, y - 1, amt, x, y);
          }
        }
        if (elev[cr] > elev[c]) {
          double amt = 0;
          if (h[cr] < wdet)
            amt = mult * lc * Tau[c] * 0.0 * time_step / p.DX;
          else
            amt = cle * etot3[c] * (elev[cr] - elev[c]) / p.DX * 0.1;
          if (amt > 0) {
            amt *= 1 - (veg1 * (1 - vlr));
            if ((elev[cr] - amt) < bedrock[cr] || y + 1 == p.jmax) amt = 0;
            if (amt > EF * 0.1) amt = EF * 0.1;
            elev_update += amt;
            elev[cr] -= amt;
            slide_GS(x, y + 1, amt, x, y);
          }
        }
        elev[c] += elev_update;
      }
    }
    // sediment leaving the four edges (LCM.cpp:4750-4827).  gtot2 is float;
    // with GCC's OpenMP array reduction each loop accumulates into a zeroed
    // private copy that is added to the shared array afterwards.
    float part[20] = {};
    const int I = p.imax, J = p.jmax;
    for (int y = 2; y < J; ++y) {
      if (h[at(I, y)] > wdet || V[at(I, y)] > 0) {
        for (int n = 1; n <= kGMax - 1; ++n) {
          if (n == 1) { part[n] += V[at(I, y)]; V[at(I, y)] = 0; }
          else part[n] += s3(sr, at(I - 1, y), n);
        }
      }
      if (h[at(1, y)] > wdet || V[at(1, y)] > 0) {
        for (int n = 1; n <= kGMax - 1; ++n) {
          if (n == 1) { part[n] += V[at(1, y)]; V[at(1, y)] = 0; }
          else part[n] += s3(sl, at(2, y), n);
        }
      }
    }
    for (int n = 0; n < 20; ++n) { gtot2[n] = gtot2[n] + part[n]; part[n] = 0; }
    for (int x = 2; x < I; ++x) {
      if (h[at(x, J)] > wdet || V[at(x, J)] > 0) {
        for (int n = 1; n <= kGMax - 1; ++n) {
          if (n == 1) { part[n] += V[at(x, J)]; V[at(x, J)] = 0; }
          else part[n] += s3(sd, at(x, J - 1), n);
        }
      }
      if (h[at(x, 1)] > wdet || V[at(x, 1)] > 0) {
        for (int n = 1; n <= kGMax - 1; ++n) {
          if (n == 1) { part[n] += V[at(x, 1)]; V[at(x, 1)] = 0; }
          else part[n] += s3(su, at(x, 2), n);
        }
      }
    }
    for (int n = 0; n < 20; ++n) gtot2[n] = gtot2[n] + part[n];
    sediQ = 0;
    for (int n = 1; n <= kGMax; ++n) {
      sediQ += gtot2[n] * p.DX * p.DX;
      globalsediq += gtot2[n] * p.DX * p.DX;
      sum_grain[n] += gtot2[n] * p.DX * p.DX;
    }
    return tempbmax;
  }

  // LCM.cpp:2277-2294
  // in_channel_landsliding (LCM.cpp:5429-5528; called every 10 iterations by
  // call_channel_landsliding, LCM.cpp:1968-1974): for each column y =
  // 2..jmax-1, the rows x of its active list (down_scan, the mask here) in
  // ascending order, x = imax taken as imax-1 and x = 1 as 2; each moves
  // material in place to every neighbour, in the order below, whose drop
  // (over 1.41 on the diagonals) exceeds tan(failure angle) DX (10000 where
  // the cell is at or below bedrock + active), cut so it stays at bedrock +
  // active, sliding the moved grain with each move.
  void in_channel_landsliding(double failureangle) {
    const double factor = std::tan((failureangle * (3.141592654 / 180))) * p.DX;
    static const int ox[8] = {1, 0, -1, -1, -1, 0, 1, 1};
    static const int oy[8] = {1, 1, 1, 0, -1, -1, -1, 0};
    for (int y = 2; y < p.jmax; ++y) {
      for (int xl = 1; xl <= p.imax; ++xl) {
        if (!mask[at(xl, y)]) continue;
        int x = xl;
        if (x == p.imax) x = p.imax - 1;
        if (x == 1) x = 2;
        const size_t c = at(x, y);
        double wet_factor = factor;
        if (elev[c] <= (bedrock[c] + p.active)) wet_factor = 10000;
        for (int k = 0; k < 8; ++k) {
          const int x2 = x + ox[k], y2 = y + oy[k];
          const size_t nb = at(x2, y2);
          const bool diag = (ox[k] != 0 && oy[k] != 0);
          const double drop = diag ? ((elev[c] - elev[nb]) / 1.41) : (elev[c] - elev[nb]);
          if (drop > wet_factor && elev[nb] > -9999) {
            double diff = drop - wet_factor;
            if ((elev[c] - diff) < (bedrock[c] + p.active)) diff = (elev[c] - (bedrock[c] + p.active));
            elev[c] -= diff;
            elev[nb] += diff;
            slide_GS(x, y, diff, x2, y2);
          }
        }
      }
    }
  }

  // creep (LCM.cpp:5293-5427; called by slope_creep, LCM.cpp:1976-1984 with
  // time = 0.028): every interior cell above its bedrock moves
  // ((z - z_nb) / D) * creep_rate * time / D to each lower neighbour
  // (D = DX on the sides, the constant root = 7.07 on the diagonals,
  // LCM.hpp:437), cut so it stays at its bedrock and at 0, in the order
  // below; moves collect in tempcreep from the unchanged elevations and are
  // added at the end; a cell with grain slides the moved grain with each
  // move (slide_GS, in the sweep's order).
  void creep(double time, double creep_rate) {
    const double root = 7.07;
    static const int ox[8] = {0, 1, 1, 1, 0, -1, -1, -1};
    static const int oy[8] = {-1, -1, 0, 1, 1, 1, 0, -1};
    std::vector<double> tempcreep(elev.size(), 0.0);
    for (int x = 2; x < p.imax; ++x) {
      for (int y = 2; y < p.jmax; ++y) {
        const size_t c = at(x, y);
        if (!(elev[c] > bedrock[c])) continue;
        for (int k = 0; k < 8; ++k) {
          const int x2 = x + ox[k], y2 = y + oy[k];
          const size_t nb = at(x2, y2);
          if (elev[nb] < elev[c] && elev[nb] > -9999) {
            const double D = (ox[k] != 0 && oy[k] != 0) ? root : p.DX;
            double temp = ((elev[c] - elev[nb]) / D) * creep_rate * time / D;
            if ((elev[c] - temp) < bedrock[c]) temp = elev[c] - bedrock[c];
            if (temp < 0) temp = 0;
            tempcreep[c] -= temp;
            tempcreep[nb] += temp;
            if (hasgrain[c]) slide_GS(x, y, temp, x2, y2);
          }
        }
      }
    }
    for (int x = 1; x <= p.imax; ++x)
      for (int y = 1; y <= p.jmax; ++y) elev[at(x, y)] += tempcreep[at(x, y)];
  }

  // global_landsliding (LCM.cpp:5529-5654; called by call_global_landsliding,
  // LCM.cpp:1986-1996): row-major sweeps over the interior (x = 2..imax-1
  // outer, y = 2..jmax-1 inner).  Each cell, with wet_factor = tan(failure
  // angle) DX (10 DX where the cell is at or below bedrock + active), moves
  // material to each of its 8 neighbours in the order below whose drop (over
  // 1.41 on the diagonals) exceeds wet_factor: diff = drop - wet_factor, cut
  // so the cell stays at bedrock + active, capped at ERODEFACTOR; total sums
  // the diffs.  Sweeps repeat while total > 0, at most 200.  The dunes
  // option (sand layer) is not part of this restatement.  Returns the number
  // of sweeps.
  int global_landsliding(double failureangle) {
    const double factor = std::tan((failureangle * (3.141592654 / 180))) * p.DX;
    static const int ox[8] = {1, 0, -1, -1, -1, 0, 1, 1};
    static const int oy[8] = {1, 1, 1, 0, -1, -1, -1, 0};
    int inc = 0;
    double total;
    do {
      total = 0;
      inc++;
      for (int x = 2; x < p.imax; ++x) {
        for (int y = 2; y < p.jmax; ++y) {
          const size_t c = at(x, y);
          double wet_factor = factor;
          if (elev[c] <= (bedrock[c] + p.active)) wet_factor = 10 * p.DX;
          for (int k = 0; k < 8; ++k) {
            const size_t nb = at(x + ox[k], y + oy[k]);
            const bool diag = (ox[k] != 0 && oy[k] != 0);
            const double drop = diag ? ((elev[c] - elev[nb]) / 1.41) : (elev[c] - elev[nb]);
            if (drop > wet_factor && elev[nb] > -9999) {
              double diff = drop - wet_factor;
              if ((elev[c] - diff) < (bedrock[c] + p.active)) diff = (elev[c] - (bedrock[c] + p.active));
              if (diff > p.ERODEFACTOR) diff = p.ERODEFACTOR;
              elev[c] -= diff;
              elev[nb] += diff;
              total += diff;
            }
          }
        }
      }
    } while (total > 0 && inc < 200);
    return inc;
  }

  void call_erosion() {
    if (counter >= erode_call) {
      erode_mult = trunc_int(p.ERODEFACTOR / erode(erode_mult));
      if (erode_mult < 1) erode_mult = 1;
      if (erode_mult > 5) erode_mult = 5;
      erode_call = counter + erode_mult;
    }
  }

  // ------------------------------------------------------- loop body ------
  // main.cpp:139-204 (hydro processes + erosion; the erosion-mode hillslope
  // processes are not part of this restatement)
  bool iterate_once(double* tt, double* cy, double* dtf) {
    previous = cycle;                                        // LCM.cpp:1873-1877
    old_cycle = std::fmod(cycle, p.output_file_save_interval);
    iod = std::fabs(waterinput - waterOut);                  // LCM.cpp:1879-1882
    set_maximum_timestep();
    counter++;                                               // LCM.cpp:1925-1933
    cycle += time_step / 60;
    new_cycle = std::fmod(cycle, p.output_file_save_interval);
    reach_water_and_sediment_input();                        // main.cpp:154
    waterinput = 0;                                          // LCM.cpp:2027-2039
    if (p.spatially_complex_rainfall) catchment_water_input_grid(get_flow_timestep());
    else catchment_water_input_and_hydrology(get_flow_timestep());
    if (dtf) *dtf = get_flow_timestep();
    flow_route();
    depth_update();
    if ((counter % p.scan_interval) == 0) scan_area();      // LCM.cpp:2019-2025
    if (!p.hydro_only) call_erosion();
    water_flux_out();
    if (!p.hydro_only) {  // main.cpp:185-187 with main.cpp:127-132's intervals
      if ((counter % 10) == 0) in_channel_landsliding(p.failureangle);  // LCM.cpp:1968-1974
      if (cycle > creep_time) {                                          // LCM.cpp:1976-1984
        creep_time += 14400;
        creep(0.028, p.creep_rate);
      }
      if (cycle > creep_time2) {                                         // LCM.cpp:1986-1996
        creep_time2 += 1440;
        global_landsliding(p.failureangle);
      }
    }
    temptotal = temptot;                                     // LCM.cpp:1952-1966
    output_data();
    if (tt) *tt = temptot;
    if (cy) *cy = cycle;
    return !(cycle / 60 < p.maxcycle);
  }
};

extern "C" {

orc_model* orc_create(const orc_params* p) {
  orc_model* m = new orc_model;
  m->p = *p;
  if (m->p.nthreads <= 0) m->p.nthreads = 1;
  if (m->p.scan_interval <= 0) m->p.scan_interval = 5;
  m->ni = p->imax + 2;
  m->nj = p->jmax + 2;
  const size_t N = static_cast<size_t>(m->ni) * m->nj;
  m->elev.assign(N, -9999);
  m->bedrock.assign(N, -9999);
  for (auto* v : {&m->h, &m->qx, &m->qy, &m->qxs, &m->qys, &m->V, &m->nman,
                  &m->area, &m->area_depth, &m->Tau, &m->Vel, &m->ss})
    v->assign(N, 0.0);
  m->vel_dir.assign(N * 9, 0.0);
  if (!p->hydro_only) {
    m->grain.assign(N * 11, 0.0);
    m->strata.assign(N * kStrataLayers * kStrataFr, 0.0);
    for (auto* v : {&m->sr, &m->sl, &m->su, &m->sd}) v->assign(N * 10, 0.0);
  }
  m->mask.assign(N, 0);
  m->hasgrain.assign(N, 0);
  m->gindex.assign(N, -9999);
  m->rfarea.assign(N, 1);
  const int rf = p->rfnum;
  m->j.assign(rf + 1, 0.000000001);
  m->jo.assign(rf + 1, 0.000000001);
  m->j_mean.assign(rf + 1, 0.0);
  m->old_j_mean.assign(rf + 1, 0.0);
  m->new_j_mean.assign(rf + 1, 0.0);
  m->nActual.assign(rf + 1, 0);
  m->input_counter.assign(rf + 1, 0);
  if (p->spatially_complex_rainfall) {
    m->rj.assign(N, 0.000000001);
    m->rjo.assign(N, 0.000000001);
    m->rj_mean.assign(N, 0.0);
    m->rold.assign(N, 0.0);
    m->rnew.assign(N, 0.0);
  }
  m->in_x.assign(static_cast<size_t>(p->imax) * p->jmax + 2, 0);
  m->in_y.assign(static_cast<size_t>(p->imax) * p->jmax + 2, 0);
  for (auto* v : {&m->sum_grain, &m->old_sum_grain, &m->Qg_step, &m->Qg_hour,
                  &m->Qg_over, &m->Qg_last})
    v->assign(kGMax + 1, 0.0);
  m->tx = p->output_file_save_interval;  // LCM.cpp:1599
  return m;
}

void orc_destroy(orc_model* m) { delete m; }

double* orc_field(orc_model* m, const char* name, long* len) {
  std::string n(name);
  std::vector<double>* v = nullptr;
  if (n == "elev") v = &m->elev;
  else if (n == "water_depth") v = &m->h;
  else if (n == "qx") v = &m->qx;
  else if (n == "qy") v = &m->qy;
  else if (n == "qxs") v = &m->qxs;
  else if (n == "qys") v = &m->qys;
  else if (n == "Vsusptot") v = &m->V;
  else if (n == "mannings") v = &m->nman;
  else if (n == "area") v = &m->area;
  else if (n == "bedrock") v = &m->bedrock;
  else if (n == "Tau") v = &m->Tau;
  else if (n == "Vel") v = &m->Vel;
  else if (n == "ss") v = &m->ss;
  else if (n == "vel_dir") v = &m->vel_dir;
  else if (n == "grain") v = &m->grain;
  else if (n == "strata") v = &m->strata;
  else if (n == "sr") v = &m->sr;
  else if (n == "sl") v = &m->sl;
  else if (n == "su") v = &m->su;
  else if (n == "sd") v = &m->sd;
  else if (n == "runoff_j_mean") v = &m->rj_mean;
  else if (n == "runoff_new_j_mean") v = &m->rnew;
  if (!v || v->empty()) { *len = 0; return nullptr; }
  *len = static_cast<long>(v->size());
  return v->data();
}

int* orc_field_int(orc_model* m, const char* name, long* len) {
  std::string n(name);
  std::vector<int>* v = nullptr;
  if (n == "mask") v = &m->mask;
  else if (n == "hasgrain") v = &m->hasgrain;
  else if (n == "index") v = &m->gindex;
  else if (n == "rfarea") v = &m->rfarea;
  if (!v) { *len = 0; return nullptr; }
  *len = static_cast<long>(v->size());
  return v->data();
}

void orc_set_reach(orc_model* m, int k, int x, int y, const double* rows, int nrows, int ncols) {
  if (k < 0 || k > 2) return;
  m->reach_x[k] = x;
  m->reach_y[k] = y;
  m->reach_rows[k].assign(nrows, std::vector<float>(ncols, 0.0f));
  for (int r = 0; r < nrows; ++r)
    for (int c = 0; c < ncols; ++c) m->reach_rows[k][r][c] = static_cast<float>(rows[static_cast<long>(r) * ncols + c]);
  if (k + 1 > m->reach_n) m->reach_n = k + 1;
}

void orc_set_rain(orc_model* m, const double* rain, int rows, int cols) {
  m->rain.assign(rows, std::vector<float>(cols, 0.0f));
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) m->rain[r][c] = static_cast<float>(rain[static_cast<long>(r) * cols + c]);
}

int orc_scalar(orc_model* m, const char* name, double* value, int set) {
  std::string n(name);
#define ORC_D(key, field)                                  \
  if (n == key) {                                          \
    if (set) m->field = *value; else *value = m->field;    \
    return 0;                                              \
  }
  ORC_D("time_step", time_step) ORC_D("cycle", cycle) ORC_D("maxdepth", maxdepth)
  ORC_D("temptot", temptot) ORC_D("waterinput", waterinput) ORC_D("waterOut", waterOut)
  ORC_D("input_output_difference", iod) ORC_D("baseflow", baseflow)
  ORC_D("time_1", time_1) ORC_D("previous", previous) ORC_D("erode_mult", erode_mult)
  ORC_D("creep_time", creep_time) ORC_D("creep_time2", creep_time2)
  ORC_D("erode_call", erode_call) ORC_D("globalsediq", globalsediq) ORC_D("sediQ", sediQ)
  ORC_D("new_cycle", new_cycle) ORC_D("old_cycle", old_cycle)
  ORC_D("Qw_newvol", Qw_newvol) ORC_D("Jw_newvol", Jw_newvol) ORC_D("Jw_overvol", Jw_overvol)
  ORC_D("Jw_hour", Jw_hour)
#undef ORC_D
  if (n == "counter") {
    if (set) m->counter = static_cast<int>(*value); else *value = m->counter;
    return 0;
  }
  if (n == "totalinputpoints") {
    if (set) m->totalinputpoints = static_cast<int>(*value); else *value = m->totalinputpoints;
    return 0;
  }
  if (n == "grain_array_tot") {
    if (set) m->grain_array_tot = static_cast<int>(*value); else *value = m->grain_array_tot;
    return 0;
  }
  if (n.rfind("j_mean", 0) == 0 || n.rfind("new_j_mean", 0) == 0 || n.rfind("sum_grain", 0) == 0) {
    // "j_mean1" etc.
    const size_t k = n.find_first_of("0123456789");
    if (k == std::string::npos) return -1;
    const int idx = std::stoi(n.substr(k));
    std::vector<double>* v = (n[0] == 'j') ? &m->j_mean : (n[0] == 'n') ? &m->new_j_mean : &m->sum_grain;
    if (idx < 0 || idx >= static_cast<int>(v->size())) return -1;
    if (set) (*v)[idx] = *value; else *value = (*v)[idx];
    return 0;
  }
  return -1;
}

// load_data / initialise_rainfall_runoff / initialise_drainage_area
// (LCM.cpp:2813-2843, 2041-2069; main.cpp:116-120).
void orc_initialise(orc_model* m) {
  const orc_params& p = m->p;
  std::fill(m->nActual.begin(), m->nActual.end(), 0);
  for (int i = 1; i < p.imax; ++i)
    for (int j = 1; j < p.jmax; ++j)
      if (m->elev[m->at(i, j)] > -9999) m->nActual[m->rfarea[m->at(i, j)]]++;
  m->time_1 = m->cycle;  // set_time_counters, LCM.cpp:1862-1871
  m->creep_time = m->cycle;
  m->creep_time2 = m->cycle;
  if (p.spatially_complex_rainfall) {
    // initialise_rainfall_runoff and initialise_drainage_area (LCM.cpp:2041-2069):
    // the per-cell TOPMODEL only; no drainage area or input points
    m->topmodel_runoff_grid(1.0);
    return;
  }
  m->topmodel_runoff(1.0);
  m->drainage_area();
  m->get_catchment_input_points();
}

void orc_flow_route(orc_model* m) { m->flow_route(); }
double orc_depth_update(orc_model* m) { return m->depth_update(); }
void orc_scan_area(orc_model* m) { m->scan_area(); }
double orc_water_flux_out(orc_model* m) { return m->water_flux_out(); }
double orc_get_flow_timestep(orc_model* m) { return m->get_flow_timestep(); }
void orc_set_maximum_timestep(orc_model* m) { m->set_maximum_timestep(); }
double orc_erode(orc_model* m, double mult) { return m->erode(mult); }
void orc_call_erosion(orc_model* m) { m->call_erosion(); }
void orc_creep(orc_model* m, double time, double creep_rate) { m->creep(time, creep_rate); }
void orc_in_channel_landsliding(orc_model* m, double failureangle) {
  m->in_channel_landsliding(failureangle);
}
int orc_global_landsliding(orc_model* m, double failureangle) {
  return m->global_landsliding(failureangle);
}
void orc_drainage_area(orc_model* m) { m->drainage_area(); }
void orc_reach_input(orc_model* m) { m->reach_water_and_sediment_input(); }
int orc_input_points(orc_model* m) { return m->get_catchment_input_points(); }

double orc_grain_call(orc_model* m, int which, int x, int y, double amount, int x2, int y2) {
  switch (which) {
    case 0: m->sort_active(x, y); return 0;
    case 1: m->addGS(x, y); return m->grain_array_tot;
    case 2: return m->sand_fraction(m->at(x, y));
    case 3: return m->d50(m->at(x, y));
    case 4: m->slide_GS(x, y, amount, x2, y2); return 0;
    default: return -1;
  }
}

int orc_iterate(orc_model* m, int n, double* tt, double* cy, double* dtf) {
  int k = 0;
  for (; k < n; ++k) {
    const bool done = m->iterate_once(tt ? tt + k : nullptr, cy ? cy + k : nullptr,
                                      dtf ? dtf + k : nullptr);
    if (done) { ++k; break; }
  }
  return k;
}

int orc_series(orc_model* m, double* out, int max_rows) {
  const int rows = static_cast<int>(m->series.size() / 14);
  const int nc = std::min(rows, max_rows);
  if (out && nc > 0) std::memcpy(out, m->series.data(), sizeof(double) * 14 * nc);
  return rows;
}

}  // extern "C"
