// C-ABI entry points of libhc.so (include/hc.h): context lifetime, host <->
// device layout conversion, and the per-method calls of the hot path.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "hc_internal.hpp"

int hc_fail(hc_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

namespace {

int planes_of(int field) {
  switch (field) {
    case HC_FIELD_VEL_DIR: return 9;
    case HC_FIELD_GRAIN: return 11;
    case HC_FIELD_STRATA: return 90;
    case HC_FIELD_SR: case HC_FIELD_SL: case HC_FIELD_SU: case HC_FIELD_SD: return 10;
    default: return 1;
  }
}

bool is_sediment_field(int f) {
  return f == HC_FIELD_VSUSP || f == HC_FIELD_QXS || f == HC_FIELD_QYS || f == HC_FIELD_VEL_DIR ||
         f == HC_FIELD_GRAIN || f == HC_FIELD_STRATA || f == HC_FIELD_HASGRAIN ||
         f == HC_FIELD_BEDROCK || f == HC_FIELD_TAU || f == HC_FIELD_VEL || f == HC_FIELD_SS ||
         f == HC_FIELD_SR || f == HC_FIELD_SL || f == HC_FIELD_SU || f == HC_FIELD_SD;
}

template <typename T>
int dalloc(hc_ctx* c, T** p, size_t n) {
  HC_HIP(c, hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)));
  HC_HIP(c, hipMemsetAsync(*p, 0, n * sizeof(T), c->stream));
  return HC_OK;
}

// Copy one plane between a host (NR x NC, stride `hstride` elements between
// consecutive cells, offset `hoff`) array of doubles and a device plane.
int put_plane(hc_ctx* c, void* dev, bool dev_float, const double* host, int hstride, int hoff) {
  const size_t n = c->ncell;
  if (dev_float) {
    std::vector<float> tmp(n, 0.0f);
    for (int x = 0; x < c->NR; ++x)
      for (int y = 0; y < c->NC; ++y)
        tmp[static_cast<size_t>(x) * c->P + y] =
            static_cast<float>(host[(static_cast<size_t>(x) * c->NC + y) * hstride + hoff]);
    HC_HIP(c, hipMemcpy(dev, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
  } else {
    if (hstride == 1) {
      HC_HIP(c, hipMemcpy2D(dev, c->P * sizeof(double), host, c->NC * sizeof(double),
                            c->NC * sizeof(double), c->NR, hipMemcpyHostToDevice));
    } else {
      std::vector<double> tmp(n, 0.0);
      for (int x = 0; x < c->NR; ++x)
        for (int y = 0; y < c->NC; ++y)
          tmp[static_cast<size_t>(x) * c->P + y] =
              host[(static_cast<size_t>(x) * c->NC + y) * hstride + hoff];
      HC_HIP(c, hipMemcpy(dev, tmp.data(), n * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  return HC_OK;
}

// Plane k of a sediment buffer (sesz-byte elements).
void* splane(const hc_ctx* c, void* base, size_t k) {
  return static_cast<char*>(base) + k * c->ncell * c->sesz;
}

// Zeroed sediment buffer of `planes` planes of sesz-byte elements.
int salloc(hc_ctx* c, void** p, size_t planes) {
  const size_t bytes = planes * c->ncell * c->sesz;
  HC_HIP(c, hipMalloc(p, bytes));
  HC_HIP(c, hipMemsetAsync(*p, 0, bytes, c->stream));
  return HC_OK;
}

// Fixed-point elevations: q = round((z - base) * 2^shift) with base the
// middle of the valid range and the largest shift (<= 40) that keeps |q| <=
// 2^30 - 1, so differences of two valid cells fit in an int32; no-data
// (z <= -9999, which the reference's `> -9999` tests exclude) is hc::kZqNodata.
int put_fixed_elev(hc_ctx* c, const double* host) {
  double lo = 0.0, hi = 0.0;
  bool any = false;
  for (size_t i = 0; i < static_cast<size_t>(c->NR) * c->NC; ++i) {
    const double z = host[i];
    if (!(z > -9999.0)) continue;
    if (!std::isfinite(z)) return hc_fail(c, HC_ERR_ARG, "upload: non-finite elevation");
    lo = any ? std::min(lo, z) : z;
    hi = any ? std::max(hi, z) : z;
    any = true;
  }
  const double base = 0.5 * (lo + hi);
  const double half = std::max(hi - base, base - lo);
  int shift = 40;
  while (shift > -60 && half * std::ldexp(1.0, shift) > 1073741823.0) --shift;
  std::vector<int32_t> tmp(c->ncell, hc::kZqNodata);
  for (int x = 0; x < c->NR; ++x)
    for (int y = 0; y < c->NC; ++y) {
      const double z = host[static_cast<size_t>(x) * c->NC + y];
      if (z > -9999.0)
        tmp[static_cast<size_t>(x) * c->P + y] =
            static_cast<int32_t>(std::llrint((z - base) * std::ldexp(1.0, shift)));
    }
  HC_HIP(c, hipMemcpy(c->zq, tmp.data(), c->ncell * sizeof(int32_t), hipMemcpyHostToDevice));
  c->zq_base = base;
  c->zq_shift = shift;
  return HC_OK;
}

int get_plane(hc_ctx* c, const void* dev, bool dev_float, double* host, int hstride, int hoff) {
  const size_t n = c->ncell;
  if (dev_float) {
    std::vector<float> tmp(n);
    HC_HIP(c, hipMemcpy(tmp.data(), dev, n * sizeof(float), hipMemcpyDeviceToHost));
    for (int x = 0; x < c->NR; ++x)
      for (int y = 0; y < c->NC; ++y)
        host[(static_cast<size_t>(x) * c->NC + y) * hstride + hoff] =
            tmp[static_cast<size_t>(x) * c->P + y];
  } else {
    if (hstride == 1) {
      HC_HIP(c, hipMemcpy2D(host, c->NC * sizeof(double), dev, c->P * sizeof(double),
                            c->NC * sizeof(double), c->NR, hipMemcpyDeviceToHost));
    } else {
      std::vector<double> tmp(n);
      HC_HIP(c, hipMemcpy(tmp.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost));
      for (int x = 0; x < c->NR; ++x)
        for (int y = 0; y < c->NC; ++y)
          host[(static_cast<size_t>(x) * c->NC + y) * hstride + hoff] =
              tmp[static_cast<size_t>(x) * c->P + y];
    }
  }
  return HC_OK;
}

int put_int_plane(hc_ctx* c, int32_t* dev, const int32_t* host) {
  std::vector<int32_t> tmp(c->ncell, 0);
  for (int x = 0; x < c->NR; ++x)
    for (int y = 0; y < c->NC; ++y)
      tmp[static_cast<size_t>(x) * c->P + y] = host[static_cast<size_t>(x) * c->NC + y];
  HC_HIP(c, hipMemcpy(dev, tmp.data(), c->ncell * sizeof(int32_t), hipMemcpyHostToDevice));
  return HC_OK;
}

// Grain rows from the reference's row numbers (index[x][y] of the padded
// grid, -9999 for none): the has-grain flags, and creation keys of epoch 0
// that keep each row's number.  rows_total is grain_array_tot, the number
// the next new row follows.
int put_row_index(hc_ctx* c, const int32_t* index, long long rows_total) {
  std::vector<int32_t> has(c->ncell, 0);
  std::vector<uint64_t> key(c->ncell, hc::kRowNone);
  std::vector<int32_t> seen;
  for (int x = 0; x < c->NR; ++x)
    for (int y = 0; y < c->NC; ++y) {
      const int32_t v = index[static_cast<size_t>(x) * c->NC + y];
      if (v == -9999) continue;
      if (v < 1 || v > rows_total)
        return hc_fail(c, HC_ERR_ARG, "grain index: a row number outside 1..rows_total");
      const size_t d = static_cast<size_t>(x) * c->P + y;
      has[d] = 1;
      key[d] = static_cast<uint64_t>(v);
      seen.push_back(v);
    }
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
    return hc_fail(c, HC_ERR_ARG, "grain index: two cells share a row number");
  HC_HIP(c, hipMemcpy(c->hasgrain, has.data(), c->ncell * sizeof(int32_t), hipMemcpyHostToDevice));
  HC_HIP(c, hipMemcpy(c->rowkey, key.data(), c->ncell * sizeof(uint64_t), hipMemcpyHostToDevice));
  c->rows_base = rows_total;
  return HC_OK;
}

int get_int_plane(hc_ctx* c, const int32_t* dev, int32_t* host) {
  std::vector<int32_t> tmp(c->ncell);
  HC_HIP(c, hipMemcpy(tmp.data(), dev, c->ncell * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int x = 0; x < c->NR; ++x)
    for (int y = 0; y < c->NC; ++y)
      host[static_cast<size_t>(x) * c->NC + y] = tmp[static_cast<size_t>(x) * c->P + y];
  return HC_OK;
}

int read_scalar(hc_ctx* c, const void* dev, double* out, bool as_bits) {
  HC_HIP(c, hipMemcpyAsync(c->h_pinned, dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HC_HIP(c, hipStreamSynchronize(c->stream));
  if (as_bits) {
    unsigned long long b;
    std::memcpy(&b, c->h_pinned, sizeof(b));
    double v;
    std::memcpy(&v, &b, sizeof(v));
    *out = v;
  } else {
    *out = c->h_pinned[0];
  }
  return HC_OK;
}

// Spatially complex rainfall's water input (LCM.cpp:3559-3573): every cell
// 1..imax-1, 1..jmax-1 takes its runoff over the step, capped at the
// erosion limit.  The runoff is the TOPMODEL interpolation of calchydrograph
// (LCM.cpp:3728-3739) formed per cell from the planes of the last TOPMODEL
// update, in the reference's operation order (no contraction here, so it is
// the value the host forms for its sums).
template <class T>
__global__ __launch_bounds__(256) void k_runoff_add(int imax, int jmax, int P, double t, double fts,
                                                     double cap, const double* __restrict__ jold,
                                                     const double* __restrict__ jnew,
                                                     T* __restrict__ h) {
  const int y = 1 + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int x = 1 + static_cast<int>(blockIdx.y);
  if (y >= jmax || x >= imax) return;
  const size_t c = static_cast<size_t>(x) * P + y;
  const double jm = jold[c] + (((jnew[c] - jold[c]) / 2) * (2 - t));
  double amt = jm * fts;
  if (amt > cap) amt = cap;
  h[c] = static_cast<T>(static_cast<double>(h[c]) + amt);
}

// Staging of hc_add_point_inputs_noalloc: the x list, the y list (int32) and
// the amounts (double) of up to HC_NOALLOC_MAX_POINTS points.  Each device the
// library runs on holds its own copy until the process ends.  In host code
// `scratch` names a host-side stand-in: only the address resolved by
// hipGetSymbolAddress (scratch_on_device) may reach a kernel or a copy.
constexpr size_t kScratchBytes = static_cast<size_t>(HC_NOALLOC_MAX_POINTS) * 16;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kScratchDevices] = {};

// This device's address of `scratch`, resolved once per device with the
// device current (the caller makes it current).  Null on failure.
unsigned char* scratch_on_device(int device) {
  if (device < 0 || device >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch_dev[device] = static_cast<unsigned char*>(p);
  }
  return g_scratch_dev[device];
}

}  // namespace

extern "C" {

int hc_create(hc_ctx** out, const hc_grid* grid, const hc_params* params) {
  if (!out || !grid || !params) return HC_ERR_ARG;
  *out = nullptr;
  if (grid->imax < 1 || grid->jmax < 1 || !(grid->DX > 0)) return HC_ERR_ARG;
  if (grid->precision != HC_FP64 && grid->precision != HC_FP32) return HC_ERR_ARG;
  if (grid->sediment && grid->precision == HC_FP32 && params->elev_fixed) return HC_ERR_ARG;
  // grain-row creation keys hold a cell coordinate in 15 bits
  if (grid->sediment && (grid->imax + 1 >= (1 << hc::kRowMinorBits) ||
                         grid->jmax + 1 >= (1 << hc::kRowMinorBits)))
    return HC_ERR_ARG;
  hc_ctx* c = new (std::nothrow) hc_ctx;
  if (!c) return HC_ERR_ARG;
  c->g = *grid;
  c->prm = *params;
  c->prec = grid->precision;
  c->esz = (c->prec == HC_FP64) ? 8 : 4;
  c->sesz = c->esz;
  c->NR = grid->imax + 2;
  c->NC = grid->jmax + 2;
  c->P = ((c->NC + 63) / 64) * 64;
  c->W = c->P / 64;
  c->ncell = static_cast<size_t>(c->NR) * c->P;
  int rc = HC_OK;
  auto fail = [&](int r) {
    hc_destroy(c);
    return r;
  };
  if (hipGetDevice(&c->device) != hipSuccess) c->device = 0;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return HC_ERR_HIP;
  }
  const size_t n = c->ncell;
  for (int k = 0; k < 2; ++k) {
    if (c->prec == HC_FP64) {
      if ((rc = dalloc(c, reinterpret_cast<double**>(&c->h[k]), n)) ||
          (rc = dalloc(c, reinterpret_cast<double**>(&c->qx[k]), n)) ||
          (rc = dalloc(c, reinterpret_cast<double**>(&c->qy[k]), n)))
        return fail(rc);
    } else {
      if ((rc = dalloc(c, reinterpret_cast<float**>(&c->h[k]), n)) ||
          (rc = dalloc(c, reinterpret_cast<float**>(&c->qx[k]), n)) ||
          (rc = dalloc(c, reinterpret_cast<float**>(&c->qy[k]), n)))
        return fail(rc);
    }
  }
  if (c->prec == HC_FP64 || grid->sediment) {  // FP64 elevations (both sediment forms)
    if ((rc = dalloc(c, reinterpret_cast<double**>(&c->z), n))) return fail(rc);
    if (params->spatially_var_mannings) {
      rc = (c->prec == HC_FP64) ? dalloc(c, reinterpret_cast<double**>(&c->nman), n)
                                : dalloc(c, reinterpret_cast<float**>(&c->nman), n);
      if (rc) return fail(rc);
    }
    if ((rc = hc_launch_fill(c, c->z, n, -9999.0, false))) return fail(rc);
  } else {
    if ((rc = dalloc(c, reinterpret_cast<float**>(&c->z), n))) return fail(rc);
    if ((rc = dalloc(c, &c->z64, n))) return fail(rc);
    if (params->elev_fixed) {
      if ((rc = dalloc(c, &c->zq, n))) return fail(rc);
      if (hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(c->zq), hc::kZqNodata, n, c->stream) !=
          hipSuccess)
        return fail(HC_ERR_HIP);
    }
    if (params->spatially_var_mannings &&
        (rc = dalloc(c, reinterpret_cast<float**>(&c->nman), n)))
      return fail(rc);
    if ((rc = hc_launch_fill(c, c->z, n, -9999.0, true))) return fail(rc);
    if ((rc = hc_launch_fill(c, c->z64, n, -9999.0, false))) return fail(rc);
  }
  if ((rc = dalloc(c, &c->mask, static_cast<size_t>(c->NR) * c->W))) return fail(rc);
  if (grid->sediment) {
    if ((rc = salloc(c, &c->V, 1)) || (rc = salloc(c, &c->qxs, 1)) ||
        (rc = salloc(c, &c->qys, 1)) || (rc = salloc(c, &c->veldir, 4)) ||
        (rc = salloc(c, &c->grain, 11)) || (rc = salloc(c, &c->strata, 90)) ||
        (rc = dalloc(c, &c->hasgrain, n)) || (rc = dalloc(c, &c->needgrain, n)) ||
        (rc = dalloc(c, &c->rowkey, n)) ||
        (rc = dalloc(c, &c->bedrock, n)) || (rc = salloc(c, &c->Tau, 1)) ||
        (rc = salloc(c, &c->Vel, 1)) || (rc = salloc(c, &c->ss, 1)) ||
        (rc = salloc(c, &c->sdir, 40)) || (rc = salloc(c, &c->etot, 1)) ||
        (rc = salloc(c, &c->etot3, 1)) || (rc = dalloc(c, &c->gtot, 20)))
      return fail(rc);
    if ((rc = hc_launch_fill(c, c->bedrock, n, -9999.0, false))) return fail(rc);
    // no cell has a grain row yet
    if (hipMemsetAsync(c->rowkey, 0xff, n * sizeof(uint64_t), c->stream) != hipSuccess)
      return fail(HC_ERR_HIP);
  }
  if ((rc = dalloc(c, &c->d_red, 8)) || (rc = dalloc(c, &c->d_partials, 1024)) ||
      (rc = dalloc(c, &c->d_count, 16)) || (rc = dalloc(c, &c->d_out, 8)))
    return fail(rc);
  if (hipHostMalloc(reinterpret_cast<void**>(&c->h_pinned), 8 * sizeof(double)) != hipSuccess)
    return fail(HC_ERR_HIP);
  // the device current here is c->device
  c->scratch = scratch_on_device(c->device);
  if (!c->scratch) return fail(HC_ERR_HIP);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(HC_ERR_HIP);
  *out = c;
  return HC_OK;
}

int hc_destroy(hc_ctx* c) {
  if (!c) return HC_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void* ptrs[] = {c->h[0], c->h[1], c->qx[0], c->qx[1], c->qy[0], c->qy[1], c->z, c->z64, c->zq,
                  c->nman, c->mask, c->V, c->qxs, c->qys, c->veldir, c->grain, c->strata,
                  c->hasgrain, c->needgrain, c->rowkey, c->bedrock, c->Tau, c->Vel, c->ss, c->sdir,
                  c->etot, c->etot3, c->gtot, c->d_red, c->d_partials, c->d_count, c->d_out,
                  c->d_pts, c->d_amt, c->d_ctl, c->d_rec, c->d_top, c->d_area, c->d_rfarea,
                  c->d_pmask, c->d_dpart, c->d_wmax, c->d_wet, c->d_slide_moves, c->d_runoff,
                  c->d_plist, c->d_pamt, c->d_dep};
  for (auto& row : c->loop_graph)
    for (auto& gexec : row)
      if (gexec) (void)hipGraphExecDestroy(gexec);
  for (auto e : c->prof_ev) (void)hipEventDestroy(e);
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->h_pinned) (void)hipHostFree(c->h_pinned);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return HC_OK;
}

const char* hc_last_error(const hc_ctx* c) { return c ? c->err.c_str() : "null context"; }

size_t hc_field_count(const hc_ctx* c, int field) {
  if (!c) return 0;
  return static_cast<size_t>(c->NR) * c->NC * planes_of(field);
}

int hc_upload(hc_ctx* c, int field, const void* host, size_t count) {
  if (!c || !host) return HC_ERR_ARG;
  if (count != hc_field_count(c, field))
    return hc_fail(c, HC_ERR_ARG, "upload: element count does not match the grid");
  if (is_sediment_field(field) && !c->g.sediment)
    return hc_fail(c, HC_ERR_STATE, "upload: sediment field on a hydro-only context");
  HC_HIP(c, hipStreamSynchronize(c->stream));
  const double* d = static_cast<const double*>(host);
  const bool f32 = (c->prec == HC_FP32);
  const bool sf = (c->sesz == 4);  // sediment planes in float
  switch (field) {
    case HC_FIELD_WATER_DEPTH: return put_plane(c, c->h[c->cur], f32, d, 1, 0);
    case HC_FIELD_QX: return put_plane(c, c->qx[c->cur], f32, d, 1, 0);
    case HC_FIELD_QY: return put_plane(c, c->qy[c->cur], f32, d, 1, 0);
    case HC_FIELD_ELEV: {
      int rc = put_plane(c, c->z, f32 && !c->g.sediment, d, 1, 0);
      if (rc == HC_OK && c->z64) rc = put_plane(c, c->z64, false, d, 1, 0);
      if (rc == HC_OK && c->zq) rc = put_fixed_elev(c, d);
      return rc;
    }
    case HC_FIELD_MANNINGS:
      if (!c->nman) return hc_fail(c, HC_ERR_STATE, "upload: context has uniform Manning n");
      return put_plane(c, c->nman, f32, d, 1, 0);
    case HC_FIELD_VSUSP: return put_plane(c, c->V, sf, d, 1, 0);
    case HC_FIELD_QXS: return put_plane(c, c->qxs, sf, d, 1, 0);
    case HC_FIELD_QYS: return put_plane(c, c->qys, sf, d, 1, 0);
    case HC_FIELD_BEDROCK: return put_plane(c, c->bedrock, false, d, 1, 0);
    case HC_FIELD_TAU: return put_plane(c, c->Tau, sf, d, 1, 0);
    case HC_FIELD_VEL: return put_plane(c, c->Vel, sf, d, 1, 0);
    case HC_FIELD_SS: return put_plane(c, c->ss, sf, d, 1, 0);
    case HC_FIELD_VEL_DIR: {
      const int dirs[4] = {1, 3, 5, 7};
      for (int k = 0; k < 4; ++k) {
        int rc = put_plane(c, splane(c, c->veldir, k), sf, d, 9, dirs[k]);
        if (rc) return rc;
      }
      return HC_OK;
    }
    case HC_FIELD_GRAIN:
      for (int k = 0; k < 11; ++k) {
        int rc = put_plane(c, splane(c, c->grain, k), sf, d, 11, k);
        if (rc) return rc;
      }
      return HC_OK;
    case HC_FIELD_STRATA:
      for (int k = 0; k < 90; ++k) {
        int rc = put_plane(c, splane(c, c->strata, k), sf, d, 90, k);
        if (rc) return rc;
      }
      return HC_OK;
    case HC_FIELD_SU: case HC_FIELD_SR: case HC_FIELD_SD: case HC_FIELD_SL: {
      const int dir = (field == HC_FIELD_SU) ? 0 : (field == HC_FIELD_SR) ? 1 : (field == HC_FIELD_SD) ? 2 : 3;
      for (int k = 0; k < 10; ++k) {
        int rc = put_plane(c, splane(c, c->sdir, dir * 10 + k), sf, d, 10, k);
        if (rc) return rc;
      }
      return HC_OK;
    }
    case HC_FIELD_HASGRAIN: {
      // rows numbered as a grain data file of one line per flagged cell in
      // row-major order would number them (hc_grain_index_set gives others)
      const int32_t* f = static_cast<const int32_t*>(host);
      std::vector<int32_t> idx(static_cast<size_t>(c->NR) * c->NC, -9999);
      long long n = 0;
      for (size_t k = 0; k < idx.size(); ++k)
        if (f[k]) idx[k] = static_cast<int32_t>(++n);
      return put_row_index(c, idx.data(), n);
    }
    case HC_FIELD_MASK: {
      const int32_t* m = static_cast<const int32_t*>(host);
      std::vector<uint64_t> bits(static_cast<size_t>(c->NR) * c->W, 0);
      for (int x = 1; x <= c->g.imax; ++x)
        for (int y = 1; y <= c->g.jmax; ++y)
          if (m[static_cast<size_t>(x) * c->NC + y])
            bits[static_cast<size_t>(x) * c->W + (y >> 6)] |= 1ull << (y & 63);
      HC_HIP(c, hipMemcpy(c->mask, bits.data(), bits.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
      return HC_OK;
    }
    default:
      return hc_fail(c, HC_ERR_ARG, "upload: unknown field");
  }
}

int hc_download(hc_ctx* c, int field, void* host, size_t count) {
  if (!c || !host) return HC_ERR_ARG;
  if (count != hc_field_count(c, field))
    return hc_fail(c, HC_ERR_ARG, "download: element count does not match the grid");
  if (is_sediment_field(field) && !c->g.sediment)
    return hc_fail(c, HC_ERR_STATE, "download: sediment field on a hydro-only context");
  HC_HIP(c, hipStreamSynchronize(c->stream));
  double* d = static_cast<double*>(host);
  const bool f32 = (c->prec == HC_FP32);
  const bool sf = (c->sesz == 4);  // sediment planes in float
  switch (field) {
    case HC_FIELD_WATER_DEPTH: return get_plane(c, c->h[c->cur], f32, d, 1, 0);
    case HC_FIELD_QX: return get_plane(c, c->qx[c->cur], f32, d, 1, 0);
    case HC_FIELD_QY: return get_plane(c, c->qy[c->cur], f32, d, 1, 0);
    case HC_FIELD_ELEV: return get_plane(c, c->z64 ? static_cast<void*>(c->z64) : c->z, false, d, 1, 0);
    case HC_FIELD_MANNINGS:
      if (!c->nman) return hc_fail(c, HC_ERR_STATE, "download: context has uniform Manning n");
      return get_plane(c, c->nman, f32, d, 1, 0);
    case HC_FIELD_VSUSP: return get_plane(c, c->V, sf, d, 1, 0);
    case HC_FIELD_QXS: return get_plane(c, c->qxs, sf, d, 1, 0);
    case HC_FIELD_QYS: return get_plane(c, c->qys, sf, d, 1, 0);
    case HC_FIELD_BEDROCK: return get_plane(c, c->bedrock, false, d, 1, 0);
    case HC_FIELD_TAU: return get_plane(c, c->Tau, sf, d, 1, 0);
    case HC_FIELD_VEL: return get_plane(c, c->Vel, sf, d, 1, 0);
    case HC_FIELD_SS: return get_plane(c, c->ss, sf, d, 1, 0);
    case HC_FIELD_VEL_DIR: {
      std::fill(d, d + count, 0.0);
      const int dirs[4] = {1, 3, 5, 7};
      for (int k = 0; k < 4; ++k) {
        int rc = get_plane(c, splane(c, c->veldir, k), sf, d, 9, dirs[k]);
        if (rc) return rc;
      }
      return HC_OK;
    }
    case HC_FIELD_GRAIN:
      for (int k = 0; k < 11; ++k) {
        int rc = get_plane(c, splane(c, c->grain, k), sf, d, 11, k);
        if (rc) return rc;
      }
      return HC_OK;
    case HC_FIELD_STRATA:
      for (int k = 0; k < 90; ++k) {
        int rc = get_plane(c, splane(c, c->strata, k), sf, d, 90, k);
        if (rc) return rc;
      }
      return HC_OK;
    case HC_FIELD_SU: case HC_FIELD_SR: case HC_FIELD_SD: case HC_FIELD_SL: {
      const int dir = (field == HC_FIELD_SU) ? 0 : (field == HC_FIELD_SR) ? 1 : (field == HC_FIELD_SD) ? 2 : 3;
      for (int k = 0; k < 10; ++k) {
        int rc = get_plane(c, splane(c, c->sdir, dir * 10 + k), sf, d, 10, k);
        if (rc) return rc;
      }
      return HC_OK;
    }
    case HC_FIELD_HASGRAIN:
      return get_int_plane(c, c->hasgrain, static_cast<int32_t*>(host));
    case HC_FIELD_MASK: {
      int32_t* m = static_cast<int32_t*>(host);
      std::vector<uint64_t> bits(static_cast<size_t>(c->NR) * c->W, 0);
      HC_HIP(c, hipMemcpy(bits.data(), c->mask, bits.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
      for (int x = 0; x < c->NR; ++x)
        for (int y = 0; y < c->NC; ++y)
          m[static_cast<size_t>(x) * c->NC + y] =
              static_cast<int32_t>((bits[static_cast<size_t>(x) * c->W + (y >> 6)] >> (y & 63)) & 1ull);
      return HC_OK;
    }
    default:
      return hc_fail(c, HC_ERR_ARG, "download: unknown field");
  }
}

int hc_flow_route(hc_ctx* c, double dt) {
  if (!c) return HC_ERR_ARG;
  return hc_launch_flow_route(c, dt);
}

int hc_depth_update(hc_ctx* c, double dt, double* maxdepth) {
  if (!c) return HC_ERR_ARG;
  int rc = hc_launch_depth_update(c, dt);
  if (rc) return rc;
  if (maxdepth) return read_scalar(c, c->d_red, maxdepth, true);
  return HC_OK;
}

int hc_flow_depth(hc_ctx* c, double dt, double* maxdepth) {
  if (!c) return HC_ERR_ARG;
  int rc = hc_launch_flow_depth(c, dt);
  if (rc) return rc;
  if (maxdepth) return read_scalar(c, c->d_red, maxdepth, true);
  return HC_OK;
}

int hc_scan_area(hc_ctx* c) {
  if (!c) return HC_ERR_ARG;
  return hc_launch_scan(c);
}

int hc_water_flux_out(hc_ctx* c, double time_step, double* temptot) {
  if (!c) return HC_ERR_ARG;
  int rc = hc_launch_flux_out(c, time_step);
  if (rc) return rc;
  if (temptot) return read_scalar(c, c->d_out, temptot, false);
  return HC_OK;
}

int hc_add_point_inputs(hc_ctx* c, const int* x, const int* y, const double* amt, int n) {
  if (!c || n < 0 || (n > 0 && (!x || !y || !amt))) return HC_ERR_ARG;
  if (n == 0) return HC_OK;
  if (n > c->pts_cap) {
    HC_HIP(c, hipStreamSynchronize(c->stream));
    if (c->d_pts) (void)hipFree(c->d_pts);
    if (c->d_amt) (void)hipFree(c->d_amt);
    c->d_pts = nullptr;
    c->d_amt = nullptr;
    const int cap = std::max(n, 1024);
    HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_pts), cap * sizeof(int)));
    HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_amt), cap * sizeof(double)));
    c->pts_cap = cap;
  }
  std::vector<int> idx(n);
  for (int k = 0; k < n; ++k) {
    if (x[k] < 0 || x[k] >= c->NR || y[k] < 0 || y[k] >= c->NC)
      return hc_fail(c, HC_ERR_ARG, "add_point_inputs: point outside the padded grid");
    idx[k] = x[k] * c->P + y[k];
  }
  // repeated cells: test-and-set one bit per padded cell, then clear the
  // bits set (O(n), the bitmap kept by the context)
  if (c->pt_bits.size() != static_cast<size_t>(c->NR) * c->W)
    c->pt_bits.assign(static_cast<size_t>(c->NR) * c->W, 0);
  bool repeats = false;
  for (int k = 0; k < n; ++k) {
    uint64_t& w = c->pt_bits[static_cast<size_t>(x[k]) * c->W + (y[k] >> 6)];
    const uint64_t b = 1ull << (y[k] & 63);
    repeats = repeats || (w & b);
    w |= b;
  }
  for (int k = 0; k < n; ++k) c->pt_bits[static_cast<size_t>(x[k]) * c->W + (y[k] >> 6)] = 0;
  HC_HIP(c, hipMemcpyAsync(c->d_pts, idx.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HC_HIP(c, hipMemcpyAsync(c->d_amt, amt, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  // distinct cells: one thread per point; repeated cells: in list order
  int rc = repeats ? hc_launch_points_ordered(c, n, c->d_pts, nullptr, nullptr, c->d_amt)
                   : hc_launch_points(c, n);
  if (rc) return rc;
  // the host vectors die at return: the copies above are from pageable
  // memory and complete before hipMemcpyAsync returns
  return HC_OK;
}

int hc_points_set(hc_ctx* c, const int* x, const int* y, const int* rain_cell, int n) {
  if (!c || n < 0 || (n > 0 && (!x || !y || !rain_cell))) return HC_ERR_ARG;
  if (c->pt_bits.size() != static_cast<size_t>(c->NR) * c->W)
    c->pt_bits.assign(static_cast<size_t>(c->NR) * c->W, 0);
  std::vector<int> buf(2 * static_cast<size_t>(n));
  int bad = 0;  // 1: outside the grid, 2: repeated cell, 3: rain cell out of range
  for (int k = 0; k < n; ++k) {
    if (x[k] < 0 || x[k] >= c->NR || y[k] < 0 || y[k] >= c->NC) { bad = 1; break; }
    if (rain_cell[k] < 0 || rain_cell[k] >= HC_RF_MAX) { bad = 3; break; }
    uint64_t& w = c->pt_bits[static_cast<size_t>(x[k]) * c->W + (y[k] >> 6)];
    const uint64_t b = 1ull << (y[k] & 63);
    if (w & b) { bad = 2; break; }
    w |= b;
    buf[k] = x[k] * c->P + y[k];
    buf[n + k] = rain_cell[k];
  }
  for (int k = 0; k < n; ++k)
    if (x[k] >= 0 && x[k] < c->NR && y[k] >= 0 && y[k] < c->NC)
      c->pt_bits[static_cast<size_t>(x[k]) * c->W + (y[k] >> 6)] = 0;
  if (bad == 1) return hc_fail(c, HC_ERR_ARG, "points_set: point outside the padded grid");
  if (bad == 2) return hc_fail(c, HC_ERR_ARG, "points_set: a cell is listed twice");
  if (bad == 3) return hc_fail(c, HC_ERR_ARG, "points_set: rain cell out of range");
  if (n > c->plist_cap || !c->d_pamt) {
    HC_HIP(c, hipStreamSynchronize(c->stream));
    if (c->d_plist) (void)hipFree(c->d_plist);
    c->d_plist = nullptr;
    const int cap = std::max(n, 1024);
    HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_plist), 2 * static_cast<size_t>(cap) * sizeof(int)));
    if (!c->d_pamt) HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_pamt), HC_RF_MAX * sizeof(double)));
    c->plist_cap = cap;
  }
  if (n > 0) {
    HC_HIP(c, hipMemcpyAsync(c->d_plist, buf.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HC_HIP(c, hipMemcpyAsync(c->d_plist + c->plist_cap, buf.data() + n, n * sizeof(int),
                             hipMemcpyHostToDevice, c->stream));
  }
  c->plist_n = n;
  return HC_OK;
}

int hc_points_add(hc_ctx* c, const double* amt, int nrf) {
  if (!c || !amt || nrf < 1 || nrf > HC_RF_MAX) return HC_ERR_ARG;
  if (c->plist_n < 0) return hc_fail(c, HC_ERR_STATE, "points_add: no point list (hc_points_set)");
  if (c->plist_n == 0) return HC_OK;
  double a[HC_RF_MAX] = {0};
  for (int k = 0; k < nrf; ++k) a[k] = amt[k];
  // pageable source: the copy has consumed `a` when it returns
  HC_HIP(c, hipMemcpyAsync(c->d_pamt, a, sizeof(a), hipMemcpyHostToDevice, c->stream));
  return hc_launch_points_rc(c, c->plist_n, c->d_plist, c->d_plist + c->plist_cap, c->d_pamt);
}

int hc_add_point_inputs_noalloc(hc_ctx* c, const int* x, const int* y, const double* amt, int n) {
  if (!c || n < 0 || (n > 0 && (!x || !y || !amt))) return HC_ERR_ARG;
  if (n > HC_NOALLOC_MAX_POINTS)
    return hc_fail(c, HC_ERR_ARG,
                   "add_point_inputs_noalloc: more than HC_NOALLOC_MAX_POINTS (16384) points");
  if (n == 0) return HC_OK;
  if (!c->scratch) return hc_fail(c, HC_ERR_STATE, "add_point_inputs_noalloc: no staging array");
  for (int k = 0; k < n; ++k)
    if (x[k] < 0 || x[k] >= c->NR || y[k] < 0 || y[k] >= c->NC)
      return hc_fail(c, HC_ERR_ARG, "add_point_inputs_noalloc: point outside the padded grid");
  int* xs = reinterpret_cast<int*>(c->scratch);
  int* ys = xs + HC_NOALLOC_MAX_POINTS;
  double* am = reinterpret_cast<double*>(c->scratch + 8 * static_cast<size_t>(HC_NOALLOC_MAX_POINTS));
  // pageable host sources: each copy has consumed the caller's array when it returns
  HC_HIP(c, hipMemcpyAsync(xs, x, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HC_HIP(c, hipMemcpyAsync(ys, y, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HC_HIP(c, hipMemcpyAsync(am, amt, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return hc_launch_points_ordered(c, n, nullptr, xs, ys, am);
}

int hc_reach_inputs(hc_ctx* c, int n, const int* x, const int* y, const double* sediment,
                    const double* water, double threshold) {
  if (!c) return HC_ERR_ARG;
  return hc_reach_inputs_impl(c, n, x, y, sediment, water, threshold);
}

int hc_runoff_set(hc_ctx* c, const double* old_j_mean, const double* new_j_mean) {
  if (!c || !old_j_mean || !new_j_mean) return HC_ERR_ARG;
  if (!c->d_runoff) {
    HC_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_runoff), 2 * c->ncell * sizeof(double)));
    HC_HIP(c, hipMemsetAsync(c->d_runoff, 0, 2 * c->ncell * sizeof(double), c->stream));
  }
  HC_HIP(c, hipStreamSynchronize(c->stream));
  int rc = put_plane(c, c->d_runoff, false, old_j_mean, 1, 0);
  if (rc == HC_OK) rc = put_plane(c, c->d_runoff + c->ncell, false, new_j_mean, 1, 0);
  return rc;
}

int hc_runoff_add(hc_ctx* c, double time, double flow_timestep, double cap) {
  if (!c) return HC_ERR_ARG;
  if (!c->d_runoff) return hc_fail(c, HC_ERR_STATE, "runoff_add: no runoff planes (hc_runoff_set)");
  if (c->g.imax < 2 || c->g.jmax < 2) return HC_OK;  // no cell 1..imax-1, 1..jmax-1
  const dim3 blk(256), grd(static_cast<unsigned>((c->g.jmax - 1 + 255) / 256),
                          static_cast<unsigned>(c->g.imax - 1));
  if (c->prec == HC_FP64)
    hipLaunchKernelGGL(k_runoff_add<double>, grd, blk, 0, c->stream, c->g.imax, c->g.jmax, c->P,
                       time, flow_timestep, cap, c->d_runoff, c->d_runoff + c->ncell,
                       static_cast<double*>(c->h[c->cur]));
  else
    hipLaunchKernelGGL(k_runoff_add<float>, grd, blk, 0, c->stream, c->g.imax, c->g.jmax, c->P,
                       time, flow_timestep, cap, c->d_runoff, c->d_runoff + c->ncell,
                       static_cast<float*>(c->h[c->cur]));
  HC_HIP(c, hipGetLastError());
  return HC_OK;
}

int hc_erode(hc_ctx* c, double mult, double* time_step, double* tempbmax, float gtot[10]) {
  if (!c || !time_step) return HC_ERR_ARG;
  if (!c->g.sediment) return hc_fail(c, HC_ERR_STATE, "erode: context has no sediment state");
  return hc_erode_impl(c, mult, time_step, tempbmax, gtot);
}

int hc_global_landsliding(hc_ctx* c, double failureangle, int* sweeps) {
  if (!c) return HC_ERR_ARG;
  if (!c->g.sediment)
    return hc_fail(c, HC_ERR_STATE, "global_landsliding: context has no sediment state");
  return hc_global_landsliding_impl(c, failureangle, sweeps);
}

int hc_creep(hc_ctx* c, double time, double creep_rate) {
  if (!c) return HC_ERR_ARG;
  if (!c->g.sediment) return hc_fail(c, HC_ERR_STATE, "creep: context has no sediment state");
  return hc_creep_impl(c, time, creep_rate);
}

int hc_in_channel_landsliding(hc_ctx* c, double failureangle) {
  if (!c) return HC_ERR_ARG;
  if (!c->g.sediment)
    return hc_fail(c, HC_ERR_STATE, "in_channel_landsliding: context has no sediment state");
  return hc_in_channel_landsliding_impl(c, failureangle);
}

int hc_synchronize(hc_ctx* c) {
  if (!c) return HC_ERR_ARG;
  HC_HIP(c, hipStreamSynchronize(c->stream));
  return HC_OK;
}

int hc_grain_index_set(hc_ctx* c, const int32_t* index, size_t count, long long rows_total) {
  if (!c || !index) return HC_ERR_ARG;
  if (!c->g.sediment) return hc_fail(c, HC_ERR_STATE, "grain index: not a sediment context");
  if (count != static_cast<size_t>(c->NR) * c->NC || rows_total < 0 || rows_total > INT32_MAX)
    return hc_fail(c, HC_ERR_ARG, "grain index: count is not (imax+2)*(jmax+2) or bad rows_total");
  HC_HIP(c, hipStreamSynchronize(c->stream));
  return put_row_index(c, index, rows_total);
}

int hc_grain_index_get(hc_ctx* c, int32_t* index, size_t count, long long* rows_total) {
  if (!c || !index) return HC_ERR_ARG;
  if (!c->g.sediment) return hc_fail(c, HC_ERR_STATE, "grain index: not a sediment context");
  if (count != static_cast<size_t>(c->NR) * c->NC)
    return hc_fail(c, HC_ERR_ARG, "grain index: count is not (imax+2)*(jmax+2)");
  HC_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<uint64_t> key(c->ncell);
  std::vector<int32_t> has(c->ncell);
  HC_HIP(c, hipMemcpy(key.data(), c->rowkey, c->ncell * sizeof(uint64_t), hipMemcpyDeviceToHost));
  HC_HIP(c, hipMemcpy(has.data(), c->hasgrain, c->ncell * sizeof(int32_t), hipMemcpyDeviceToHost));
  // rows of the initial data keep their numbers; later rows follow in key order
  std::vector<std::pair<uint64_t, size_t>> made;
  for (int x = 0; x < c->NR; ++x)
    for (int y = 0; y < c->NC; ++y) {
      const size_t d = static_cast<size_t>(x) * c->P + y, h = static_cast<size_t>(x) * c->NC + y;
      const uint64_t k = key[d];
      if (!has[d]) {
        index[h] = -9999;
      } else if (k == hc::kRowNone) {
        return hc_fail(c, HC_ERR_STATE, "grain index: a grain row of unknown creation order");
      } else if (k < hc::row_epoch_base(1)) {
        index[h] = static_cast<int32_t>(k);
      } else {
        made.emplace_back(k, h);
      }
    }
  std::sort(made.begin(), made.end());
  if (c->rows_base + static_cast<long long>(made.size()) > INT32_MAX)
    return hc_fail(c, HC_ERR_STATE, "grain index: more rows than an int holds");
  for (size_t r = 0; r < made.size(); ++r)
    index[made[r].second] = static_cast<int32_t>(c->rows_base + 1 + static_cast<long long>(r));
  if (rows_total) *rows_total = c->rows_base + static_cast<long long>(made.size());
  return HC_OK;
}

int hc_check_health(hc_ctx* c, long long* bad) {
  if (!c) return HC_ERR_ARG;
  int rc = hc_launch_health(c);
  if (rc) return rc;
  HC_HIP(c, hipMemcpyAsync(c->h_pinned, c->d_red + 2, sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c->stream));
  HC_HIP(c, hipStreamSynchronize(c->stream));
  unsigned long long n;
  std::memcpy(&n, c->h_pinned, sizeof(n));
  if (bad) *bad = static_cast<long long>(n);
  if (n) return hc_fail(c, HC_ERR_UNSTABLE, "health: non-finite or negative water depth");
  return HC_OK;
}

}  // extern "C"
