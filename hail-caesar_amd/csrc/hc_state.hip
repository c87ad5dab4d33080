// hc_state.hip - exact binary save and load of a context's device state.
//
// The reference has no checkpoint: it hot-starts from rasters written at 6
// significant figures (water_init_from_raster_on, LCM.cpp:331-373,1258-1262)
// and a Grainz file (read_in_graindata_from_file, LCM.cpp:590-683), and does
// not restore qx, qy, vel_dir or its clocks.  A state image here is every
// persistent device buffer in its stored layout (both generations of the
// ping-pong pairs, padding included), the small device scalars and the host
// values that decide how the next step computes (generation, tile shape,
// fixed-point elevation frame), so a context loaded from it continues
// bit for bit as the one that saved it.
//
// Image layout (native byte order, this library's version only):
//   StateHeader, then the buffers of state_buffers() in order, each as raw
//   bytes.  Scratch (partials, counters, drain slots, hillslope moves) is not
//   part of the image: every kernel that reads it writes it first.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "hc_internal.hpp"

namespace {

constexpr char kMagic[8] = {'H', 'C', 'S', 'T', 'A', 'T', 'E', '2'};

struct StateHeader {
  char magic[8];
  int32_t prec, imax, jmax, P, W, sediment, has_nman, has_zq;
  int32_t cur, fused_rows, fused_strip, zq_shift;
  double DX, zq_base;
  uint64_t red[8];       // d_red: maxdepth / tempbmax bits, health and sweep counts
  int64_t rows_base;     // grain rows of the initial data
  uint32_t row_epoch, pad0;
  uint64_t payload;      // bytes after the header
};

struct Buf {
  void* p;
  size_t bytes;
};

std::vector<Buf> state_buffers(hc_ctx* c) {
  const size_t n = c->ncell, e = c->esz;
  std::vector<Buf> b;
  for (int k = 0; k < 2; ++k) {
    b.push_back({c->h[k], n * e});
    b.push_back({c->qx[k], n * e});
    b.push_back({c->qy[k], n * e});
  }
  b.push_back({c->z, n * (c->g.sediment ? sizeof(double) : e)});  // FP64 with sediment
  if (c->z64) b.push_back({c->z64, n * sizeof(double)});
  if (c->zq) b.push_back({c->zq, n * sizeof(int32_t)});
  if (c->nman) b.push_back({c->nman, n * e});
  b.push_back({c->mask, static_cast<size_t>(c->NR) * c->W * sizeof(uint64_t)});
  if (c->g.sediment) {
    const size_t s = c->sesz;  // sediment planes: the context's storage type
    b.push_back({c->V, n * s});
    b.push_back({c->qxs, n * s});
    b.push_back({c->qys, n * s});
    b.push_back({c->veldir, 4 * n * s});
    b.push_back({c->grain, 11 * n * s});
    b.push_back({c->strata, 90 * n * s});
    b.push_back({c->hasgrain, n * sizeof(int32_t)});
    b.push_back({c->needgrain, n * sizeof(int32_t)});
    b.push_back({c->rowkey, n * sizeof(uint64_t)});
    b.push_back({c->bedrock, n * sizeof(double)});
    b.push_back({c->Tau, n * s});
    b.push_back({c->Vel, n * s});
    b.push_back({c->ss, n * s});
    b.push_back({c->sdir, 40 * n * s});
    b.push_back({c->etot, n * s});
    b.push_back({c->etot3, n * s});
    b.push_back({c->gtot, 20 * sizeof(float)});
  }
  return b;
}

size_t payload_bytes(const std::vector<Buf>& b) {
  size_t s = 0;
  for (const Buf& x : b) s += x.bytes;
  return s;
}

StateHeader header_of(const hc_ctx* c) {
  StateHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kMagic, sizeof(kMagic));
  h.prec = c->prec;
  h.imax = c->g.imax;
  h.jmax = c->g.jmax;
  h.P = c->P;
  h.W = c->W;
  h.sediment = c->g.sediment ? 1 : 0;
  h.has_nman = c->nman ? 1 : 0;
  h.has_zq = c->zq ? 1 : 0;
  h.cur = c->cur;
  h.fused_rows = c->fused_rows;
  h.fused_strip = c->fused_strip;
  h.zq_shift = c->zq_shift;
  h.DX = c->g.DX;
  h.zq_base = c->zq_base;
  h.rows_base = c->rows_base;
  h.row_epoch = c->row_epoch;
  return h;
}

int refuse_loop(hc_ctx* c, const char* what) {
  if (!c->loop_ready) return HC_OK;
  return hc_fail(c, HC_ERR_STATE,
                 std::string(what) +
                     ": the device-resident loop keeps its control state on the device; "
                     "save or load the state of a host-stepped context");
}

}  // namespace

extern "C" {

int hc_state_bytes(hc_ctx* c, size_t* bytes) {
  if (!c || !bytes) return HC_ERR_ARG;
  *bytes = sizeof(StateHeader) + payload_bytes(state_buffers(c));
  return HC_OK;
}

int hc_state_save(hc_ctx* c, void* host, size_t bytes) {
  if (!c || !host) return HC_ERR_ARG;
  if (int rc = refuse_loop(c, "state save")) return rc;
  const std::vector<Buf> b = state_buffers(c);
  const size_t need = sizeof(StateHeader) + payload_bytes(b);
  if (bytes != need) return hc_fail(c, HC_ERR_ARG, "state save: buffer size is not hc_state_bytes()");
  StateHeader h = header_of(c);
  h.payload = need - sizeof(StateHeader);
  HC_HIP(c, hipStreamSynchronize(c->stream));
  HC_HIP(c, hipMemcpy(h.red, c->d_red, sizeof(h.red), hipMemcpyDeviceToHost));
  unsigned char* out = static_cast<unsigned char*>(host);
  std::memcpy(out, &h, sizeof(h));
  size_t off = sizeof(h);
  for (const Buf& x : b) {
    HC_HIP(c, hipMemcpy(out + off, x.p, x.bytes, hipMemcpyDeviceToHost));
    off += x.bytes;
  }
  return HC_OK;
}

int hc_state_load(hc_ctx* c, const void* host, size_t bytes) {
  if (!c || !host) return HC_ERR_ARG;
  if (int rc = refuse_loop(c, "state load")) return rc;
  if (bytes < sizeof(StateHeader)) return hc_fail(c, HC_ERR_ARG, "state load: image too short");
  StateHeader h;
  std::memcpy(&h, host, sizeof(h));
  if (std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0)
    return hc_fail(c, HC_ERR_ARG, "state load: not a state image of this version");
  const StateHeader mine = header_of(c);
  if (h.prec != mine.prec || h.imax != mine.imax || h.jmax != mine.jmax || h.P != mine.P ||
      h.W != mine.W || h.sediment != mine.sediment || h.has_nman != mine.has_nman ||
      h.has_zq != mine.has_zq || !(h.DX == mine.DX))
    return hc_fail(c, HC_ERR_STATE,
                   "state load: the image was saved from a context of another grid, precision or "
                   "configuration");
  const std::vector<Buf> b = state_buffers(c);
  const size_t need = sizeof(StateHeader) + payload_bytes(b);
  if (bytes != need || h.payload != need - sizeof(StateHeader))
    return hc_fail(c, HC_ERR_ARG, "state load: image size does not match this context");
  if (h.cur != 0 && h.cur != 1) return hc_fail(c, HC_ERR_ARG, "state load: corrupt header");
  if (hc_set_tiles(c, h.fused_rows, h.fused_strip) != HC_OK)
    return hc_fail(c, HC_ERR_ARG, "state load: corrupt header (tile shape)");
  HC_HIP(c, hipStreamSynchronize(c->stream));
  const unsigned char* in = static_cast<const unsigned char*>(host);
  size_t off = sizeof(h);
  for (const Buf& x : b) {
    HC_HIP(c, hipMemcpy(x.p, in + off, x.bytes, hipMemcpyHostToDevice));
    off += x.bytes;
  }
  HC_HIP(c, hipMemcpy(c->d_red, h.red, sizeof(h.red), hipMemcpyHostToDevice));
  c->cur = h.cur;
  c->zq_shift = h.zq_shift;
  c->zq_base = h.zq_base;
  c->rows_base = h.rows_base;
  c->row_epoch = h.row_epoch;
  return HC_OK;
}

}  // extern "C"
