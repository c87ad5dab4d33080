// Command-line driver with the reference's interface
// (src/main.cpp:49-218): hail-caesar-mi355x <path> <param file> [--fp32] ...
// With --checkpoint FILE the host-stepped run writes an exact binary
// checkpoint to FILE every --checkpoint-every iterations (default 1000) and
// at the end; --restart FILE continues a run from such a file (the same
// path and parameter file as the run that wrote it).  --max-iterations N
// ends the run after N iterations (a batch job's time slice), checkpointing
// first when --checkpoint is given; a later --restart carries on.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

#include "../../../include/hc.h"
#include "../../../include/hc_model.h"

int main(int argc, char* argv[]) {
  if (argc < 3) {
    std::cout << "usage: " << argv[0]
              << " <path> <parameter file> [--fp32] [--fixed-elev] [--separate] [--device-loop] [--graph]"
                 " [--checkpoint FILE] [--checkpoint-every N] [--restart FILE] [--max-iterations N]\n";
    return 0;
  }
  hcm_options opt{};
  opt.precision = HC_FP64;
  opt.fused = 1;
  opt.quiet = 0;
  opt.write_files = 1;
  opt.flat_dz = 0.0;
  opt.device_loop = 0;
  opt.use_graph = 0;
  opt.elev_fixed = 0;
  const char* ckpt = nullptr;
  const char* restart = nullptr;
  long every = 1000, max_iters = -1;
  for (int a = 3; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--fp32")) opt.precision = HC_FP32;
    else if (!std::strcmp(argv[a], "--device-loop")) opt.device_loop = 1;
    else if (!std::strcmp(argv[a], "--graph")) { opt.device_loop = 1; opt.use_graph = 1; }
    else if (!std::strcmp(argv[a], "--separate")) opt.fused = 0;
    else if (!std::strcmp(argv[a], "--fixed-elev")) opt.elev_fixed = 1;
    else if (!std::strcmp(argv[a], "--checkpoint") && a + 1 < argc) ckpt = argv[++a];
    else if (!std::strcmp(argv[a], "--restart") && a + 1 < argc) restart = argv[++a];
    else if (!std::strcmp(argv[a], "--checkpoint-every") && a + 1 < argc) every = std::atol(argv[++a]);
    else if (!std::strcmp(argv[a], "--max-iterations") && a + 1 < argc) max_iters = std::atol(argv[++a]);
    else {
      std::cerr << "error: unknown or incomplete option " << argv[a] << "\n";
      return 1;
    }
  }
  if ((ckpt || restart) && opt.device_loop) {
    std::cerr << "error: --checkpoint and --restart need the host-stepped loop (no --device-loop, --graph)\n";
    return 1;
  }
  if (every < 1) {
    std::cerr << "error: --checkpoint-every needs a positive count\n";
    return 1;
  }
  const auto t0 = std::chrono::steady_clock::now();
  hcm_model* m = nullptr;
  if (hcm_create(&m, argv[1], argv[2], &opt) != HC_OK) {
    std::cerr << "error: " << hcm_last_error(m) << "\n";
    hcm_destroy(m);
    return 1;
  }
  auto fail = [&]() {
    std::cerr << "error: " << hcm_last_error(m) << "\n";
    hcm_destroy(m);
    return 1;
  };
  if (restart && hcm_restart(m, restart) != HC_OK) return fail();
  std::cout << "Entering main model loop..." << std::endl;
  long iters = 0;
  int finished = 0;
  const int chunk = ckpt ? static_cast<int>(std::min(every, 1000L)) : 1000;
  long since = 0;
  bool sliced = false;
  while (!finished) {
    long want = chunk;
    if (ckpt) want = std::min(want, every - since);
    if (max_iters >= 0) want = std::min(want, max_iters - iters);
    if (want <= 0) {
      sliced = true;
      break;
    }
    int done = 0;
    if (hcm_iterate(m, static_cast<int>(want), nullptr, nullptr, nullptr, &done, &finished) != HC_OK)
      return fail();
    iters += done;
    since += done;
    if (ckpt && (since >= every || finished)) {
      if (hcm_checkpoint(m, ckpt) != HC_OK) return fail();
      since = 0;
    }
  }
  if (sliced && ckpt && since > 0 && hcm_checkpoint(m, ckpt) != HC_OK) return fail();
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (sliced) {
    std::cout << "stopped after " << iters << " iterations" << (ckpt ? " (checkpoint written)" : "")
              << std::endl;
    hcm_destroy(m);
    return 0;
  }
  std::cout << "THE SIMULATION IS FINISHED!" << std::endl;
  std::cout << "iterations: " << iters << "  wall: " << secs << " s" << std::endl;
  hcm_destroy(m);
  return 0;
}
