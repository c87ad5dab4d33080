// Command-line driver with the reference's interface
// (src/main.cpp:49-218): hail-caesar-mi355x <path> <param file> [--fp32]
#include <chrono>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

#include "../../../include/hc.h"
#include "../../../include/hc_model.h"

int main(int argc, char* argv[]) {
  if (argc < 3) {
    std::cout << "usage: " << argv[0]
              << " <path> <parameter file> [--fp32] [--fixed-elev] [--separate] [--device-loop] [--graph]\n";
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
  for (int a = 3; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--fp32")) opt.precision = HC_FP32;
    else if (!std::strcmp(argv[a], "--device-loop")) opt.device_loop = 1;
    else if (!std::strcmp(argv[a], "--graph")) { opt.device_loop = 1; opt.use_graph = 1; }
    else if (!std::strcmp(argv[a], "--separate")) opt.fused = 0;
    else if (!std::strcmp(argv[a], "--fixed-elev")) opt.elev_fixed = 1;
  }
  const auto t0 = std::chrono::steady_clock::now();
  hcm_model* m = nullptr;
  if (hcm_create(&m, argv[1], argv[2], &opt) != HC_OK) {
    std::cerr << "error: " << hcm_last_error(m) << "\n";
    hcm_destroy(m);
    return 1;
  }
  std::cout << "Entering main model loop..." << std::endl;
  long iters = 0;
  int finished = 0;
  while (!finished) {
    int done = 0;
    if (hcm_iterate(m, 1000, nullptr, nullptr, nullptr, &done, &finished) != HC_OK) {
      std::cerr << "error: " << hcm_last_error(m) << "\n";
      hcm_destroy(m);
      return 1;
    }
    iters += done;
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cout << "THE SIMULATION IS FINISHED!" << std::endl;
  std::cout << "iterations: " << iters << "  wall: " << secs << " s" << std::endl;
  hcm_destroy(m);
  return 0;
}
