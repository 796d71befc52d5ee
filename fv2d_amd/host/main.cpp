// main.cpp -- the fv2d executable: fv2d <settings.ini> [--fp32] [--per-step]
//
// Reads the run's .ini file (the reference's keys and defaults), sets up the initial state
// (built-in problem or restart file), runs the explicit hydro solver on the GPU through the
// C ABI and writes HDF5/XDMF snapshots, like the reference's main.cpp. Errors print
// "ERROR: <reason>" on stderr and exit with status 1.
#include <cstring>
#include <iostream>
#include <string>

#include "driver.h"
#include "params.h"

int main(int argc, char **argv) {
  if (argc < 2) {
    std::cerr << "usage: " << argv[0] << " <settings.ini> [--fp32] [--per-step]" << std::endl;
    return 2;
  }
  fv2dh::DriverOptions opt;
  for (int k = 2; k < argc; ++k) {
    if (!std::strcmp(argv[k], "--fp32")) {
      opt.precision = FV2D_FP32;
    } else if (!std::strcmp(argv[k], "--per-step")) {
      opt.per_step = true;
    } else {
      std::cerr << "unknown option " << argv[k] << std::endl;
      return 2;
    }
  }
  std::cout << "fv2d -- explicit 2D compressible hydrodynamics on MI355X (gfx950)" << std::endl;
  try {
    const fv2dh::RunParams p = fv2dh::read_ini(argv[1], std::cerr);
    fv2dh::run_simulation(p, opt, std::cout);
  } catch (const std::exception &e) {
    std::cerr << "ERROR: " << e.what() << std::endl;
    return 1;
  }
  std::cout << std::endl << "All done." << std::endl;
  return 0;
}
