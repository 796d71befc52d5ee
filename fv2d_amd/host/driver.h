// driver.h -- the time loop of the reference's main.cpp (main.cpp:26-88) over the C ABI.
#pragma once

#include <ostream>

#include "params.h"

namespace fv2dh {

struct DriverOptions {
  int precision = FV2D_FP64;
  // true: one fv2d_compute_dt + fv2d_step per iteration, exactly main.cpp's sequence;
  // false: between log lines and snapshots the steps run in the device-resident loop
  // (fv2d_advance), same steps and dts, reset counts reported per run of steps.
  bool per_step = false;
};

// Runs the simulation of p (initial state or restart, snapshots, log). Returns 0, or throws
// std::runtime_error with the reason.
int run_simulation(const RunParams &p, const DriverOptions &opt, std::ostream &out);

}  // namespace fv2dh
