// stage_fp32.hip -- FP32-form instantiations of the stage kernel: reconstruction, Riemann
// fluxes and the parabolic / gravity stencils in FP32, U stored in FP32, the conserved
// update accumulated in FP64.
#include "fv2d_device.h"

namespace fv2d {

hipError_t launch_stage_fp32(const KParams &kp, const StageIO<float> &io, const StripGeom &g, int recon, int riem,
                             bool tc, bool visc, hipStream_t st) {
  return launch_stage<float, float>(kp, io, g, recon, riem, tc, visc, st);
}

namespace {
struct BatchLaunch {
  const BatchJob *jobs;
  const BatchItem *items;
  const int *range;
  int *counter;
  int stage, grid;
  hipStream_t st;
  template <int RC, int RM, bool TC, bool VI, bool XG>
  hipError_t operator()() {
    return launch_batch_x<float, float, RC, RM, TC, VI, XG>(jobs, items, range, counter, stage, grid, st);
  }
};
struct BatchOcc {
  template <int RC, int RM, bool TC, bool VI, bool XG>
  int operator()() {
    return batch_occupancy_x<float, float, RC, RM, TC, VI, XG>();
  }
};
}  // namespace

hipError_t launch_batch_fp32(const StageClass &c, const BatchJob *jobs, const BatchItem *items, const int *range,
                             int *counter, int stage, int grid, hipStream_t st) {
  BatchLaunch f{jobs, items, range, counter, stage, grid, st};
  return with_class(c, f);
}

int batch_occupancy_fp32(const StageClass &c) {
  BatchOcc f;
  return with_class(c, f);
}

}  // namespace fv2d
