// stage_fp32.hip -- FP32-form instantiations of the stage kernel: reconstruction, Riemann
// fluxes and the parabolic / gravity stencils in FP32, the conserved update accumulated in
// FP64, U stored in FP32 (or in FP64: FV2D_FP32_STATE64).
#include "fv2d_device.h"

namespace fv2d {

hipError_t launch_stage_fp32(const KParams &kp, const StageIO<float> &io, const StripGeom &g, int recon, int riem,
                             bool tc, bool visc, hipStream_t st) {
  return launch_stage<float, float, DIVF_RUNTIME>(kp, io, g, recon, riem, tc, visc, st);
}

// FP32 arithmetic on a state stored in FP64 (FV2D_FP32_STATE64).
hipError_t launch_stage_fp32s64(const KParams &kp, const StageIO<double> &io, const StripGeom &g, int recon,
                                int riem, bool tc, bool visc, hipStream_t st) {
  return launch_stage<float, double, DIVF_RUNTIME>(kp, io, g, recon, riem, tc, visc, st);
}

namespace {
struct BatchLaunch {
  const BatchJob *jobs;
  const BatchItem *items;
  const int *range;
  int *counter;
  int stage, grid;
  hipStream_t st;
  int nw;
  template <int RC, int RM, bool TC, bool VI, bool XG>
  hipError_t operator()() {
    if (nw == 8) return launch_batch_x<float, float, RC, RM, TC, VI, XG, 8>(jobs, items, range, counter, stage, grid, st);
    return launch_batch_x<float, float, RC, RM, TC, VI, XG, 4>(jobs, items, range, counter, stage, grid, st);
  }
};
struct BatchOcc {
  int nw;
  template <int RC, int RM, bool TC, bool VI, bool XG>
  int operator()() {
    if (nw == 8) return batch_occupancy_x<float, float, RC, RM, TC, VI, XG, 8>();
    return batch_occupancy_x<float, float, RC, RM, TC, VI, XG, 4>();
  }
};
}  // namespace

hipError_t launch_batch_fp32(const StageClass &c, const BatchJob *jobs, const BatchItem *items, const int *range,
                             int *counter, int stage, int grid, hipStream_t st) {
  BatchLaunch f{jobs, items, range, counter, stage, grid, st, c.nw};
  return with_class(c, f);
}

int batch_occupancy_fp32(const StageClass &c) {
  BatchOcc f{c.nw};
  return with_class(c, f);
}

}  // namespace fv2d
