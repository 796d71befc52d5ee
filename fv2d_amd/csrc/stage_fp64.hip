// stage_fp64.hip -- FP64 (parity form) instantiations of the stage kernel: the single-grid
// kernels with the general division form and the batch kernels; the exact-division
// single-grid kernels are in stage_fp64_exact.hip (compiled in parallel).
// Built with -ffp-contract=off: every a*b+c stays a separate multiply and add, as in the
// reference's CPU build, so results match it bit for bit.
#include "fv2d_device.h"

namespace fv2d {

hipError_t launch_stage_fp64(const KParams &kp, const StageIO<double> &io, const StripGeom &g, int recon, int riem,
                             bool tc, bool visc, hipStream_t st) {
  if (kp.divx == DIV_EXACT && kp.divy == DIV_EXACT)
    return launch_stage_fp64_exact(kp, io, g, recon, riem, tc, visc, st);
  return launch_stage<double, double, DIVF_GENERAL>(kp, io, g, recon, riem, tc, visc, st);
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
    if (nw == 8) return launch_batch_x<double, double, RC, RM, TC, VI, XG, 8>(jobs, items, range, counter, stage, grid, st);
    return launch_batch_x<double, double, RC, RM, TC, VI, XG, 4>(jobs, items, range, counter, stage, grid, st);
  }
};
struct BatchOcc {
  int nw;
  template <int RC, int RM, bool TC, bool VI, bool XG>
  int operator()() {
    if (nw == 8) return batch_occupancy_x<double, double, RC, RM, TC, VI, XG, 8>();
    return batch_occupancy_x<double, double, RC, RM, TC, VI, XG, 4>();
  }
};
}  // namespace

hipError_t launch_batch_fp64(const StageClass &c, const BatchJob *jobs, const BatchItem *items, const int *range,
                             int *counter, int stage, int grid, hipStream_t st) {
  BatchLaunch f{jobs, items, range, counter, stage, grid, st, c.nw};
  return with_class(c, f);
}

int batch_occupancy_fp64(const StageClass &c) {
  BatchOcc f{c.nw};
  return with_class(c, f);
}

}  // namespace fv2d
