// stage_fp64.hip -- FP64 (parity form) instantiations of the stage kernel.
// Built with -ffp-contract=off: every a*b+c stays a separate multiply and add, as in the
// reference's CPU build, so results match it bit for bit.
#include "fv2d_device.h"

namespace fv2d {

hipError_t launch_stage_fp64(const KParams &kp, const StageIO<double> &io, const StripGeom &g, int recon, int riem,
                             bool tc, bool visc, hipStream_t st) {
  return launch_stage<double, double>(kp, io, g, recon, riem, tc, visc, st);
}

}  // namespace fv2d
