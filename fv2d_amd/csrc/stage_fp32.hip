// stage_fp32.hip -- FP32-form instantiations of the stage kernel: reconstruction, Riemann
// fluxes and the parabolic / gravity stencils in FP32, U stored in FP32, the conserved
// update accumulated in FP64.
#include "fv2d_device.h"

namespace fv2d {

hipError_t launch_stage_fp32(const KParams &kp, const StageIO<float> &io, const StripGeom &g, int recon, int riem,
                             bool tc, bool visc, hipStream_t st) {
  return launch_stage<float, float>(kp, io, g, recon, riem, tc, visc, st);
}

}  // namespace fv2d
