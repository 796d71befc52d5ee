// stage_fp64_exact.hip -- FP64 (parity form) single-grid stage kernels for grids whose cell
// sizes are both powers of two: the divisions by dx and dy are exact multiplies by 1/dx and
// 1/dy, fixed at compile time (DIVF_EXACT). Built with -ffp-contract=off, as stage_fp64.hip.
#include "fv2d_device.h"

namespace fv2d {

hipError_t launch_stage_fp64_exact(const KParams &kp, const StageIO<double> &io, const StripGeom &g, int recon,
                                   int riem, bool tc, bool visc, hipStream_t st) {
  return launch_stage<double, double, DIVF_EXACT>(kp, io, g, recon, riem, tc, visc, st);
}

}  // namespace fv2d
