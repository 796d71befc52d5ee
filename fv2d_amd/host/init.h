// init.h -- the reference's initial-condition library (InitFunctor, Init.h:20-359) on the
// host: fills the domain primitives of a run in snapshot order (4 planes rho, u, v, prs;
// j-outer, i-inner; Nx*Ny each), the input of fv2d_set_prim.
//
// Cells start at zero, as the reference's freshly allocated views do, and each problem
// sets the fields its initialiser sets. Cell centres follow getPos (SimInfo.h:495-499).
// H84 and C91 draw their perturbation from a Kokkos random pool in the reference, whose
// stream depends on the thread schedule and cannot be reproduced; here it comes from a
// counter-based generator (SplitMix64 keyed by seed and cell index), drand(-0.5, 0.5) =
// -0.5 + U[0,1).
#pragma once

#include <vector>

#include "params.h"

namespace fv2dh {

// q: 4*Nx*Ny doubles.
void init_problem(const RunParams &p, double *q);
std::vector<double> init_problem(const RunParams &p);

// U[0,1) from the top 53 bits of SplitMix64(seed * golden + idx).
double uniform01(unsigned long long seed, unsigned long long idx);

}  // namespace fv2dh
