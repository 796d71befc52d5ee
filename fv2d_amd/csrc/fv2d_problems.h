// fv2d_problems.h -- initial conditions evaluated per cell (host and device).
//
// The workloads of the benchmark and of the serving stream: the Sod tube along x, the
// Sedov-type point blast and the Kelvin-Helmholtz shear layer, plus the counter-based
// random numbers they use (SplitMix64 keyed by seed and cell index, so every cell's noise
// is reproducible on its own). fv2d_amd/problems.py generates the same states in numpy;
// the integer RNG is bit-identical, the transcendental functions agree to rounding.
#pragma once

#include "fv2d_math.h"

namespace fv2d {

enum : int { PROB_SOD_X = 0, PROB_SEDOV = 1, PROB_KH = 2 };

FV2D_HD unsigned long long splitmix64(unsigned long long x) {
  unsigned long long z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// U[0,1) from the top 53 bits of splitmix64(seed * golden + idx).
FV2D_HD double uniform01(unsigned long long seed, unsigned long long idx) {
  const unsigned long long z = splitmix64(seed * 0x9E3779B97F4A7C15ull + idx);
  return (double)(z >> 11) * 0x1p-53;
}

struct ProblemArgs {
  int kind;
  int Nx, Ny;
  double xmin, ymin, dx, dy, xmid, ymid;
  double gamma0;
  unsigned long long seed;
  double amp;  // KH: amplitude of the sin(4 pi x) perturbation of v
};

// Primitive state (rho, u, v, p) of cell (i, j); centres as getPos (SimInfo.h:495-499).
FV2D_HD void problem_cell(const ProblemArgs &a, int i, int j, double q[4]) {
  const double x = a.xmin + (i + 0.5) * a.dx;
  const double y = a.ymin + (j + 0.5) * a.dy;
  q[0] = 1.0;
  q[1] = 0.0;
  q[2] = 0.0;
  q[3] = 1.0;
  if (a.kind == PROB_SOD_X) {
    const bool left = x < 0.5;
    q[0] = left ? 1.0 : 0.125;
    q[3] = left ? 1.0 : 0.1;
  } else if (a.kind == PROB_SEDOV) {
    // energy E = 1 deposited uniformly in cells within r < 0.02 of the centre
    const double xr = x - a.xmid, yr = y - a.ymid;
    const double r = sqrt(xr * xr + yr * yr);
    const double r0 = 0.02;
    q[3] = (r < r0) ? (a.gamma0 - 1.0) * 1.0 / (M_PI * r0 * r0) : 1.0e-5;
  } else {
    // shear band 0.25 < y < 0.75 (rho 2, u +0.5) in rho 1, u -0.5; v = amp sin(4 pi x) + noise
    const bool band = (y > 0.25) && (y < 0.75);
    q[0] = band ? 2.0 : 1.0;
    q[1] = band ? 0.5 : -0.5;
    const unsigned long long idx = (unsigned long long)j * (unsigned long long)a.Nx + (unsigned long long)i;
    q[2] = a.amp * sin(4.0 * M_PI * x) + 1.0e-3 * (2.0 * uniform01(a.seed, idx) - 1.0);
    q[3] = 2.5;
  }
}

}  // namespace fv2d
