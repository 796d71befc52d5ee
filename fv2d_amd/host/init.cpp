// init.cpp -- initial conditions of Init.h restated per cell (see init.h).
#include "init.h"

#include <cmath>

namespace fv2dh {

namespace {

unsigned long long splitmix64(unsigned long long x) {
  unsigned long long z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Cell {
  double rho = 0, u = 0, v = 0, p = 0;
};

// initSodX / initSodY (Init.h:20-35, 83-98): x (or y) <= 0.5 is the high state
void sod(double s, Cell &c) {
  if (s <= 0.5) {
    c.rho = 1.0;
    c.p = 1.0;
  } else {
    c.rho = 0.125;
    c.p = 0.1;
  }
  c.u = 0.0;
}

}  // namespace

double uniform01(unsigned long long seed, unsigned long long idx) {
  const unsigned long long z = splitmix64(seed * 0x9E3779B97F4A7C15ull + idx);
  return (double)(z >> 11) * 0x1p-53;
}

void init_problem(const RunParams &p, double *q) {
  const InitType kind = init_type(p.problem);
  const long long n = (long long)p.Nx * p.Ny;
  const double xmid = 0.5 * (p.xmin + p.xmax);
  const double ymid = 0.5 * (p.ymin + p.ymax);
  for (int j = 0; j < p.Ny; ++j) {
    for (int i = 0; i < p.Nx; ++i) {
      const double x = p.xmin + (i + 0.5) * p.dx;
      const double y = p.ymin + (j + 0.5) * p.dy;
      const unsigned long long idx = (unsigned long long)j * p.Nx + i;
      Cell c;
      switch (kind) {
      case SOD_X:
        sod(x, c);
        break;
      case SOD_Y:
        sod(y, c);
        break;
      case BLAST: {  // initBlast (Init.h:103-131)
        const double xr = xmid - x, yr = ymid - y;
        const double r = std::sqrt(xr * xr + yr * yr);
        if (r < 0.2) {
          c.rho = 1.0;
          c.p = 10.0;
        } else {
          c.rho = 1.2;
          c.p = 0.1;
        }
        break;
      }
      case DIFFUSION: {  // initDiffusion (Init.h:185-206)
        const double x0 = x - xmid, y0 = y - ymid;
        const double r = std::sqrt(x0 * x0 + y0 * y0);
        c.rho = (r < 0.2) ? 1.0 : 0.1;
        c.p = 1.0;
        c.u = 1.0;
        c.v = 1.0;
        break;
      }
      case RAYLEIGH_TAYLOR: {  // initRayleighTaylor (Init.h:211-237)
        const double P0 = 2.5;
        c.rho = (y < ymid) ? 1.0 : 2.0;
        c.u = 0.0;
        c.p = P0 + 0.1 * p.gy * y;
        if (y > -1.0 / 3.0 && y < 1.0 / 3.0)
          c.v = 0.01 * (1.0 + std::cos(4 * M_PI * x)) * (1 + std::cos(3.0 * M_PI * y)) / 4.0;
        break;
      }
      case H84: {  // initH84 (Init.h:136-156)
        const double pert = p.h84_pert * (-0.5 + uniform01((unsigned long long)p.seed, idx));
        c.rho = std::pow(y, p.m1);
        c.u = 0.0;
        c.v = pert;
        c.p = std::pow(y, p.m1 + 1.0);
        break;
      }
      case C91: {  // initC91 (Init.h:159-182)
        const double T = 1.0 + p.theta1 * y;
        const double pert = p.c91_pert * (-0.5 + uniform01((unsigned long long)p.seed, idx));
        c.rho = std::pow(T, p.m1);
        c.p = std::pow(T, p.m1 + 1.0) * (1.0 + pert);
        c.u = 0.0;
        c.v = 0.0;
        break;
      }
      case KELVIN_HELMHOLTZ: {  // initKelvinHelmholtz (Init.h:245-265), Lecoanet et al. 2016
        const double q1 = std::tanh((y - p.kh_y1) / p.kh_a);
        const double q2 = std::tanh((y - p.kh_y2) / p.kh_a);
        const double s2 = p.kh_sigma * p.kh_sigma;
        const double dy1 = (y - p.kh_y1) * (y - p.kh_y1);
        const double dy2 = (y - p.kh_y2) * (y - p.kh_y2);
        c.rho = 1.0 + p.kh_rho_fac * 0.5 * (q1 - q2);
        c.u = p.kh_uflow * (q1 - q2 - 1.0);
        c.v = p.kh_amp * std::sin(2.0 * M_PI * x) * (std::exp(-dy1 / s2) + std::exp(-dy2 / s2));
        c.p = p.kh_P0;
        break;
      }
      case GRESHO_VORTEX: {  // initGreshoVortex (Init.h:42-80), Miczek et al. 2015
        const double xr = x - xmid, yr = y - ymid;
        const double r = std::sqrt(xr * xr + yr * yr);
        const double p0 = p.gresho_density / (p.gamma0 * p.gresho_Mach * p.gresho_Mach);
        double uphi;
        c.rho = p.gresho_density;
        if (r < 0.2) {
          uphi = 5.0 * r;
          c.p = p0 + 12.5 * r * r;
        } else if (r < 0.4) {
          uphi = 2.0 - 5.0 * r;
          c.p = p0 + 12.5 * r * r + 4.0 * (1.0 - 5.0 * r + std::log(5.0 * r));
        } else {
          uphi = 0.0;
          c.p = p0 - 2.0 + 4.0 * std::log(2.0);
        }
        const double xnr = xr / r, ynr = yr / r;
        c.u = -ynr * uphi;
        c.v = xnr * uphi;
        break;
      }
      }
      q[0 * n + idx] = c.rho;
      q[1 * n + idx] = c.u;
      q[2 * n + idx] = c.v;
      q[3 * n + idx] = c.p;
    }
  }
}

std::vector<double> init_problem(const RunParams &p) {
  std::vector<double> q(4 * (size_t)p.Nx * p.Ny);
  init_problem(p, q.data());
  return q;
}

}  // namespace fv2dh
