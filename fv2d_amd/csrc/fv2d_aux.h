// fv2d_aux.h -- auxiliary kernels shared by the single-simulation handle and the batch
// engine: state upload, device-generated initial conditions, the dt reduction, loop
// control, read-back and the deterministic conservation sums.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/fv2d_amd.h"
#include "fv2d_device.h"
#include "fv2d_problems.h"

namespace fv2d {

constexpr int AUX_THREADS = 256;

// U = primToCons(q) and Q0 = q on the domain (SimInfo.h:589-600). q: compact planes.
template <typename S>
__global__ void init_state_kernel(KParams kp, const double *q, S *U, S *Q0) {
  const long long n = (long long)kp.Nx * kp.Ny;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(c / kp.Nx), i = (int)(c % kp.Nx);
    St<double> qq;
#pragma unroll
    for (int v = 0; v < 4; ++v) qq.v[v] = q[v * n + c];
    const St<double> uu = prim_to_cons(qq, kp.gm1);
    const long long k = (long long)j * kp.pitch + i;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      U[v * kp.plane + k] = S(uu.v[v]);
      Q0[v * kp.plane + k] = S(qq.v[v]);
    }
  }
}

// Max hyperbolic inverse dt over Q0 or over reset(consToPrim(U)): workgroup b covers a
// contiguous run of cells and writes partials[b] (the stage epilogue's partial layout).
template <typename R, typename S>
__global__ void __launch_bounds__(256) dt_cells_kernel(KParams kp, const S *src, int from_q0, double *partials,
                                                       int nparts) {
  const long long n = (long long)kp.Nx * kp.Ny;
  const long long per = (n + nparts - 1) / nparts;
  const long long c0 = blockIdx.x * per;
  const long long c1 = (c0 + per < n) ? c0 + per : n;
  double m = 0.0;
  for (long long c = c0 + threadIdx.x; c < c1; c += 256) {
    const int j = (int)(c / kp.Nx), i = (int)(c % kp.Nx);
    const long long k = (long long)j * kp.pitch + i;
    St<R> q;
#pragma unroll
    for (int v = 0; v < 4; ++v) q.v[v] = R(src[v * kp.plane + k]);
    if (!from_q0) {
      q = cons_to_prim(q, R(kp.gm1));
      reset_negative(q, R(kp.eps_reset));
    }
    const double inv = inv_dt_hyp(q, kp);
    if (inv > m) m = inv;
  }
  __shared__ double red[4];
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Loop control of main.cpp:62-83 for one step, given the reduced hyperbolic inverse dt:
// while t + epsilon < tend: dt = CFL / max(inv_hyp, inv_tc, inv_visc) (ComputeDt.h:64), take
// the step, t += dt. Sets ctl->active = 0 when the loop has ended.
__device__ inline void advance_control(StepCtl *ctl, double inv_hyp, const KParams &kp, double *dts_log) {
  const bool go = (ctl->t + kp.epsilon < ctl->tend) && (ctl->max_steps <= 0 || ctl->nsteps < ctl->max_steps);
  if (!go) {
    ctl->active = 0;
    return;
  }
  double mm = inv_hyp;  // std::max({hyp, tc, visc}) keeps the first of equal maxima
  if (mm < kp.inv_tc) mm = kp.inv_tc;
  if (mm < kp.inv_visc) mm = kp.inv_visc;
  const double dt = kp.CFL / mm;
  ctl->dt = dt;
  ctl->dt_dx = dt / kp.dx;
  ctl->dt_dy = dt / kp.dy;
  ctl->t = ctl->t + dt;
  if (dts_log) dts_log[ctl->nsteps] = dt;
  ctl->nsteps = ctl->nsteps + 1;
  ctl->active = 1;
}

// Reduce the partial maxima into ctl->inv_hyp; in advance mode also run the loop control.
static __global__ __attribute__((unused)) void __launch_bounds__(1024) reduce_kernel(StepCtl *ctl, const double *partials, int nparts,
                                                             KParams kp, int advance, double *dts_log) {
  double m = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) m = fmax(m, partials[i]);
  __shared__ double red[16];
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double b = red[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); ++i) b = fmax(b, red[i]);
  ctl->inv_hyp = b;
  if (advance) advance_control(ctl, b, kp, dts_log);
}

static __global__ __attribute__((unused)) void set_dt_kernel(StepCtl *ctl, double dt, double dx, double dy) {
  ctl->dt = dt;
  ctl->dt_dx = dt / dx;
  ctl->dt_dy = dt / dy;
  ctl->active = 1;
}

static __global__ __attribute__((unused)) void set_loop_kernel(StepCtl *ctl, double t, double tend, int max_steps) {
  ctl->t = t;
  ctl->tend = tend;
  ctl->max_steps = max_steps;
  ctl->nsteps = 0;
  ctl->active = 1;
}

static __global__ __attribute__((unused)) void clear_ctl_kernel(StepCtl *ctl) {
  ctl->dt = ctl->dt_dx = ctl->dt_dy = 0.0;
  ctl->t = ctl->tend = 0.0;
  ctl->inv_hyp = 0.0;
  ctl->active = 0;
  ctl->nsteps = ctl->max_steps = 0;
  ctl->fresh = 0;
  ctl->neg_rho = ctl->neg_prs = ctl->nan_cnt = 0ull;
}

// Domain primitives (snapshot order) from Q0 or from reset(consToPrim(U)) (main.cpp:80-81).
template <typename S>
__global__ void get_prim_kernel(KParams kp, const S *src, int from_q0, double *q) {
  const long long n = (long long)kp.Nx * kp.Ny;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(c / kp.Nx), i = (int)(c % kp.Nx);
    const long long k = (long long)j * kp.pitch + i;
    St<double> s;
#pragma unroll
    for (int v = 0; v < 4; ++v) s.v[v] = (double)src[v * kp.plane + k];
    if (!from_q0) {
      s = cons_to_prim(s, kp.gm1);
      reset_negative(s, kp.eps_reset);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) q[v * n + c] = s.v[v];
  }
}

template <typename S>
__global__ void get_cons_kernel(KParams kp, const S *src, double *u) {
  const long long n = (long long)kp.Nx * kp.Ny;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(c / kp.Nx), i = (int)(c % kp.Nx);
    const long long k = (long long)j * kp.pitch + i;
#pragma unroll
    for (int v = 0; v < 4; ++v) u[v * n + c] = (double)src[v * kp.plane + k];
  }
}

// Row sums in a fixed order: one workgroup per (row, field), fixed tree over the lanes.
template <typename S>
__global__ void __launch_bounds__(256) row_sum_kernel(KParams kp, const S *src, double *rowsum) {
  const int j = blockIdx.x, v = blockIdx.y;
  const S *row = src + v * kp.plane + (long long)j * kp.pitch;
  double s = 0.0;
  for (int i = threadIdx.x; i < kp.Nx; i += 256) s += (double)row[i];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) rowsum[(long long)v * kp.Ny + j] = red[0];
}

static __global__ __attribute__((unused)) void __launch_bounds__(256) total_sum_kernel(KParams kp, const double *rowsum, double *out) {
  const int v = blockIdx.x;
  double s = 0.0;
  for (int j = threadIdx.x; j < kp.Ny; j += 256) s += rowsum[(long long)v * kp.Ny + j];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[v] = red[0] * kp.dx * kp.dy;
}

// U = primToCons(q) and Q0 = q from a problem evaluated per cell on the device.
template <typename S>
__global__ void problem_ic_kernel(KParams kp, ProblemArgs a, S *U, S *Q0) {
  const long long n = (long long)kp.Nx * kp.Ny;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(c / kp.Nx), i = (int)(c % kp.Nx);
    St<double> qq;
    problem_cell(a, i, j, qq.v);
    const St<double> uu = prim_to_cons(qq, kp.gm1);
    const long long k = (long long)j * kp.pitch + i;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      U[v * kp.plane + k] = S(uu.v[v]);
      Q0[v * kp.plane + k] = S(qq.v[v]);
    }
  }
}

// ProblemArgs of a problem on the grid of p (cell sizes and centre as the solver's).
inline int problem_args(const fv2d_params &p, const KParams &kp, const fv2d_problem &prob, ProblemArgs &a) {
  if (prob.kind != FV2D_PROB_SOD_X && prob.kind != FV2D_PROB_BLAST && prob.kind != FV2D_PROB_KH) return FV2D_E_INVALID;
  if (!(prob.amplitude == prob.amplitude) || prob.amplitude > 1e300 || prob.amplitude < -1e300) return FV2D_E_INVALID;
  a.kind = prob.kind;
  a.Nx = p.Nx;
  a.Ny = p.Ny;
  a.xmin = p.xmin;
  a.ymin = p.ymin;
  a.dx = kp.dx;
  a.dy = kp.dy;
  a.xmid = 0.5 * (p.xmin + p.xmax);
  a.ymid = 0.5 * (p.ymin + p.ymax);
  a.gamma0 = p.gamma0;
  a.seed = prob.seed;
  a.amp = prob.amplitude;
  return FV2D_OK;
}

inline int grid_1d(long long n) {
  long long b = (n + AUX_THREADS - 1) / AUX_THREADS;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace fv2d
