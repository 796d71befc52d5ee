// fv2d_device.h -- HIP kernels of the stage (gfx950) and their launch dispatch.
#pragma once

#include <hip/hip_runtime.h>

#include "fv2d_stage.h"

namespace fv2d {

// Tile geometry of the stage kernel: 32 x 8 cells, one cell per thread, 4 wave64s.
constexpr int STAGE_TX = 32;
constexpr int STAGE_TY = 8;

inline int stage_grid_x(int Nx) { return (Nx + STAGE_TX - 1) / STAGE_TX; }
inline int stage_grid_y(int Ny) { return (Ny + STAGE_TY - 1) / STAGE_TY; }

// Wave64 max of a non-negative double.
__device__ inline double wave_max(double m) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  return m;
}

// Epilogue commit: reset counters (rare, atomics) and the workgroup's max inverse dt into
// partials[block] (no atomics on the hot value: the reduction kernel reads all partials).
template <int NT>
__device__ inline void commit_epilogue(const CellRegs &r, StepCtl *ctl, double *partials) {
  __shared__ double red[NT / 64];
  if (r.neg_rho) atomicAdd(&ctl->neg_rho, (unsigned long long)r.neg_rho);
  if (r.neg_prs) atomicAdd(&ctl->neg_prs, (unsigned long long)r.neg_prs);
  if (r.nan_cnt) atomicAdd(&ctl->nan_cnt, (unsigned long long)r.nan_cnt);
  const double m = wave_max(r.inv);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = red[0];
#pragma unroll
    for (int i = 1; i < NT / 64; ++i) b = fmax(b, red[i]);
    partials[blockIdx.x + blockIdx.y * gridDim.x] = b;
  }
}

template <class K, typename S>
__global__ void __launch_bounds__(K::NT) stage_kernel(const KParams kp, const StageIO<S> io) {
  __shared__ typename K::Smem sm;
  const StepCtl *ctl = io.ctl;
  if (!ctl->active) return;  // uniform: the device loop has ended
  const double dt = ctl->dt, dt_dx = ctl->dt_dx, dt_dy = ctl->dt_dy;
  CellRegs r;
  const int tid = threadIdx.x;
  K::phase_load(tid, blockIdx.x, blockIdx.y, sm, kp, io, r);
  __syncthreads();
  K::phase_flux(tid, sm, kp);
  __syncthreads();
  K::phase_update(tid, blockIdx.x, blockIdx.y, sm, kp, io, dt, dt_dx, dt_dy, r);
  if (io.epilogue) commit_epilogue<K::NT>(r, io.ctl, io.partials);
}

template <typename R, typename S, int RECON, int RIEM, bool TC, bool VISC>
hipError_t launch_stage_t(const KParams &kp, const StageIO<S> &io, hipStream_t st) {
  typedef TileStage<R, S, STAGE_TX, STAGE_TY, RECON, RIEM, TC, VISC> K;
  const dim3 grid(stage_grid_x(kp.Nx), stage_grid_y(kp.Ny));
  hipLaunchKernelGGL((stage_kernel<K, S>), grid, dim3(K::NT), 0, st, kp, io);
  return hipGetLastError();
}

template <typename R, typename S, int RECON, int RIEM>
hipError_t launch_stage_rr(const KParams &kp, const StageIO<S> &io, bool tc, bool visc, hipStream_t st) {
  if (tc) {
    if (visc) return launch_stage_t<R, S, RECON, RIEM, true, true>(kp, io, st);
    return launch_stage_t<R, S, RECON, RIEM, true, false>(kp, io, st);
  }
  if (visc) return launch_stage_t<R, S, RECON, RIEM, false, true>(kp, io, st);
  return launch_stage_t<R, S, RECON, RIEM, false, false>(kp, io, st);
}

template <typename R, typename S, int RECON>
hipError_t launch_stage_r(const KParams &kp, const StageIO<S> &io, int riem, bool tc, bool visc, hipStream_t st) {
  switch (riem) {
  case HLL:
    return launch_stage_rr<R, S, RECON, HLL>(kp, io, tc, visc, st);
  case FSLP:
    return launch_stage_rr<R, S, RECON, FSLP>(kp, io, tc, visc, st);
  default:
    return launch_stage_rr<R, S, RECON, HLLC>(kp, io, tc, visc, st);
  }
}

// PCM_WB computes a well-balanced face and then falls through to PCM (Update.h:27-33).
template <typename R, typename S>
hipError_t launch_stage(const KParams &kp, const StageIO<S> &io, int recon, int riem, bool tc, bool visc,
                        hipStream_t st) {
  if (recon == PLM) return launch_stage_r<R, S, PLM>(kp, io, riem, tc, visc, st);
  return launch_stage_r<R, S, PCM>(kp, io, riem, tc, visc, st);
}

// Defined in stage_fp64.hip / stage_fp32.hip.
hipError_t launch_stage_fp64(const KParams &kp, const StageIO<double> &io, int recon, int riem, bool tc, bool visc,
                             hipStream_t st);
hipError_t launch_stage_fp32(const KParams &kp, const StageIO<float> &io, int recon, int riem, bool tc, bool visc,
                             hipStream_t st);

}  // namespace fv2d
