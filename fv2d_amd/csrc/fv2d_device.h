// fv2d_device.h -- HIP kernel of the fused stage (gfx950) and its launch dispatch.
#pragma once

#include <hip/hip_runtime.h>

#include "fv2d_stage.h"

namespace fv2d {

// Lane l receives lane l-1's value (DPP wave_shr:1, a VALU move, no LDS); lane 0 gets 0.
__device__ inline int lane_shr1_i(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, false); }
__device__ inline double lane_shr1(double x) {
  return __hiloint2double(lane_shr1_i(__double2hiint(x)), lane_shr1_i(__double2loint(x)));
}
__device__ inline float lane_shr1(float x) { return __int_as_float(lane_shr1_i(__float_as_int(x))); }

// Wave64 max of a non-negative double.
__device__ inline double wave_max(double m) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  return m;
}

// Epilogue commit: reset counters (rare, atomics) and the workgroup's max inverse dt into
// partials[blk] (no atomics on the hot value: the reduction kernel reads all partials).
template <int NT>
__device__ inline void commit_epilogue(unsigned nr, unsigned np, unsigned nn, double inv, StepCtl *ctl,
                                       double *partials, int blk) {
  __shared__ double red[NT / 64];
  if (nr) atomicAdd(&ctl->neg_rho, (unsigned long long)nr);
  if (np) atomicAdd(&ctl->neg_prs, (unsigned long long)np);
  if (nn) atomicAdd(&ctl->nan_cnt, (unsigned long long)nn);
  const double m = wave_max(inv);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = red[0];
#pragma unroll
    for (int i = 1; i < NT / 64; ++i) b = fmax(b, red[i]);
    partials[blk] = b;
  }
}

#ifndef FV2D_WAVES_PER_EU
#define FV2D_WAVES_PER_EU 4
#endif

// Ask the compiler for 4 waves per SIMD (at most 128 VGPRs): with a 40 KB LDS image, 4 workgroups
// of 4 waves fill a CU; the FP64 KH variant would otherwise take 133 VGPRs and 3 waves.
template <class K, typename R, typename S>
__global__ void __launch_bounds__(K::NT) __attribute__((amdgpu_waves_per_eu(FV2D_WAVES_PER_EU))) strip_kernel(const KParams kp, const StageIO<S> io, const StripGeom g) {
  __shared__ typename K::Smem sm;
  const StepCtl *ctl = io.ctl;
  if (!ctl->active) return;  // uniform: the device loop has ended
  const double dt = ctl->dt, dt_dx = ctl->dt_dx, dt_dy = ctl->dt_dy;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: ring slots in SGPRs
  int sx, sy;
  strip_of_block(g, blockIdx.x, sx, sy);
  const int i0 = sx * STRIP_OUT, jb = sy * g.seg;
  const int je = (jb + g.seg < kp.Ny) ? jb + g.seg : kp.Ny;

  LaneRegs<R> r;
  r.neg_rho = r.neg_prs = r.nan_cnt = 0;
  r.inv = 0.0;

  K::pro_load(lane, w, i0, jb, sm, kp, io);
  __syncthreads();
  K::pro_slopes(lane, w, jb, sm);
  __syncthreads();
  K::pro_flux(lane, w, jb, sm, kp);
  __syncthreads();

  int xw = 0;
  for (int jc = jb; jc < je; jc += STRIP_ROWS) {
    K::chunk_load(lane, w, xw, i0, jb, jc, sm, kp, io, r);
    __syncthreads();
    K::pass1_slopes(lane, w, jb, jc, sm, r);
#pragma unroll
    for (int v = 0; v < 4; ++v) r.sxl[v] = lane_shr1(r.sx[v]);
    K::pass1_flux(lane, w, jb, jc, sm, kp, r);
    __syncthreads();
    K::pass2_flux(lane, w, jb, jc, sm, kp, r);
    __syncthreads();
    K::update_faces(lane, w, jb, jc, sm, kp, r);
#pragma unroll
    for (int v = 0; v < 4; ++v) r.fxl[v] = lane_shr1(r.fx[v]);
    if (K::kTC) r.tcxl = lane_shr1(r.tcx);
    if (K::kVISC) {
#pragma unroll
      for (int v = 0; v < 3; ++v) r.vxl[v] = lane_shr1(r.vx[v]);
    }
    K::update_cell(lane, w, i0, jb, jc, je, sm, kp, io, dt, dt_dx, dt_dy, r);
    __syncthreads();
    xw = (xw + 1) & (STRIP_ROWS - 1);
  }
  if (io.epilogue) commit_epilogue<K::NT>(r.neg_rho, r.neg_prs, r.nan_cnt, r.inv, io.ctl, io.partials, blockIdx.x);
}

template <typename R, typename S, int RECON, int RIEM, bool TC, bool VISC, bool XG>
hipError_t launch_stage_x(const KParams &kp, const StageIO<S> &io, const StripGeom &g, hipStream_t st) {
  typedef StripStage<R, S, RECON, RIEM, TC, VISC, XG> K;
  hipLaunchKernelGGL((strip_kernel<K, R, S>), dim3(g.nblocks), dim3(K::NT), 0, st, kp, io, g);
  return hipGetLastError();
}

// Gravity, the well-balanced wall flux and wall conduction modes need the XG variant.
inline bool needs_xg(const KParams &kp) { return kp.grav != GRAV_NONE || kp.wb || kp.bctc_ymin || kp.bctc_ymax; }

template <typename R, typename S, int RECON, int RIEM, bool TC, bool VISC>
hipError_t launch_stage_t(const KParams &kp, const StageIO<S> &io, const StripGeom &g, hipStream_t st) {
  if (needs_xg(kp)) return launch_stage_x<R, S, RECON, RIEM, TC, VISC, true>(kp, io, g, st);
  return launch_stage_x<R, S, RECON, RIEM, TC, VISC, false>(kp, io, g, st);
}

template <typename R, typename S, int RECON, int RIEM>
hipError_t launch_stage_rr(const KParams &kp, const StageIO<S> &io, const StripGeom &g, bool tc, bool visc,
                           hipStream_t st) {
  if (tc) {
    if (visc) return launch_stage_t<R, S, RECON, RIEM, true, true>(kp, io, g, st);
    return launch_stage_t<R, S, RECON, RIEM, true, false>(kp, io, g, st);
  }
  if (visc) return launch_stage_t<R, S, RECON, RIEM, false, true>(kp, io, g, st);
  return launch_stage_t<R, S, RECON, RIEM, false, false>(kp, io, g, st);
}

template <typename R, typename S, int RECON>
hipError_t launch_stage_r(const KParams &kp, const StageIO<S> &io, const StripGeom &g, int riem, bool tc, bool visc,
                          hipStream_t st) {
  switch (riem) {
  case HLL:
    return launch_stage_rr<R, S, RECON, HLL>(kp, io, g, tc, visc, st);
  case FSLP:
    return launch_stage_rr<R, S, RECON, FSLP>(kp, io, g, tc, visc, st);
  default:
    return launch_stage_rr<R, S, RECON, HLLC>(kp, io, g, tc, visc, st);
  }
}

// PCM_WB computes a well-balanced face and then falls through to PCM (Update.h:27-33).
template <typename R, typename S>
hipError_t launch_stage(const KParams &kp, const StageIO<S> &io, const StripGeom &g, int recon, int riem, bool tc,
                        bool visc, hipStream_t st) {
  if (recon == PLM) return launch_stage_r<R, S, PLM>(kp, io, g, riem, tc, visc, st);
  return launch_stage_r<R, S, PCM>(kp, io, g, riem, tc, visc, st);
}

// Defined in stage_fp64.hip / stage_fp32.hip.
hipError_t launch_stage_fp64(const KParams &kp, const StageIO<double> &io, const StripGeom &g, int recon, int riem,
                             bool tc, bool visc, hipStream_t st);
hipError_t launch_stage_fp32(const KParams &kp, const StageIO<float> &io, const StripGeom &g, int recon, int riem,
                             bool tc, bool visc, hipStream_t st);

}  // namespace fv2d
