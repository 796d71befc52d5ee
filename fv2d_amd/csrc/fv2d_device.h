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

// Values that are the same in every lane (control-block fields): read once into scalar
// registers, so that branches on them stay uniform and they take no VGPRs.
__device__ inline int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ inline double uni(double x) {
  return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(x)),
                          __builtin_amdgcn_readfirstlane(__double2loint(x)));
}

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

// Waves per SIMD asked of the compiler: K::kWPE / K::kWPEBatch per kernel class (fv2d_stage.h), or one value
// for every class when FV2D_WAVES_PER_EU is defined (build variants for measurements).
#ifdef FV2D_WAVES_PER_EU
#define FV2D_WPE(K) FV2D_WAVES_PER_EU
#define FV2D_WPE_BATCH(K) FV2D_WAVES_PER_EU
#else
#define FV2D_WPE(K) K::kWPE
#define FV2D_WPE_BATCH(K) K::kWPEBatch
#endif

// Rows [jb, je) of the strip starting at column i0 - 2: the segment prologue, then the chunk
// loop.
template <class K, typename R, typename S>
__device__ inline void strip_rows(const KParams &kp, const StageIO<S> &io, int i0, int jb, int je, double dt,
                                  double dt_dx, double dt_dy, typename K::Smem &sm, typename K::Regs &r) {
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: ring slots in SGPRs
  RowPrefetch<S> pf;
  if (K::kPF && jb < je) K::chunk_prefetch(lane, w, 0, i0, jb, kp, io, pf);
  K::pro_load(lane, w, i0, jb, sm, kp, io);
  __syncthreads();
  K::pro_slopes(lane, w, jb, sm);
  __syncthreads();
  K::pro_flux(lane, w, jb, sm, kp);
  __syncthreads();

  int xw = 0;
  for (int jc = jb; jc < je; jc += K::NW) {
    if (K::kPF) {
      K::chunk_commit(lane, w, xw, i0, jb, jc, sm, kp, io, pf, r);
      if (!K::kLateOwn) K::load_own(lane, jc + w, i0, kp, io, r);
      if (jc + K::NW < je) K::chunk_prefetch(lane, w, (xw + 1) % K::NW, i0, jc + K::NW, kp, io, pf);
    } else {
      K::chunk_load(lane, w, xw, i0, jb, jc, sm, kp, io, r);
      if (!K::kLateOwn) K::load_own(lane, jc + w, i0, kp, io, r);
    }
    __syncthreads();
    K::pass1_slopes(lane, w, jb, jc, sm, r);
#pragma unroll
    for (int v = 0; v < 4; ++v) r.sxl[v] = lane_shr1(r.sx[v]);
    __syncthreads();
    // the x and the y interface and the x-face conduction / viscous terms in one phase:
    // independent dependency chains the compiler can interleave (the x-face work reads only
    // registers and the primitive ring, which is stable since the load barrier)
    K::pass1_flux(lane, w, jb, jc, sm, kp, r);
    K::pass2_flux(lane, w, jb, jc, sm, kp, r);
    K::update_faces(lane, w, jb, jc, sm, kp, r);
    if (K::kLateOwn) K::load_own(lane, jc + w, i0, kp, io, r);
#pragma unroll
    for (int v = 0; v < 4; ++v) r.fxl[v] = lane_shr1(r.fx[v]);
    if (K::kTC) r.tcxl = lane_shr1(r.tcx);
    if (K::kVISC) {
#pragma unroll
      for (int v = 0; v < 3; ++v) r.vxl[v] = lane_shr1(r.vx[v]);
    }
    __syncthreads();
    K::update_cell(lane, w, i0, jb, jc, je, sm, kp, io, dt, dt_dx, dt_dy, r);
    __syncthreads();
    xw = (xw + 1) % K::NW;
  }
}

// One workgroup's share of a stage: block `blk` of the simulation described by kp, io, g
// (its strip and rows from block_rows; a block without rows only reports a zero dt partial).
// Shared by the single-simulation kernel and the batched (persistent, work-list) kernel.
template <class K, typename R, typename S>
__device__ inline void strip_block(const KParams &kp, const StageIO<S> &io, const StripGeom &g, int blk, double dt,
                                   double dt_dx, double dt_dy, typename K::Smem &sm) {
  int sx, jb, je;
  block_rows(g, kp.Ny, blk, sx, jb, je);
  const int i0 = sx * STRIP_OUT;

  typename K::Regs r;
  r.neg_rho = r.neg_prs = r.nan_cnt = 0;
  r.inv = 0.0;

  if (jb < je) strip_rows<K, R, S>(kp, io, i0, jb, je, dt, dt_dx, dt_dy, sm, r);
  if (io.epilogue) commit_epilogue<K::NT>(r.neg_rho, r.neg_prs, r.nan_cnt, r.inv, io.ctl, io.partials, blk);
}

template <class K, typename R, typename S>
__global__ void __launch_bounds__(K::NT) __attribute__((amdgpu_waves_per_eu(FV2D_WPE(K))))
strip_kernel(const KParams kp, const StageIO<S> io, const StripGeom g) {
  __shared__ typename K::Smem sm;
  const StepCtl *ctl = io.ctl;
  if (!uni(ctl->active)) return;  // the device loop has ended
  strip_block<K, R, S>(kp, io, g, blockIdx.x, uni(ctl->dt), uni(ctl->dt_dx), uni(ctl->dt_dy), sm);
}

// Batched stage: a persistent grid pulls (job, item) pairs of one kernel class from its run
// [range[0], range[0] + range[1]) of the work list through an atomic counter and runs strip_block on them with each job's own
// parameters, buffers and dt. stage = 1 or 2 (RK2); an Euler job does its whole step in
// stage 1, with U / Us alternating by step parity. A job's arithmetic does not depend on
// the other jobs, so it is bitwise the same alone or in the mix.
template <class K, typename R, typename S>
__global__ void __launch_bounds__(K::NT) __attribute__((amdgpu_waves_per_eu(FV2D_WPE_BATCH(K))))
batch_stage_kernel(const BatchJob *__restrict__ jobs, const BatchItem *__restrict__ items, const int *range,
                   int *counter, int stage) {
  __shared__ typename K::Smem sm;
  __shared__ int s_item;
  const int base = range[0], n = range[1];  // this class's run of the work list
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);  // uniform: descriptor loads stay scalar
    __syncthreads();
    if (it >= n) break;
    const BatchItem wi = items[base + it];
    const BatchJob &J = jobs[wi.job];
    StepCtl *ctl = J.ctl;
    if (!uni(ctl->active)) continue;
    const int fresh = uni(ctl->fresh), nsteps = uni(ctl->nsteps);
    if (!J.rk2 && stage == 2) continue;
    StageIO<S> io;
    io.q0 = (const S *)J.Q0;
    io.grav_row = J.grav_row;
    io.ctl = ctl;
    io.partials = J.partials;
    if (J.rk2) {
      if (stage == 1) {
        io.src = (const S *)J.U;
        io.qsrc = fresh ? QSRC_Q0 : QSRC_U_RESET;
        io.avg = nullptr;
        io.out = (S *)J.Us;
        io.epilogue = 0;
      } else {
        io.src = (const S *)J.Us;
        io.qsrc = QSRC_U;
        io.avg = (const S *)J.U;
        io.out = (S *)J.U;
        io.epilogue = 1;
      }
    } else {
      const bool odd = ((nsteps - 1) & 1) != 0;  // nsteps counts the step being taken
      io.src = (const S *)(odd ? J.Us : J.U);
      io.out = (S *)(odd ? J.U : J.Us);
      io.qsrc = fresh ? QSRC_Q0 : QSRC_U_RESET;
      io.avg = nullptr;
      io.epilogue = 1;
    }
    StripGeom g;
    g.nsx = J.nsx;
    g.nsy = J.nsy;
    g.seg = J.seg;
    g.nblocks = J.nblocks;
    g.nw = K::NW;
    g.sched = 0;  // uniform segments: the work list holds one item per (strip, segment)
    strip_block<K, R, S>(J.kp, io, g, wi.blk, uni(ctl->dt), uni(ctl->dt_dx), uni(ctl->dt_dy), sm);
    __syncthreads();
  }
}

// Kernel class of a simulation: the template switches of the stage and the chunk height.
struct StageClass {
  int recon, riem, tc, visc, xg, nw;
  bool operator==(const StageClass &o) const {
    return recon == o.recon && riem == o.riem && tc == o.tc && visc == o.visc && xg == o.xg && nw == o.nw;
  }
};

template <typename R, typename S, int RECON, int RIEM, bool TC, bool VISC, bool XG, int NW>
hipError_t launch_batch_x(const BatchJob *jobs, const BatchItem *items, const int *range, int *counter, int stage,
                          int grid, hipStream_t st) {
  typedef StripStage<R, S, RECON, RIEM, TC, VISC, XG, NW> K;
  hipLaunchKernelGGL((batch_stage_kernel<K, R, S>), dim3(grid), dim3(K::NT), 0, st, jobs, items, range, counter,
                     stage);
  return hipGetLastError();
}

// Resident workgroups per CU of a class's batched kernel (for the persistent grid size).
template <typename R, typename S, int RECON, int RIEM, bool TC, bool VISC, bool XG, int NW>
int batch_occupancy_x() {
  typedef StripStage<R, S, RECON, RIEM, TC, VISC, XG, NW> K;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_stage_kernel<K, R, S>, K::NT, 0) != hipSuccess) n = 1;
  return n > 0 ? n : 1;
}

// Calls f.template operator()<RECON, RIEM, TC, VISC, XG>() for the class c.
template <class F>
auto with_class(const StageClass &c, F &&f) {
#define FV2D_CLS(RC, RM)                                                                                               \
  if (c.tc) {                                                                                                          \
    if (c.visc) return c.xg ? f.template operator()<RC, RM, true, true, true>()                                        \
                            : f.template operator()<RC, RM, true, true, false>();                                      \
    return c.xg ? f.template operator()<RC, RM, true, false, true>() : f.template operator()<RC, RM, true, false, false>(); \
  }                                                                                                                    \
  if (c.visc) return c.xg ? f.template operator()<RC, RM, false, true, true>()                                         \
                          : f.template operator()<RC, RM, false, true, false>();                                       \
  return c.xg ? f.template operator()<RC, RM, false, false, true>() : f.template operator()<RC, RM, false, false, false>();
  if (c.recon == PLM) {
    if (c.riem == HLL) {
      FV2D_CLS(PLM, HLL)
    } else if (c.riem == FSLP) {
      FV2D_CLS(PLM, FSLP)
    } else {
      FV2D_CLS(PLM, HLLC)
    }
  } else {
    if (c.riem == HLL) {
      FV2D_CLS(PCM, HLL)
    } else if (c.riem == FSLP) {
      FV2D_CLS(PCM, FSLP)
    } else {
      FV2D_CLS(PCM, HLLC)
    }
  }
#undef FV2D_CLS
}

// Single-grid launch chain: run-time switches -> template arguments. DIVF is fixed by the
// caller (stage_fp64.hip / stage_fp64_exact.hip build one division form each, so the two
// forms compile in separate units).
template <typename R, typename S, int RECON, int RIEM, bool TC, bool VISC, bool XG, int NW, int DIVF>
hipError_t launch_stage_n(const KParams &kp, const StageIO<S> &io, const StripGeom &g, hipStream_t st) {
  typedef StripStage<R, S, RECON, RIEM, TC, VISC, XG, NW, DIVF> K;
  hipLaunchKernelGGL((strip_kernel<K, R, S>), dim3(g.nblocks), dim3(K::NT), 0, st, kp, io, g);
  return hipGetLastError();
}

template <typename R, typename S, int RECON, int RIEM, bool TC, bool VISC, bool XG, int DIVF>
hipError_t launch_stage_x(const KParams &kp, const StageIO<S> &io, const StripGeom &g, hipStream_t st) {
  switch (g.nw) {
  case 4:
    return launch_stage_n<R, S, RECON, RIEM, TC, VISC, XG, 4, DIVF>(kp, io, g, st);
  case 8:
    return launch_stage_n<R, S, RECON, RIEM, TC, VISC, XG, 8, DIVF>(kp, io, g, st);
  default:
    return hipErrorInvalidValue;
  }
}

// Gravity, the well-balanced wall flux and wall conduction modes need the XG variant.
inline bool needs_xg(const KParams &kp) { return kp.grav != GRAV_NONE || kp.wb || kp.bctc_ymin || kp.bctc_ymax; }

template <typename R, typename S, int RECON, int RIEM, bool TC, bool VISC, int DIVF>
hipError_t launch_stage_t(const KParams &kp, const StageIO<S> &io, const StripGeom &g, hipStream_t st) {
  if (needs_xg(kp)) return launch_stage_x<R, S, RECON, RIEM, TC, VISC, true, DIVF>(kp, io, g, st);
  return launch_stage_x<R, S, RECON, RIEM, TC, VISC, false, DIVF>(kp, io, g, st);
}

template <typename R, typename S, int RECON, int RIEM, int DIVF>
hipError_t launch_stage_rr(const KParams &kp, const StageIO<S> &io, const StripGeom &g, bool tc, bool visc,
                           hipStream_t st) {
  if (tc) {
    if (visc) return launch_stage_t<R, S, RECON, RIEM, true, true, DIVF>(kp, io, g, st);
    return launch_stage_t<R, S, RECON, RIEM, true, false, DIVF>(kp, io, g, st);
  }
  if (visc) return launch_stage_t<R, S, RECON, RIEM, false, true, DIVF>(kp, io, g, st);
  return launch_stage_t<R, S, RECON, RIEM, false, false, DIVF>(kp, io, g, st);
}

template <typename R, typename S, int RECON, int DIVF>
hipError_t launch_stage_r(const KParams &kp, const StageIO<S> &io, const StripGeom &g, int riem, bool tc, bool visc,
                          hipStream_t st) {
  switch (riem) {
  case HLL:
    return launch_stage_rr<R, S, RECON, HLL, DIVF>(kp, io, g, tc, visc, st);
  case FSLP:
    return launch_stage_rr<R, S, RECON, FSLP, DIVF>(kp, io, g, tc, visc, st);
  default:
    return launch_stage_rr<R, S, RECON, HLLC, DIVF>(kp, io, g, tc, visc, st);
  }
}

// PCM_WB computes a well-balanced face and then falls through to PCM (Update.h:27-33).
template <typename R, typename S, int DIVF>
hipError_t launch_stage(const KParams &kp, const StageIO<S> &io, const StripGeom &g, int recon, int riem, bool tc,
                        bool visc, hipStream_t st) {
  if (recon == PLM) return launch_stage_r<R, S, PLM, DIVF>(kp, io, g, riem, tc, visc, st);
  return launch_stage_r<R, S, PCM, DIVF>(kp, io, g, riem, tc, visc, st);
}

// Defined in stage_fp64.hip / stage_fp64_exact.hip / stage_fp32.hip.
hipError_t launch_batch_fp64(const StageClass &c, const BatchJob *jobs, const BatchItem *items, const int *range,
                             int *counter, int stage, int grid, hipStream_t st);
hipError_t launch_batch_fp32(const StageClass &c, const BatchJob *jobs, const BatchItem *items, const int *range,
                             int *counter, int stage, int grid, hipStream_t st);
int batch_occupancy_fp64(const StageClass &c);
int batch_occupancy_fp32(const StageClass &c);
hipError_t launch_stage_fp64(const KParams &kp, const StageIO<double> &io, const StripGeom &g, int recon, int riem,
                             bool tc, bool visc, hipStream_t st);
hipError_t launch_stage_fp64_exact(const KParams &kp, const StageIO<double> &io, const StripGeom &g, int recon,
                                   int riem, bool tc, bool visc, hipStream_t st);
hipError_t launch_stage_fp32(const KParams &kp, const StageIO<float> &io, const StripGeom &g, int recon, int riem,
                             bool tc, bool visc, hipStream_t st);
hipError_t launch_stage_fp32s64(const KParams &kp, const StageIO<double> &io, const StripGeom &g, int recon,
                                int riem, bool tc, bool visc, hipStream_t st);

}  // namespace fv2d
