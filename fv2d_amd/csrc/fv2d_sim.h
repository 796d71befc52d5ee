// fv2d_sim.h -- the step sequence and the device-resident time loop of one simulation, on
// device buffers its owner provides: the handle's allocations (fv2d_capi.hip) or the
// no-allocation scratch (fv2d_noalloc.hip). Both entry points run exactly this code.
#pragma once

#include <hip/hip_runtime.h>

#include <string.h>

#include "../../include/fv2d_amd.h"
#include "fv2d_aux.h"
#include "fv2d_device.h"

namespace fv2d {

// Device buffers and geometry of one simulation.
struct SimDev {
  fv2d_params p;
  KParams kp;
  StripGeom geom;
  int nparts = 0;  // per-workgroup partial maxima of the dt reduction (= geom.nblocks)
  hipStream_t stream = nullptr;
  void *U = nullptr, *Us = nullptr, *Q0 = nullptr;  // SoA planes, row pitch kp.pitch
  StepCtl *ctl = nullptr;
  double *partials = nullptr;
  const float *grav_row = nullptr;  // analytical gravity per row, else null
  bool fresh_q0 = false;            // the next first stage reads Q0 (main.cpp:57-58)
  // optional hook run before and after every stage launch (kernel timing)
  hipError_t (*mark)(void *ctx) = nullptr;
  void *mark_ctx = nullptr;
};

// Sets the message fv2d_last_error(NULL) returns (fv2d_capi.hip).
void set_global_error(const char *msg);

// The state is stored in FP32 (FV2D_FP32), else in FP64.
inline bool sim_fp32(const SimDev &d) { return d.p.precision == FV2D_FP32; }
// FP32 arithmetic on an FP64 state (FV2D_FP32_STATE64).
inline bool sim_fp32s64(const SimDev &d) { return d.p.precision == FV2D_FP32_STATE64; }

// Max hyperbolic inverse dt of Q0 (from_q0) or of reset(consToPrim(src)) into the partials.
inline hipError_t sim_launch_dt_tiles(SimDev &d, const void *src, int from_q0) {
  const dim3 grid(d.nparts);
  if (sim_fp32(d))
    hipLaunchKernelGGL((dt_cells_kernel<float, float>), grid, dim3(256), 0, d.stream, d.kp, (const float *)src, from_q0,
                       d.partials, d.nparts);
  else if (sim_fp32s64(d))
    hipLaunchKernelGGL((dt_cells_kernel<float, double>), grid, dim3(256), 0, d.stream, d.kp, (const double *)src,
                       from_q0, d.partials, d.nparts);
  else
    hipLaunchKernelGGL((dt_cells_kernel<double, double>), grid, dim3(256), 0, d.stream, d.kp, (const double *)src,
                       from_q0, d.partials, d.nparts);
  return hipGetLastError();
}

inline hipError_t sim_launch_reduce(SimDev &d, int advance, double *dts_log) {
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, d.stream, d.ctl, d.partials, d.nparts, d.kp, advance,
                     dts_log);
  return hipGetLastError();
}

// One stage: out = [avg +] src + dt L(Q(src or Q0)).
inline hipError_t sim_launch_stage(SimDev &d, const void *src, int qsrc, const void *avg, void *out, int epilogue) {
  const int recon = (d.p.reconstruction == FV2D_PLM) ? PLM : PCM;
  hipError_t e;
  if (d.mark && (e = d.mark(d.mark_ctx)) != hipSuccess) return e;
  if (sim_fp32(d)) {
    StageIO<float> io;
    io.src = (const float *)src;
    io.q0 = (const float *)d.Q0;
    io.avg = (const float *)avg;
    io.out = (float *)out;
    io.grav_row = d.grav_row;
    io.ctl = d.ctl;
    io.partials = d.partials;
    io.qsrc = qsrc;
    io.epilogue = epilogue;
    e = launch_stage_fp32(d.kp, io, d.geom, recon, d.p.riemann_solver, d.p.tc_active != 0, d.p.visc_active != 0,
                          d.stream);
  } else {
    StageIO<double> io;
    io.src = (const double *)src;
    io.q0 = (const double *)d.Q0;
    io.avg = (const double *)avg;
    io.out = (double *)out;
    io.grav_row = d.grav_row;
    io.ctl = d.ctl;
    io.partials = d.partials;
    io.qsrc = qsrc;
    io.epilogue = epilogue;
    if (sim_fp32s64(d))
      e = launch_stage_fp32s64(d.kp, io, d.geom, recon, d.p.riemann_solver, d.p.tc_active != 0,
                               d.p.visc_active != 0, d.stream);
    else
      e = launch_stage_fp64(d.kp, io, d.geom, recon, d.p.riemann_solver, d.p.tc_active != 0,
                            d.p.visc_active != 0, d.stream);
  }
  if (e != hipSuccess) return e;
  if (d.mark && (e = d.mark(d.mark_ctx)) != hipSuccess) return e;
  return hipSuccess;
}

// The stages of one step (UpdateFunctor::update, Update.h:193-222) plus the fused
// c2p / reset / dt epilogue. Euler ping-pongs U and Us; RK2 writes U* to Us and the final
// average in place into U.
inline hipError_t sim_enqueue_step(SimDev &d, bool first_from_q0, bool &swapped) {
  const int q1 = first_from_q0 ? QSRC_Q0 : QSRC_U_RESET;
  hipError_t e;
  swapped = false;
  if (d.p.time_stepping == FV2D_TS_EULER) {
    if ((e = sim_launch_stage(d, d.U, q1, nullptr, d.Us, 1)) != hipSuccess) return e;
    void *t = d.U;
    d.U = d.Us;
    d.Us = t;
    swapped = true;
  } else {
    if ((e = sim_launch_stage(d, d.U, q1, nullptr, d.Us, 0)) != hipSuccess) return e;
    if ((e = sim_launch_stage(d, d.Us, QSRC_U, d.U, d.U, 1)) != hipSuccess) return e;
  }
  return hipSuccess;
}

// The time loop of main.cpp:62-84 without I/O, dt and t on the device: steps are queued in
// chunks of 64 (steps past the end of the loop are no-ops on the device), the control block
// is read back after each chunk. dts_log (device, >= max_steps entries, or null) receives
// the dt of every step. On return c holds the final control block and d.U the state.
inline hipError_t sim_advance(SimDev &d, double t0, double tend, int max_steps, double *dts_log, StepCtl &c) {
  hipError_t e;
  hipLaunchKernelGGL(set_loop_kernel, dim3(1), dim3(1), 0, d.stream, d.ctl, t0, tend, max_steps);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  double *log = (max_steps > 0) ? dts_log : nullptr;
  const int chunk = 64;
  int done = 0;
  memset(&c, 0, sizeof c);
  for (;;) {
    const int n = (max_steps > 0) ? ((max_steps - done < chunk) ? max_steps - done : chunk) : chunk;
    void *U_start = d.U;
    int swaps = 0;
    for (int k = 0; k < n; ++k) {
      if ((e = sim_launch_reduce(d, 1, log)) != hipSuccess) return e;
      bool swapped;
      if ((e = sim_enqueue_step(d, d.fresh_q0 && done + k == 0, swapped)) != hipSuccess) return e;
      swaps += swapped ? 1 : 0;
    }
    if ((e = hipMemcpyAsync(&c, d.ctl, sizeof c, hipMemcpyDeviceToHost, d.stream)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(d.stream)) != hipSuccess) return e;
    const int taken = c.nsteps - done;
    if (swaps) {
      // Euler ping-pong: the state is in U_start after an even number of taken steps.
      if (taken % 2 == 0) {
        if (d.U != U_start) {
          d.Us = d.U;
          d.U = U_start;
        }
      } else if (d.U == U_start) {
        void *x = d.U;
        d.U = d.Us;
        d.Us = x;
      }
    }
    if (c.nsteps > 0) d.fresh_q0 = false;
    done = c.nsteps;
    if (!c.active || (max_steps > 0 && done >= max_steps)) break;
  }
  // reduce the last epilogue so the next dt computation sees the final state
  if ((e = sim_launch_reduce(d, 0, nullptr)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(&c, d.ctl, sizeof c, hipMemcpyDeviceToHost, d.stream)) != hipSuccess) return e;
  return hipStreamSynchronize(d.stream);
}

}  // namespace fv2d
