// fv2d_noalloc.hip -- fv2d_run_noalloc: a whole run (state upload from device primitives,
// the device-resident time loop, primitives back to device memory) that allocates nothing.
//
// The simulation's buffers (U, U*, the initial primitives, the control block, the dt
// partials, the dt log and the analytical-gravity rows) live in `scratch`, a file-scope
// device array of fixed size that every GPU holds, at its own address, until the process
// ends. Host code never passes the array's name to a kernel or a copy (on the host it is a
// stand-in, not a device address): the address of this GPU's copy comes from
// hipGetSymbolAddress, looked up once per device and kept in g_scratch[device].
//
// Calls on one GPU share its array and keep the last call's contents, so everything a run
// reads is written first: the control block is cleared by a kernel, U and Q0 come from the
// caller's primitives, the partials are written before they are reduced. A per-device mutex
// serialises the calls of one process (each call returns only when its work is done);
// calls from several processes each have their own copy of the array.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>

#include "../../include/fv2d_amd.h"
#include "fv2d_params.h"
#include "fv2d_sim.h"

using namespace fv2d;

namespace {

// Layout of the scratch array (byte offsets, each region 256-byte aligned).
constexpr size_t NA_CTL = 0;                                                          // StepCtl
constexpr size_t NA_PARTS = 256;                                                      // dt partials
constexpr size_t NA_DTS = NA_PARTS + sizeof(double) * FV2D_NOALLOC_MAX_WORKGROUPS;    // dt log
constexpr size_t NA_GRAV = NA_DTS + sizeof(double) * FV2D_NOALLOC_MAX_STEPS;          // gravity rows
constexpr size_t NA_STATE = NA_GRAV + sizeof(float) * FV2D_NOALLOC_MAX_ROWS;          // U, U*, Q0
constexpr size_t NA_STATE_BYTES = 4 * sizeof(double) * (size_t)FV2D_NOALLOC_MAX_PLANE;  // one state, FP64
constexpr size_t NA_BYTES = NA_STATE + 3 * NA_STATE_BYTES;
static_assert(NA_DTS % 256 == 0 && NA_GRAV % 256 == 0 && NA_STATE % 256 == 0, "scratch regions must stay aligned");
static_assert(sizeof(StepCtl) <= NA_PARTS, "control block region too small");

}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_lock[kMaxDevices];
unsigned char *g_scratch[kMaxDevices];  // this process's address of `scratch` on each device

// The message goes where fv2d_last_error(NULL) reads it.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
  char b[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(b, sizeof b, fmt, ap);
  va_end(ap);
  fv2d::set_global_error(b);
  return code;
}

// The scratch address on the current device (device `dev`), looked up on the first call.
// Called with g_lock[dev] held.
int scratch_of(int dev, unsigned char **out) {
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(FV2D_E_HIP, "fv2d_run_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch[dev] = (unsigned char *)ptr;
  }
  *out = g_scratch[dev];
  return FV2D_OK;
}

}  // namespace

extern "C" {

int fv2d_run_noalloc(const fv2d_params *p, const double *q_in, double *q_out, double *t, double tend,
                     int32_t max_steps, double *dts, int32_t *nsteps, fv2d_step_stats *stats, void *stream) {
  if (!p || !q_in || !q_out || !t) return fail(FV2D_E_INVALID, "fv2d_run_noalloc: p, q_in, q_out and t must not be NULL");
  std::string why;
  int rc = check_params(p, why);
  if (rc) return fail(rc, "fv2d_run_noalloc: %s", why.c_str());
  if (max_steps <= 0 && dts) return fail(FV2D_E_INVALID, "fv2d_run_noalloc: dts needs max_steps > 0");
  if (max_steps <= 0 && !(tend < FV2D_TEND_NONE))
    return fail(FV2D_E_INVALID, "fv2d_run_noalloc: no end: give tend < 1e300 or max_steps > 0");
  if (max_steps > FV2D_NOALLOC_MAX_STEPS)
    return fail(FV2D_E_INVALID, "fv2d_run_noalloc: max_steps %d exceeds the limit %d", max_steps,
                FV2D_NOALLOC_MAX_STEPS);
  const int pitch = (p->Nx + 63) / 64 * 64;
  const long long plane = (long long)pitch * p->Ny;
  if (plane > FV2D_NOALLOC_MAX_PLANE || p->Ny > FV2D_NOALLOC_MAX_ROWS)
    return fail(FV2D_E_INVALID,
                "fv2d_run_noalloc: grid %d x %d exceeds the scratch limit (row pitch %d x %d rows = %lld cells, at "
                "most %d)",
                p->Nx, p->Ny, pitch, p->Ny, plane, FV2D_NOALLOC_MAX_PLANE);

  SimDev d;
  d.p = *p;
  build_kparams(p, pitch, d.kp);
  d.geom = sim_geom(p);
  d.nparts = d.geom.nblocks;
  if (d.nparts > FV2D_NOALLOC_MAX_WORKGROUPS)
    return fail(FV2D_E_INVALID, "fv2d_run_noalloc: %d workgroups exceed the partials capacity %d (raise seg_rows)",
                d.nparts, FV2D_NOALLOC_MAX_WORKGROUPS);
  d.stream = (hipStream_t)stream;

  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(FV2D_E_HIP, "fv2d_run_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(FV2D_E_INVALID, "fv2d_run_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_lock[dev]);
  unsigned char *base = nullptr;
  if ((rc = scratch_of(dev, &base))) return rc;

  const size_t esize = (p->precision == FV2D_FP32) ? 4 : 8;
  const size_t state_bytes = 4 * esize * (size_t)plane;
  d.ctl = (StepCtl *)(base + NA_CTL);
  d.partials = (double *)(base + NA_PARTS);
  double *dts_log = (double *)(base + NA_DTS);
  d.U = base + NA_STATE;
  d.Us = base + NA_STATE + state_bytes;
  d.Q0 = base + NA_STATE + 2 * state_bytes;

#define NA_HIP(call)                                                                                                   \
  do {                                                                                                                 \
    hipError_t _e = (call);                                                                                            \
    if (_e != hipSuccess) return fail(FV2D_E_HIP, "fv2d_run_noalloc: %s: %s", #call, hipGetErrorString(_e));         \
  } while (0)

  if (p->gravity_mode == FV2D_GRAV_ANALYTICAL) {
    // the row table (host libm, through float) in pieces from a stack buffer
    float *rows = (float *)(base + NA_GRAV);
    float buf[256];
    for (int j0 = 0; j0 < p->Ny; j0 += 256) {
      const int j1 = (j0 + 256 < p->Ny) ? j0 + 256 : p->Ny;
      gravity_rows(*p, d.kp, j0, j1, buf);
      NA_HIP(hipMemcpyAsync(rows + j0, buf, sizeof(float) * (j1 - j0), hipMemcpyHostToDevice, d.stream));
      NA_HIP(hipStreamSynchronize(d.stream));
    }
    d.grav_row = rows;
  }

  // fv2d_set_prim: U = primToCons(q_in), Q0 = q_in, the first dt from Q0
  const long long n = (long long)p->Nx * p->Ny;
  if (esize == 4)
    hipLaunchKernelGGL(init_state_kernel<float>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, d.stream, d.kp, q_in,
                       (float *)d.U, (float *)d.Q0);
  else
    hipLaunchKernelGGL(init_state_kernel<double>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, d.stream, d.kp, q_in,
                       (double *)d.U, (double *)d.Q0);
  NA_HIP(hipGetLastError());
  hipLaunchKernelGGL(clear_ctl_kernel, dim3(1), dim3(1), 0, d.stream, d.ctl);
  NA_HIP(hipGetLastError());
  NA_HIP(sim_launch_dt_tiles(d, d.Q0, 1));
  NA_HIP(sim_launch_reduce(d, 0, nullptr));
  d.fresh_q0 = true;

  // fv2d_advance
  StepCtl c;
  NA_HIP(sim_advance(d, *t, tend, max_steps, dts_log, c));
  if (dts && c.nsteps > 0)
    NA_HIP(hipMemcpyAsync(dts, dts_log, sizeof(double) * c.nsteps, hipMemcpyDeviceToHost, d.stream));

  // fv2d_get_prim
  const void *src = d.fresh_q0 ? d.Q0 : d.U;
  if (esize == 4)
    hipLaunchKernelGGL(get_prim_kernel<float>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, d.stream, d.kp,
                       (const float *)src, d.fresh_q0 ? 1 : 0, q_out);
  else
    hipLaunchKernelGGL(get_prim_kernel<double>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, d.stream, d.kp,
                       (const double *)src, d.fresh_q0 ? 1 : 0, q_out);
  NA_HIP(hipGetLastError());
  NA_HIP(hipStreamSynchronize(d.stream));
#undef NA_HIP
  *t = c.t;
  if (nsteps) *nsteps = c.nsteps;
  if (stats) {
    stats->negative_density = c.neg_rho;
    stats->negative_pressure = c.neg_prs;
    stats->nan_count = c.nan_cnt;
  }
  return FV2D_OK;
}

}  // extern "C"
