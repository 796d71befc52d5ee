// fv2d_capi.hip -- C ABI of include/fv2d_amd.h: handle, device memory, the step sequence
// and the auxiliary kernels (state upload, dt reduction, loop control, read-back, sums).
//
// Device layout of one simulation: SoA planes [rho, mx, my, E] of domain cells only,
// row pitch rounded up to 64 elements (no ghosts are stored: the stage kernel derives them
// from the boundary rules). Buffers: U (master), Us (stage output / ping-pong), Q0 (the
// primitives given to fv2d_set_prim, used by the first stage only), a control block, the
// per-workgroup partial maxima of the dt reduction, and a staging buffer for host I/O.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/fv2d_amd.h"
#include "fv2d_aux.h"
#include "fv2d_device.h"
#include "fv2d_params.h"
#include "fv2d_sim.h"

using namespace fv2d;

struct fv2d_handle {
  SimDev d;  // parameters, geometry, stream, state buffers, control block, dt partials
  int device = 0;
  size_t esize = 8;           // bytes per stored element
  float *grav_tab = nullptr;  // analytical gravity per row (owned; d.grav_row points here)
  double *staging = nullptr;  // 4 * Nx * Ny doubles
  double *rowsum = nullptr;   // 4 * Ny doubles
  double *sums = nullptr;     // 4 doubles
  double *dts_log = nullptr;
  int dts_cap = 0;
  bool have_state = false;
  unsigned long long last_counts[3] = {0, 0, 0};
  // stage-kernel timing (fv2d_profile): event pairs recorded since the last read
  bool prof = false;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  std::string err;
};

static thread_local std::string g_err;

#define FAILH(h, code, ...)                                                                                            \
  do {                                                                                                                 \
    char _b[512];                                                                                                      \
    snprintf(_b, sizeof(_b), __VA_ARGS__);                                                                             \
    (h)->err = _b;                                                                                                     \
    return (code);                                                                                                     \
  } while (0)

#define HIPH(h, call)                                                                                                  \
  do {                                                                                                                 \
    hipError_t _e = (call);                                                                                            \
    if (_e != hipSuccess) FAILH(h, FV2D_E_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, __LINE__); \
  } while (0)

// ------------------------------------------------------------------ host helpers


// The message fv2d_last_error(NULL) returns (failing create / check / no-allocation run).
void fv2d::set_global_error(const char *msg) { g_err = msg; }

static bool is_fp32(const fv2d_handle *h) { return sim_fp32(h->d); }

// hipError_t of a shared step-sequence call -> the handle's error code and message
#define SIMH(h, call)                                                                                                  \
  do {                                                                                                                 \
    hipError_t _e = (call);                                                                                            \
    if (_e != hipSuccess) FAILH(h, FV2D_E_HIP, "%s: %s", #call, hipGetErrorString(_e));                              \
  } while (0)

// fv2d_profile: an event on the handle's stream before and after every stage launch. The
// events only time the launches (they are read after a stream synchronization), so they skip
// the system-scope memory fence a default event performs when it completes.
static const unsigned kProfEventFlags = hipEventDisableSystemFence;
static hipError_t prof_mark(void *ctx) {
  fv2d_handle *h = (fv2d_handle *)ctx;
  if (h->ev_used == h->ev.size()) {
    hipEvent_t e;
    hipError_t r = hipEventCreateWithFlags(&e, kProfEventFlags);
    if (r != hipSuccess) return r;
    h->ev.push_back(e);
  }
  return hipEventRecord(h->ev[h->ev_used++], h->d.stream);
}

// ------------------------------------------------------------------ C ABI

extern "C" {

int fv2d_abi_version(void) { return FV2D_ABI_VERSION; }

void fv2d_params_default(fv2d_params *p) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  p->Nx = 32;
  p->Ny = 32;
  p->Ng = 2;
  p->xmin = 0.0;
  p->xmax = 1.0;
  p->ymin = 0.0;
  p->ymax = 1.0;
  p->CFL = 0.8;
  p->reconstruction = FV2D_PCM;
  p->riemann_solver = FV2D_HLLC;
  p->time_stepping = FV2D_TS_EULER;
  p->boundary_x = FV2D_BC_REFLECTING;
  p->boundary_y = FV2D_BC_REFLECTING;
  p->gamma0 = 5.0 / 3.0;
  p->well_balanced_flux_at_y_bc = 0;
  p->fslp_K = 1.1;
  p->gravity_mode = FV2D_GRAV_NONE;
  p->gx = 0.0;
  p->gy = 0.0;
  p->hot_bubble_g0 = 0.0;
  p->tc_active = 0;
  p->tc_mode = FV2D_TCM_CONSTANT;
  p->kappa = 0.0;
  p->bctc_ymin = FV2D_BCTC_NONE;
  p->bctc_ymax = FV2D_BCTC_NONE;
  p->bctc_ymin_value = 1.0;
  p->bctc_ymax_value = 1.0;
  p->visc_active = 0;
  p->mu = 0.0;
  p->epsilon = 1.0e-6;
  p->epsilon_reset_negative = 1.0e-8;
  p->precision = FV2D_FP64;
  p->seg_rows = 0;
}

int fv2d_params_check(const fv2d_params *p) {
  std::string why;
  int rc = check_params(p, why);
  if (rc) set_global_error(why.c_str());
  return rc;
}

const char *fv2d_last_error(const fv2d_handle *h) {
  if (!h) return g_err.c_str();
  return h->err.c_str();
}

void fv2d_destroy(fv2d_handle *h) {
  if (!h) return;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(h->device);
  if (h->d.stream) (void)hipStreamSynchronize(h->d.stream);
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  void *bufs[] = {h->d.U, h->d.Us, h->d.Q0, h->d.ctl, h->d.partials, h->grav_tab, h->staging, h->rowsum, h->sums, h->dts_log};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (h->d.stream) (void)hipStreamDestroy(h->d.stream);
  (void)hipSetDevice(cur);
  delete h;
}

int fv2d_create(const fv2d_params *p, fv2d_handle **out) {
  if (!out) {
    set_global_error("fv2d_create: out is NULL");
    return FV2D_E_INVALID;
  }
  *out = nullptr;
  std::string why;
  int rc = check_params(p, why);
  if (rc) {
    set_global_error(why.c_str());
    return rc;
  }
  fv2d_handle *h = new fv2d_handle();
  h->d.p = *p;
  h->esize = (p->precision == FV2D_FP32) ? 4 : 8;
  const int pitch = (p->Nx + 63) / 64 * 64;
  build_kparams(p, pitch, h->d.kp);
  h->d.geom = sim_geom(p);
  h->d.nparts = h->d.geom.nblocks;

  auto fail = [&](const char *what, hipError_t e) {
    char b[512];
    snprintf(b, sizeof b, "fv2d_create: %s: %s", what, hipGetErrorString(e));
    set_global_error(b);
    fv2d_destroy(h);
    return (e == hipErrorOutOfMemory) ? FV2D_E_NOMEM : FV2D_E_HIP;
  };
  hipError_t e;
  if ((e = hipGetDevice(&h->device)) != hipSuccess) return fail("hipGetDevice", e);
  if ((e = hipStreamCreateWithFlags(&h->d.stream, hipStreamNonBlocking)) != hipSuccess) return fail("stream", e);
  const size_t state_bytes = 4 * (size_t)h->d.kp.plane * h->esize;
  if ((e = hipMalloc(&h->d.U, state_bytes)) != hipSuccess) return fail("hipMalloc U", e);
  if ((e = hipMalloc(&h->d.Us, state_bytes)) != hipSuccess) return fail("hipMalloc Us", e);
  if ((e = hipMalloc(&h->d.Q0, state_bytes)) != hipSuccess) return fail("hipMalloc Q0", e);
  if ((e = hipMalloc((void **)&h->d.ctl, sizeof(StepCtl))) != hipSuccess) return fail("hipMalloc ctl", e);
  if ((e = hipMalloc((void **)&h->d.partials, sizeof(double) * h->d.nparts)) != hipSuccess)
    return fail("hipMalloc partials", e);
  if ((e = hipMalloc((void **)&h->rowsum, sizeof(double) * 4 * (size_t)p->Ny)) != hipSuccess)
    return fail("hipMalloc rowsum", e);
  if ((e = hipMalloc((void **)&h->sums, sizeof(double) * 4)) != hipSuccess) return fail("hipMalloc sums", e);
  if (p->gravity_mode == FV2D_GRAV_ANALYTICAL) {
    // getAnalyticalGravity (Gravity.h:15-30) depends on the row only; evaluated here with
    // the host libm and passed through float, as getGravity returns float.
    float *tab = new float[p->Ny];
    gravity_rows(*p, h->d.kp, 0, p->Ny, tab);
    e = hipMalloc((void **)&h->grav_tab, sizeof(float) * p->Ny);
    if (e == hipSuccess) e = hipMemcpy(h->grav_tab, tab, sizeof(float) * p->Ny, hipMemcpyHostToDevice);
    delete[] tab;
    if (e != hipSuccess) return fail("gravity table", e);
    h->d.grav_row = h->grav_tab;
  }
  hipLaunchKernelGGL(clear_ctl_kernel, dim3(1), dim3(1), 0, h->d.stream, h->d.ctl);
  if ((e = hipStreamSynchronize(h->d.stream)) != hipSuccess) return fail("init", e);
  *out = h;
  return FV2D_OK;
}

static int ensure_staging(fv2d_handle *h) {
  if (h->staging) return FV2D_OK;
  HIPH(h, hipMalloc((void **)&h->staging, sizeof(double) * 4 * (size_t)h->d.p.Nx * h->d.p.Ny));
  return FV2D_OK;
}

int fv2d_set_prim(fv2d_handle *h, const double *q) {
  if (!h) return FV2D_E_INVALID;
  if (!q) FAILH(h, FV2D_E_INVALID, "fv2d_set_prim: q is NULL");
  HIPH(h, hipSetDevice(h->device));
  int rc;
  if ((rc = ensure_staging(h))) return rc;
  const long long n = (long long)h->d.p.Nx * h->d.p.Ny;
  HIPH(h, hipMemcpyAsync(h->staging, q, sizeof(double) * 4 * n, hipMemcpyHostToDevice, h->d.stream));
  if (is_fp32(h))
    hipLaunchKernelGGL(init_state_kernel<float>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->d.stream, h->d.kp, h->staging,
                       (float *)h->d.U, (float *)h->d.Q0);
  else
    hipLaunchKernelGGL(init_state_kernel<double>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->d.stream, h->d.kp,
                       h->staging, (double *)h->d.U, (double *)h->d.Q0);
  HIPH(h, hipGetLastError());
  hipLaunchKernelGGL(clear_ctl_kernel, dim3(1), dim3(1), 0, h->d.stream, h->d.ctl);
  SIMH(h, sim_launch_dt_tiles(h->d, h->d.Q0, 1));
  SIMH(h, sim_launch_reduce(h->d, 0, nullptr));
  HIPH(h, hipStreamSynchronize(h->d.stream));
  h->have_state = true;
  h->d.fresh_q0 = true;
  memset(h->last_counts, 0, sizeof h->last_counts);
  return FV2D_OK;
}

int fv2d_set_problem(fv2d_handle *h, const fv2d_problem *prob) {
  if (!h) return FV2D_E_INVALID;
  if (!prob) FAILH(h, FV2D_E_INVALID, "fv2d_set_problem: prob is NULL");
  ProblemArgs a;
  if (problem_args(h->d.p, h->d.kp, *prob, a)) FAILH(h, FV2D_E_INVALID, "fv2d_set_problem: unknown kind %d or bad amplitude", prob->kind);
  HIPH(h, hipSetDevice(h->device));
  const long long n = (long long)h->d.p.Nx * h->d.p.Ny;
  if (is_fp32(h))
    hipLaunchKernelGGL(problem_ic_kernel<float>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->d.stream, h->d.kp, a,
                       (float *)h->d.U, (float *)h->d.Q0);
  else
    hipLaunchKernelGGL(problem_ic_kernel<double>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->d.stream, h->d.kp, a,
                       (double *)h->d.U, (double *)h->d.Q0);
  HIPH(h, hipGetLastError());
  hipLaunchKernelGGL(clear_ctl_kernel, dim3(1), dim3(1), 0, h->d.stream, h->d.ctl);
  SIMH(h, sim_launch_dt_tiles(h->d, h->d.Q0, 1));
  SIMH(h, sim_launch_reduce(h->d, 0, nullptr));
  HIPH(h, hipStreamSynchronize(h->d.stream));
  h->have_state = true;
  h->d.fresh_q0 = true;
  memset(h->last_counts, 0, sizeof h->last_counts);
  return FV2D_OK;
}

int fv2d_compute_dt(fv2d_handle *h, double *dt, double inv[3]) {
  if (!h) return FV2D_E_INVALID;
  if (!h->have_state) FAILH(h, FV2D_E_STATE, "fv2d_compute_dt: no state (call fv2d_set_prim first)");
  HIPH(h, hipSetDevice(h->device));
  StepCtl c;
  HIPH(h, hipMemcpyAsync(&c, h->d.ctl, sizeof c, hipMemcpyDeviceToHost, h->d.stream));
  HIPH(h, hipStreamSynchronize(h->d.stream));
  const double a = c.inv_hyp, b = h->d.kp.inv_tc, d = h->d.kp.inv_visc;
  double m = a;
  if (m < b) m = b;
  if (m < d) m = d;
  if (dt) *dt = h->d.p.CFL / m;
  if (inv) {
    inv[0] = a;
    inv[1] = b;
    inv[2] = d;
  }
  return FV2D_OK;
}

static int read_counts(fv2d_handle *h, unsigned long long tot[3]) {
  StepCtl c;
  HIPH(h, hipMemcpyAsync(&c, h->d.ctl, sizeof c, hipMemcpyDeviceToHost, h->d.stream));
  HIPH(h, hipStreamSynchronize(h->d.stream));
  tot[0] = c.neg_rho;
  tot[1] = c.neg_prs;
  tot[2] = c.nan_cnt;
  return FV2D_OK;
}

int fv2d_step(fv2d_handle *h, double dt, fv2d_step_stats *stats) {
  if (!h) return FV2D_E_INVALID;
  if (!h->have_state) FAILH(h, FV2D_E_STATE, "fv2d_step: no state (call fv2d_set_prim first)");
  HIPH(h, hipSetDevice(h->device));
  hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(1), 0, h->d.stream, h->d.ctl, dt, h->d.kp.dx, h->d.kp.dy);
  HIPH(h, hipGetLastError());
  bool swapped;
  int rc;
  SIMH(h, sim_enqueue_step(h->d, h->d.fresh_q0, swapped));
  SIMH(h, sim_launch_reduce(h->d, 0, nullptr));
  h->d.fresh_q0 = false;
  unsigned long long tot[3];
  if ((rc = read_counts(h, tot))) return rc;
  if (stats) {
    stats->negative_density = tot[0] - h->last_counts[0];
    stats->negative_pressure = tot[1] - h->last_counts[1];
    stats->nan_count = tot[2] - h->last_counts[2];
  }
  memcpy(h->last_counts, tot, sizeof tot);
  return FV2D_OK;
}

int fv2d_advance(fv2d_handle *h, double *t, double tend, int32_t max_steps, double *dts, int32_t *nsteps,
                 fv2d_step_stats *stats) {
  if (!h) return FV2D_E_INVALID;
  if (!t) FAILH(h, FV2D_E_INVALID, "fv2d_advance: t is NULL");
  if (!h->have_state) FAILH(h, FV2D_E_STATE, "fv2d_advance: no state (call fv2d_set_prim first)");
  if (max_steps <= 0 && dts) FAILH(h, FV2D_E_INVALID, "fv2d_advance: dts needs max_steps > 0");
  if (max_steps <= 0 && !(tend < FV2D_TEND_NONE))
    FAILH(h, FV2D_E_INVALID, "fv2d_advance: no end: give tend < 1e300 or max_steps > 0");
  HIPH(h, hipSetDevice(h->device));
  // the device log of the dts exists only when the caller asks for them (no allocation
  // otherwise)
  if (dts && max_steps > h->dts_cap) {
    if (h->dts_log) HIPH(h, hipFree(h->dts_log));
    h->dts_log = nullptr;
    h->dts_cap = 0;
    HIPH(h, hipMalloc((void **)&h->dts_log, sizeof(double) * (size_t)max_steps));
    h->dts_cap = max_steps;
  }
  StepCtl c;
  SIMH(h, sim_advance(h->d, *t, tend, max_steps, dts ? h->dts_log : nullptr, c));
  const int done = c.nsteps;
  if (dts && done > 0) {
    HIPH(h, hipMemcpyAsync(dts, h->dts_log, sizeof(double) * done, hipMemcpyDeviceToHost, h->d.stream));
    HIPH(h, hipStreamSynchronize(h->d.stream));
  }
  *t = c.t;
  if (nsteps) *nsteps = done;
  unsigned long long tot[3] = {c.neg_rho, c.neg_prs, c.nan_cnt};
  if (stats) {
    stats->negative_density = tot[0] - h->last_counts[0];
    stats->negative_pressure = tot[1] - h->last_counts[1];
    stats->nan_count = tot[2] - h->last_counts[2];
  }
  memcpy(h->last_counts, tot, sizeof tot);
  return FV2D_OK;
}

int fv2d_get_prim(fv2d_handle *h, double *q) {
  if (!h) return FV2D_E_INVALID;
  if (!q) FAILH(h, FV2D_E_INVALID, "fv2d_get_prim: q is NULL");
  if (!h->have_state) FAILH(h, FV2D_E_STATE, "fv2d_get_prim: no state");
  HIPH(h, hipSetDevice(h->device));
  int rc;
  if ((rc = ensure_staging(h))) return rc;
  const long long n = (long long)h->d.p.Nx * h->d.p.Ny;
  const void *src = h->d.fresh_q0 ? h->d.Q0 : h->d.U;
  if (is_fp32(h))
    hipLaunchKernelGGL(get_prim_kernel<float>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->d.stream, h->d.kp,
                       (const float *)src, h->d.fresh_q0 ? 1 : 0, h->staging);
  else
    hipLaunchKernelGGL(get_prim_kernel<double>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->d.stream, h->d.kp,
                       (const double *)src, h->d.fresh_q0 ? 1 : 0, h->staging);
  HIPH(h, hipGetLastError());
  HIPH(h, hipMemcpyAsync(q, h->staging, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, h->d.stream));
  HIPH(h, hipStreamSynchronize(h->d.stream));
  return FV2D_OK;
}

int fv2d_get_cons(fv2d_handle *h, double *u) {
  if (!h) return FV2D_E_INVALID;
  if (!u) FAILH(h, FV2D_E_INVALID, "fv2d_get_cons: u is NULL");
  if (!h->have_state) FAILH(h, FV2D_E_STATE, "fv2d_get_cons: no state");
  HIPH(h, hipSetDevice(h->device));
  int rc;
  if ((rc = ensure_staging(h))) return rc;
  const long long n = (long long)h->d.p.Nx * h->d.p.Ny;
  if (is_fp32(h))
    hipLaunchKernelGGL(get_cons_kernel<float>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->d.stream, h->d.kp,
                       (const float *)h->d.U, h->staging);
  else
    hipLaunchKernelGGL(get_cons_kernel<double>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->d.stream, h->d.kp,
                       (const double *)h->d.U, h->staging);
  HIPH(h, hipGetLastError());
  HIPH(h, hipMemcpyAsync(u, h->staging, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, h->d.stream));
  HIPH(h, hipStreamSynchronize(h->d.stream));
  return FV2D_OK;
}

int fv2d_sums(fv2d_handle *h, double out[4]) {
  if (!h) return FV2D_E_INVALID;
  if (!out) FAILH(h, FV2D_E_INVALID, "fv2d_sums: out is NULL");
  if (!h->have_state) FAILH(h, FV2D_E_STATE, "fv2d_sums: no state");
  HIPH(h, hipSetDevice(h->device));
  if (is_fp32(h))
    hipLaunchKernelGGL(row_sum_kernel<float>, dim3(h->d.p.Ny, 4), dim3(256), 0, h->d.stream, h->d.kp, (const float *)h->d.U,
                       h->rowsum);
  else
    hipLaunchKernelGGL(row_sum_kernel<double>, dim3(h->d.p.Ny, 4), dim3(256), 0, h->d.stream, h->d.kp,
                       (const double *)h->d.U, h->rowsum);
  HIPH(h, hipGetLastError());
  hipLaunchKernelGGL(total_sum_kernel, dim3(4), dim3(256), 0, h->d.stream, h->d.kp, h->rowsum, h->sums);
  HIPH(h, hipGetLastError());
  HIPH(h, hipMemcpyAsync(out, h->sums, sizeof(double) * 4, hipMemcpyDeviceToHost, h->d.stream));
  HIPH(h, hipStreamSynchronize(h->d.stream));
  return FV2D_OK;
}

void *fv2d_get_stream(fv2d_handle *h) { return h ? (void *)h->d.stream : nullptr; }

int fv2d_profile(fv2d_handle *h, int enable) {
  if (!h) return FV2D_E_INVALID;
  if (enable < 0) FAILH(h, FV2D_E_INVALID, "fv2d_profile: enable=%d must be >= 0", enable);
  HIPH(h, hipSetDevice(h->device));
  HIPH(h, hipStreamSynchronize(h->d.stream));
  // the events of the launches to come are created here, not while they are enqueued: a
  // hipEventCreate inside a device loop holds the host back and starves the GPU
  const size_t want = enable > 1 ? 2 * (size_t)enable : (enable ? 256 : 0);
  while (h->ev.size() < want) {
    hipEvent_t e;
    HIPH(h, hipEventCreateWithFlags(&e, kProfEventFlags));
    h->ev.push_back(e);
  }
  h->prof = enable != 0;
  h->d.mark = h->prof ? prof_mark : nullptr;
  h->d.mark_ctx = h;
  h->ev_used = 0;
  return FV2D_OK;
}

int fv2d_profile_read(fv2d_handle *h, double *stage_ms, int64_t *stage_launches) {
  if (!h) return FV2D_E_INVALID;
  HIPH(h, hipSetDevice(h->device));
  HIPH(h, hipStreamSynchronize(h->d.stream));
  double ms = 0.0;
  for (size_t k = 0; k + 1 < h->ev_used; k += 2) {
    float e = 0.0f;
    HIPH(h, hipEventElapsedTime(&e, h->ev[k], h->ev[k + 1]));
    ms += e;
  }
  if (stage_ms) *stage_ms = ms;
  if (stage_launches) *stage_launches = (int64_t)(h->ev_used / 2);
  h->ev_used = 0;
  return FV2D_OK;
}

int fv2d_synchronize(fv2d_handle *h) {
  if (!h) return FV2D_E_INVALID;
  HIPH(h, hipSetDevice(h->device));
  HIPH(h, hipStreamSynchronize(h->d.stream));
  return FV2D_OK;
}

}  // extern "C"
