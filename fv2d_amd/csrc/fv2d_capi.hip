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

using namespace fv2d;

struct fv2d_handle {
  fv2d_params p;
  KParams kp;
  int device = 0;
  hipStream_t stream = nullptr;
  size_t esize = 8;  // bytes per stored element
  void *U = nullptr, *Us = nullptr, *Q0 = nullptr;
  StepCtl *ctl = nullptr;
  double *partials = nullptr;
  int nparts = 0;
  StripGeom geom;
  float *grav_row = nullptr;
  double *staging = nullptr;  // 4 * Nx * Ny doubles
  double *rowsum = nullptr;   // 4 * Ny doubles
  double *sums = nullptr;     // 4 doubles
  double *dts_log = nullptr;
  int dts_cap = 0;
  bool have_state = false;
  bool fresh_q0 = false;
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


static void set_global_error(const char *msg) { g_err = msg; }

static bool is_fp32(const fv2d_handle *h) { return h->p.precision == FV2D_FP32; }

static int launch_dt_tiles(fv2d_handle *h, const void *src, int from_q0) {
  const dim3 grid(h->nparts);
  if (is_fp32(h))
    hipLaunchKernelGGL((dt_cells_kernel<float, float>), grid, dim3(256), 0, h->stream, h->kp, (const float *)src,
                       from_q0, h->partials, h->nparts);
  else
    hipLaunchKernelGGL((dt_cells_kernel<double, double>), grid, dim3(256), 0, h->stream, h->kp, (const double *)src,
                       from_q0, h->partials, h->nparts);
  HIPH(h, hipGetLastError());
  return FV2D_OK;
}

static int launch_reduce(fv2d_handle *h, int advance, double *dts_log) {
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, h->stream, h->ctl, h->partials, h->nparts, h->kp, advance,
                     dts_log);
  HIPH(h, hipGetLastError());
  return FV2D_OK;
}

// One stage: out = [avg +] src + dt L(Q(src or Q0)).
static int prof_mark(fv2d_handle *h) {
  if (h->ev_used == h->ev.size()) {
    hipEvent_t e;
    HIPH(h, hipEventCreate(&e));
    h->ev.push_back(e);
  }
  HIPH(h, hipEventRecord(h->ev[h->ev_used++], h->stream));
  return FV2D_OK;
}

static int launch_one_stage(fv2d_handle *h, const void *src, int qsrc, const void *avg, void *out, int epilogue) {
  const int recon = (h->p.reconstruction == FV2D_PLM) ? PLM : PCM;
  hipError_t e;
  int rc;
  if (h->prof && (rc = prof_mark(h))) return rc;
  if (is_fp32(h)) {
    StageIO<float> io;
    io.src = (const float *)src;
    io.q0 = (const float *)h->Q0;
    io.avg = (const float *)avg;
    io.out = (float *)out;
    io.grav_row = h->grav_row;
    io.ctl = h->ctl;
    io.partials = h->partials;
    io.qsrc = qsrc;
    io.epilogue = epilogue;
    e = launch_stage_fp32(h->kp, io, h->geom, recon, h->p.riemann_solver, h->p.tc_active != 0,
                          h->p.visc_active != 0, h->stream);
  } else {
    StageIO<double> io;
    io.src = (const double *)src;
    io.q0 = (const double *)h->Q0;
    io.avg = (const double *)avg;
    io.out = (double *)out;
    io.grav_row = h->grav_row;
    io.ctl = h->ctl;
    io.partials = h->partials;
    io.qsrc = qsrc;
    io.epilogue = epilogue;
    e = launch_stage_fp64(h->kp, io, h->geom, recon, h->p.riemann_solver, h->p.tc_active != 0,
                          h->p.visc_active != 0, h->stream);
  }
  if (e != hipSuccess) FAILH(h, FV2D_E_HIP, "stage launch: %s", hipGetErrorString(e));
  if (h->prof && (rc = prof_mark(h))) return rc;
  return FV2D_OK;
}

// The stages of one step (UpdateFunctor::update, Update.h:193-222) plus the fused
// c2p / reset / dt epilogue. Euler ping-pongs U and Us; RK2 writes U* to Us and the final
// average in place into U.
static int enqueue_step(fv2d_handle *h, bool first_from_q0, bool &swapped) {
  const int q1 = first_from_q0 ? QSRC_Q0 : QSRC_U_RESET;
  int rc;
  swapped = false;
  if (h->p.time_stepping == FV2D_TS_EULER) {
    if ((rc = launch_one_stage(h, h->U, q1, nullptr, h->Us, 1))) return rc;
    void *t = h->U;
    h->U = h->Us;
    h->Us = t;
    swapped = true;
  } else {
    if ((rc = launch_one_stage(h, h->U, q1, nullptr, h->Us, 0))) return rc;
    if ((rc = launch_one_stage(h, h->Us, QSRC_U, h->U, h->U, 1))) return rc;
  }
  return FV2D_OK;
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
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  void *bufs[] = {h->U, h->Us, h->Q0, h->ctl, h->partials, h->grav_row, h->staging, h->rowsum, h->sums, h->dts_log};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (h->stream) (void)hipStreamDestroy(h->stream);
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
  h->p = *p;
  h->esize = (p->precision == FV2D_FP32) ? 4 : 8;
  const int pitch = (p->Nx + 63) / 64 * 64;
  build_kparams(p, pitch, h->kp);
  h->geom = strip_geom(p->Nx, p->Ny, pick_seg_rows(p));
  h->nparts = h->geom.nblocks;

  auto fail = [&](const char *what, hipError_t e) {
    char b[512];
    snprintf(b, sizeof b, "fv2d_create: %s: %s", what, hipGetErrorString(e));
    set_global_error(b);
    fv2d_destroy(h);
    return (e == hipErrorOutOfMemory) ? FV2D_E_NOMEM : FV2D_E_HIP;
  };
  hipError_t e;
  if ((e = hipGetDevice(&h->device)) != hipSuccess) return fail("hipGetDevice", e);
  if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess) return fail("stream", e);
  const size_t state_bytes = 4 * (size_t)h->kp.plane * h->esize;
  if ((e = hipMalloc(&h->U, state_bytes)) != hipSuccess) return fail("hipMalloc U", e);
  if ((e = hipMalloc(&h->Us, state_bytes)) != hipSuccess) return fail("hipMalloc Us", e);
  if ((e = hipMalloc(&h->Q0, state_bytes)) != hipSuccess) return fail("hipMalloc Q0", e);
  if ((e = hipMalloc((void **)&h->ctl, sizeof(StepCtl))) != hipSuccess) return fail("hipMalloc ctl", e);
  if ((e = hipMalloc((void **)&h->partials, sizeof(double) * h->nparts)) != hipSuccess)
    return fail("hipMalloc partials", e);
  if ((e = hipMalloc((void **)&h->rowsum, sizeof(double) * 4 * (size_t)p->Ny)) != hipSuccess)
    return fail("hipMalloc rowsum", e);
  if ((e = hipMalloc((void **)&h->sums, sizeof(double) * 4)) != hipSuccess) return fail("hipMalloc sums", e);
  if (p->gravity_mode == FV2D_GRAV_ANALYTICAL) {
    // getAnalyticalGravity (Gravity.h:15-30) depends on the row only; evaluated here with
    // the host libm and passed through float, as getGravity returns float.
    float *tab = new float[p->Ny];
    for (int j = 0; j < p->Ny; ++j) {
      const double y = p->ymin + (j + 0.5) * h->kp.dy;
      const double g = p->hot_bubble_g0 * sin(y * M_PI * 2.0 / p->ymax);
      tab[j] = (float)g;
    }
    e = hipMalloc((void **)&h->grav_row, sizeof(float) * p->Ny);
    if (e == hipSuccess) e = hipMemcpy(h->grav_row, tab, sizeof(float) * p->Ny, hipMemcpyHostToDevice);
    delete[] tab;
    if (e != hipSuccess) return fail("gravity table", e);
  }
  hipLaunchKernelGGL(clear_ctl_kernel, dim3(1), dim3(1), 0, h->stream, h->ctl);
  if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return fail("init", e);
  *out = h;
  return FV2D_OK;
}

static int ensure_staging(fv2d_handle *h) {
  if (h->staging) return FV2D_OK;
  HIPH(h, hipMalloc((void **)&h->staging, sizeof(double) * 4 * (size_t)h->p.Nx * h->p.Ny));
  return FV2D_OK;
}

int fv2d_set_prim(fv2d_handle *h, const double *q) {
  if (!h) return FV2D_E_INVALID;
  if (!q) FAILH(h, FV2D_E_INVALID, "fv2d_set_prim: q is NULL");
  HIPH(h, hipSetDevice(h->device));
  int rc;
  if ((rc = ensure_staging(h))) return rc;
  const long long n = (long long)h->p.Nx * h->p.Ny;
  HIPH(h, hipMemcpyAsync(h->staging, q, sizeof(double) * 4 * n, hipMemcpyHostToDevice, h->stream));
  if (is_fp32(h))
    hipLaunchKernelGGL(init_state_kernel<float>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->stream, h->kp, h->staging,
                       (float *)h->U, (float *)h->Q0);
  else
    hipLaunchKernelGGL(init_state_kernel<double>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->stream, h->kp,
                       h->staging, (double *)h->U, (double *)h->Q0);
  HIPH(h, hipGetLastError());
  hipLaunchKernelGGL(clear_ctl_kernel, dim3(1), dim3(1), 0, h->stream, h->ctl);
  if ((rc = launch_dt_tiles(h, h->Q0, 1))) return rc;
  if ((rc = launch_reduce(h, 0, nullptr))) return rc;
  HIPH(h, hipStreamSynchronize(h->stream));
  h->have_state = true;
  h->fresh_q0 = true;
  memset(h->last_counts, 0, sizeof h->last_counts);
  return FV2D_OK;
}

int fv2d_set_problem(fv2d_handle *h, const fv2d_problem *prob) {
  if (!h) return FV2D_E_INVALID;
  if (!prob) FAILH(h, FV2D_E_INVALID, "fv2d_set_problem: prob is NULL");
  ProblemArgs a;
  if (problem_args(h->p, h->kp, *prob, a)) FAILH(h, FV2D_E_INVALID, "fv2d_set_problem: unknown kind %d or bad amplitude", prob->kind);
  HIPH(h, hipSetDevice(h->device));
  const long long n = (long long)h->p.Nx * h->p.Ny;
  if (is_fp32(h))
    hipLaunchKernelGGL(problem_ic_kernel<float>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->stream, h->kp, a,
                       (float *)h->U, (float *)h->Q0);
  else
    hipLaunchKernelGGL(problem_ic_kernel<double>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->stream, h->kp, a,
                       (double *)h->U, (double *)h->Q0);
  HIPH(h, hipGetLastError());
  hipLaunchKernelGGL(clear_ctl_kernel, dim3(1), dim3(1), 0, h->stream, h->ctl);
  int rc;
  if ((rc = launch_dt_tiles(h, h->Q0, 1))) return rc;
  if ((rc = launch_reduce(h, 0, nullptr))) return rc;
  HIPH(h, hipStreamSynchronize(h->stream));
  h->have_state = true;
  h->fresh_q0 = true;
  memset(h->last_counts, 0, sizeof h->last_counts);
  return FV2D_OK;
}

int fv2d_compute_dt(fv2d_handle *h, double *dt, double inv[3]) {
  if (!h) return FV2D_E_INVALID;
  if (!h->have_state) FAILH(h, FV2D_E_STATE, "fv2d_compute_dt: no state (call fv2d_set_prim first)");
  HIPH(h, hipSetDevice(h->device));
  StepCtl c;
  HIPH(h, hipMemcpyAsync(&c, h->ctl, sizeof c, hipMemcpyDeviceToHost, h->stream));
  HIPH(h, hipStreamSynchronize(h->stream));
  const double a = c.inv_hyp, b = h->kp.inv_tc, d = h->kp.inv_visc;
  double m = a;
  if (m < b) m = b;
  if (m < d) m = d;
  if (dt) *dt = h->p.CFL / m;
  if (inv) {
    inv[0] = a;
    inv[1] = b;
    inv[2] = d;
  }
  return FV2D_OK;
}

static int read_counts(fv2d_handle *h, unsigned long long tot[3]) {
  StepCtl c;
  HIPH(h, hipMemcpyAsync(&c, h->ctl, sizeof c, hipMemcpyDeviceToHost, h->stream));
  HIPH(h, hipStreamSynchronize(h->stream));
  tot[0] = c.neg_rho;
  tot[1] = c.neg_prs;
  tot[2] = c.nan_cnt;
  return FV2D_OK;
}

int fv2d_step(fv2d_handle *h, double dt, fv2d_step_stats *stats) {
  if (!h) return FV2D_E_INVALID;
  if (!h->have_state) FAILH(h, FV2D_E_STATE, "fv2d_step: no state (call fv2d_set_prim first)");
  HIPH(h, hipSetDevice(h->device));
  hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(1), 0, h->stream, h->ctl, dt, h->kp.dx, h->kp.dy);
  HIPH(h, hipGetLastError());
  bool swapped;
  int rc;
  if ((rc = enqueue_step(h, h->fresh_q0, swapped))) return rc;
  if ((rc = launch_reduce(h, 0, nullptr))) return rc;
  h->fresh_q0 = false;
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
  HIPH(h, hipSetDevice(h->device));
  if (max_steps > h->dts_cap) {
    if (h->dts_log) HIPH(h, hipFree(h->dts_log));
    h->dts_log = nullptr;
    h->dts_cap = 0;
    HIPH(h, hipMalloc((void **)&h->dts_log, sizeof(double) * (size_t)max_steps));
    h->dts_cap = max_steps;
  }
  hipLaunchKernelGGL(set_loop_kernel, dim3(1), dim3(1), 0, h->stream, h->ctl, *t, tend, max_steps);
  HIPH(h, hipGetLastError());
  double *log = (max_steps > 0) ? h->dts_log : nullptr;
  const int chunk = 64;
  int done = 0;
  StepCtl c;
  memset(&c, 0, sizeof c);
  for (;;) {
    // A chunk of steps; steps past the end of the loop are no-ops on the device.
    const int n = (max_steps > 0) ? ((max_steps - done < chunk) ? max_steps - done : chunk) : chunk;
    void *U_start = h->U;
    int swaps = 0;
    for (int k = 0; k < n; ++k) {
      int rc;
      if ((rc = launch_reduce(h, 1, log))) return rc;
      bool swapped;
      if ((rc = enqueue_step(h, h->fresh_q0 && done + k == 0, swapped))) return rc;
      swaps += swapped ? 1 : 0;
    }
    HIPH(h, hipMemcpyAsync(&c, h->ctl, sizeof c, hipMemcpyDeviceToHost, h->stream));
    HIPH(h, hipStreamSynchronize(h->stream));
    const int taken = c.nsteps - done;
    if (swaps) {
      // Euler ping-pong: the state is in U_start after an even number of taken steps.
      if (taken % 2 == 0) {
        if (h->U != U_start) {
          h->Us = h->U;
          h->U = U_start;
        }
      } else if (h->U == U_start) {
        void *x = h->U;
        h->U = h->Us;
        h->Us = x;
      }
    }
    if (c.nsteps > 0) h->fresh_q0 = false;
    done = c.nsteps;
    if (!c.active || (max_steps > 0 && done >= max_steps)) break;
  }
  // reduce the last epilogue so compute_dt sees the final state
  int rc;
  if ((rc = launch_reduce(h, 0, nullptr))) return rc;
  if (dts && done > 0) HIPH(h, hipMemcpyAsync(dts, h->dts_log, sizeof(double) * done, hipMemcpyDeviceToHost, h->stream));
  HIPH(h, hipMemcpyAsync(&c, h->ctl, sizeof c, hipMemcpyDeviceToHost, h->stream));
  HIPH(h, hipStreamSynchronize(h->stream));
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
  const long long n = (long long)h->p.Nx * h->p.Ny;
  const void *src = h->fresh_q0 ? h->Q0 : h->U;
  if (is_fp32(h))
    hipLaunchKernelGGL(get_prim_kernel<float>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->stream, h->kp,
                       (const float *)src, h->fresh_q0 ? 1 : 0, h->staging);
  else
    hipLaunchKernelGGL(get_prim_kernel<double>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->stream, h->kp,
                       (const double *)src, h->fresh_q0 ? 1 : 0, h->staging);
  HIPH(h, hipGetLastError());
  HIPH(h, hipMemcpyAsync(q, h->staging, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, h->stream));
  HIPH(h, hipStreamSynchronize(h->stream));
  return FV2D_OK;
}

int fv2d_get_cons(fv2d_handle *h, double *u) {
  if (!h) return FV2D_E_INVALID;
  if (!u) FAILH(h, FV2D_E_INVALID, "fv2d_get_cons: u is NULL");
  if (!h->have_state) FAILH(h, FV2D_E_STATE, "fv2d_get_cons: no state");
  HIPH(h, hipSetDevice(h->device));
  int rc;
  if ((rc = ensure_staging(h))) return rc;
  const long long n = (long long)h->p.Nx * h->p.Ny;
  if (is_fp32(h))
    hipLaunchKernelGGL(get_cons_kernel<float>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->stream, h->kp,
                       (const float *)h->U, h->staging);
  else
    hipLaunchKernelGGL(get_cons_kernel<double>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, h->stream, h->kp,
                       (const double *)h->U, h->staging);
  HIPH(h, hipGetLastError());
  HIPH(h, hipMemcpyAsync(u, h->staging, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, h->stream));
  HIPH(h, hipStreamSynchronize(h->stream));
  return FV2D_OK;
}

int fv2d_sums(fv2d_handle *h, double out[4]) {
  if (!h) return FV2D_E_INVALID;
  if (!out) FAILH(h, FV2D_E_INVALID, "fv2d_sums: out is NULL");
  if (!h->have_state) FAILH(h, FV2D_E_STATE, "fv2d_sums: no state");
  HIPH(h, hipSetDevice(h->device));
  if (is_fp32(h))
    hipLaunchKernelGGL(row_sum_kernel<float>, dim3(h->p.Ny, 4), dim3(256), 0, h->stream, h->kp, (const float *)h->U,
                       h->rowsum);
  else
    hipLaunchKernelGGL(row_sum_kernel<double>, dim3(h->p.Ny, 4), dim3(256), 0, h->stream, h->kp,
                       (const double *)h->U, h->rowsum);
  HIPH(h, hipGetLastError());
  hipLaunchKernelGGL(total_sum_kernel, dim3(4), dim3(256), 0, h->stream, h->kp, h->rowsum, h->sums);
  HIPH(h, hipGetLastError());
  HIPH(h, hipMemcpyAsync(out, h->sums, sizeof(double) * 4, hipMemcpyDeviceToHost, h->stream));
  HIPH(h, hipStreamSynchronize(h->stream));
  return FV2D_OK;
}

void *fv2d_get_stream(fv2d_handle *h) { return h ? (void *)h->stream : nullptr; }

int fv2d_profile(fv2d_handle *h, int enable) {
  if (!h) return FV2D_E_INVALID;
  HIPH(h, hipSetDevice(h->device));
  HIPH(h, hipStreamSynchronize(h->stream));
  h->prof = enable != 0;
  h->ev_used = 0;
  return FV2D_OK;
}

int fv2d_profile_read(fv2d_handle *h, double *stage_ms, int64_t *stage_launches) {
  if (!h) return FV2D_E_INVALID;
  HIPH(h, hipSetDevice(h->device));
  HIPH(h, hipStreamSynchronize(h->stream));
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
  HIPH(h, hipStreamSynchronize(h->stream));
  return FV2D_OK;
}

}  // extern "C"
