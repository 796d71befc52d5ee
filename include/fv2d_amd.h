/*
 * fv2d_amd.h -- C ABI of the MI355X (gfx950) explicit 2D compressible-hydro step.
 *
 * This header is the drop-in boundary for the hot path of mdelorme/fv2d: one explicit
 * time step (ghost fill -> PCM/PLM-minmod reconstruction -> HLL/HLLC flux -> conservative
 * update, with constant gravity, viscosity and thermal conduction, Euler or SSP-RK2), the
 * conversion back to primitives with the negative-value reset, and the CFL dt reduction.
 *
 * The reference has no FFI: its path sits behind the functor calls made from main.cpp.
 * Each entry point below names the reference call it replaces (file:line in fv2d).
 *
 * Conventions
 *  - Every function returns 0 on success or a negative FV2D_E* code; no C++ exception
 *    crosses this boundary. fv2d_last_error() gives the message of the last failure.
 *  - Host arrays of primitives/conservatives use the snapshot order of the reference's
 *    IOManager (IOManager.h:155-175): four planes [rho, u, v, prs] (or [rho, rho*u, rho*v,
 *    E]), each Nx*Ny doubles, j-outer / i-inner, domain cells only (no ghosts).
 *  - One handle owns one simulation, its device state and one HIP stream. A handle must
 *    be used by one host thread at a time; distinct handles are independent.
 *  - Numerical trouble is not an error, as in the reference (SimInfo.h:602-646): negative
 *    density/pressure are clamped and counted, NaNs are counted, and the step succeeds.
 */
#ifndef FV2D_AMD_H
#define FV2D_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FV2D_ABI_VERSION 1

/* Error codes. */
#define FV2D_OK 0
#define FV2D_E_INVALID (-1)     /* bad argument or parameter combination */
#define FV2D_E_UNSUPPORTED (-2) /* a reference option this build rejects (message says why) */
#define FV2D_E_HIP (-3)         /* HIP runtime failure */
#define FV2D_E_NOMEM (-4)       /* device or host allocation failed */
#define FV2D_E_STATE (-5)       /* call made in the wrong order (e.g. step before set_prim) */

/* Option values: same members and order as the reference enums (SimInfo.h:41-97). */
enum fv2d_riemann { FV2D_HLL = 0, FV2D_HLLC = 1, FV2D_FSLP = 2 };
enum fv2d_boundary { FV2D_BC_ABSORBING = 0, FV2D_BC_REFLECTING = 1, FV2D_BC_PERIODIC = 2 };
enum fv2d_time_stepping { FV2D_TS_EULER = 0, FV2D_TS_RK2 = 1 };
enum fv2d_reconstruction { FV2D_PCM = 0, FV2D_PCM_WB = 1, FV2D_PLM = 2 };
enum fv2d_gravity_mode { FV2D_GRAV_NONE = 0, FV2D_GRAV_CONSTANT = 1, FV2D_GRAV_ANALYTICAL = 2 };
enum fv2d_tc_mode { FV2D_TCM_CONSTANT = 0, FV2D_TCM_B02 = 1 };
enum fv2d_bctc_mode { FV2D_BCTC_NONE = 0, FV2D_BCTC_FIXED_TEMPERATURE = 1, FV2D_BCTC_FIXED_GRADIENT = 2 };
/* Arithmetic form. FP64 is the parity form. FP32 runs reconstruction, Riemann fluxes and the
 * parabolic/gravity stencils in FP32, stores the state in FP32 and keeps the conserved-state
 * accumulation U += dt*dF, the conservation sums and the dt reduction in FP64. */
enum fv2d_precision { FV2D_FP64 = 0, FV2D_FP32 = 1 };

/*
 * Solver configuration: the hot-path subset of the reference's DeviceParams/Params
 * (SimInfo.h:266-492), with the same names and meaning. dx, dy, Ntx, ... are derived
 * (SimInfo.h:368-376). Fill with fv2d_params_default() first: it sets the reference's
 * .ini defaults (SimInfo.h:360-459, 536-551).
 */
typedef struct fv2d_params {
  /* [mesh] */
  int32_t Nx, Ny;   /* domain cells */
  int32_t Ng;       /* ghost layers (>= 2 for PLM); results do not depend on it */
  double xmin, xmax, ymin, ymax;
  /* [solvers] */
  double CFL;
  int32_t reconstruction; /* fv2d_reconstruction */
  int32_t riemann_solver; /* fv2d_riemann */
  int32_t time_stepping;  /* fv2d_time_stepping */
  /* [run] */
  int32_t boundary_x, boundary_y; /* fv2d_boundary */
  /* [physics] */
  double gamma0;
  int32_t well_balanced_flux_at_y_bc;
  double fslp_K;
  /* [gravity] */
  int32_t gravity_mode; /* fv2d_gravity_mode */
  double gx, gy;
  double hot_bubble_g0; /* analytical mode "hot_bubble" ([hot_bubble] g0) */
  /* [thermal_conduction] */
  int32_t tc_active;
  int32_t tc_mode; /* fv2d_tc_mode */
  double kappa;
  int32_t bctc_ymin, bctc_ymax; /* fv2d_bctc_mode */
  double bctc_ymin_value, bctc_ymax_value;
  /* [viscosity] */
  int32_t visc_active;
  double mu;
  /* [misc] */
  double epsilon;                /* loop-end tolerance and inactive parabolic inverse dt */
  double epsilon_reset_negative; /* value written over negative density / pressure */
  /* this build */
  int32_t precision; /* fv2d_precision */
  int32_t seg_rows;  /* rows per workgroup segment of the stage kernel, multiple of 4 (0: automatic) */
  int32_t reserved[6];
} fv2d_params;

/* Per-step counters of the negative-value reset (SimInfo.h:604-606). */
typedef struct fv2d_step_stats {
  uint64_t negative_density;
  uint64_t negative_pressure;
  uint64_t nan_count;
} fv2d_step_stats;

typedef struct fv2d_handle fv2d_handle;

int fv2d_abi_version(void);

/* Reference .ini defaults (SimInfo.h:360-459 and 536-551). */
void fv2d_params_default(fv2d_params *p);

/* Check a configuration without creating anything. Rejected options return
 * FV2D_E_UNSUPPORTED with the reason in fv2d_last_error(NULL). */
int fv2d_params_check(const fv2d_params *p);

/* Replaces: allocation of U and Q (main.cpp:33-34) and the construction of UpdateFunctor
 * (Update.h:50-56) and ComputeDtFunctor (ComputeDt.h:15) from Params. Allocates the device
 * state on the current HIP device. On failure *h is NULL and fv2d_last_error(NULL) says why. */
int fv2d_create(const fv2d_params *p, fv2d_handle **h);

/* Frees the handle and its device memory. NULL is allowed. */
void fv2d_destroy(fv2d_handle *h);

/* Message of the last failing call on h (h == NULL: the last failing create/check). */
const char *fv2d_last_error(const fv2d_handle *h);

/* Replaces: the primitive state produced by InitFunctor::init (Init.h:310-358) or
 * IOManager::loadSnapshot (IOManager.h:358-379), followed by primToCons (main.cpp:58,
 * SimInfo.h:589-600). q: 4*Nx*Ny doubles in snapshot order. Ghosts are derived from the
 * boundary rules (BoundaryConditions.h:82-147). The next step's first stage and the next
 * fv2d_compute_dt use q exactly as given (not c2p(p2c(q))), as the reference does. */
int fv2d_set_prim(fv2d_handle *h, const double *q);

/* Replaces: ComputeDtFunctor::computeDt (ComputeDt.h:18-65) on the current primitives.
 * dt = CFL / max(inv[0], inv[1], inv[2]) with inv = {hyperbolic, thermal conduction,
 * viscosity} inverse time steps (inactive terms give misc.epsilon). inv may be NULL. */
int fv2d_compute_dt(fv2d_handle *h, double *dt, double inv[3]);

/* Replaces: UpdateFunctor::update (Update.h:193-222), then consToPrim (SimInfo.h:576-587)
 * and checkNegatives (SimInfo.h:602-646), i.e. main.cpp:79-81. One full Euler or SSP-RK2
 * step with the given dt. stats may be NULL. Synchronous: returns when the step is done. */
int fv2d_step(fv2d_handle *h, double dt, fv2d_step_stats *stats);

/* Replaces: the time loop of main.cpp:62-84 without I/O. Runs steps with the dt kept on the
 * device (dt_n = computeDt(Q_n), t += dt_n) while t + epsilon < tend and fewer than
 * max_steps steps were taken in this call (max_steps <= 0: no step limit). Writes the dt of
 * every step into dts (capacity max_steps entries; may be NULL), the number of steps into
 * *nsteps and the accumulated reset counters into stats (may be NULL). *t is the start time
 * on entry and the end time on return. */
int fv2d_advance(fv2d_handle *h, double *t, double tend, int32_t max_steps, double *dts,
                 int32_t *nsteps, fv2d_step_stats *stats);

/* Replaces: the device-to-host copy of Q for a snapshot (IOManager.h:151-170 and 244-262).
 * q: 4*Nx*Ny doubles [rho, u, v, prs] in snapshot order, after the negative-value reset. */
int fv2d_get_prim(fv2d_handle *h, double *q);

/* Conservative state U (4*Nx*Ny doubles [rho, rho*u, rho*v, E], snapshot order). No
 * reference counterpart on the path: used by parity tests and restart checks. */
int fv2d_get_cons(fv2d_handle *h, double *u);

/* Sum over the domain of each conservative field times dx*dy, in a fixed order (the
 * mass/energy totals of python/plot_energy_evolution.py:28-47). out: 4 doubles. */
int fv2d_sums(fv2d_handle *h, double out[4]);

/* Waits for all work queued on the handle's stream. */
int fv2d_synchronize(fv2d_handle *h);

/* The hipStream_t the handle launches its kernels on (as void*). No reference counterpart:
 * lets a caller order its own work or record HIP events against the solver's stream. */
void *fv2d_get_stream(fv2d_handle *h);

/* Kernel timing, the counterpart of the reference's Kokkos kernel labels used by Kokkos
 * Tools (e.g. "Update", Update.h:102). While enabled, every stage-kernel launch is bracketed
 * by HIP events on the handle's stream; fv2d_profile_read returns the summed duration (ms)
 * and the number of timed stage launches since the last read, then resets them. */
int fv2d_profile(fv2d_handle *h, int enable);
int fv2d_profile_read(fv2d_handle *h, double *stage_ms, int64_t *stage_launches);

#ifdef __cplusplus
}
#endif

#endif /* FV2D_AMD_H */
