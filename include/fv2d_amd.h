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
#define FV2D_E_BUSY (-6)        /* batch: no free slot or arena space for the job right now */

/* A loop end time at or above this value (or NaN) counts as "no end time": fv2d_advance and
 * fv2d_batch_submit then need max_steps > 0 and return FV2D_E_INVALID otherwise. */
#define FV2D_TEND_NONE 1e300

/* Option values: same members and order as the reference enums (SimInfo.h:41-97). */
enum fv2d_riemann { FV2D_HLL = 0, FV2D_HLLC = 1, FV2D_FSLP = 2 };
enum fv2d_boundary { FV2D_BC_ABSORBING = 0, FV2D_BC_REFLECTING = 1, FV2D_BC_PERIODIC = 2 };
enum fv2d_time_stepping { FV2D_TS_EULER = 0, FV2D_TS_RK2 = 1 };
enum fv2d_reconstruction { FV2D_PCM = 0, FV2D_PCM_WB = 1, FV2D_PLM = 2 };
enum fv2d_gravity_mode { FV2D_GRAV_NONE = 0, FV2D_GRAV_CONSTANT = 1, FV2D_GRAV_ANALYTICAL = 2 };
enum fv2d_tc_mode { FV2D_TCM_CONSTANT = 0, FV2D_TCM_B02 = 1 };
enum fv2d_bctc_mode { FV2D_BCTC_NONE = 0, FV2D_BCTC_FIXED_TEMPERATURE = 1, FV2D_BCTC_FIXED_GRADIENT = 2 };
/* Arithmetic form. FP64 is the parity form. FP32 runs reconstruction, Riemann fluxes and the
 * parabolic/gravity stencils in FP32 and stores the state in FP32; the step's increment dt*dF
 * is formed in FP32 and added to U, with the RK2 average, in FP32 (each rounded once, like
 * the storage); the dt reduction over cells and the conservation sums are FP64.
 * FP32_STATE64 is the FP32 form with the state U stored in FP64 (the increment is added to
 * the FP64 state in FP64, and the state is never rounded to FP32): twice the state bytes of
 * FP32, the FP32 form's arithmetic. Batches take FP64 or FP32. */
enum fv2d_precision { FV2D_FP64 = 0, FV2D_FP32 = 1, FV2D_FP32_STATE64 = 2 };

/*
 * Solver configuration: the hot-path subset of the reference's DeviceParams/Params
 * (SimInfo.h:266-492), with the same names and meaning. dx, dy, Ntx, ... are derived
 * (SimInfo.h:368-376). Fill with fv2d_params_default() first: it sets the reference's
 * .ini defaults (SimInfo.h:360-459, 536-551).
 */
typedef struct fv2d_params {
  /* [mesh] */
  int32_t Nx, Ny;   /* domain cells: 2 <= Nx <= 2^26, Ny >= 2, Nx * Ny <= 2^30 */
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
  int32_t seg_rows;   /* rows per workgroup segment of the stage kernel, >= 4 (0: automatic); on
                         grids of 2048 or more segments these are the long segments, and the
                         last seg_rows rows of each eighth of the grid run in segments of a
                         quarter of it (the launch then ends on short workgroups) */
  int32_t chunk_rows; /* rows per chunk = waves per workgroup of the stage kernel: 4 or 8 (0: automatic) */
  int32_t reserved[5];
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

/* Message of the last failing call on h (h == NULL: the last failing create / check /
 * fv2d_run_noalloc of the calling thread). */
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
 * max_steps steps were taken in this call (max_steps <= 0: no step limit). A call with
 * neither end (max_steps <= 0 and tend >= FV2D_TEND_NONE or NaN) returns FV2D_E_INVALID.
 * Writes the dt of every step into dts (capacity max_steps entries; may be NULL), the number
 * of steps into *nsteps and the accumulated reset counters into stats (may be NULL). *t is
 * the start time on entry and the end time on return. */
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
 * and the number of timed stage launches since the last read, then resets them. enable: 0
 * off; 1 on, with events for 128 launches created before returning; n > 1 on, with events
 * for n launches created before returning (more are created as the launches need them, which
 * costs host time inside the caller's timed region). FV2D_E_INVALID for enable < 0. */
int fv2d_profile(fv2d_handle *h, int enable);
int fv2d_profile_read(fv2d_handle *h, double *stage_ms, int64_t *stage_launches);

/* ------------------------------------------------------------------------------------
 * Initial conditions generated on the device.
 *
 * The problems of the benchmark and serving workloads, evaluated per cell on the GPU so a
 * job needs no host upload: FV2D_PROB_SOD_X (initSodX, Init.h:21-41: rho 1 / 0.125, p 1 /
 * 0.1, split at x = 0.5), FV2D_PROB_BLAST (the Sedov-type point blast: rho 1, p 1e-5, and
 * p = (gamma-1) E / (pi r0^2) with E = 1, r0 = 0.02 in cells whose centre lies within r0 of
 * the domain centre; initBlast, Init.h:104-135, puts a fixed pressure in a disc instead) and
 * FV2D_PROB_KH (Kelvin-Helmholtz: band 0.25 < y < 0.75 with rho 2, u 0.5, else rho 1,
 * u -0.5, p 2.5, v = amplitude sin(4 pi x) + 1e-3 (2 U - 1), U a counter-based uniform
 * keyed by seed and cell; initKelvinHelmholtz, Init.h:246-287, draws its noise from a
 * Kokkos random pool instead). Cell centres as getPos (SimInfo.h:495-499).
 * ------------------------------------------------------------------------------------ */
enum fv2d_problem_kind { FV2D_PROB_SOD_X = 0, FV2D_PROB_BLAST = 1, FV2D_PROB_KH = 2 };

typedef struct fv2d_problem {
  int32_t kind; /* fv2d_problem_kind */
  int32_t reserved0;
  uint64_t seed;    /* FV2D_PROB_KH noise */
  double amplitude; /* FV2D_PROB_KH: amplitude of the sin(4 pi x) perturbation of v */
  double reserved[5];
} fv2d_problem;

/* Replaces: InitFunctor::init (Init.h:310-358) followed by primToCons (main.cpp:58) for the
 * problems above, on the device. Same contract as fv2d_set_prim afterwards. */
int fv2d_set_problem(fv2d_handle *h, const fv2d_problem *prob);

/* ------------------------------------------------------------------------------------
 * Batched serving (no reference counterpart: the reference runs one simulation per
 * process). A batch holds many independent simulations ("jobs") of one precision on one
 * GPU. Each job keeps its state, its dt and its step counter on the device until its end
 * (t_end or a step count); a launch advances every running job by up to a fixed number of
 * steps, whatever step each job is at. The stage kernels are persistent: their workgroups
 * pull (job, strip segment) items of one kernel class from a device work list, so jobs of
 * any size share one launch. Jobs are admitted while a slot and room in the batch's HBM
 * arena are free (FV2D_E_BUSY otherwise). A job's results do not depend on the other jobs
 * in the batch: they equal the same job run alone, bit for bit.
 *
 * Threading: one host thread drives a batch. fv2d_batch_launch queues work and returns;
 * fv2d_batch_wait blocks until it is done. Submit and collect may be called while a launch
 * is in flight: their device work is ordered after it, and the collected job has finished.
 * ------------------------------------------------------------------------------------ */
typedef struct fv2d_batch fv2d_batch;

typedef struct fv2d_batch_config {
  int32_t precision;        /* fv2d_precision of every job in the batch */
  int32_t max_jobs;         /* job slots */
  int64_t arena_bytes;      /* device memory for job state; admission bound */
  int32_t steps_per_launch; /* steps per launch (captured once as a HIP graph) */
  int32_t blocks_per_cu;    /* persistent workgroups per CU (0: the occupancy limit) */
  int32_t max_items;        /* work-list capacity in (job, segment) items */
  int32_t use_graph;        /* 1: replay a captured graph; 0: launch the kernels directly */
  int32_t reserved[8];
} fv2d_batch_config;

enum fv2d_job_status { FV2D_JOB_FREE = 0, FV2D_JOB_RUNNING = 1, FV2D_JOB_DONE = 2 };

typedef struct fv2d_job_info {
  int64_t id;
  int32_t status; /* fv2d_job_status */
  int32_t nsteps; /* steps taken */
  double t;       /* simulation time reached */
  double dt_last; /* last dt taken */
  uint64_t negative_density, negative_pressure, nan_count;
} fv2d_job_info;

void fv2d_batch_config_default(fv2d_batch_config *c);

/* Device bytes one job of these parameters takes from a batch arena. */
int64_t fv2d_job_bytes(const fv2d_params *p);

int fv2d_batch_create(const fv2d_batch_config *c, fv2d_batch **b);
void fv2d_batch_destroy(fv2d_batch *b);
const char *fv2d_batch_last_error(const fv2d_batch *b);

/* Admit a job: parameters p (precision must be the batch's), initial state from prob
 * (device-generated) or from q (4*Nx*Ny primitives, snapshot order; exactly one of the two
 * non-NULL), run until t + epsilon >= tend or max_steps steps (<= 0: no limit), the loop
 * rule of main.cpp:62-84; a job with neither end (max_steps <= 0 and tend >= FV2D_TEND_NONE or
 * NaN) returns FV2D_E_INVALID. FV2D_E_BUSY if no slot or arena space is free now. */
int fv2d_batch_submit(fv2d_batch *b, const fv2d_params *p, const fv2d_problem *prob, const double *q,
                      double tend, int32_t max_steps, int64_t *job_id);

/* Queue `launches` launches (each up to steps_per_launch steps of every running job). */
int fv2d_batch_launch(fv2d_batch *b, int32_t launches);

/* Wait for queued launches; *running = jobs still running (may be NULL). */
int fv2d_batch_wait(fv2d_batch *b, int32_t *running);

/* Ids of finished, not yet collected jobs (as of the last wait): up to cap into ids, count
 * into *n. */
int fv2d_batch_poll(fv2d_batch *b, int64_t *ids, int32_t cap, int32_t *n);

/* Status of a job (as of the last wait). */
int fv2d_batch_job_info(fv2d_batch *b, int64_t id, fv2d_job_info *info);

/* Results of a finished job: primitives q (4*Nx*Ny, snapshot order; may be NULL), the sums
 * of fv2d_sums (may be NULL) and its info (may be NULL). Frees the job's slot and memory. */
int fv2d_batch_collect(fv2d_batch *b, int64_t id, double *q, double sums[4], fv2d_job_info *info);

/* Drop a job (running or finished) without results. */
int fv2d_batch_cancel(fv2d_batch *b, int64_t id);

/* Duration (ms) of the last waited-for launches, from HIP events on the batch stream. */
int fv2d_batch_last_launch_ms(fv2d_batch *b, double *ms);

/* ------------------------------------------------------------------------------------
 * Runs that allocate nothing.
 *
 * No counterpart in the reference. For callers that cannot allocate device memory during
 * a run: code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc. fv2d_run_noalloc does what fv2d_create + fv2d_set_prim + fv2d_advance +
 * fv2d_get_prim + fv2d_destroy do, with the same kernels and bit for bit the same results,
 * but calls no hipMalloc / hipFree: the simulation's buffers (state, stage state, initial
 * primitives, control block, dt partials, dt log, gravity rows) live in a fixed device
 * array of the library (about 96 MiB plus 0.3 MiB) that every GPU the function runs on
 * holds until the process ends.
 *
 * q_in, q_out: DEVICE pointers, 4*Nx*Ny doubles each in snapshot order (the caller
 * allocates them); q_out may equal q_in. *t: start time on entry, end time on return.
 * tend, max_steps, dts (host, capacity max_steps), nsteps, stats: as fv2d_advance (stats
 * counts this run). stream: the hipStream_t to run on (NULL: the null stream); the call
 * returns when its work is done.
 *
 * Limits (FV2D_E_INVALID beyond them): row pitch (Nx rounded up to 64) times Ny at most
 * FV2D_NOALLOC_MAX_PLANE cells, Ny at most FV2D_NOALLOC_MAX_ROWS, max_steps at most
 * FV2D_NOALLOC_MAX_STEPS, at most FV2D_NOALLOC_MAX_WORKGROUPS stage workgroups.
 * Concurrency: calls on one GPU share its array; calls from threads of one process are
 * serialised by the library, calls that overlapped on one GPU would race.
 * ------------------------------------------------------------------------------------ */
#define FV2D_NOALLOC_MAX_PLANE (1 << 20)
#define FV2D_NOALLOC_MAX_ROWS 16384
#define FV2D_NOALLOC_MAX_STEPS 16384
#define FV2D_NOALLOC_MAX_WORKGROUPS 16384

int fv2d_run_noalloc(const fv2d_params *p, const double *q_in, double *q_out, double *t, double tend,
                     int32_t max_steps, double *dts, int32_t *nsteps, fv2d_step_stats *stats, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FV2D_AMD_H */
