/*
 * fv2d_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the reference (mdelorme/fv2d) hot path, used as the parity
 * reference of the HIP path and as the timed CPU baseline. Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it. The product never links or calls it.
 *
 * The oracle keeps the reference's data model: full arrays with Ng ghost layers
 * (Ntx x Nty cells), Q and U kept separately, slopes in their own arrays, and one pass per
 * reference kernel, in the reference order (the time loop of main.cpp:62-84). Layout is
 * SoA [var][j][i] with x fastest.
 */
#ifndef FV2D_ORACLE_H
#define FV2D_ORACLE_H

#include <stdint.h>

#include "../include/fv2d_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* Simulation object. Returns NULL and fills err (may be NULL, 256 bytes) on a rejected
 * configuration. */
orc_sim *orc_create(const fv2d_params *p, char *err);
void orc_destroy(orc_sim *s);

/* Number of OpenMP threads used by the array passes (<= 0: OpenMP default). */
void orc_set_threads(int n);
int orc_get_threads(void);

/* Domain primitives in snapshot order -> Q (domain), fillBoundaries(Q), U = p2c(Q) on the
 * full range (Init.h:355-357, main.cpp:58). */
void orc_set_prim(orc_sim *s, const double *q);
/* Q / U domain in snapshot order. */
void orc_get_prim(const orc_sim *s, double *q);
void orc_get_cons(const orc_sim *s, double *u);
/* Full Q / U including ghosts: 4 * Ntx * Nty doubles [var][j][i]. */
void orc_get_prim_full(const orc_sim *s, double *q);
void orc_get_cons_full(const orc_sim *s, double *u);
/* Overwrites the domain of Q only (no ghost fill, U untouched): failure-path tests. */
void orc_poke_prim(orc_sim *s, const double *q);

/* ComputeDtFunctor::computeDt (ComputeDt.h:18-65). inv may be NULL. */
double orc_compute_dt(const orc_sim *s, double inv[3]);
/* UpdateFunctor::update + consToPrim + checkNegatives (main.cpp:79-81). counts may be NULL. */
void orc_step(orc_sim *s, double dt, uint64_t counts[3]);
/* checkNegatives alone (SimInfo.h:602-646). */
void orc_check_negatives(orc_sim *s, uint64_t counts[3]);
/* Time loop of main.cpp:62-84 without I/O; returns steps taken, dts may be NULL. */
int orc_run(orc_sim *s, double *t, double tend, int max_steps, double *dts, uint64_t counts[3]);
/* BoundaryManager::fillBoundaries on Q (BoundaryConditions.h:82-147). */
void orc_fill_boundaries(orc_sim *s);
/* Sum over the domain of each U field times dx*dy, rows summed in order. */
void orc_sums(const orc_sim *s, double out[4]);

/* Point functions, exposed for unit tests. States are (rho, u, v, p) or (rho, mx, my, E). */
void orc_prim_to_cons(const double q[4], double gamma0, double u[4]);
void orc_cons_to_prim(const double u[4], double gamma0, double q[4]);
double orc_speed_of_sound(const double q[4], double gamma0);
void orc_hll(const double qL[4], const double qR[4], double gamma0, double flux[4], double *pout);
void orc_hllc(const double qL[4], const double qR[4], double gamma0, double flux[4], double *pout);
void orc_fslp(const double qL[4], const double qR[4], double gamma0, double fslp_K, double gdx,
              double flux[4], double *pout);
double orc_minmod(double dL, double dR);
/* getGravity (Gravity.h:39-57) at cell (i, j) for direction dir (0 = x, 1 = y). */
double orc_gravity(const orc_sim *s, int i, int j, int dir);

/* Host memory bandwidth: best STREAM-triad GB/s over reps passes of n doubles per array. */
double orc_stream_triad(long n, int reps);

/* Mesh facts. */
void orc_mesh(const orc_sim *s, int32_t out_int[8], double out_real[2]);

#ifdef __cplusplus
}
#endif

#endif /* FV2D_ORACLE_H */
