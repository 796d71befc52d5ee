/*
 * fv2d_oracle.c -- TEST INFRASTRUCTURE ONLY (parity reference and CPU baseline).
 *
 * A plain-C, FP64 restatement of the reference hot path of mdelorme/fv2d. Each function
 * cites the reference file:line it follows and keeps the reference's operation order, so
 * that, compiled with -ffp-contract=off and without -ffast-math, it produces the reference's
 * results bit for bit. It also keeps the reference's pass structure (separate ghost fill,
 * slopes, flux/update, thermal-conduction, viscosity, c2p, negative-check and dt passes, and
 * per-step U0/U* copies for RK2), which makes it the "restated reference CPU path" that is
 * timed as the CPU baseline.
 *
 * Planted-mistake builds: compiling with -DORACLE_PLANT=n (n > 0) introduces one deliberate,
 * plausible mistake (see the PLANT(n) sites). The test suite builds such variants and checks
 * that its pins (closed forms, invariants, known answers) reject every one of them.
 */
#include "fv2d_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif
#define PLANT(n) (ORACLE_PLANT == (n))

enum { IR = 0, IU = 1, IV = 2, IP = 3, IE = 3 };
enum { IX = 0, IY = 1 };

typedef struct {
  double v[4];
} state;

struct orc_sim {
  fv2d_params p;
  int Nx, Ny, Ng, Ntx, Nty, ibeg, iend, jbeg, jend;
  double dx, dy;
  size_t plane;
  double *Q, *U, *slopesX, *slopesY, *U0, *Ustar;
};

static int g_threads = 0;

void orc_set_threads(int n) { g_threads = n; }

int orc_get_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static) num_threads(orc_get_threads())")
#else
#define OMP_FOR
#endif

/* ---------------------------------------------------------------- array access */

static inline size_t at(const orc_sim *s, int i, int j) { return (size_t)j * (size_t)s->Ntx + (size_t)i; }

/* getStateFromArray / setStateInArray (States.h:6-17) */
static inline state get_state(const orc_sim *s, const double *A, int i, int j) {
  state r;
  size_t k = at(s, i, j);
  for (int v = 0; v < 4; ++v) r.v[v] = A[(size_t)v * s->plane + k];
  return r;
}

static inline void set_state(const orc_sim *s, double *A, int i, int j, state st) {
  size_t k = at(s, i, j);
  for (int v = 0; v < 4; ++v) A[(size_t)v * s->plane + k] = st.v[v];
}

static inline double aget(const orc_sim *s, const double *A, int v, int i, int j) {
  return A[(size_t)v * s->plane + at(s, i, j)];
}

/* ---------------------------------------------------------------- state algebra */

/* primToCons (States.h:19-30) */
static inline state prim_to_cons(state q, double gamma0) {
  state r;
  r.v[IR] = q.v[IR];
  r.v[IU] = q.v[IR] * q.v[IU];
  r.v[IV] = q.v[IR] * q.v[IV];
  double Ek = 0.5 * (r.v[IU] * r.v[IU] + r.v[IV] * r.v[IV]) / q.v[IR];
  r.v[IE] = (Ek + q.v[IP] / (gamma0 - 1.0));
  return r;
}

/* consToPrim (States.h:32-43) */
static inline state cons_to_prim(state u, double gamma0) {
  state r;
  r.v[IR] = u.v[IR];
  r.v[IU] = u.v[IU] / u.v[IR];
  r.v[IV] = u.v[IV] / u.v[IR];
  double Ek = 0.5 * r.v[IR] * (r.v[IU] * r.v[IU] + r.v[IV] * r.v[IV]);
#if PLANT(8)
  r.v[IP] = (u.v[IE] - Ek) * gamma0;
#else
  r.v[IP] = (u.v[IE] - Ek) * (gamma0 - 1.0);
#endif
  return r;
}

/* speedOfSound (States.h:45-46) */
static inline double speed_of_sound(state q, double gamma0) { return sqrt(q.v[IP] * gamma0 / q.v[IR]); }

/* swap_component (States.h:103-110) */
static inline state swap_component(state q, int dir) {
  if (dir == IX) return q;
  state r = {{q.v[IR], q.v[IV], q.v[IU], q.v[IP]}};
  return r;
}

/* ---------------------------------------------------------------- Riemann solvers */

/* hll (RiemannSolvers.h:6-51) */
static void hll(state qL, state qR, state *flux, double *pout, double gamma0) {
  const double aL = speed_of_sound(qL, gamma0);
  const double aR = speed_of_sound(qR, gamma0);
  /* Davis estimates (RiemannSolvers.h:12-19) */
  const double sminL = qL.v[IU] - aL;
  const double smaxL = qL.v[IU] + aL;
  const double sminR = qR.v[IU] - aR;
  const double smaxR = qR.v[IU] + aR;
  const double SL = fmin(sminL, sminR);
  const double SR = fmax(smaxL, smaxR);

  /* physical flux (RiemannSolvers.h:21-29) */
  state FL, FR;
  {
    const state *qs[2] = {&qL, &qR};
    state *fs[2] = {&FL, &FR};
    for (int k = 0; k < 2; ++k) {
      const state q = *qs[k];
      const double Ek = 0.5 * q.v[IR] * (q.v[IU] * q.v[IU] + q.v[IV] * q.v[IV]);
      const double E = (q.v[IP] / (gamma0 - 1.0) + Ek);
      state f = {{q.v[IR] * q.v[IU], q.v[IR] * q.v[IU] * q.v[IU] + q.v[IP], q.v[IR] * q.v[IU] * q.v[IV],
                  (q.v[IP] + E) * q.v[IU]}};
      *fs[k] = f;
    }
  }

  if (SL >= 0.0) {
    *flux = FL;
    *pout = qL.v[IP];
  } else if (SR <= 0.0) {
    *flux = FR;
    *pout = qR.v[IP];
  } else {
    /* (SR*FL - SL*FR + SL*SR*(uR - uL)) / (SR - SL)  (RiemannSolvers.h:46-49) */
    state uL = prim_to_cons(qL, gamma0);
    state uR = prim_to_cons(qR, gamma0);
    *pout = 0.5 * (qL.v[IP] + qR.v[IP]);
    const double SLSR = SL * SR;
    for (int k = 0; k < 4; ++k) {
#if PLANT(1)
      flux->v[k] = (FL.v[k] * SR - FR.v[k] * SL - (uR.v[k] - uL.v[k]) * SLSR) / (SR - SL);
#else
      flux->v[k] = (FL.v[k] * SR - FR.v[k] * SL + (uR.v[k] - uL.v[k]) * SLSR) / (SR - SL);
#endif
    }
  }
}

/* hllc (RiemannSolvers.h:53-128) */
static void hllc(state qL, state qR, state *flux, double *pout, double gamma0) {
  const double rL = qL.v[IR], uL = qL.v[IU], vL = qL.v[IV], pL = qL.v[IP];
  const double rR = qR.v[IR], uR = qR.v[IU], vR = qR.v[IV], pR = qR.v[IP];

  const double entho = 1.0 / (gamma0 - 1.0);

  const double ekL = 0.5 * rL * (uL * uL + vL * vL);
  const double EL = ekL + pL * entho;
  const double ekR = 0.5 * rR * (uR * uR + vR * vR);
  const double ER = ekR + pR * entho;

  const double cfastL = speed_of_sound(qL, gamma0);
  const double cfastR = speed_of_sound(qR, gamma0);

  const double SL = fmin(uL, uR) - fmax(cfastL, cfastR);
  const double SR = fmax(uL, uR) + fmax(cfastL, cfastR);

  const double rcL = rL * (uL - SL);
  const double rcR = rR * (SR - uR);

  const double uS = (rcR * uR + rcL * uL + (pL - pR)) / (rcR + rcL);
  const double pS = (rcR * pL + rcL * pR + rcL * rcR * (uL - uR)) / (rcR + rcL);

  const double rSL = rL * (SL - uL) / (SL - uS);
  const double ESL = ((SL - uL) * EL - pL * uL + pS * uS) / (SL - uS);

  const double rSR = rR * (SR - uR) / (SR - uS);
  const double ESR = ((SR - uR) * ER - pR * uR + pS * uS) / (SR - uS);

  state st;
  double E;
  if (SL > 0.0) {
    st = qL;
    E = EL;
    *pout = pL;
  } else if (uS > 0.0) {
    st.v[IR] = rSL;
    st.v[IU] = uS;
    st.v[IV] = qL.v[IV];
    st.v[IP] = pS;
    E = ESL;
    *pout = pS;
  } else if (SR > 0.0) {
    st.v[IR] = rSR;
    st.v[IU] = uS;
#if PLANT(2)
    st.v[IV] = qL.v[IV];
#else
    st.v[IV] = qR.v[IV];
#endif
    st.v[IP] = pS;
    E = ESR;
    *pout = pS;
  } else {
    st = qR;
    E = ER;
    *pout = pR;
  }

  flux->v[IR] = st.v[IR] * st.v[IU];
  flux->v[IU] = st.v[IR] * st.v[IU] * st.v[IU] + st.v[IP];
  flux->v[IV] = flux->v[IR] * st.v[IV];
  flux->v[IE] = (E + st.v[IP]) * st.v[IU];
}

/* fslp (RiemannSolvers.h:137-171). Kokkos::max/min/abs follow the (a < b ? b : a) forms. */
static inline double kmax(double a, double b) { return (a < b) ? b : a; }
static inline double kmin(double a, double b) { return (b < a) ? b : a; }

static void fslp(state qL, state qR, state *flux, double *pout, double gdx, double gamma0, double K) {
  const double rhoL = qL.v[IR], uL = qL.v[IU], pL = qL.v[IP];
  const double csL = sqrt(gamma0 * pL / rhoL);
  const double rhoR = qR.v[IR], uR = qR.v[IU], pR = qR.v[IP];
  const double csR = sqrt(gamma0 * pR / rhoR);

#if PLANT(14)
  const double ai = K * kmin(rhoL * csL, rhoR * csR);
#else
  const double ai = K * kmax(rhoL * csL, rhoR * csR);
#endif
#if PLANT(13)
  const double theta = kmin(1.0, kmin(fabs(uL) / csL, fabs(uR) / csR));
#else
  const double theta = kmin(1.0, kmax(fabs(uL) / csL, fabs(uR) / csR));
#endif

#if PLANT(15)
  const double ustar = 0.5 * (uR + uL) - 0.5 / ai * (pR - pL + 0.5 * (rhoL + rhoR) * gdx);
#else
  const double ustar = 0.5 * (uR + uL) - 0.5 / ai * (pR - pL - 0.5 * (rhoL + rhoR) * gdx);
#endif
  const double Pi = 0.5 * (pR + pL) - theta * 0.5 * ai * (uR - uL);

  const state qs = (ustar > 0 ? qL : qR);
  const double Ekstar = 0.5 * qs.v[IR] * (qs.v[IU] * qs.v[IU] + qs.v[IV] * qs.v[IV]);
  const double Estar = Ekstar + qs.v[IP] / (gamma0 - 1.0);

  flux->v[IR] = ustar * qs.v[IR];
  flux->v[IU] = ustar * qs.v[IR] * qs.v[IU] + Pi;
  flux->v[IV] = ustar * qs.v[IR] * qs.v[IV];
  flux->v[IE] = ustar * (Estar + Pi);
  *pout = Pi;
}

/* ---------------------------------------------------------------- physics helpers */

/* getPos (SimInfo.h:495-499) */
static inline void get_pos(const orc_sim *s, int i, int j, double *x, double *y) {
  *x = s->p.xmin + (i - s->ibeg + 0.5) * s->dx;
  *y = s->p.ymin + (j - s->jbeg + 0.5) * s->dy;
}

/* getGravity (Gravity.h:39-57): the value goes through `float` (the return type). */
static inline double get_gravity(const orc_sim *s, int i, int j, int dir) {
  double g;
  switch (s->p.gravity_mode) {
  case FV2D_GRAV_CONSTANT:
    g = (dir == IX ? s->p.gx : s->p.gy);
    break;
  case FV2D_GRAV_ANALYTICAL: {
    /* getAnalyticalGravity (Gravity.h:15-30), hot bubble, also returned as float */
    double x, y;
    get_pos(s, i, j, &x, &y);
#if PLANT(16)
    double ga = s->p.hot_bubble_g0 * sin(x * M_PI * 2.0 / s->p.ymax);
#elif PLANT(18)
    double ga = s->p.hot_bubble_g0 * sin(y * M_PI * 2.0 / (s->p.ymax - s->p.ymin));
#else
    double ga = s->p.hot_bubble_g0 * sin(y * M_PI * 2.0 / s->p.ymax);
#endif
#if PLANT(17)
    return ga;
#endif
    g = (double)(float)ga;
    break;
  }
  default:
    g = 0.0;
  }
#if PLANT(5)
  return g;
#else
  return (double)(float)g;
#endif
}

double orc_gravity(const orc_sim *s, int i, int j, int dir) { return get_gravity(s, i, j, dir); }

/* computeMu (Viscosity.h:8-16) and computeKappa (ThermalConduction.h:8-26), constant modes */
static inline double compute_mu(const orc_sim *s) { return s->p.mu; }
static inline double compute_kappa(const orc_sim *s) { return s->p.kappa; }

/* ---------------------------------------------------------------- boundaries */

/* fillAbsorbing / fillReflecting / fillPeriodic (BoundaryConditions.h:15-71), x direction */
static inline state fill_x(const orc_sim *s, const double *Q, int i, int j, int iref) {
  switch (s->p.boundary_x) {
  case FV2D_BC_REFLECTING: {
    int ipiv = (i < iref ? s->ibeg : s->iend);
    int isym = 2 * ipiv - i - 1;
    state q = get_state(s, Q, isym, j);
#if !PLANT(3)
    q.v[IU] *= -1.0;
#endif
    return q;
  }
  case FV2D_BC_PERIODIC: {
    int ii = (i < s->ibeg) ? i + s->Nx : i - s->Nx;
    return get_state(s, Q, ii, j);
  }
  default:
    return get_state(s, Q, iref, j);
  }
}

static inline state fill_y(const orc_sim *s, const double *Q, int i, int j, int jref) {
  switch (s->p.boundary_y) {
  case FV2D_BC_REFLECTING: {
    int jpiv = (j < jref ? s->jbeg : s->jend);
    int jsym = 2 * jpiv - j - 1;
    state q = get_state(s, Q, i, jsym);
    q.v[IV] *= -1.0;
    return q;
  }
  case FV2D_BC_PERIODIC: {
    int jj = (j < s->jbeg) ? j + s->Ny : j - s->Ny;
    return get_state(s, Q, i, jj);
  }
  default:
    return get_state(s, Q, i, jref);
  }
}

/* BoundaryManager::fillBoundaries (BoundaryConditions.h:82-147): X kernel on the interior
 * rows, then Y kernel over the full width (corners come from the X ghosts). */
static void fill_boundaries(orc_sim *s, double *Q) {
  OMP_FOR
  for (int j = s->jbeg; j < s->jend; ++j) {
    for (int i = 0; i < s->Ng; ++i) {
      int ileft = i, iright = s->iend + i;
      set_state(s, Q, ileft, j, fill_x(s, Q, ileft, j, s->ibeg));
      set_state(s, Q, iright, j, fill_x(s, Q, iright, j, s->iend - 1));
    }
  }
  OMP_FOR
  for (int j = 0; j < s->Ng; ++j) {
    for (int i = 0; i < s->Ntx; ++i) {
      int jtop = j, jbot = s->jend + j;
      set_state(s, Q, i, jtop, fill_y(s, Q, i, jtop, s->jbeg));
      set_state(s, Q, i, jbot, fill_y(s, Q, i, jbot, s->jend - 1));
    }
  }
}

/* ---------------------------------------------------------------- array kernels */

/* consToPrim over range_tot (SimInfo.h:576-587) */
static void cons_to_prim_array(orc_sim *s, const double *U, double *Q) {
  OMP_FOR
  for (int j = 0; j < s->Nty; ++j)
    for (int i = 0; i < s->Ntx; ++i) set_state(s, Q, i, j, cons_to_prim(get_state(s, U, i, j), s->p.gamma0));
}

/* primToCons over range_tot (SimInfo.h:589-600) */
static void prim_to_cons_array(orc_sim *s, const double *Q, double *U) {
  OMP_FOR
  for (int j = 0; j < s->Nty; ++j)
    for (int i = 0; i < s->Ntx; ++i) set_state(s, U, i, j, prim_to_cons(get_state(s, Q, i, j), s->p.gamma0));
}

/* minmod (Update.h:77-85): product test, then the smaller magnitude (ties -> dR) */
static inline double minmod(double dL, double dR) {
  if (dL * dR < 0.0)
    return 0.0;
#if PLANT(4)
  else if (fabs(dL) > fabs(dR))
#else
  else if (fabs(dL) < fabs(dR))
#endif
    return dL;
  else
    return dR;
}

double orc_minmod(double dL, double dR) { return minmod(dL, dR); }

/* computeSlopes (Update.h:59-91) over range_slopes */
static void compute_slopes(orc_sim *s, const double *Q) {
  OMP_FOR
  for (int j = s->jbeg - 1; j < s->jend + 1; ++j) {
    for (int i = s->ibeg - 1; i < s->iend + 1; ++i) {
      for (int v = 0; v < 4; ++v) {
        double dL = aget(s, Q, v, i, j) - aget(s, Q, v, i - 1, j);
        double dR = aget(s, Q, v, i + 1, j) - aget(s, Q, v, i, j);
        double dU = aget(s, Q, v, i, j) - aget(s, Q, v, i, j - 1);
        double dD = aget(s, Q, v, i, j + 1) - aget(s, Q, v, i, j);
        size_t k = (size_t)v * s->plane + at(s, i, j);
#if PLANT(10)
        s->slopesX[k] = minmod(dU, dD);
        s->slopesY[k] = minmod(dL, dR);
#else
        s->slopesX[k] = minmod(dL, dR);
        s->slopesY[k] = minmod(dU, dD);
#endif
      }
    }
  }
}

/* reconstruct (Update.h:16-37). PCM_WB computes a value and then falls through to the
 * piecewise-constant default (no `break`), so it is PCM. */
static inline state reconstruct(const orc_sim *s, const double *Q, const double *slopes, int i, int j, double sign,
                                int dir) {
  state q = get_state(s, Q, i, j);
  state r;
  if (s->p.reconstruction == FV2D_PLM) {
    state sl = get_state(s, slopes, i, j);
    for (int v = 0; v < 4; ++v) r.v[v] = q.v[v] + sl.v[v] * sign * 0.5;
  } else {
    r = q;
  }
  return swap_component(r, dir);
}

static inline void riemann(const orc_sim *s, state qL, state qR, state *flux, double *pout, double gdx) {
  switch (s->p.riemann_solver) {
  case FV2D_HLL:
    hll(qL, qR, flux, pout, s->p.gamma0);
    break;
  case FV2D_FSLP:
    fslp(qL, qR, flux, pout, gdx, s->p.gamma0, s->p.fslp_K);
    break;
  default:
    hllc(qL, qR, flux, pout, s->p.gamma0);
  }
}

/* updateAlongDir (Update.h:110-166) for one cell */
static void update_along_dir(orc_sim *s, const double *Q, double *U, double dt, int i, int j, int dir) {
  const double *slopes = (dir == IX ? s->slopesX : s->slopesY);
  int dxm = (dir == IX ? -1 : 0), dxp = (dir == IX ? 1 : 0);
  int dym = (dir == IY ? -1 : 0), dyp = (dir == IY ? 1 : 0);

  state qCL = reconstruct(s, Q, slopes, i, j, -1.0, dir);
  state qCR = reconstruct(s, Q, slopes, i, j, 1.0, dir);
  state qL = reconstruct(s, Q, slopes, i + dxm, j + dym, 1.0, dir);
  state qR = reconstruct(s, Q, slopes, i + dxp, j + dyp, -1.0, dir);

  const double gdx = (dir == IX ? s->p.gx * s->dx : s->p.gy * s->dy);

  state fluxL, fluxR;
  double poutL, poutR;
  riemann(s, qL, qCL, &fluxL, &poutL, gdx);
  riemann(s, qCR, qR, &fluxR, &poutR, gdx);

  fluxL = swap_component(fluxL, dir);
  fluxR = swap_component(fluxR, dir);

  /* well-balanced wall flux (Update.h:149-156) */
  if (s->p.well_balanced_flux_at_y_bc && (j == s->jbeg || j == s->jend - 1) && dir == IY) {
    double g = get_gravity(s, i, j, dir);
    if (j == s->jbeg) {
#if PLANT(21)
      state f = {{0.0, 0.0, poutR + aget(s, Q, IR, i, j) * g * s->dy, 0.0}};
#elif PLANT(22)
      state f = {{0.0, 0.0, poutL - aget(s, Q, IR, i, j) * g * s->dy, 0.0}};
#else
      state f = {{0.0, 0.0, poutR - aget(s, Q, IR, i, j) * g * s->dy, 0.0}};
#endif
      fluxL = f;
    } else {
      state f = {{0.0, 0.0, poutL + aget(s, Q, IR, i, j) * g * s->dy, 0.0}};
      fluxR = f;
    }
  }

  state un = get_state(s, U, i, j);
  const double d = (dir == IX ? s->dx : s->dy);
  for (int v = 0; v < 4; ++v) un.v[v] += (fluxL.v[v] - fluxR.v[v]) * dt / d;

  /* gravity source, always into the y momentum (Update.h:161-166) */
  if (s->p.gravity_mode != FV2D_GRAV_NONE) {
    double g = get_gravity(s, i, j, dir);
    un.v[IV] += dt * aget(s, Q, IR, i, j) * g;
#if PLANT(6)
    un.v[IE] += dt * 0.5 * (fluxL.v[IR] - fluxR.v[IR]) * g;
#else
    un.v[IE] += dt * 0.5 * (fluxL.v[IR] + fluxR.v[IR]) * g;
#endif
  }
  set_state(s, U, i, j, un);
}

/* computeFluxesAndUpdate (Update.h:93-174): X then Y per cell */
static void compute_fluxes_and_update(orc_sim *s, const double *Q, double *U, double dt) {
  OMP_FOR
  for (int j = s->jbeg; j < s->jend; ++j)
    for (int i = s->ibeg; i < s->iend; ++i) {
      update_along_dir(s, Q, U, dt, i, j, IX);
      update_along_dir(s, Q, U, dt, i, j, IY);
    }
}

/* ThermalConductionFunctor::applyThermalConduction (ThermalConduction.h:36-107), constant
 * kappa. computeKappa's neighbour calls pass doubles to int parameters; with a constant
 * kappa they return kappa, so kappa_face = 0.5*(kappa + kappa). */
static void apply_thermal_conduction(orc_sim *s, const double *Q, double *U, double dt) {
  const double dx = s->dx, dy = s->dy;
  OMP_FOR
  for (int j = s->jbeg; j < s->jend; ++j)
    for (int i = s->ibeg; i < s->iend; ++i) {
      const double kc = compute_kappa(s);
      double kappaL = 0.5 * (kc + compute_kappa(s));
      double kappaR = 0.5 * (kc + compute_kappa(s));
      double kappaU = 0.5 * (kc + compute_kappa(s));
      double kappaD = 0.5 * (kc + compute_kappa(s));

      double TC = aget(s, Q, IP, i, j) / aget(s, Q, IR, i, j);
      double TL = aget(s, Q, IP, i - 1, j) / aget(s, Q, IR, i - 1, j);
      double TR = aget(s, Q, IP, i + 1, j) / aget(s, Q, IR, i + 1, j);
      double TU = aget(s, Q, IP, i, j - 1) / aget(s, Q, IR, i, j - 1);
      double TD = aget(s, Q, IP, i, j + 1) / aget(s, Q, IR, i, j + 1);

      double FL = kappaL * (TC - TL) / dx;
      double FR = kappaR * (TR - TC) / dx;
      double FU = kappaU * (TC - TU) / dy;
      double FD = kappaD * (TD - TC) / dy;

      /* wall modes replace the x fluxes FL / FR on the first / last row
       * (ThermalConduction.h:75-103) */
#if PLANT(20)
      /* the physically expected form: the wall modes replace the y fluxes FU / FD */
      if (j == s->jbeg && s->p.bctc_ymin != FV2D_BCTC_NONE) {
        if (s->p.bctc_ymin == FV2D_BCTC_FIXED_TEMPERATURE)
          FU = kappaU * 2.0 * (TC - s->p.bctc_ymin_value) / dy;
        else if (s->p.bctc_ymin == FV2D_BCTC_FIXED_GRADIENT)
          FU = kappaU * s->p.bctc_ymin_value;
      }
      if (j == s->jend - 1 && s->p.bctc_ymax != FV2D_BCTC_NONE) {
        if (s->p.bctc_ymax == FV2D_BCTC_FIXED_TEMPERATURE)
          FD = kappaD * 2.0 * (s->p.bctc_ymax_value - TC) / dy;
        else if (s->p.bctc_ymax == FV2D_BCTC_FIXED_GRADIENT)
          FD = kappaD * s->p.bctc_ymax_value;
      }
#else
      if (j == s->jbeg && s->p.bctc_ymin != FV2D_BCTC_NONE) {
        if (s->p.bctc_ymin == FV2D_BCTC_FIXED_TEMPERATURE)
#if PLANT(19)
          FL = kappaL * (TC - s->p.bctc_ymin_value) / dy;
#else
          FL = kappaL * 2.0 * (TC - s->p.bctc_ymin_value) / dy;
#endif
        else if (s->p.bctc_ymin == FV2D_BCTC_FIXED_GRADIENT)
          FL = kappaL * s->p.bctc_ymin_value;
      }
      if (j == s->jend - 1 && s->p.bctc_ymax != FV2D_BCTC_NONE) {
        if (s->p.bctc_ymax == FV2D_BCTC_FIXED_TEMPERATURE)
          FR = kappaR * 2.0 * (s->p.bctc_ymax_value - TC) / dy;
        else if (s->p.bctc_ymax == FV2D_BCTC_FIXED_GRADIENT)
          FR = kappaR * s->p.bctc_ymax_value;
      }
#endif

      size_t k = (size_t)IE * s->plane + at(s, i, j);
#if PLANT(7)
      U[k] += dt / dx * (FR - FL) + dt / dy * (FU - FD);
#else
      U[k] += dt / dx * (FR - FL) + dt / dy * (FD - FU);
#endif
    }
}

/* ViscosityFunctor::applyViscosity (Viscosity.h:27-118). The face-flux difference is not
 * divided by dx or dy, as in the reference. */
static void apply_viscosity(orc_sim *s, const double *Q, double *U, double dt) {
  const double dx = s->dx, dy = s->dy;
  OMP_FOR
  for (int j = s->jbeg; j < s->jend; ++j)
    for (int i = s->ibeg; i < s->iend; ++i) {
      const double four_thirds = 4.0 / 3.0;
      const double two_thirds = 2.0 / 3.0;
      state st[3][3];
      for (int di = -1; di < 2; ++di)
        for (int dj = -1; dj < 2; ++dj) st[dj + 1][di + 1] = get_state(s, Q, i + di, j + dj);

      state vf[2];
      for (int dir = 0; dir < 2; ++dir) {
        state flux = {{0.0, 0.0, 0.0, 0.0}};
        const double one_over_dx = 1.0 / dx;
        const double one_over_dy = 1.0 / dy;
        const double mu = compute_mu(s);
        for (int side = 1; side < 3; ++side) {
          double sign = (side == 1 ? -1.0 : 1.0);
          if (dir == IX) {
            double qiu = (st[1][side].v[IU] + st[1][side - 1].v[IU]) * 0.5;
            double qiv = (st[1][side].v[IV] + st[1][side - 1].v[IV]) * 0.5;
            double dudx = one_over_dx * (st[1][side].v[IU] - st[1][side - 1].v[IU]);
            double dvdx = one_over_dx * (st[1][side].v[IV] - st[1][side - 1].v[IV]);
            double dudy = 0.25 * one_over_dy *
                          (st[2][side].v[IU] - st[0][side].v[IU] + st[2][side - 1].v[IU] - st[0][side - 1].v[IU]);
            double dvdy = 0.25 * one_over_dy *
                          (st[2][side].v[IV] - st[0][side].v[IV] + st[2][side - 1].v[IV] - st[0][side - 1].v[IV]);
            const double tau_xx = four_thirds * dudx - two_thirds * dvdy;
            const double tau_xy = dvdx + dudy;
            flux.v[IU] += sign * mu * tau_xx;
            flux.v[IV] += sign * mu * tau_xy;
            flux.v[IE] += sign * mu * (tau_xx * qiu + tau_xy * qiv);
          } else {
            double qiu = (st[side][1].v[IU] + st[side - 1][1].v[IU]) * 0.5;
            double qiv = (st[side][1].v[IV] + st[side - 1][1].v[IV]) * 0.5;
            double dudy = one_over_dy * (st[side][1].v[IU] - st[side - 1][1].v[IU]);
            double dvdy = one_over_dy * (st[side][1].v[IV] - st[side - 1][1].v[IV]);
            double dudx = 0.25 * one_over_dx *
                          (st[side][2].v[IU] - st[side][0].v[IU] + st[side - 1][2].v[IU] - st[side - 1][0].v[IU]);
            double dvdx = 0.25 * one_over_dx *
                          (st[side][2].v[IV] - st[side][0].v[IV] + st[side - 1][2].v[IV] - st[side - 1][0].v[IV]);
            const double tau_yy = four_thirds * dvdy - two_thirds * dudx;
            const double tau_xy = dvdx + dudy;
            flux.v[IU] += sign * mu * tau_xy;
            flux.v[IV] += sign * mu * tau_yy;
            flux.v[IE] += sign * mu * (tau_xy * qiu + tau_yy * qiv);
          }
        }
        vf[dir] = flux;
      }

      state un = get_state(s, U, i, j);
      for (int v = 0; v < 4; ++v) {
#if PLANT(9)
        un.v[v] += (vf[0].v[v] + vf[1].v[v]) * dt / dx;
#else
        un.v[v] += (vf[0].v[v] + vf[1].v[v]) * dt;
#endif
      }
      set_state(s, U, i, j, un);
    }
}

/* euler_step (Update.h:176-191) */
static void euler_step(orc_sim *s, double *Q, double *U, double dt) {
  fill_boundaries(s, Q);
  if (s->p.reconstruction == FV2D_PLM) compute_slopes(s, Q);
  compute_fluxes_and_update(s, Q, U, dt);
  if (s->p.tc_active) apply_thermal_conduction(s, Q, U, dt);
  if (s->p.visc_active) apply_viscosity(s, Q, U, dt);
}

static void copy_array(orc_sim *s, double *dst, const double *src) {
  const size_t n = 4 * s->plane;
  OMP_FOR
  for (long long k = 0; k < (long long)n; ++k) dst[k] = src[k];
}

/* UpdateFunctor::update (Update.h:193-222) */
static void update(orc_sim *s, double *Q, double *U, double dt) {
  if (s->p.time_stepping == FV2D_TS_EULER) {
    euler_step(s, Q, U, dt);
    return;
  }
  /* RK2: U0 = U, U* = U (the reference allocates and zero-fills both every step) */
  memset(s->U0, 0, 4 * s->plane * sizeof(double));
  memset(s->Ustar, 0, 4 * s->plane * sizeof(double));
  copy_array(s, s->U0, U);
  copy_array(s, s->Ustar, U);
  euler_step(s, Q, s->Ustar, dt);
  copy_array(s, U, s->Ustar);
  cons_to_prim_array(s, s->Ustar, Q);
  euler_step(s, Q, U, dt);
  /* RK2 Correct on the domain (Update.h:214-220) */
  OMP_FOR
  for (int j = s->jbeg; j < s->jend; ++j)
    for (int i = s->ibeg; i < s->iend; ++i)
      for (int v = 0; v < 4; ++v) {
        size_t k = (size_t)v * s->plane + at(s, i, j);
#if PLANT(11)
        U[k] = 0.5 * (s->Ustar[k] + U[k]);
#else
        U[k] = 0.5 * (s->U0[k] + U[k]);
#endif
      }
}

/* checkNegatives (SimInfo.h:602-646) */
void orc_check_negatives(orc_sim *s, uint64_t counts[3]) {
  const double eps = s->p.epsilon_reset_negative;
  unsigned long long nr = 0, np = 0, nn = 0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : nr, np, nn) schedule(static) num_threads(orc_get_threads())
#endif
  for (int j = s->jbeg; j < s->jend; ++j)
    for (int i = s->ibeg; i < s->iend; ++i) {
      size_t kr = (size_t)IR * s->plane + at(s, i, j);
      size_t kp = (size_t)IP * s->plane + at(s, i, j);
      if (s->Q[kr] < 0) {
        s->Q[kr] = eps;
        nr++;
      }
      if (s->Q[kp] < 0) {
        s->Q[kp] = eps;
        np++;
      }
      for (int v = 0; v < 4; ++v)
        if (isnan(s->Q[(size_t)v * s->plane + at(s, i, j)])) nn++;
    }
  if (counts) {
    counts[0] = nr;
    counts[1] = np;
    counts[2] = nn;
  }
}

/* ComputeDtFunctor::computeDt (ComputeDt.h:18-65). The Max reductions ignore NaN (fmax). */
double orc_compute_dt(const orc_sim *s, double inv_out[3]) {
  double inv_hyp = -DBL_MAX, inv_tc = -DBL_MAX, inv_visc = -DBL_MAX;
  const double dx = s->dx, dy = s->dy;
#ifdef _OPENMP
#pragma omp parallel for reduction(max : inv_hyp, inv_tc, inv_visc) schedule(static) num_threads(orc_get_threads())
#endif
  for (int j = s->jbeg; j < s->jend; ++j)
    for (int i = s->ibeg; i < s->iend; ++i) {
      state q = get_state(s, s->Q, i, j);
      double cs = speed_of_sound(q, s->p.gamma0);
#if PLANT(12)
      double hyp = (cs + fabs(q.v[IU])) / dx + (cs + fabs(q.v[IV])) / dx;
#else
      double hyp = (cs + fabs(q.v[IU])) / dx + (cs + fabs(q.v[IV])) / dy;
#endif
      double tc = s->p.epsilon;
      if (s->p.tc_active) tc = fmax(2.0 * compute_kappa(s) / (dx * dx), 2.0 * compute_kappa(s) / (dy * dy));
      double vi = s->p.epsilon;
      if (s->p.visc_active) vi = fmax(2.0 * compute_mu(s) / (dx * dx), 2.0 * compute_mu(s) / (dy * dy));
      inv_hyp = fmax(inv_hyp, hyp);
      inv_tc = fmax(inv_tc, tc);
      inv_visc = fmax(inv_visc, vi);
    }
  if (inv_out) {
    inv_out[0] = inv_hyp;
    inv_out[1] = inv_tc;
    inv_out[2] = inv_visc;
  }
  /* std::max({a, b, c}) keeps the first of equal maxima */
  double m = inv_hyp;
  if (m < inv_tc) m = inv_tc;
  if (m < inv_visc) m = inv_visc;
  return s->p.CFL / m;
}

/* main.cpp:79-81 */
void orc_step(orc_sim *s, double dt, uint64_t counts[3]) {
  update(s, s->Q, s->U, dt);
  cons_to_prim_array(s, s->U, s->Q);
  orc_check_negatives(s, counts);
}

/* main.cpp:62-84 without I/O */
int orc_run(orc_sim *s, double *t, double tend, int max_steps, double *dts, uint64_t counts[3]) {
  int n = 0;
  uint64_t acc[3] = {0, 0, 0};
  while (*t + s->p.epsilon < tend && (max_steps <= 0 || n < max_steps)) {
    double dt = orc_compute_dt(s, NULL);
    if (dts) dts[n] = dt;
    uint64_t c[3];
    orc_step(s, dt, c);
    for (int k = 0; k < 3; ++k) acc[k] += c[k];
    *t += dt;
    ++n;
  }
  if (counts)
    for (int k = 0; k < 3; ++k) counts[k] = acc[k];
  return n;
}

void orc_fill_boundaries(orc_sim *s) { fill_boundaries(s, s->Q); }

/* ---------------------------------------------------------------- object */

orc_sim *orc_create(const fv2d_params *p, char *err) {
#define FAIL(msg)                                                                                                      \
  do {                                                                                                                 \
    if (err) snprintf(err, 256, "%s", msg);                                                                            \
    return NULL;                                                                                                       \
  } while (0)
  if (p->Nx < 2 || p->Ny < 2) FAIL("Nx and Ny must be >= 2");
  if (p->Ng < 2) FAIL("Nghosts must be >= 2");
  if (p->tc_active && p->tc_mode != FV2D_TCM_CONSTANT) FAIL("thermal conduction mode B02 is not supported");
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
  if (!s) FAIL("out of memory");
  s->p = *p;
  s->Nx = p->Nx;
  s->Ny = p->Ny;
  s->Ng = p->Ng;
  /* SimInfo.h:368-376 */
  s->Ntx = s->Nx + 2 * s->Ng;
  s->Nty = s->Ny + 2 * s->Ng;
  s->ibeg = s->Ng;
  s->iend = s->Ng + s->Nx;
  s->jbeg = s->Ng;
  s->jend = s->Ng + s->Ny;
  s->dx = (p->xmax - p->xmin) / s->Nx;
  s->dy = (p->ymax - p->ymin) / s->Ny;
  s->plane = (size_t)s->Ntx * (size_t)s->Nty;
  size_t bytes = 4 * s->plane * sizeof(double);
  s->Q = (double *)calloc(1, bytes);
  s->U = (double *)calloc(1, bytes);
  s->slopesX = (double *)calloc(1, bytes);
  s->slopesY = (double *)calloc(1, bytes);
  s->U0 = (double *)calloc(1, bytes);
  s->Ustar = (double *)calloc(1, bytes);
  if (!s->Q || !s->U || !s->slopesX || !s->slopesY || !s->U0 || !s->Ustar) {
    orc_destroy(s);
    FAIL("out of memory");
  }
  return s;
#undef FAIL
}

void orc_destroy(orc_sim *s) {
  if (!s) return;
  free(s->Q);
  free(s->U);
  free(s->slopesX);
  free(s->slopesY);
  free(s->U0);
  free(s->Ustar);
  free(s);
}

void orc_set_prim(orc_sim *s, const double *q) {
  const size_t n = (size_t)s->Nx * (size_t)s->Ny;
  OMP_FOR
  for (int j = 0; j < s->Ny; ++j)
    for (int i = 0; i < s->Nx; ++i)
      for (int v = 0; v < 4; ++v)
        s->Q[(size_t)v * s->plane + at(s, i + s->ibeg, j + s->jbeg)] = q[(size_t)v * n + (size_t)j * s->Nx + i];
  fill_boundaries(s, s->Q);
  prim_to_cons_array(s, s->Q, s->U);
}

void orc_poke_prim(orc_sim *s, const double *q) {
  const size_t n = (size_t)s->Nx * (size_t)s->Ny;
  for (int j = 0; j < s->Ny; ++j)
    for (int i = 0; i < s->Nx; ++i)
      for (int v = 0; v < 4; ++v)
        s->Q[(size_t)v * s->plane + at(s, i + s->ibeg, j + s->jbeg)] = q[(size_t)v * n + (size_t)j * s->Nx + i];
}

static void get_domain(const orc_sim *s, const double *A, double *out) {
  const size_t n = (size_t)s->Nx * (size_t)s->Ny;
  for (int j = 0; j < s->Ny; ++j)
    for (int i = 0; i < s->Nx; ++i)
      for (int v = 0; v < 4; ++v)
        out[(size_t)v * n + (size_t)j * s->Nx + i] = A[(size_t)v * s->plane + at(s, i + s->ibeg, j + s->jbeg)];
}

void orc_get_prim(const orc_sim *s, double *q) { get_domain(s, s->Q, q); }
void orc_get_cons(const orc_sim *s, double *u) { get_domain(s, s->U, u); }
void orc_get_prim_full(const orc_sim *s, double *q) { memcpy(q, s->Q, 4 * s->plane * sizeof(double)); }
void orc_get_cons_full(const orc_sim *s, double *u) { memcpy(u, s->U, 4 * s->plane * sizeof(double)); }

void orc_sums(const orc_sim *s, double out[4]) {
  for (int v = 0; v < 4; ++v) {
    double tot = 0.0;
    for (int j = s->jbeg; j < s->jend; ++j) {
      double row = 0.0;
      for (int i = s->ibeg; i < s->iend; ++i) row += aget(s, s->U, v, i, j);
      tot += row;
    }
    out[v] = tot * s->dx * s->dy;
  }
}

void orc_mesh(const orc_sim *s, int32_t oi[8], double orl[2]) {
  oi[0] = s->Ntx;
  oi[1] = s->Nty;
  oi[2] = s->ibeg;
  oi[3] = s->iend;
  oi[4] = s->jbeg;
  oi[5] = s->jend;
  oi[6] = s->Nx;
  oi[7] = s->Ny;
  orl[0] = s->dx;
  orl[1] = s->dy;
}

/* ---------------------------------------------------------------- point exports */

static state load4(const double *a) {
  state s = {{a[0], a[1], a[2], a[3]}};
  return s;
}
static void store4(double *a, state s) {
  for (int v = 0; v < 4; ++v) a[v] = s.v[v];
}

void orc_prim_to_cons(const double q[4], double gamma0, double u[4]) { store4(u, prim_to_cons(load4(q), gamma0)); }
void orc_cons_to_prim(const double u[4], double gamma0, double q[4]) { store4(q, cons_to_prim(load4(u), gamma0)); }
double orc_speed_of_sound(const double q[4], double gamma0) { return speed_of_sound(load4(q), gamma0); }

void orc_hll(const double qL[4], const double qR[4], double gamma0, double flux[4], double *pout) {
  state f;
  hll(load4(qL), load4(qR), &f, pout, gamma0);
  store4(flux, f);
}

void orc_hllc(const double qL[4], const double qR[4], double gamma0, double flux[4], double *pout) {
  state f;
  hllc(load4(qL), load4(qR), &f, pout, gamma0);
  store4(flux, f);
}

void orc_fslp(const double qL[4], const double qR[4], double gamma0, double fslp_K, double gdx, double flux[4],
              double *pout) {
  state f;
  fslp(load4(qL), load4(qR), &f, pout, gdx, gamma0, fslp_K);
  store4(flux, f);
}

/* ---------------------------------------------------------------- host bandwidth */

/* STREAM triad a[i] = b[i] + s * c[i] over n doubles per array with orc_get_threads()
 * threads (first touch by the same static schedule), best of reps passes, in GB/s at the
 * STREAM convention of 24 bytes per element. Context for the CPU baseline: the restated
 * reference path is bound by host memory bandwidth. */
double orc_stream_triad(long n, int reps) {
  double *a = (double *)malloc(sizeof(double) * (size_t)n);
  double *b = (double *)malloc(sizeof(double) * (size_t)n);
  double *c = (double *)malloc(sizeof(double) * (size_t)n);
  if (!a || !b || !c) {
    free(a);
    free(b);
    free(c);
    return 0.0;
  }
  OMP_FOR
  for (long i = 0; i < n; ++i) {
    a[i] = 0.0;
    b[i] = 1.0;
    c[i] = 2.0;
  }
  double best = 0.0;
  const double s = 3.0;
  for (int r = 0; r < reps; ++r) {
#ifdef _OPENMP
    const double t0 = omp_get_wtime();
#endif
    OMP_FOR
    for (long i = 0; i < n; ++i) a[i] = b[i] + s * c[i];
#ifdef _OPENMP
    const double el = omp_get_wtime() - t0;
    const double gbs = 24.0 * (double)n / el / 1e9;
    if (gbs > best) best = gbs;
#endif
  }
  volatile double sink = a[n / 2];
  (void)sink;
  free(a);
  free(b);
  free(c);
  return best;
}
