// fv2d_math.h -- per-cell numerics of the hydro step, shared by the HIP kernels and the
// host emulation of the tile kernel.
//
// Every function keeps the reference's operation order (cited file:line), so that in
// double precision, compiled without FMA contraction, it reproduces the reference bit for
// bit. The real type R is double (parity form) or float (FP32 form); Ops<R> picks the
// division and square root: IEEE-exact for double, hardware approximations for float.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FV2D_HD __host__ __device__ inline __attribute__((always_inline))
#else
#define FV2D_HD inline __attribute__((always_inline))
#endif

namespace fv2d {

enum : int { IR = 0, IU = 1, IV = 2, IP = 3, IE = 3 };

// Option values, same as include/fv2d_amd.h (and the reference enums, SimInfo.h:41-97).
enum : int { HLL = 0, HLLC = 1, FSLP = 2 };
enum : int { BC_ABSORBING = 0, BC_REFLECTING = 1, BC_PERIODIC = 2 };
enum : int { PCM = 0, PCM_WB = 1, PLM = 2 };
enum : int { GRAV_NONE = 0, GRAV_CONSTANT = 1, GRAV_ANALYTICAL = 2 };
enum : int { BCTC_NONE = 0, BCTC_FIXED_TEMPERATURE = 1, BCTC_FIXED_GRADIENT = 2 };

template <typename R>
struct St {
  R v[4];
};

template <typename R>
struct Ops;

// a / d for a divisor d fixed over the grid (dx or dy), given rd = RN(1/d), in one of three
// modes chosen on the host by div_mode(d):
//  DIV_EXACT (2): d is a power of two whose reciprocal is a normal double, so rd = 1/d
//    exactly and a * rd is the correctly rounded a / d for every a (both round the same
//    real number; zeros, infinities and NaNs included): one multiply.
//  DIV_MARKSTEIN (1): Markstein's correction (P. Markstein, IBM J. Res. Dev. 34(1), 1990):
//    with q0 = RN(a*rd) within one ulp of a/d, r = a - d*q0 is exact in one FMA and
//    RN(q0 + r*rd) is the correctly rounded quotient, barring underflow and overflow. The
//    fast path is taken only for 2^-960 <= |a| <= 2^960 and 2^-60 <= d <= 2^60, so every
//    intermediate stays normal; zeros return a itself (a / d keeps the sign of a); anything
//    else (tiny, huge, inf, NaN) uses the IEEE division.
//  DIV_IEEE (0): the IEEE division.
enum : int { DIV_IEEE = 0, DIV_MARKSTEIN = 1, DIV_EXACT = 2 };

FV2D_HD int div_mode(double d) {
  if (d >= 0x1p-1022 && d <= 0x1p+1022) {
    int e;
    if (frexp(d, &e) == 0.5) return DIV_EXACT;
  }
  return (d >= 0x1p-60 && d <= 0x1p+60) ? DIV_MARKSTEIN : DIV_IEEE;
}

FV2D_HD double markstein_div(double a, double d, double rd, int mode) {
  if (mode == DIV_EXACT) return a * rd;
  const double aa = fabs(a);
  if (mode == DIV_MARKSTEIN) {
    if (aa >= 0x1p-960 && aa <= 0x1p+960) {
      const double q0 = a * rd;
      const double r = fma(-d, q0, a);
      return fma(r, rd, q0);
    }
    if (aa == 0.0) return a;
  }
  return a / d;
}

// IEEE square root of a double. On the device (gfx950) the compiler's sequence scales inputs
// below 2^-767 by 2^256 (and the result back by 2^-128), runs v_rsq_f64 and two Newton
// steps (10 instructions), and returns +-0 and +inf unchanged. For x in [2^-767, DBL_MAX]
// the scale factors are 2^0 and the special-value test fails, so the core alone gives the
// same bits; other inputs take the full sequence.
FV2D_HD double sqrt_ieee(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (x >= 0x1p-767 && x <= 0x1.fffffffffffffp+1023) {
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = r * 0.5;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
  }
#endif
  return sqrt(x);
}

template <>
struct Ops<double> {
  static FV2D_HD double div(double a, double b) { return a / b; }
  static FV2D_HD double div_c(double a, double d, double rd, int mode) { return markstein_div(a, d, rd, mode); }
  static FV2D_HD double sqrt_(double a) { return sqrt_ieee(a); }
  static FV2D_HD double min_(double a, double b) { return fmin(a, b); }
  static FV2D_HD double max_(double a, double b) { return fmax(a, b); }
};

template <>
struct Ops<float> {
  static FV2D_HD float div(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return a * __builtin_amdgcn_rcpf(b);
#else
    return a / b;
#endif
  }
  static FV2D_HD float div_c(float a, float d, float rd, int mode) {
    (void)d;
    (void)mode;
    return a * rd;
  }
  static FV2D_HD float sqrt_(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(a);
#else
    return sqrtf(a);
#endif
  }
  static FV2D_HD float min_(float a, float b) { return fminf(a, b); }
  static FV2D_HD float max_(float a, float b) { return fmaxf(a, b); }
};

// Quotients a_k / b of several numerators by one divisor (the two velocities of
// consToPrim and T = p / rho, uS and pS of HLLC, its star density and energy, the four HLL
// star fluxes). The compiler's FP64 division on the device is: ds = v_div_scale(b, for a),
// r = v_rcp_f64(ds), two Newton steps on r (4 FMAs), then ns = v_div_scale(a), m = ns * r,
// one FMA correction, v_div_fmas and v_div_fixup. The first half sees a only through the
// scaling of b, which is the identity unless a, b or a / b lie near the ends of the exponent
// range. DivBy runs it once for the first numerator; for each further numerator it
// recomputes the scaled divisor (one instruction) and reuses the refined reciprocal when the
// two are bitwise equal, so every quotient has the bits of a / b (and is a / b when they
// differ). On the host and in the FP32 form (where division is a * rcp(b), the reciprocal
// shared by the compiler) it is plain division.
template <typename R>
struct DivBy {
  R b, a0;
  FV2D_HD DivBy(R b_, R a0_) : b(b_), a0(a0_) {}
  FV2D_HD R first() const { return Ops<R>::div(a0, b); }
  FV2D_HD R operator()(R a) const { return Ops<R>::div(a, b); }
};

#if defined(__HIP_DEVICE_COMPILE__)
template <>
struct DivBy<double> {
  double b, a0, ds, rf;
  __device__ DivBy(double b_, double a0_) : b(b_), a0(a0_) {
    bool scale;
    ds = __builtin_amdgcn_div_scale(a0, b, false, &scale);
    const double r = __builtin_amdgcn_rcp(ds);
    const double r1 = fma(r, fma(-ds, r, 1.0), r);
    rf = fma(r1, fma(-ds, r1, 1.0), r1);
  }
  // the second half of the division, for a numerator whose scaled divisor is ds
  __device__ double quot(double a) const {
    bool scale;
    const double ns = __builtin_amdgcn_div_scale(a, b, true, &scale);
    const double m = ns * rf;
    return __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(fma(-ds, m, ns), rf, m, scale), b, a);
  }
  __device__ double first() const { return quot(a0); }
  __device__ double operator()(double a) const {
    bool scale;
    if (__builtin_amdgcn_div_scale(a, b, false, &scale) == ds) return quot(a);
    return a / b;
  }
};
#endif

// Physics constants of one simulation, precomputed on the host with the reference's own
// expressions (e.g. gm1 = gamma0 - 1.0, entho = 1.0 / (gamma0 - 1.0)), so the kernels
// see the same bits the reference computes inline.
struct KParams {
  int Nx, Ny;
  int pitch;          // elements per row of a field plane
  int bcx, bcy;       // boundary types
  int grav;           // gravity mode
  int wb;             // well_balanced_flux_at_y_bc
  int bctc_ymin, bctc_ymax;
  int divx, divy;     // markstein_div mode of the divisions by dx and by dy (div_mode)
  long long plane;    // elements per field plane (pitch * Ny)
  double gamma0, gm1, entho;
  double dx, dy, inv_dx, inv_dy;  // inv_* = 1.0 / d (Viscosity.h:57-58), also RN(1/d) for markstein_div
  double gxf, gyf;                // (double)(float)g: getGravity returns float (Gravity.h:39)
  double gdx_x, gdx_y;            // gx*dx, gy*dy (Update.h:131, for FSLP)
  double fslp_K;
  double kappa, mu;
  double bctc_ymin_value, bctc_ymax_value;
  double eps_reset;               // misc.epsilon_reset_negative
  double inv_tc, inv_visc;        // parabolic inverse dts (constant over the grid)
  double CFL, epsilon;
};

// ------------------------------------------------------------------ state algebra

// primToCons (States.h:19-30)
template <typename R>
FV2D_HD St<R> prim_to_cons(const St<R> &q, R gm1) {
  typedef Ops<R> O;
  St<R> r;
  r.v[IR] = q.v[IR];
  r.v[IU] = q.v[IR] * q.v[IU];
  r.v[IV] = q.v[IR] * q.v[IV];
  R Ek = O::div(R(0.5) * (r.v[IU] * r.v[IU] + r.v[IV] * r.v[IV]), q.v[IR]);
  r.v[IE] = Ek + O::div(q.v[IP], gm1);
  return r;
}

// consToPrim (States.h:32-43)
template <typename R>
FV2D_HD St<R> cons_to_prim(const St<R> &u, R gm1, const DivBy<R> &byr) {
  St<R> r;
  r.v[IR] = u.v[IR];
  r.v[IU] = byr.first();
  r.v[IV] = byr(u.v[IV]);
  R Ek = R(0.5) * r.v[IR] * (r.v[IU] * r.v[IU] + r.v[IV] * r.v[IV]);
  r.v[IP] = (u.v[IE] - Ek) * gm1;
  return r;
}
template <typename R>
FV2D_HD St<R> cons_to_prim(const St<R> &u, R gm1) {
  return cons_to_prim(u, gm1, DivBy<R>(u.v[IR], u.v[IU]));
}

// Negative density / pressure clamp of checkNegatives (SimInfo.h:617-626), on one cell.
template <typename R>
FV2D_HD void reset_negative(St<R> &q, R eps) {
  if (q.v[IR] < R(0)) q.v[IR] = eps;
  if (q.v[IP] < R(0)) q.v[IP] = eps;
}

// speedOfSound (States.h:45-46)
template <typename R>
FV2D_HD R speed_of_sound(const St<R> &q, R gamma0) {
  return Ops<R>::sqrt_(Ops<R>::div(q.v[IP] * gamma0, q.v[IR]));
}

// swap_component (States.h:103-110) for the y sweep
template <typename R>
FV2D_HD St<R> swap_uv(const St<R> &q) {
  St<R> r;
  r.v[IR] = q.v[IR];
  r.v[IU] = q.v[IV];
  r.v[IV] = q.v[IU];
  r.v[IP] = q.v[IP];
  return r;
}

// minmod (Update.h:77-85): product sign test, then the smaller magnitude (ties -> dR).
template <typename R>
FV2D_HD R minmod(R dL, R dR) {
  if (dL * dR < R(0)) return R(0);
  return (fabs(dL) < fabs(dR)) ? dL : dR;
}
// FP32 form: the median of (dL, dR, 0), one v_med3_f32. It is the value above except for the
// sign of a zero result and for a product that underflows to zero (the median is then 0,
// not the smaller magnitude); the parity form keeps the reference's expression.
template <>
FV2D_HD float minmod(float dL, float dR) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_fmed3f(dL, dR, 0.0f);
#else
  return fmaxf(fminf(dL, dR), fminf(fmaxf(dL, dR), 0.0f));
#endif
}

// PLM face value q + ((slope * sign) * 0.5) (Update.h:25, via the State operators
// States.h:64-92). sign = +1 for the right face, -1 for the left face.
template <typename R>
FV2D_HD R plm_face(R q, R slope, R sign) {
  return q + slope * sign * R(0.5);
}

// ------------------------------------------------------------------ Riemann solvers
// All take primitive face states in the sweep frame (u normal) and return the flux in
// that frame plus the interface pressure pout.

// hll (RiemannSolvers.h:6-51), Davis wave-speed estimates
template <typename R>
FV2D_HD void hll(const St<R> &qL, const St<R> &qR, R gamma0, R gm1, St<R> &flux, R &pout) {
  typedef Ops<R> O;
  const R aL = speed_of_sound(qL, gamma0);
  const R aR = speed_of_sound(qR, gamma0);
  const R SL = O::min_(qL.v[IU] - aL, qR.v[IU] - aR);
  const R SR = O::max_(qL.v[IU] + aL, qR.v[IU] + aR);

  // physical flux (RiemannSolvers.h:21-29)
  St<R> FL, FR;
  {
    const R Ek = R(0.5) * qL.v[IR] * (qL.v[IU] * qL.v[IU] + qL.v[IV] * qL.v[IV]);
    const R E = O::div(qL.v[IP], gm1) + Ek;
    FL.v[0] = qL.v[IR] * qL.v[IU];
    FL.v[1] = qL.v[IR] * qL.v[IU] * qL.v[IU] + qL.v[IP];
    FL.v[2] = qL.v[IR] * qL.v[IU] * qL.v[IV];
    FL.v[3] = (qL.v[IP] + E) * qL.v[IU];
  }
  {
    const R Ek = R(0.5) * qR.v[IR] * (qR.v[IU] * qR.v[IU] + qR.v[IV] * qR.v[IV]);
    const R E = O::div(qR.v[IP], gm1) + Ek;
    FR.v[0] = qR.v[IR] * qR.v[IU];
    FR.v[1] = qR.v[IR] * qR.v[IU] * qR.v[IU] + qR.v[IP];
    FR.v[2] = qR.v[IR] * qR.v[IU] * qR.v[IV];
    FR.v[3] = (qR.v[IP] + E) * qR.v[IU];
  }

  if (SL >= R(0)) {
    flux = FL;
    pout = qL.v[IP];
  } else if (SR <= R(0)) {
    flux = FR;
    pout = qR.v[IP];
  } else {
    // (SR*FL - SL*FR + SL*SR*(uR - uL)) / (SR - SL), uL/uR from primToCons
    const St<R> uL = prim_to_cons(qL, gm1);
    const St<R> uR = prim_to_cons(qR, gm1);
    pout = R(0.5) * (qL.v[IP] + qR.v[IP]);
    const R SLSR = SL * SR;
    const R den = SR - SL;
    R num[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) num[k] = FL.v[k] * SR - FR.v[k] * SL + (uR.v[k] - uL.v[k]) * SLSR;
    const DivBy<R> byd(den, num[0]);
    flux.v[0] = byd.first();
#pragma unroll
    for (int k = 1; k < 4; ++k) flux.v[k] = byd(num[k]);
  }
}

// hllc (RiemannSolvers.h:53-128). The reference evaluates both star states and both total
// energies and then picks one of four branches (SL > 0, uS > 0, SR > 0, else, at
// RiemannSolvers.h:93-122); here the side X of the picked branch is found first (the left
// side when SL > 0 or uS > 0), and only that side's total energy and, in the star branches,
// its star state are computed, by the same expressions on the same operands: the value of
// the picked branch is the reference's bit for bit.
template <typename R>
FV2D_HD void hllc(const St<R> &qL, const St<R> &qR, R gamma0, R entho, St<R> &flux, R &pout) {
  typedef Ops<R> O;
  const R rL = qL.v[IR], uL = qL.v[IU], vL = qL.v[IV], pL = qL.v[IP];
  const R rR = qR.v[IR], uR = qR.v[IU], vR = qR.v[IV], pR = qR.v[IP];

  const R cfastL = speed_of_sound(qL, gamma0);
  const R cfastR = speed_of_sound(qR, gamma0);

  const R cmax = O::max_(cfastL, cfastR);
  const R SL = O::min_(uL, uR) - cmax;
  const R SR = O::max_(uL, uR) + cmax;

  const R rcL = rL * (uL - SL);
  const R rcR = rR * (SR - uR);

  const R den = rcR + rcL;
  const DivBy<R> byd(den, rcR * uR + rcL * uL + (pL - pR));
  const R uS = byd.first();
  const R pS = byd(rcR * pL + rcL * pR + rcL * rcR * (uL - uR));

  // branches SL > 0 (left state), uS > 0 (left star), SR > 0 (right star), else (right state)
  const bool left = (SL > R(0)) || (uS > R(0));
  const bool star = !(SL > R(0)) && ((uS > R(0)) || (SR > R(0)));
  const R rX = left ? rL : rR;
  const R uX = left ? uL : uR;
  const R vX = left ? vL : vR;
  const R pX = left ? pL : pR;
  // EL / ER (RiemannSolvers.h:71-74) of the picked side
  const R EX = R(0.5) * rX * (uX * uX + vX * vX) + pX * entho;
  R sr, su, sp, E;
  if (star) {
    const R SX = left ? SL : SR;
    const DivBy<R> bys(SX - uS, rX * (SX - uX));
    sr = bys.first();
    E = bys((SX - uX) * EX - pX * uX + pS * uS);
    su = uS;
    sp = pS;
  } else {
    sr = rX;
    su = uX;
    sp = pX;
    E = EX;
  }
  pout = sp;

  flux.v[IR] = sr * su;
  flux.v[IU] = sr * su * su + sp;
  flux.v[IV] = flux.v[IR] * vX;
  flux.v[IE] = (E + sp) * su;
}

// fslp (RiemannSolvers.h:137-171); Kokkos::max/min are (a < b ? b : a) / (b < a ? b : a).
template <typename R>
FV2D_HD void fslp(const St<R> &qL, const St<R> &qR, R gamma0, R gm1, R K, R gdx, St<R> &flux, R &pout) {
  typedef Ops<R> O;
  const R rhoL = qL.v[IR], uL = qL.v[IU], pL = qL.v[IP];
  const R csL = O::sqrt_(O::div(gamma0 * pL, rhoL));
  const R rhoR = qR.v[IR], uR = qR.v[IU], pR = qR.v[IP];
  const R csR = O::sqrt_(O::div(gamma0 * pR, rhoR));

  const R a1 = rhoL * csL, a2 = rhoR * csR;
  const R ai = K * ((a1 < a2) ? a2 : a1);
  const R t1 = O::div(fabs(uL), csL), t2 = O::div(fabs(uR), csR);
  const R tm = (t1 < t2) ? t2 : t1;
  const R theta = (tm < R(1)) ? tm : R(1);

  const R ustar = R(0.5) * (uR + uL) - O::div(R(0.5), ai) * (pR - pL - R(0.5) * (rhoL + rhoR) * gdx);
  const R Pi = R(0.5) * (pR + pL) - theta * R(0.5) * ai * (uR - uL);

  const bool up = ustar > R(0);
  const R sr = up ? qL.v[IR] : qR.v[IR];
  const R su = up ? qL.v[IU] : qR.v[IU];
  const R sv = up ? qL.v[IV] : qR.v[IV];
  const R sp = up ? qL.v[IP] : qR.v[IP];
  const R Ekstar = R(0.5) * sr * (su * su + sv * sv);
  const R Estar = Ekstar + O::div(sp, gm1);

  flux.v[IR] = ustar * sr;
  flux.v[IU] = ustar * sr * su + Pi;
  flux.v[IV] = ustar * sr * sv;
  flux.v[IE] = ustar * (Estar + Pi);
  pout = Pi;
}

template <int RIEM, typename R>
FV2D_HD void riemann(const St<R> &qL, const St<R> &qR, R gamma0, R gm1, R entho, R K, R gdx, St<R> &flux,
                     R &pout) {
  if (RIEM == HLL)
    hll(qL, qR, gamma0, gm1, flux, pout);
  else if (RIEM == FSLP)
    fslp(qL, qR, gamma0, gm1, K, gdx, flux, pout);
  else
    hllc(qL, qR, gamma0, entho, flux, pout);
}

// ------------------------------------------------------------------ boundaries

// Ghost cell g (domain-relative, -2..N+1) -> source cell in [0, N) and a flag telling
// whether the normal velocity is negated. Composes the X-then-Y fill of
// BoundaryManager::fillBoundaries (BoundaryConditions.h:82-147): absorbing copies the
// edge cell, reflecting mirrors 2*piv - i - 1 and negates the normal velocity, periodic
// wraps by N. Indices beyond the ghost layers (only in partial edge tiles, never used by a
// domain cell) are clamped into the domain.
FV2D_HD int map_ghost(int g, int N, int bc, bool &flip) {
  flip = false;
  if (g >= 0 && g < N) return g;
  int s;
  if (bc == BC_PERIODIC) {
    s = (g < 0) ? g + N : g - N;
  } else if (bc == BC_REFLECTING) {
    s = (g < 0) ? -g - 1 : 2 * N - g - 1;
    flip = true;
  } else {
    s = (g < 0) ? 0 : N - 1;
  }
  s = (s < 0) ? 0 : ((s >= N) ? N - 1 : s);
  return s;
}

// The hyperbolic inverse time step of one cell (ComputeDt.h:31-34), given its sound speed;
// divx / divy: the division modes of dx and dy (kp.divx / kp.divy, or those values known at
// compile time).
template <typename R>
FV2D_HD double inv_dt_hyp_cs(const St<R> &q, R cs, const KParams &kp, int divx, int divy) {
  if (sizeof(R) == 8) {
    return markstein_div((double)(cs + fabs(q.v[IU])), kp.dx, kp.inv_dx, divx) +
           markstein_div((double)(cs + fabs(q.v[IV])), kp.dy, kp.inv_dy, divy);
  } else {
    // FP32 form: the per-cell value in FP32, the reduction over cells in FP64
    return (double)((cs + fabs(q.v[IU])) * R(kp.inv_dx) + (cs + fabs(q.v[IV])) * R(kp.inv_dy));
  }
}
template <typename R>
FV2D_HD double inv_dt_hyp_cs(const St<R> &q, R cs, const KParams &kp) {
  return inv_dt_hyp_cs(q, cs, kp, kp.divx, kp.divy);
}

template <typename R>
FV2D_HD double inv_dt_hyp(const St<R> &q, const KParams &kp) {
  return inv_dt_hyp_cs(q, speed_of_sound(q, R(kp.gamma0)), kp);
}

}  // namespace fv2d
