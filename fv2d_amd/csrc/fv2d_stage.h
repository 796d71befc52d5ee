// fv2d_stage.h -- the fused stage of the hydro step, tiled for LDS.
//
// One workgroup owns a TX x TY tile of domain cells, one cell per thread. A stage is
// three barrier-separated phases:
//   load   : the primitives of the tile plus a 2-cell halo go to LDS. They are computed on
//            the fly from U (consToPrim + negative reset) or taken from the initial Q, and
//            ghost cells are synthesised from the boundary rules (map_ghost), so no ghost
//            is ever stored in HBM. Each thread also keeps its own cell's U in registers.
//   flux   : every x and y interface of the tile is solved once (PLM-minmod slopes and
//            face states from LDS, then HLL / HLLC / FSLP); fluxes go to LDS.
//   update : each thread applies, in the reference's accumulation order, the x flux
//            difference (+ x gravity), the y flux difference (+ y gravity and the
//            well-balanced wall flux), thermal conduction and viscosity to its U; RK2
//            stage 2 averages with U^n; the result is stored. The optional epilogue
//            converts the new U to primitives, counts negative / NaN values and forms the
//            cell's hyperbolic inverse dt for the next step's reduction.
// Interface fluxes are computed once per interface, whereas the reference solves each
// interface twice with the same arguments (Update.h:142-143): same bits, half the work.
//
// The phase functions are plain host/device code; the HIP kernel calls them with
// __syncthreads() between phases, the host emulation calls them thread by thread.
#pragma once

#include "fv2d_math.h"

namespace fv2d {

// Device-resident control block of one simulation.
struct StepCtl {
  double dt, dt_dx, dt_dy;  // dt of the current step, dt/dx, dt/dy
  double t, tend;           // device-side time and end time (advance mode)
  double inv_hyp;           // max hyperbolic inverse dt of the current primitives
  int active;               // stages run only while set
  int nsteps;               // steps taken in the current advance call
  int max_steps;            // advance limit (<= 0: none)
  int pad;
  unsigned long long neg_rho, neg_prs, nan_cnt;  // running totals of the reset
};

enum : int { QSRC_U_RESET = 0, QSRC_U = 1, QSRC_Q0 = 2 };

template <typename S>
struct StageIO {
  const S *src;           // U planes: halo source and base of the update
  const S *q0;            // initial primitives (QSRC_Q0), else null
  const S *avg;           // U^n planes for the RK2 average, else null
  S *out;                 // output U planes
  const float *grav_row;  // analytical gravity per row (Ny), else null
  StepCtl *ctl;
  double *partials;       // per-workgroup max inverse dt (epilogue)
  int qsrc;               // QSRC_*
  int epilogue;
};

// Per-thread state carried between phases, and the epilogue results.
struct CellRegs {
  double u[4];
  double u0[4];
  unsigned neg_rho, neg_prs, nan_cnt;
  double inv;
};

template <typename R, typename S, int TX_, int TY_, int RECON, int RIEM, bool TC, bool VISC>
struct TileStage {
  static constexpr int TX = TX_, TY = TY_, NT = TX_ * TY_;
  static constexpr int HX = TX + 4, HY = TY + 4;
  static constexpr int NXF = (TX + 1) * TY;  // x interfaces of the tile
  static constexpr int NYF = TX * (TY + 1);  // y interfaces of the tile
  static constexpr int RING = 4 * HX + 4 * TY;

  struct Smem {
    R q[4][HY][HX];
    R t[TC ? HY : 1][TC ? HX : 1];
    R fx[4][TY][TX + 1];
    R fy[4][TY + 1][TX];
    R py[TY + 1][TX];
  };

  static FV2D_HD void put(Smem &sm, int lx, int ly, const St<R> &q) {
#pragma unroll
    for (int v = 0; v < 4; ++v) sm.q[v][ly][lx] = q.v[v];
    if (TC) sm.t[TC ? ly : 0][TC ? lx : 0] = Ops<R>::div(q.v[IP], q.v[IR]);  // T = p / rho
  }

  // Primitive state of domain-relative cell (gi, gj), ghosts included.
  static FV2D_HD void fetch_q(int gi, int gj, const KParams &kp, const StageIO<S> &io, St<R> &q) {
    bool fx, fy;
    const int si = map_ghost(gi, kp.Nx, kp.bcx, fx);
    const int sj = map_ghost(gj, kp.Ny, kp.bcy, fy);
    const long long k = (long long)sj * kp.pitch + si;
    if (io.qsrc == QSRC_Q0) {
#pragma unroll
      for (int v = 0; v < 4; ++v) q.v[v] = R(io.q0[v * kp.plane + k]);
    } else {
      St<R> u;
#pragma unroll
      for (int v = 0; v < 4; ++v) u.v[v] = R(io.src[v * kp.plane + k]);
      q = cons_to_prim(u, R(kp.gm1));
      if (io.qsrc == QSRC_U_RESET) reset_negative(q, R(kp.eps_reset));
    }
    if (fx) q.v[IU] = -q.v[IU];
    if (fy) q.v[IV] = -q.v[IV];
  }

  static FV2D_HD void ring_pos(int k, int &lx, int &ly) {
    if (k < 2 * HX) {
      ly = k / HX;
      lx = k % HX;
    } else if (k < 4 * HX) {
      const int kk = k - 2 * HX;
      ly = TY + 2 + kk / HX;
      lx = kk % HX;
    } else {
      const int kk = k - 4 * HX;
      ly = 2 + kk / 4;
      const int c = kk % 4;
      lx = (c < 2) ? c : TX + c;
    }
  }

  // ---------------------------------------------------------------- phase 1: load
  static FV2D_HD void phase_load(int tid, int bx, int by, Smem &sm, const KParams &kp, const StageIO<S> &io,
                                 CellRegs &r) {
    const int i0 = bx * TX, j0 = by * TY;
    {
      const int xl = tid % TX, yl = tid / TX;
      const int gi = i0 + xl, gj = j0 + yl;
      St<R> q;
      if (gi < kp.Nx && gj < kp.Ny) {
        const long long k = (long long)gj * kp.pitch + gi;
        St<R> u;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const S s = io.src[v * kp.plane + k];
          r.u[v] = (double)s;
          u.v[v] = R(s);
        }
        if (io.avg) {
#pragma unroll
          for (int v = 0; v < 4; ++v) r.u0[v] = (double)io.avg[v * kp.plane + k];
        }
        if (io.qsrc == QSRC_Q0) {
#pragma unroll
          for (int v = 0; v < 4; ++v) q.v[v] = R(io.q0[v * kp.plane + k]);
        } else {
          q = cons_to_prim(u, R(kp.gm1));
          if (io.qsrc == QSRC_U_RESET) reset_negative(q, R(kp.eps_reset));
        }
      } else {
        fetch_q(gi, gj, kp, io, q);
      }
      put(sm, xl + 2, yl + 2, q);
    }
    for (int k = tid; k < RING; k += NT) {
      int lx, ly;
      ring_pos(k, lx, ly);
      St<R> q;
      fetch_q(i0 + lx - 2, j0 + ly - 2, kp, io, q);
      put(sm, lx, ly, q);
    }
  }

  // Face states of an interface from the four cells a | b || c | d along the sweep:
  // left state = b's right face, right state = c's left face (reconstruct, Update.h:16-37,
  // with the slopes of computeSlopes, Update.h:59-91).
  static FV2D_HD void faces(R a, R b, R c, R d, R &fl, R &fr) {
    if (RECON == PLM) {
      const R sb = minmod(b - a, c - b);
      const R sc = minmod(c - b, d - c);
      fl = plm_face(b, sb, R(1));
      fr = plm_face(c, sc, R(-1));
    } else {
      fl = b;
      fr = c;
    }
  }

  // ---------------------------------------------------------------- phase 2: fluxes
  static FV2D_HD void phase_flux(int tid, Smem &sm, const KParams &kp) {
    const R gamma0 = R(kp.gamma0), gm1 = R(kp.gm1), entho = R(kp.entho), K = R(kp.fslp_K);
    for (int k = tid; k < NXF + NYF; k += NT) {
      if (k < NXF) {
        // interface between local cells xi-1 and xi of row yl
        const int xi = k % (TX + 1), yl = k / (TX + 1);
        const int row = yl + 2, c0 = xi;  // halo columns c0 .. c0+3 hold a b c d
        St<R> qL, qR;
#pragma unroll
        for (int v = 0; v < 4; ++v)
          faces(sm.q[v][row][c0], sm.q[v][row][c0 + 1], sm.q[v][row][c0 + 2], sm.q[v][row][c0 + 3], qL.v[v],
                qR.v[v]);
        St<R> f;
        R pout;
        riemann<RIEM>(qL, qR, gamma0, gm1, entho, K, R(kp.gdx_x), f, pout);
#pragma unroll
        for (int v = 0; v < 4; ++v) sm.fx[v][yl][xi] = f.v[v];
      } else {
        // interface between local rows yi-1 and yi of column xl
        const int kk = k - NXF;
        const int xl = kk % TX, yi = kk / TX;
        const int col = xl + 2, r0 = yi;
        St<R> fL, fR;
#pragma unroll
        for (int v = 0; v < 4; ++v)
          faces(sm.q[v][r0][col], sm.q[v][r0 + 1][col], sm.q[v][r0 + 2][col], sm.q[v][r0 + 3][col], fL.v[v],
                fR.v[v]);
        const St<R> qL = swap_uv(fL), qR = swap_uv(fR);
        St<R> f;
        R pout;
        riemann<RIEM>(qL, qR, gamma0, gm1, entho, K, R(kp.gdx_y), f, pout);
        const St<R> fb = swap_uv(f);
#pragma unroll
        for (int v = 0; v < 4; ++v) sm.fy[v][yi][xl] = fb.v[v];
        sm.py[yi][xl] = pout;
      }
    }
  }

  // U += ((FL - FR) * dt) / d (Update.h:158-159; States.h:49-101 fix the order). In the
  // FP32 form the flux difference is applied in FP64 with a precomputed dt / d.
  static FV2D_HD void add_flux(double acc[4], const St<R> &FL, const St<R> &FR, double dt, double dt_d, double d) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (sizeof(R) == 8)
        acc[v] = acc[v] + ((double)(FL.v[v] - FR.v[v]) * dt) / d;
      else
        acc[v] = acc[v] + ((double)FL.v[v] - (double)FR.v[v]) * dt_d;
    }
  }

  // Gravity source (Update.h:161-166): always into the y momentum; energy from the mass
  // fluxes of the sweep.
  static FV2D_HD void add_gravity(double acc[4], R rho, R FL0, R FR0, double dt, double g) {
    acc[IV] = acc[IV] + dt * (double)rho * g;
    acc[IE] = acc[IE] + dt * 0.5 * ((double)FL0 + (double)FR0) * g;
  }

  // ---------------------------------------------------------------- phase 3: update
  static FV2D_HD void phase_update(int tid, int bx, int by, Smem &sm, const KParams &kp, const StageIO<S> &io,
                                   double dt, double dt_dx, double dt_dy, CellRegs &r) {
    const int xl = tid % TX, yl = tid / TX;
    const int gi = bx * TX + xl, gj = by * TY + yl;
    r.neg_rho = r.neg_prs = r.nan_cnt = 0;
    r.inv = 0.0;
    if (gi >= kp.Nx || gj >= kp.Ny) return;
    const int cx = xl + 2, cy = yl + 2;
    const R rho = sm.q[IR][cy][cx];

    double acc[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] = r.u[v];

    double gX = 0.0, gY = 0.0;
    if (kp.grav == GRAV_CONSTANT) {
      gX = kp.gxf;
      gY = kp.gyf;
    } else if (kp.grav == GRAV_ANALYTICAL) {
      gX = gY = (double)io.grav_row[gj];
    }

    // x sweep (Update.h:171)
    {
      St<R> FL, FR;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        FL.v[v] = sm.fx[v][yl][xl];
        FR.v[v] = sm.fx[v][yl][xl + 1];
      }
      add_flux(acc, FL, FR, dt, dt_dx, kp.dx);
      if (kp.grav != GRAV_NONE) add_gravity(acc, rho, FL.v[IR], FR.v[IR], dt, gX);
    }
    // y sweep (Update.h:172)
    {
      St<R> FL, FR;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        FL.v[v] = sm.fy[v][yl][xl];
        FR.v[v] = sm.fy[v][yl + 1][xl];
      }
      // well-balanced wall flux (Update.h:149-156), with the pout of the other face
      if (kp.wb && (gj == 0 || gj == kp.Ny - 1)) {
        const R g = R(gY);
        if (gj == 0) {
          FL.v[0] = R(0);
          FL.v[1] = R(0);
          FL.v[2] = sm.py[yl + 1][xl] - rho * g * R(kp.dy);
          FL.v[3] = R(0);
        } else {
          FR.v[0] = R(0);
          FR.v[1] = R(0);
          FR.v[2] = sm.py[yl][xl] + rho * g * R(kp.dy);
          FR.v[3] = R(0);
        }
      }
      add_flux(acc, FL, FR, dt, dt_dy, kp.dy);
      if (kp.grav != GRAV_NONE) add_gravity(acc, rho, FL.v[IR], FR.v[IR], dt, gY);
    }

    // thermal conduction into E (ThermalConduction.h:42-106)
    if (TC) {
      const R kc = R(kp.kappa);
      const R kf = R(0.5) * (kc + kc);
      const R dx = R(kp.dx), dy = R(kp.dy);
      const R Tc = sm.t[TC ? cy : 0][TC ? cx : 0];
      const R TL = sm.t[TC ? cy : 0][TC ? cx - 1 : 0];
      const R TR = sm.t[TC ? cy : 0][TC ? cx + 1 : 0];
      const R TU = sm.t[TC ? cy - 1 : 0][TC ? cx : 0];
      const R TD = sm.t[TC ? cy + 1 : 0][TC ? cx : 0];
      R FL = Ops<R>::div(kf * (Tc - TL), dx);
      R FR = Ops<R>::div(kf * (TR - Tc), dx);
      const R FU = Ops<R>::div(kf * (Tc - TU), dy);
      const R FD = Ops<R>::div(kf * (TD - Tc), dy);
      // wall modes replace the x fluxes on the first / last row (ThermalConduction.h:75-103)
      if (gj == 0 && kp.bctc_ymin != BCTC_NONE) {
        if (kp.bctc_ymin == BCTC_FIXED_TEMPERATURE)
          FL = Ops<R>::div(kf * R(2) * (Tc - R(kp.bctc_ymin_value)), dy);
        else if (kp.bctc_ymin == BCTC_FIXED_GRADIENT)
          FL = kf * R(kp.bctc_ymin_value);
      }
      if (gj == kp.Ny - 1 && kp.bctc_ymax != BCTC_NONE) {
        if (kp.bctc_ymax == BCTC_FIXED_TEMPERATURE)
          FR = Ops<R>::div(kf * R(2) * (R(kp.bctc_ymax_value) - Tc), dy);
        else if (kp.bctc_ymax == BCTC_FIXED_GRADIENT)
          FR = kf * R(kp.bctc_ymax_value);
      }
      acc[IE] = acc[IE] + (dt_dx * (double)(FR - FL) + dt_dy * (double)(FD - FU));
    }

    // viscosity (Viscosity.h:33-117); stencil[r][c] = Q(i + c - 1, j + r - 1)
    if (VISC) {
      const R mu = R(kp.mu);
      const R odx = R(kp.inv_dx), ody = R(kp.inv_dy);
      const R four_thirds = R(4.0 / 3.0), two_thirds = R(2.0 / 3.0);
#define FV2D_SQ(c, r, v) sm.q[v][cy + (r)-1][cx + (c)-1]
      R fxu = R(0), fxv = R(0), fxe = R(0);
      R fyu = R(0), fyv = R(0), fye = R(0);
#pragma unroll
      for (int side = 1; side < 3; ++side) {
        const R sign = (side == 1) ? R(-1) : R(1);
        {  // x faces
          const R qiu = (FV2D_SQ(side, 1, IU) + FV2D_SQ(side - 1, 1, IU)) * R(0.5);
          const R qiv = (FV2D_SQ(side, 1, IV) + FV2D_SQ(side - 1, 1, IV)) * R(0.5);
          const R dudx = odx * (FV2D_SQ(side, 1, IU) - FV2D_SQ(side - 1, 1, IU));
          const R dvdx = odx * (FV2D_SQ(side, 1, IV) - FV2D_SQ(side - 1, 1, IV));
          const R dudy = R(0.25) * ody *
                         (FV2D_SQ(side, 2, IU) - FV2D_SQ(side, 0, IU) + FV2D_SQ(side - 1, 2, IU) -
                          FV2D_SQ(side - 1, 0, IU));
          const R dvdy = R(0.25) * ody *
                         (FV2D_SQ(side, 2, IV) - FV2D_SQ(side, 0, IV) + FV2D_SQ(side - 1, 2, IV) -
                          FV2D_SQ(side - 1, 0, IV));
          const R tau_xx = four_thirds * dudx - two_thirds * dvdy;
          const R tau_xy = dvdx + dudy;
          fxu = fxu + sign * mu * tau_xx;
          fxv = fxv + sign * mu * tau_xy;
          fxe = fxe + sign * mu * (tau_xx * qiu + tau_xy * qiv);
        }
        {  // y faces
          const R qiu = (FV2D_SQ(1, side, IU) + FV2D_SQ(1, side - 1, IU)) * R(0.5);
          const R qiv = (FV2D_SQ(1, side, IV) + FV2D_SQ(1, side - 1, IV)) * R(0.5);
          const R dudy = ody * (FV2D_SQ(1, side, IU) - FV2D_SQ(1, side - 1, IU));
          const R dvdy = ody * (FV2D_SQ(1, side, IV) - FV2D_SQ(1, side - 1, IV));
          const R dudx = R(0.25) * odx *
                         (FV2D_SQ(2, side, IU) - FV2D_SQ(0, side, IU) + FV2D_SQ(2, side - 1, IU) -
                          FV2D_SQ(0, side - 1, IU));
          const R dvdx = R(0.25) * odx *
                         (FV2D_SQ(2, side, IV) - FV2D_SQ(0, side, IV) + FV2D_SQ(2, side - 1, IV) -
                          FV2D_SQ(0, side - 1, IV));
          const R tau_yy = four_thirds * dvdy - two_thirds * dudx;
          const R tau_xy = dvdx + dudy;
          fyu = fyu + sign * mu * tau_xy;
          fyv = fyv + sign * mu * tau_yy;
          fye = fye + sign * mu * (tau_xy * qiu + tau_yy * qiv);
        }
      }
#undef FV2D_SQ
      // U += dt * (vf_x + vf_y) (Viscosity.h:116); the density components are zero
      acc[IR] = acc[IR] + (double)(R(0) + R(0)) * dt;
      acc[IU] = acc[IU] + (double)(fxu + fyu) * dt;
      acc[IV] = acc[IV] + (double)(fxv + fyv) * dt;
      acc[IE] = acc[IE] + (double)(fxe + fye) * dt;
    }

    // SSP-RK2 average (Update.h:214-220)
    if (io.avg) {
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] = 0.5 * (r.u0[v] + acc[v]);
    }

    const long long k = (long long)gj * kp.pitch + gi;
    St<R> un;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const S s = S(acc[v]);
      io.out[v * kp.plane + k] = s;
      un.v[v] = R(s);
    }

    // epilogue: consToPrim + checkNegatives (SimInfo.h:576-646) and the next dt's cell term
    if (io.epilogue) {
      St<R> q = cons_to_prim(un, R(kp.gm1));
      const R eps = R(kp.eps_reset);
      if (q.v[IR] < R(0)) {
        q.v[IR] = eps;
        r.neg_rho = 1;
      }
      if (q.v[IP] < R(0)) {
        q.v[IP] = eps;
        r.neg_prs = 1;
      }
      unsigned nn = 0;
#pragma unroll
      for (int v = 0; v < 4; ++v) nn += (q.v[v] != q.v[v]) ? 1u : 0u;
      r.nan_cnt = nn;
      const double inv = inv_dt_hyp(q, kp);
      r.inv = (inv == inv) ? inv : 0.0;  // the reference's fmax-based Max ignores NaN
    }
  }
};

}  // namespace fv2d
