// fv2d_params.h -- host-side validation of fv2d_params and the derived kernel constants.
// Shared by the C-ABI layer and the host emulation of the kernels.
#pragma once

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <string>

#include "../../include/fv2d_amd.h"
#include "fv2d_stage.h"

namespace fv2d {

// Reject what this build does not implement, with the reason.
inline int check_params(const fv2d_params *p, std::string &why) {
  char b[512];
  if (!p) {
    why = "params is NULL";
    return FV2D_E_INVALID;
  }
  if (p->Nx < 2 || p->Ny < 2) {
    snprintf(b, sizeof b, "Nx and Ny must be >= 2 (got %d x %d)", p->Nx, p->Ny);
    why = b;
    return FV2D_E_INVALID;
  }
  if (p->Ng < 1 || (p->reconstruction == FV2D_PLM && p->Ng < 2)) {
    snprintf(b, sizeof b, "Nghosts=%d is too small: the stencils reach 2 cells (PLM) or 1 (PCM)", p->Ng);
    why = b;
    return FV2D_E_INVALID;
  }
  if ((long long)p->Nx * p->Ny > (1ll << 30)) {
    why = "grid larger than 2^30 cells";
    return FV2D_E_INVALID;
  }
  // the stage kernel addresses a row's cells with 32-bit byte offsets (URow, fv2d_stage.h)
  if (p->Nx > (1 << 26)) {
    snprintf(b, sizeof b, "Nx=%d is wider than 2^26 cells", p->Nx);
    why = b;
    return FV2D_E_INVALID;
  }
  if (!(p->xmax > p->xmin) || !(p->ymax > p->ymin)) {
    why = "empty domain (xmax <= xmin or ymax <= ymin)";
    return FV2D_E_INVALID;
  }
  if (p->reconstruction < 0 || p->reconstruction > 2 || p->riemann_solver < 0 || p->riemann_solver > 2 ||
      p->time_stepping < 0 || p->time_stepping > 1 || p->boundary_x < 0 || p->boundary_x > 2 || p->boundary_y < 0 ||
      p->boundary_y > 2 || p->gravity_mode < 0 || p->gravity_mode > 2 || p->bctc_ymin < 0 || p->bctc_ymin > 2 ||
      p->bctc_ymax < 0 || p->bctc_ymax > 2 || p->precision < 0 || p->precision > 2) {
    why = "an option value is out of range";
    return FV2D_E_INVALID;
  }
  if (p->chunk_rows != 0 && !chunk_rows_supported(p->chunk_rows)) {
    snprintf(b, sizeof b, "chunk_rows=%d must be 0, 4 or 8", p->chunk_rows);
    why = b;
    return FV2D_E_INVALID;
  }
  if (p->seg_rows != 0 && p->seg_rows < 4) {
    snprintf(b, sizeof b, "seg_rows=%d must be 0 or at least 4", p->seg_rows);
    why = b;
    return FV2D_E_INVALID;
  }
  if (p->tc_active && p->tc_mode == FV2D_TCM_B02) {
    why = "thermal_conduction.conductivity_mode=B02 is not supported: the reference never reads its b02_* "
          "parameters from the .ini (SimInfo.h:307-310), so its values are undefined";
    return FV2D_E_UNSUPPORTED;
  }
  if (p->tc_active && p->tc_mode != FV2D_TCM_CONSTANT) {
    why = "unknown thermal conduction mode";
    return FV2D_E_INVALID;
  }
  return FV2D_OK;
}

inline void build_kparams(const fv2d_params *p, int pitch, KParams &kp) {
  memset(&kp, 0, sizeof kp);
  kp.Nx = p->Nx;
  kp.Ny = p->Ny;
  kp.pitch = pitch;
  kp.plane = (long long)pitch * p->Ny;
  kp.bcx = p->boundary_x;
  kp.bcy = p->boundary_y;
  kp.grav = p->gravity_mode;
  kp.wb = p->well_balanced_flux_at_y_bc ? 1 : 0;
  kp.bctc_ymin = p->tc_active ? p->bctc_ymin : 0;
  kp.bctc_ymax = p->tc_active ? p->bctc_ymax : 0;
  kp.gamma0 = p->gamma0;
  kp.gm1 = p->gamma0 - 1.0;
  kp.entho = 1.0 / (p->gamma0 - 1.0);
  kp.dx = (p->xmax - p->xmin) / p->Nx;  // SimInfo.h:375-376
  kp.dy = (p->ymax - p->ymin) / p->Ny;
  kp.inv_dx = 1.0 / kp.dx;
  kp.inv_dy = 1.0 / kp.dy;
  kp.divx = div_mode(kp.dx);
  kp.divy = div_mode(kp.dy);
  kp.gxf = (double)(float)p->gx;
  kp.gyf = (double)(float)p->gy;
  kp.gdx_x = p->gx * kp.dx;
  kp.gdx_y = p->gy * kp.dy;
  kp.fslp_K = p->fslp_K;
  kp.kappa = p->kappa;
  kp.mu = p->mu;
  kp.bctc_ymin_value = p->bctc_ymin_value;
  kp.bctc_ymax_value = p->bctc_ymax_value;
  kp.eps_reset = p->epsilon_reset_negative;
  // ComputeDt.h:35-44: inactive terms give epsilon; constant kappa / mu give one value
  kp.inv_tc = p->epsilon;
  if (p->tc_active) kp.inv_tc = fmax(2.0 * p->kappa / (kp.dx * kp.dx), 2.0 * p->kappa / (kp.dy * kp.dy));
  kp.inv_visc = p->epsilon;
  if (p->visc_active) kp.inv_visc = fmax(2.0 * p->mu / (kp.dx * kp.dx), 2.0 * p->mu / (kp.dy * kp.dy));
  kp.CFL = p->CFL;
  kp.epsilon = p->epsilon;
}

// getAnalyticalGravity (Gravity.h:15-30) for rows [j0, j1): it depends on the row only; it is
// evaluated with the host libm and passed through float, as getGravity returns float.
inline void gravity_rows(const fv2d_params &p, const KParams &kp, int j0, int j1, float *out) {
  for (int j = j0; j < j1; ++j) {
    const double y = p.ymin + (j + 0.5) * kp.dy;
    const double g = p.hot_bubble_g0 * sin(y * M_PI * 2.0 / p.ymax);
    out[j - j0] = (float)g;
  }
}

// Rows per chunk (= waves per workgroup) of the single-grid stage kernel.
inline int pick_chunk_rows(const fv2d_params *p) {
  return p->chunk_rows > 0 ? p->chunk_rows : CHUNK_ROWS_DEFAULT;
}

// Rows per workgroup segment: about 64 by default, a multiple of the chunk height (a 4-row
// prologue per segment, enough workgroups to fill 256 CUs at 4096^2), fewer on short grids so
// that small grids still spread over the chip.
inline int pick_seg_rows(const fv2d_params *p, int nw) {
  if (p->seg_rows > 0) return p->seg_rows;
  const int step = nw < 8 ? 8 : nw;
  int seg = 64 / step * step;
  const long long strips = (p->Nx + 61) / 62;
  while (seg > step && strips * ((p->Ny + seg - 1) / seg) < 2048) seg -= step;
  return seg;
}

// Strip geometry of the single-grid kernel: the banded long-then-short schedule when the
// uniform segments make at least 2048 workgroups (two rounds of 4 per CU on 256 CUs), else
// the uniform one.
inline StripGeom sim_geom(const fv2d_params *p) {
  const int nw = pick_chunk_rows(p);
  const int seg = pick_seg_rows(p, nw);
  const StripGeom u = strip_geom(p->Nx, p->Ny, seg, nw);
  return (u.nblocks >= 2048) ? strip_geom_banded(p->Nx, p->Ny, seg, nw) : u;
}

}  // namespace fv2d
