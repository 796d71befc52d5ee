// tile_emu.cpp -- TEST INFRASTRUCTURE: host emulation of the stage kernel.
//
// Runs the phase functions of fv2d_stage.h (the code the HIP kernel runs) on the CPU,
// workgroup by workgroup and lane by lane, with the kernel's barriers as phase boundaries
// and its DPP lane shifts as copies between lane records, and with the same step sequence
// as the C-ABI layer (set_prim -> dt cells -> reduce; step -> stages -> reduce). The LDS
// image of every workgroup is filled with NaN and the lane records with garbage before
// use, so a read of anything not written first shows up in the results. Built with
// -fsanitize=address,undefined it also checks every global and LDS index. Used by the CPU
// test suite to compare the kernel logic with the oracle before any GPU run.
#include "tile_emu.h"

using namespace fv2d;
using namespace emu;

namespace {

template <typename R, typename S>
void stage(Sim &s, const S *src, int qsrc, const S *q0, const S *avg, S *out, int epilogue) {
  StageIO<S> io;
  io.src = src;
  io.q0 = q0;
  io.avg = avg;
  io.out = out;
  io.grav_row = s.grav_row.empty() ? nullptr : s.grav_row.data();
  io.ctl = &s.ctl;
  io.partials = s.partials.data();
  io.qsrc = qsrc;
  io.epilogue = epilogue;
  switch (s.g.nw) {
  case 8: stage_nw<R, S, 8>(s, io); break;
  default: stage_nw<R, S, 4>(s, io); break;
  }
}

void reduce(Sim &s) {
  double m = 0.0;
  for (double v : s.partials) m = fmax(m, v);
  s.ctl.inv_hyp = m;
}

template <typename R, typename S>
void dt_tiles(Sim &s, const S *src, int from_q0) {
  const KParams &kp = s.kp;
  const long long n = (long long)kp.Nx * kp.Ny;
  const int nparts = s.g.nblocks;
  const long long per = (n + nparts - 1) / nparts;
  for (int b = 0; b < nparts; ++b) {
    double m = 0.0;
    for (long long c = b * per; c < n && c < (b + 1) * per; ++c) {
      const int j = (int)(c / kp.Nx), i = (int)(c % kp.Nx);
      const long long k = (long long)j * kp.pitch + i;
      St<R> q;
      for (int v = 0; v < 4; ++v) q.v[v] = R(src[v * kp.plane + k]);
      if (!from_q0) {
        q = cons_to_prim(q, R(kp.gm1));
        reset_negative(q, R(kp.eps_reset));
      }
      const double inv = inv_dt_hyp(q, kp);
      if (inv > m) m = inv;
    }
    s.partials[b] = m;
  }
}

template <typename R, typename S>
void set_prim_t(Sim &s, std::vector<S> &U, std::vector<S> &Q0, const double *q) {
  const KParams &kp = s.kp;
  const long long n = (long long)kp.Nx * kp.Ny;
  for (long long c = 0; c < n; ++c) {
    const int j = (int)(c / kp.Nx), i = (int)(c % kp.Nx);
    St<double> qq;
    for (int v = 0; v < 4; ++v) qq.v[v] = q[v * n + c];
    const St<double> uu = prim_to_cons(qq, kp.gm1);
    const long long k = (long long)j * kp.pitch + i;
    for (int v = 0; v < 4; ++v) {
      U[v * kp.plane + k] = S(uu.v[v]);
      Q0[v * kp.plane + k] = S(qq.v[v]);
    }
  }
  memset(&s.ctl, 0, sizeof s.ctl);
  dt_tiles<R, S>(s, Q0.data(), 1);
  reduce(s);
  s.fresh = true;
  memset(s.last, 0, sizeof s.last);
}

template <typename R, typename S>
void step_t(Sim &s, std::vector<S> &U, std::vector<S> &Us, std::vector<S> &Q0, double dt) {
  s.ctl.dt = dt;
  s.ctl.dt_dx = dt / s.kp.dx;
  s.ctl.dt_dy = dt / s.kp.dy;
  s.ctl.active = 1;
  const int q1 = s.fresh ? QSRC_Q0 : QSRC_U_RESET;
  if (s.p.time_stepping == FV2D_TS_EULER) {
    stage<R, S>(s, U.data(), q1, Q0.data(), nullptr, Us.data(), 1);
    U.swap(Us);
  } else {
    stage<R, S>(s, U.data(), q1, Q0.data(), nullptr, Us.data(), 0);
    stage<R, S>(s, Us.data(), QSRC_U, Q0.data(), U.data(), U.data(), 1);
  }
  reduce(s);
  s.fresh = false;
}

template <typename S>
void get_prim_t(const Sim &s, const std::vector<S> &U, const std::vector<S> &Q0, double *q) {
  const KParams &kp = s.kp;
  const long long n = (long long)kp.Nx * kp.Ny;
  for (long long c = 0; c < n; ++c) {
    const int j = (int)(c / kp.Nx), i = (int)(c % kp.Nx);
    const long long k = (long long)j * kp.pitch + i;
    St<double> v4;
    const std::vector<S> &src = s.fresh ? Q0 : U;
    for (int v = 0; v < 4; ++v) v4.v[v] = (double)src[v * kp.plane + k];
    if (!s.fresh) {
      v4 = cons_to_prim(v4, kp.gm1);
      reset_negative(v4, kp.eps_reset);
    }
    for (int v = 0; v < 4; ++v) q[v * n + c] = v4.v[v];
  }
}

template <typename S>
void get_cons_t(const Sim &s, const std::vector<S> &U, double *u) {
  const KParams &kp = s.kp;
  const long long n = (long long)kp.Nx * kp.Ny;
  for (long long c = 0; c < n; ++c) {
    const int j = (int)(c / kp.Nx), i = (int)(c % kp.Nx);
    const long long k = (long long)j * kp.pitch + i;
    for (int v = 0; v < 4; ++v) u[v * n + c] = (double)U[v * kp.plane + k];
  }
}

}  // namespace

extern "C" {

int emu_div_mode(double d) { return div_mode(d); }

// Number of a[k] for which markstein_div(a, d, RN(1/d), mode) differs from a / d in any bit;
// mode < 0: the mode the solver picks for d (div_mode).
long long emu_markstein_mismatches(const double *a, long long n, double d, int mode) {
  const double rd = 1.0 / d;
  if (mode < 0) mode = div_mode(d);
  long long bad = 0;
  for (long long k = 0; k < n; ++k) {
    const double x = markstein_div(a[k], d, rd, mode), y = a[k] / d;
    if (memcmp(&x, &y, sizeof x) != 0 && !(x != x && y != y)) ++bad;
  }
  return bad;
}

void *emu_create(const fv2d_params *p, char *err) {
  std::string why;
  if (check_params(p, why)) {
    if (err) snprintf(err, 256, "%s", why.c_str());
    return nullptr;
  }
  if (p->precision == FV2D_FP32_STATE64) {
    if (err) snprintf(err, 256, "the FP32 form with an FP64 state is not emulated");
    return nullptr;
  }
  Sim *s = new Sim();
  s->p = *p;
  s->fp32 = p->precision == FV2D_FP32;
  const int pitch = (p->Nx + 63) / 64 * 64;
  build_kparams(p, pitch, s->kp);
  s->g = sim_geom(p);
  s->partials.assign((size_t)s->g.nblocks, 0.0);
  const size_t n = 4 * (size_t)s->kp.plane;
  // unused padding columns hold NaN: a read of them would poison the results
  const double dn = std::numeric_limits<double>::quiet_NaN();
  const float fn = std::numeric_limits<float>::quiet_NaN();
  if (s->fp32) {
    s->Uf.assign(n, fn);
    s->Usf.assign(n, fn);
    s->Q0f.assign(n, fn);
  } else {
    s->Ud.assign(n, dn);
    s->Usd.assign(n, dn);
    s->Q0d.assign(n, dn);
  }
  if (p->gravity_mode == FV2D_GRAV_ANALYTICAL) {
    s->grav_row.resize(p->Ny);
    for (int j = 0; j < p->Ny; ++j) {
      const double y = p->ymin + (j + 0.5) * s->kp.dy;
      s->grav_row[j] = (float)(p->hot_bubble_g0 * sin(y * M_PI * 2.0 / p->ymax));
    }
  }
  memset(&s->ctl, 0, sizeof s->ctl);
  return s;
}

void emu_destroy(void *h) { delete (Sim *)h; }

// Replaces the simulation's geometry by the banded long-then-short schedule with long
// segments of seg rows (whatever the grid size), so that small grids exercise it.
void emu_use_banded(void *h, int seg) {
  Sim &s = *(Sim *)h;
  s.g = strip_geom_banded(s.p.Nx, s.p.Ny, seg, s.g.nw);
  s.partials.assign((size_t)s.g.nblocks, 0.0);
}

// Geometry of block_rows for a test of the mapping: returns nblocks and fills, per block,
// (sx, jb, je) into out (3 * nblocks ints, when out is not NULL).
int emu_block_rows(int Nx, int Ny, int seg, int nw, int banded, int *out) {
  const StripGeom g = banded ? strip_geom_banded(Nx, Ny, seg, nw) : strip_geom(Nx, Ny, seg, nw);
  if (out)
    for (int b = 0; b < g.nblocks; ++b) block_rows(g, Ny, b, out[3 * b], out[3 * b + 1], out[3 * b + 2]);
  return g.nblocks;
}

// Geometry the single-grid kernel uses for these parameters: (nblocks, seg, sched, lseg,
// sseg, tail) into out[6].
void emu_sim_geom(const fv2d_params *p, int *out) {
  const StripGeom g = sim_geom(p);
  out[0] = g.nblocks;
  out[1] = g.seg;
  out[2] = g.sched;
  out[3] = g.lseg;
  out[4] = g.sseg;
  out[5] = g.tail;
}

void emu_set_prim(void *h, const double *q) {
  Sim &s = *(Sim *)h;
  if (s.fp32) set_prim_t<float, float>(s, s.Uf, s.Q0f, q);
  else set_prim_t<double, double>(s, s.Ud, s.Q0d, q);
}

double emu_compute_dt(void *h, double inv[3]) {
  Sim &s = *(Sim *)h;
  double m = s.ctl.inv_hyp;
  if (m < s.kp.inv_tc) m = s.kp.inv_tc;
  if (m < s.kp.inv_visc) m = s.kp.inv_visc;
  if (inv) {
    inv[0] = s.ctl.inv_hyp;
    inv[1] = s.kp.inv_tc;
    inv[2] = s.kp.inv_visc;
  }
  return s.p.CFL / m;
}

void emu_step(void *h, double dt, uint64_t counts[3]) {
  Sim &s = *(Sim *)h;
  if (s.fp32) step_t<float, float>(s, s.Uf, s.Usf, s.Q0f, dt);
  else step_t<double, double>(s, s.Ud, s.Usd, s.Q0d, dt);
  const unsigned long long tot[3] = {s.ctl.neg_rho, s.ctl.neg_prs, s.ctl.nan_cnt};
  if (counts)
    for (int k = 0; k < 3; ++k) counts[k] = tot[k] - s.last[k];
  memcpy(s.last, tot, sizeof tot);
}

void emu_get_prim(void *h, double *q) {
  Sim &s = *(Sim *)h;
  if (s.fp32) get_prim_t<float>(s, s.Uf, s.Q0f, q);
  else get_prim_t<double>(s, s.Ud, s.Q0d, q);
}

void emu_get_cons(void *h, double *u) {
  Sim &s = *(Sim *)h;
  if (s.fp32) get_cons_t<float>(s, s.Uf, u);
  else get_cons_t<double>(s, s.Ud, u);
}

}  // extern "C"
