// tile_emu.cpp -- TEST INFRASTRUCTURE: host emulation of the stage kernel.
//
// Runs the phase functions of fv2d_stage.h (the code the HIP kernel runs) on the CPU,
// workgroup by workgroup and thread by thread, with the same step sequence as the C-ABI
// layer (set_prim -> dt tiles -> reduce; step -> stages -> reduce). The LDS image of every
// workgroup is filled with NaN before use, so a read of a slot that no thread wrote shows
// up as a NaN in the results. Built with -fsanitize=address,undefined it also checks every
// global and LDS index. Used by the CPU test suite to compare the kernel logic with the
// oracle before any GPU run.
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <limits>
#include <string>
#include <vector>

#include "../../fv2d_amd/csrc/fv2d_params.h"
#include "../../fv2d_amd/csrc/fv2d_stage.h"

using namespace fv2d;

namespace {

constexpr int TX = 32, TY = 8;

struct Sim {
  fv2d_params p;
  KParams kp;
  bool fp32;
  std::vector<double> Ud, Usd, Q0d;
  std::vector<float> Uf, Usf, Q0f;
  std::vector<float> grav_row;
  std::vector<double> partials;
  int gx, gy;
  StepCtl ctl;
  bool fresh = false;
  unsigned long long last[3] = {0, 0, 0};
};

template <class K, typename S>
void run_stage(Sim &s, const StageIO<S> &io) {
  const KParams &kp = s.kp;
  std::vector<CellRegs> regs(K::NT);
  typename K::Smem *sm = (typename K::Smem *)malloc(sizeof(typename K::Smem));
  typedef decltype(sm->q[0][0][0]) RR;
  typedef typename std::remove_reference<RR>::type R;
  for (int by = 0; by < s.gy; ++by)
    for (int bx = 0; bx < s.gx; ++bx) {
      R *raw = (R *)sm;
      for (size_t k = 0; k < sizeof(typename K::Smem) / sizeof(R); ++k) raw[k] = std::numeric_limits<R>::quiet_NaN();
      for (int t = 0; t < K::NT; ++t) K::phase_load(t, bx, by, *sm, kp, io, regs[t]);
      for (int t = 0; t < K::NT; ++t) K::phase_flux(t, *sm, kp);
      for (int t = 0; t < K::NT; ++t)
        K::phase_update(t, bx, by, *sm, kp, io, s.ctl.dt, s.ctl.dt_dx, s.ctl.dt_dy, regs[t]);
      if (io.epilogue) {
        double m = 0.0;
        for (int t = 0; t < K::NT; ++t) {
          m = fmax(m, regs[t].inv);
          s.ctl.neg_rho += regs[t].neg_rho;
          s.ctl.neg_prs += regs[t].neg_prs;
          s.ctl.nan_cnt += regs[t].nan_cnt;
        }
        s.partials[bx + by * s.gx] = m;
      }
    }
  free(sm);
}

template <typename R, typename S, int RECON, int RIEM>
void stage_rr(Sim &s, const StageIO<S> &io) {
  const bool tc = s.p.tc_active, vi = s.p.visc_active;
  if (tc && vi) run_stage<TileStage<R, S, TX, TY, RECON, RIEM, true, true>, S>(s, io);
  else if (tc) run_stage<TileStage<R, S, TX, TY, RECON, RIEM, true, false>, S>(s, io);
  else if (vi) run_stage<TileStage<R, S, TX, TY, RECON, RIEM, false, true>, S>(s, io);
  else run_stage<TileStage<R, S, TX, TY, RECON, RIEM, false, false>, S>(s, io);
}

template <typename R, typename S, int RECON>
void stage_r(Sim &s, const StageIO<S> &io) {
  switch (s.p.riemann_solver) {
  case HLL: stage_rr<R, S, RECON, HLL>(s, io); break;
  case FSLP: stage_rr<R, S, RECON, FSLP>(s, io); break;
  default: stage_rr<R, S, RECON, HLLC>(s, io); break;
  }
}

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
  if (s.p.reconstruction == FV2D_PLM) stage_r<R, S, PLM>(s, io);
  else stage_r<R, S, PCM>(s, io);
}

void reduce(Sim &s) {
  double m = 0.0;
  for (double v : s.partials) m = fmax(m, v);
  s.ctl.inv_hyp = m;
}

template <typename R, typename S>
void dt_tiles(Sim &s, const S *src, int from_q0) {
  const KParams &kp = s.kp;
  for (int by = 0; by < s.gy; ++by)
    for (int bx = 0; bx < s.gx; ++bx) {
      double m = 0.0;
      for (int t = 0; t < TX * TY; ++t) {
        const int gi = bx * TX + t % TX, gj = by * TY + t / TX;
        if (gi >= kp.Nx || gj >= kp.Ny) continue;
        const long long k = (long long)gj * kp.pitch + gi;
        St<R> q;
        for (int v = 0; v < 4; ++v) q.v[v] = R(src[v * kp.plane + k]);
        if (!from_q0) {
          q = cons_to_prim(q, R(kp.gm1));
          reset_negative(q, R(kp.eps_reset));
        }
        const double inv = inv_dt_hyp(q, kp);
        m = fmax(m, (inv == inv) ? inv : 0.0);
      }
      s.partials[bx + by * s.gx] = m;
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

void *emu_create(const fv2d_params *p, char *err) {
  std::string why;
  if (check_params(p, why)) {
    if (err) snprintf(err, 256, "%s", why.c_str());
    return nullptr;
  }
  Sim *s = new Sim();
  s->p = *p;
  s->fp32 = p->precision == FV2D_FP32;
  const int pitch = (p->Nx + 63) / 64 * 64;
  build_kparams(p, pitch, s->kp);
  s->gx = (p->Nx + TX - 1) / TX;
  s->gy = (p->Ny + TY - 1) / TY;
  s->partials.assign((size_t)s->gx * s->gy, 0.0);
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
