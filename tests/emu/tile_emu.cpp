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

struct Sim {
  fv2d_params p;
  KParams kp;
  bool fp32;
  std::vector<double> Ud, Usd, Q0d;
  std::vector<float> Uf, Usf, Q0f;
  std::vector<float> grav_row;
  std::vector<double> partials;
  StripGeom g;
  StepCtl ctl;
  bool fresh = false;
  unsigned long long last[3] = {0, 0, 0};
};

// Lane l receives lane l-1's value; lane 0 gets 0 (the DPP wave_shr:1 of the kernel).
template <typename R, typename F>
void shift(std::vector<LaneRegs<R>> &regs, F field) {
  for (int w = 0; w < STRIP_ROWS; ++w)
    for (int l = 63; l >= 0; --l) {
      LaneRegs<R> &dst = regs[w * 64 + l];
      field(dst, l ? &regs[w * 64 + l - 1] : nullptr);
    }
}

template <class K, typename S>
void run_stage(Sim &s, const StageIO<S> &io) {
  const KParams &kp = s.kp;
  typedef decltype(K::Smem::q[0][0][0]) RR;
  typedef typename std::remove_const<typename std::remove_reference<RR>::type>::type R;
  std::vector<LaneRegs<R>> regs(K::NT);
  typename K::Smem *sm = (typename K::Smem *)malloc(sizeof(typename K::Smem));
  const R nan = std::numeric_limits<R>::quiet_NaN();
  for (int blk = 0; blk < s.g.nblocks; ++blk) {
    int sx, sy;
    strip_of_block(s.g, blk, sx, sy);
    const int i0 = sx * STRIP_OUT, jb = sy * s.g.seg;
    const int je = (jb + s.g.seg < kp.Ny) ? jb + s.g.seg : kp.Ny;
    R *raw = (R *)sm;
    for (size_t k = 0; k < sizeof(typename K::Smem) / sizeof(R); ++k) raw[k] = nan;
    for (auto &r : regs) {
      memset(&r, 0xff, sizeof r);  // garbage registers: every use must follow a definition
      r.neg_rho = r.neg_prs = r.nan_cnt = 0;
      r.inv = 0.0;
    }
#define LANES(stmt)                                                                                                    \
  for (int t = 0; t < K::NT; ++t) {                                                                                    \
    const int lane = t & 63, w = t >> 6;                                                                               \
    LaneRegs<R> &r = regs[t];                                                                                          \
    (void)lane;                                                                                                        \
    (void)w;                                                                                                           \
    (void)r;                                                                                                           \
    stmt;                                                                                                              \
  }
    std::vector<RowPrefetch<S>> pfs(K::NT);
    for (auto &p : pfs) memset(&p, 0xff, sizeof p);  // garbage until loaded
    if (K::kPF && jb < je) LANES(K::chunk_prefetch(lane, w, 0, i0, jb, kp, io, pfs[t]));
    LANES(K::pro_load(lane, w, i0, jb, *sm, kp, io));
    LANES(K::pro_slopes(lane, w, jb, *sm));
    LANES(K::pro_flux(lane, w, jb, *sm, kp));
    int xw = 0;
    for (int jc = jb; jc < je; jc += STRIP_ROWS) {
      if (K::kPF) {
        LANES(K::chunk_commit(lane, w, xw, i0, jb, jc, *sm, kp, io, pfs[t], r));
        if (jc + STRIP_ROWS < je)
          LANES(K::chunk_prefetch(lane, w, (xw + 1) & (STRIP_ROWS - 1), i0, jc + STRIP_ROWS, kp, io, pfs[t]));
      } else {
        LANES(K::chunk_load(lane, w, xw, i0, jb, jc, *sm, kp, io, r));
      }
      LANES(K::pass1_slopes(lane, w, jb, jc, *sm, r));
      shift(regs, [](LaneRegs<R> &d, const LaneRegs<R> *s_) {
        for (int v = 0; v < 4; ++v) d.sxl[v] = s_ ? s_->sx[v] : R(0);
      });
      LANES(K::pass1_flux(lane, w, jb, jc, *sm, kp, r));
      LANES(K::pass2_flux(lane, w, jb, jc, *sm, kp, r));
      LANES(K::update_faces(lane, w, jb, jc, *sm, kp, r));
      shift(regs, [](LaneRegs<R> &d, const LaneRegs<R> *s_) {
        for (int v = 0; v < 4; ++v) d.fxl[v] = s_ ? s_->fx[v] : R(0);
        if (K::kTC) d.tcxl = s_ ? s_->tcx : R(0);
        if (K::kVISC)
          for (int v = 0; v < 3; ++v) d.vxl[v] = s_ ? s_->vx[v] : R(0);
      });
      LANES(K::update_cell(lane, w, i0, jb, jc, je, *sm, kp, io, s.ctl.dt, s.ctl.dt_dx, s.ctl.dt_dy, r));
      xw = (xw + 1) & (STRIP_ROWS - 1);
    }
#undef LANES
    if (io.epilogue) {
      double m = 0.0;
      for (auto &r : regs) {
        m = fmax(m, r.inv);
        s.ctl.neg_rho += r.neg_rho;
        s.ctl.neg_prs += r.neg_prs;
        s.ctl.nan_cnt += r.nan_cnt;
      }
      s.partials[blk] = m;
    }
  }
  free(sm);
}

template <typename R, typename S, int RECON, int RIEM, bool XG>
void stage_rrx(Sim &s, const StageIO<S> &io) {
  const bool tc = s.p.tc_active, vi = s.p.visc_active;
  if (tc && vi) run_stage<StripStage<R, S, RECON, RIEM, true, true, XG>, S>(s, io);
  else if (tc) run_stage<StripStage<R, S, RECON, RIEM, true, false, XG>, S>(s, io);
  else if (vi) run_stage<StripStage<R, S, RECON, RIEM, false, true, XG>, S>(s, io);
  else run_stage<StripStage<R, S, RECON, RIEM, false, false, XG>, S>(s, io);
}

template <typename R, typename S, int RECON, int RIEM>
void stage_rr(Sim &s, const StageIO<S> &io) {
  const KParams &kp = s.kp;
  if (kp.grav != GRAV_NONE || kp.wb || kp.bctc_ymin || kp.bctc_ymax) stage_rrx<R, S, RECON, RIEM, true>(s, io);
  else stage_rrx<R, S, RECON, RIEM, false>(s, io);
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

// Number of a[k] for which markstein_div(a, d, RN(1/d)) differs from a / d in any bit.
long long emu_markstein_mismatches(const double *a, long long n, double d) {
  const double rd = 1.0 / d;
  const int fast = (d >= 0x1p-60 && d <= 0x1p+60) ? 1 : 0;
  long long bad = 0;
  for (long long k = 0; k < n; ++k) {
    const double x = markstein_div(a[k], d, rd, fast), y = a[k] / d;
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
  Sim *s = new Sim();
  s->p = *p;
  s->fp32 = p->precision == FV2D_FP32;
  const int pitch = (p->Nx + 63) / 64 * 64;
  build_kparams(p, pitch, s->kp);
  s->g = strip_geom(p->Nx, p->Ny, pick_seg_rows(p));
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
