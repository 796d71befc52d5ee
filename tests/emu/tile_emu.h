// tile_emu.h -- TEST INFRASTRUCTURE: the templates of the host emulation of the stage
// kernel (tile_emu.cpp), shared by its instantiation units (tile_emu_inst.cpp).
#pragma once

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <limits>
#include <string>
#include <vector>

#include "../../fv2d_amd/csrc/fv2d_params.h"
#include "../../fv2d_amd/csrc/fv2d_stage.h"

namespace emu {
using namespace fv2d;
}

namespace emu {

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
  const int nw = (int)regs.size() / 64;
  for (int w = 0; w < nw; ++w)
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
    int sx, jb, je;
    block_rows(s.g, kp.Ny, blk, sx, jb, je);
    const int i0 = sx * STRIP_OUT;
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
    if (jb < je) {
      if (K::kPF) LANES(K::chunk_prefetch(lane, w, 0, i0, jb, kp, io, pfs[t]));
      LANES(K::pro_load(lane, w, i0, jb, *sm, kp, io));
      LANES(K::pro_slopes(lane, w, jb, *sm));
      LANES(K::pro_flux(lane, w, jb, *sm, kp));
    }
    int xw = 0;
    for (int jc = jb; jc < je; jc += K::NW) {
      if (K::kPF) {
        LANES(K::chunk_commit(lane, w, xw, i0, jb, jc, *sm, kp, io, pfs[t], r));
        if (!K::kLateOwn) LANES(K::load_own(lane, jc + w, i0, kp, io, r));
        if (jc + K::NW < je)
          LANES(K::chunk_prefetch(lane, w, (xw + 1) % K::NW, i0, jc + K::NW, kp, io, pfs[t]));
      } else {
        LANES(K::chunk_load(lane, w, xw, i0, jb, jc, *sm, kp, io, r));
        if (!K::kLateOwn) LANES(K::load_own(lane, jc + w, i0, kp, io, r));
      }
      LANES(K::pass1_slopes(lane, w, jb, jc, *sm, r));
      shift(regs, [](LaneRegs<R> &d, const LaneRegs<R> *s_) {
        for (int v = 0; v < 4; ++v) d.sxl[v] = s_ ? s_->sx[v] : R(0);
      });
      LANES(K::pass1_flux(lane, w, jb, jc, *sm, kp, r));
      LANES(K::pass2_flux(lane, w, jb, jc, *sm, kp, r));
      LANES(K::update_faces(lane, w, jb, jc, *sm, kp, r));
      if (K::kLateOwn) LANES(K::load_own(lane, jc + w, i0, kp, io, r));
      shift(regs, [](LaneRegs<R> &d, const LaneRegs<R> *s_) {
        for (int v = 0; v < 4; ++v) d.fxl[v] = s_ ? s_->fx[v] : R(0);
        if (K::kTC) d.tcxl = s_ ? s_->tcx : R(0);
        if (K::kVISC)
          for (int v = 0; v < 3; ++v) d.vxl[v] = s_ ? s_->vx[v] : R(0);
      });
      LANES(K::update_cell(lane, w, i0, jb, jc, je, *sm, kp, io, s.ctl.dt, s.ctl.dt_dx, s.ctl.dt_dy, r));
      xw = (xw + 1) % K::NW;
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

template <typename R, typename S, int RECON, int RIEM, bool XG, int NW>
void stage_rrxn(Sim &s, const StageIO<S> &io) {
  const bool tc = s.p.tc_active, vi = s.p.visc_active;
  if (tc && vi) run_stage<StripStage<R, S, RECON, RIEM, true, true, XG, NW>, S>(s, io);
  else if (tc) run_stage<StripStage<R, S, RECON, RIEM, true, false, XG, NW>, S>(s, io);
  else if (vi) run_stage<StripStage<R, S, RECON, RIEM, false, true, XG, NW>, S>(s, io);
  else run_stage<StripStage<R, S, RECON, RIEM, false, false, XG, NW>, S>(s, io);
}

template <typename R, typename S, int RECON, int RIEM, int NW>
void stage_rr(Sim &s, const StageIO<S> &io) {
  const KParams &kp = s.kp;
  if (kp.grav != GRAV_NONE || kp.wb || kp.bctc_ymin || kp.bctc_ymax) stage_rrxn<R, S, RECON, RIEM, true, NW>(s, io);
  else stage_rrxn<R, S, RECON, RIEM, false, NW>(s, io);
}

template <typename R, typename S, int RECON, int NW>
void stage_r(Sim &s, const StageIO<S> &io) {
  switch (s.p.riemann_solver) {
  case HLL: stage_rr<R, S, RECON, HLL, NW>(s, io); break;
  case FSLP: stage_rr<R, S, RECON, FSLP, NW>(s, io); break;
  default: stage_rr<R, S, RECON, HLLC, NW>(s, io); break;
  }
}

// One stage of every kernel class with NW-row chunks; each (R, NW) pair is instantiated in a
// translation unit of its own (tile_emu_inst.cpp), so the classes compile in parallel.
template <typename R, typename S, int NW>
void stage_nw(Sim &s, const StageIO<S> &io) {
  if (s.p.reconstruction == FV2D_PLM) stage_r<R, S, PLM, NW>(s, io);
  else stage_r<R, S, PCM, NW>(s, io);
}

#define EMU_DECLARE(R, NW) extern template void stage_nw<R, R, NW>(Sim &, const StageIO<R> &);
EMU_DECLARE(double, 4)
EMU_DECLARE(double, 8)
EMU_DECLARE(float, 4)
EMU_DECLARE(float, 8)
#undef EMU_DECLARE

}  // namespace emu
