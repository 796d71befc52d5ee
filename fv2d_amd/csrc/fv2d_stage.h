// fv2d_stage.h -- the fused stage of the hydro step: column-strip marching, one wave per
// row, one lane per column.
//
// A workgroup of 4 wave64s owns a strip of 62 domain columns and a segment of rows, and
// marches up the segment 4 rows (one "chunk") at a time. Lane l of every wave stands for
// strip column l (columns 0..65 of the strip are domain columns i0-2 .. i0+63; 2..63 are
// the 62 output columns), wave w for row jc+w of the current chunk. Per chunk:
//   load   wave w turns the U of row jc+2+w into primitives (consToPrim + negative
//          reset, or the initial Q) and stores them, and T = p/rho, in a 7-row LDS ring.
//          Ghost cells are synthesised from the boundary rules (map_ghost); no ghost is
//          stored in HBM.
//   pass 1 x slope of column l+1 (minmod), shifted one lane right (DPP) to get column l's;
//          the x interface (l, l+1) is solved in registers. The y slope of row jc+w+1 is
//          computed once and put in an LDS ring for the wave above.
//   pass 2 the y interface (jc+w, jc+w+1) is solved with the two y slopes; its flux stays
//          in registers (top face of this wave's cell) and goes to an LDS ring (bottom face
//          of the next row's cell).
//   update the x-face fluxes of the left neighbour come from the lane to the left (DPP);
//          the hydro, gravity, conduction and viscosity terms are applied in the reference's
//          accumulation order; RK2 stage 2 averages with U^n. The optional epilogue forms
//          the next step's primitives (reset counts) and hyperbolic inverse dt.
// Every interface flux, slope and x-face conduction / viscous flux is computed once;
// the reference solves each interface twice with the same arguments (Update.h:142-143)
// and evaluates the same face terms from both sides, with the same bits.
//
// The phase functions are plain host/device code. The HIP kernel runs them with barriers
// and DPP lane shifts in between; the host emulation runs them lane by lane.
#pragma once

#include <type_traits>

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
  int fresh;                // batch mode: the next first stage reads Q0
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

// One job of a batch (device-resident descriptor). All jobs of a batch share the
// precision; buffers are typed by it.
struct BatchJob {
  KParams kp;
  int nblocks;   // workgroup items of this job (strip geometry)
  int seg;       // rows per segment
  int nsx, nsy;
  void *U, *Us, *Q0;
  const float *grav_row;
  StepCtl *ctl;
  double *partials;
  int rk2;       // SSP-RK2 (else Euler)
  int used;      // slot holds a job
  int cls;       // kernel class (index into the batch's class table)
};

struct BatchItem {
  int job;  // slot index
  int blk;  // workgroup item of the job
};

// Accessor of one row that is the same for the whole wave (a ring row, the own cells, the
// output row), read or written at per-lane columns. On the GPU the row address goes into a
// raw buffer resource (scalar registers: untyped, stride 0, 2^31 - 1 bytes, the gfx9
// 32-bit data-format word) and each lane supplies a 32-bit byte offset, so the four fields
// of a cell cost no vector address arithmetic; on the host it is a plain pointer. The row
// must be wave-uniform, and its columns within 2^28 of it.
template <typename T>
struct URow {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __amdgpu_buffer_rsrc_t r;
  __device__ explicit URow(const T *row)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(row), 0, 0x7fffffff, 0x00020000)) {}
  __device__ T operator[](int col) const {
    if constexpr (sizeof(T) == 4)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, col * 4, 0, 0));
    else
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, col * 8, 0, 0));
  }
  __device__ void store(int col, T x) const {
    if constexpr (sizeof(T) == 4)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, x), r, col * 4, 0, 0);
    else
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, col * 8, 0, 0);
  }
#else
  T *p;
  explicit URow(const T *row) : p(const_cast<T *>(row)) {}
  T operator[](int col) const { return p[col]; }
  void store(int col, T x) const { p[col] = x; }
#endif
};

// Strip / segment geometry shared by the launcher, the kernel and the emulation.
//  sched 0 (batch work items, small grids): uniform segments of `seg` rows; block b is
//    (strip, segment) item strip_of_block(b).
//  sched 1 (single grids of two or more rounds of workgroups): block b runs on XCD b % 8
//    (the hardware's round-robin) as that XCD's (b / 8)-th block. XCD x owns the row band
//    [x * nchunks / 8, (x + 1) * nchunks / 8) of chunks; its blocks, in dispatch order, take
//    the band's long segments (lseg chunks) top-down and then its last `tail` chunks as short
//    segments (sseg chunks), each segment across all nsx strips. The last workgroups of a
//    launch are then short ones, so the launch drains in a short workgroup's time.
struct StripGeom {
  int nsx, nsy;  // strips across x, segments along y (sched 0)
  int seg;       // rows per segment (a multiple of nw keeps every chunk full)
  int nblocks;
  int nw;        // rows per chunk = waves per workgroup (a CHUNK_ROWS value)
  int sched;     // 0 uniform, 1 banded long-then-short (above)
  int nchunks;   // chunks (nw rows) covering Ny (sched 1)
  int lseg, sseg, tail;  // long / short segment and tail heights in chunks (sched 1)
  int per_xcd;   // blocks per XCD (sched 1); blocks past a band's segments are empty
};

constexpr int STRIP_OUT = 62;  // output columns per strip
constexpr int STRIP_W = 64;    // lanes = strip columns held in registers

// Rows per chunk (= waves per workgroup) the kernels are built for: 4 or 8. Larger chunks
// spread the rings' fixed rows over more waves (LDS: FP64 6.2 KB per wave + 13.6 KB per
// workgroup) and halve the barriers per row, but every barrier then waits for more waves.
constexpr int CHUNK_ROWS_DEFAULT = 4;
FV2D_HD bool chunk_rows_supported(int nw) { return nw == 4 || nw == 8; }

FV2D_HD StripGeom strip_geom(int Nx, int Ny, int seg, int nw) {
  StripGeom g;
  g.seg = seg;
  g.nw = nw;
  g.nsx = (Nx + STRIP_OUT - 1) / STRIP_OUT;
  g.nsy = (Ny + seg - 1) / seg;
  g.nblocks = g.nsx * g.nsy;
  g.sched = 0;
  g.nchunks = g.lseg = g.sseg = g.tail = g.per_xcd = 0;
  return g;
}

// Band of XCD x in chunks, and the number of its long and short segments (sched 1).
FV2D_HD void xcd_band(const StripGeom &g, int x, int &c0, int &c1, int &nl, int &ns) {
  c0 = (int)((long long)x * g.nchunks / 8);
  c1 = (int)((long long)(x + 1) * g.nchunks / 8);
  const int h = c1 - c0, t = h < g.tail ? h : g.tail;
  nl = (h - t + g.lseg - 1) / g.lseg;
  ns = (t + g.sseg - 1) / g.sseg;
}

// The banded long-then-short schedule (sched 1) for a grid whose uniform segmentation
// already spans at least two rounds of workgroups: long segments of `seg` rows, short ones
// of a quarter of that but at least two chunks (a short segment still pays the 4-row
// prologue), a tail of one long segment per band.
FV2D_HD StripGeom strip_geom_banded(int Nx, int Ny, int seg, int nw) {
  StripGeom g = strip_geom(Nx, Ny, seg, nw);
  g.sched = 1;
  g.nchunks = (Ny + nw - 1) / nw;
  g.lseg = (seg + nw - 1) / nw;
  g.sseg = g.lseg >= 8 ? g.lseg / 4 : (g.lseg < 2 ? g.lseg : 2);
  g.tail = g.lseg;
  g.per_xcd = 0;
  for (int x = 0; x < 8; ++x) {
    int c0, c1, nl, ns;
    xcd_band(g, x, c0, c1, nl, ns);
    const int n = g.nsx * (nl + ns);
    if (n > g.per_xcd) g.per_xcd = n;
  }
  g.nblocks = 8 * g.per_xcd;
  return g;
}

// Workgroup id -> (strip, segment). Consecutive workgroup ids are dispatched round-robin
// over the 8 XCDs (id % 8); XCD x gets the contiguous run of logical items
// [x*q + min(x, r), ...) of its q or q+1 workgroups (nblocks = 8q + r), so that
// x-neighbouring strips (which share halo columns) run on the same L2.
FV2D_HD void strip_of_block(const StripGeom &g, int blk, int &sx, int &sy) {
  const int q = g.nblocks >> 3, r = g.nblocks & 7;
  const int x = blk & 7, k = blk >> 3;
  const int logical = x * q + (x < r ? x : r) + k;
  sx = logical % g.nsx;
  sy = logical / g.nsx;
}

// Workgroup id -> strip sx and its output rows [jb, je) (je == jb: nothing to do).
FV2D_HD void block_rows(const StripGeom &g, int Ny, int blk, int &sx, int &jb, int &je) {
  if (g.sched == 0) {
    int sy;
    strip_of_block(g, blk, sx, sy);
    jb = sy * g.seg;
    je = (jb + g.seg < Ny) ? jb + g.seg : Ny;
    return;
  }
  const int x = blk & 7, k = blk >> 3;
  int c0, c1, nl, ns;
  xcd_band(g, x, c0, c1, nl, ns);
  const int m = k / g.nsx;
  sx = k - m * g.nsx;
  const int cl = c1 - (c1 - c0 < g.tail ? c1 - c0 : g.tail);  // first chunk of the tail
  int a, b;
  if (m < nl) {
    a = c0 + m * g.lseg;
    b = (a + g.lseg < cl) ? a + g.lseg : cl;
  } else if (m < nl + ns) {
    a = cl + (m - nl) * g.sseg;
    b = (a + g.sseg < c1) ? a + g.sseg : c1;
  } else {
    a = b = 0;
  }
  jb = a * g.nw;
  je = (b * g.nw < Ny) ? b * g.nw : Ny;
  if (je < jb) je = jb;
}

// Raw values (U or Q0, storage type) of the next chunk's ring row and halo columns, loaded
// one chunk ahead so that the load latency overlaps the current chunk's work.
template <typename S>
struct RowPrefetch {
  S row[4];
  S extra[4];
  int flags;  // ghost sign flips: bit 0/1 row u/v, bit 2/3 extra u/v
};

// Per-lane state carried between phases (registers on the GPU). S: storage type of U.
template <typename R, typename S = R>
struct LaneRegs {
  S u[4];        // own cell U (base of the update)
  S u0[4];       // own cell U^n (RK2 average)
  R sx[4];       // x slope of strip column l+1
  R sxl[4];      // x slope of column l (from lane l-1)
  R syu[4];      // y slope of (l, row+1)
  R fx[4];       // flux at x interface (l, l+1)
  R fxl[4];      // flux at x interface (l-1, l) (from lane l-1)
  R fy[4];       // flux at the top y interface of the cell
  R py;          // its interface pressure
  R tcx, tcxl;   // conduction flux at x faces (l, l+1) and (l-1, l)
  R vx[3], vxl[3];  // viscous x-face terms mu*tau_xx, mu*tau_xy, mu*(tau.u) at (l,l+1), (l-1,l)
  unsigned neg_rho, neg_prs, nan_cnt;
  double inv;
};

// XG: gravity, the well-balanced wall flux or wall conduction modes are on; without it
// their code is compiled out.
// NW_: rows per chunk = waves per workgroup (a chunk_rows_supported value).
// DIVF_: how the update phases divide by dx and dy (FP64): DIVF_RUNTIME picks the exact-multiply
// or the general form per call from kp (emulation, batch kernels, FP32 form); the single-grid
// FP64 kernels are built once with DIVF_EXACT (both cell sizes powers of two) and once with
// DIVF_GENERAL, so each holds only the form it runs.
enum : int { DIVF_RUNTIME = 0, DIVF_EXACT = 1, DIVF_GENERAL = 2 };
template <typename R, typename S, int RECON, int RIEM, bool TC, bool VISC, bool XG, int NW_,
          int DIVF_ = DIVF_RUNTIME>
struct StripStage {
  static constexpr int NW = NW_;
  static constexpr int DIVF = DIVF_;
  static constexpr int NT = 64 * NW;
  static constexpr int HX = STRIP_W + 2;  // loaded columns: 64 lanes + 2 right halo columns
  // Ring rows. While chunk jc runs (rows jc .. jc+NW-1): primitives of rows jc-1 .. jc+NW+1,
  // slopes / fluxes of rows jc-1 .. jc+NW, viscous terms of faces jc-1 .. jc+NW-1.
  static constexpr int NQ = NW + 3;  // primitive / temperature ring rows
  static constexpr int NY = NW + 2;  // y-slope / y-flux ring rows
  static constexpr int NV = NW + 1;  // viscous y-face ring rows
  static constexpr bool kTC = TC, kVISC = VISC;
  // Waves per SIMD asked of the compiler (its VGPR budget). FP64: 4 (at most 128 VGPRs; the
  // 39 KB LDS image allows 4 workgroups of 4 waves per CU). FP32: 5 (96 VGPRs); the
  // single-grid kernel of the classes without conduction, viscosity and the gravity / wall
  // terms fits 6 (80 VGPRs), where the batch kernel (work-list state) and the other classes
  // would spill. Each measured faster than one wave fewer.
  static constexpr int kWPE = sizeof(R) == 8 ? 4 : ((TC || VISC || XG) ? 5 : 6);
  static constexpr int kWPEBatch = sizeof(R) == 8 ? 4 : 5;
#ifndef FV2D_PREFETCH_FP64
#define FV2D_PREFETCH_FP64 0
#endif
  // Register prefetch of the next chunk's rows (FP64: only when it fits the VGPR budget).
  static constexpr bool kPF = sizeof(S) == 4 || FV2D_PREFETCH_FP64;
  // the halo columns too (FP32: 4 more registers; FP64 loads them in chunk_commit)
  static constexpr bool kPFX = kPF && sizeof(S) == 4;
  typedef LaneRegs<R, S> Regs;

  struct Smem {
    R q[4][NQ][HX];
    R t[TC ? NQ : 1][TC ? HX : 1];
    // Slot of row j holds the y slope of row j until pass 2 of row j, then the flux at the
    // top face of row j (written by the same wave, right after it reads the slope).
    R yr[4][NY][STRIP_W];
    // Viscous terms mu*tau_xy, mu*tau_yy, mu*(tau.u) at the top face of row j.
    R vy[VISC ? 3 : 1][VISC ? NV : 1][STRIP_W];
    R pw[STRIP_W];  // interface pressure at the top of row Ny-2 (well-balanced wall flux)
  };

  // Ring slots. Rows are domain-relative (ghost rows negative / >= Ny); jb = segment start.
  static FV2D_HD int qslot(int j, int jb) { return (j - jb + 2) % NQ; }
  static FV2D_HD int yslot(int j, int jb) { return (j - jb + 1) % NY; }
  static FV2D_HD int vslot(int j, int jb) { return (j - jb + 1) % NV; }

  // Storage values of domain-relative cell (gi, gj), ghosts included (mapped to their source
  // cell; fx / fy: the normal velocity changes sign). UNI: gj is the same for the whole wave.
  template <bool UNI>
  static FV2D_HD void fetch_raw(int gi, int gj, const KParams &kp, const StageIO<S> &io, S raw[4], bool &fx,
                                bool &fy) {
    fx = false;
    fy = false;
    const int si = (gi >= 0 && gi < kp.Nx) ? gi : map_ghost(gi, kp.Nx, kp.bcx, fx);
    const int sj = (gj >= 0 && gj < kp.Ny) ? gj : map_ghost(gj, kp.Ny, kp.bcy, fy);
    const S *base = (io.qsrc == QSRC_Q0) ? io.q0 : io.src;
    if (UNI) {
      const S *row = base + (long long)sj * kp.pitch;
#pragma unroll
      for (int v = 0; v < 4; ++v) raw[v] = URow<S>(row + (long long)v * kp.plane)[si];
    } else {
      const long long k = (long long)sj * kp.pitch + si;
#pragma unroll
      for (int v = 0; v < 4; ++v) raw[v] = base[v * kp.plane + k];
    }
  }

  // Primitives from storage values: Q0 as given, else consToPrim (+ reset); with conduction
  // also T = p / rho (ThermalConduction.h:56-60), which shares consToPrim's divisor rho
  // unless the reset changed it.
  static FV2D_HD void raw_to_q(const S raw[4], bool fx, bool fy, const KParams &kp, const StageIO<S> &io,
                               St<R> &q, R &t) {
    if (io.qsrc == QSRC_Q0) {
#pragma unroll
      for (int v = 0; v < 4; ++v) q.v[v] = R(raw[v]);
      if (TC) t = Ops<R>::div(q.v[IP], q.v[IR]);
    } else {
      St<R> u;
#pragma unroll
      for (int v = 0; v < 4; ++v) u.v[v] = R(raw[v]);
      const DivBy<R> byr(u.v[IR], u.v[IU]);
      q = cons_to_prim(u, R(kp.gm1), byr);
      if (io.qsrc == QSRC_U_RESET) reset_negative(q, R(kp.eps_reset));
      if (TC) {
        if (q.v[IR] == u.v[IR])
          t = byr(q.v[IP]);
        else
          t = Ops<R>::div(q.v[IP], q.v[IR]);
      }
    }
    if (fx) q.v[IU] = -q.v[IU];
    if (fy) q.v[IV] = -q.v[IV];
  }

  // Primitive state (and T) of domain-relative cell (gi, gj), ghosts included.
  template <bool UNI>
  static FV2D_HD void fetch_q(int gi, int gj, const KParams &kp, const StageIO<S> &io, St<R> &q, R &t) {
    S raw[4];
    bool fx, fy;
    fetch_raw<UNI>(gi, gj, kp, io, raw, fx, fy);
    raw_to_q(raw, fx, fy, kp, io, q, t);
  }

  static FV2D_HD void put(Smem &sm, int slot, int col, const St<R> &q, R t) {
#pragma unroll
    for (int v = 0; v < 4; ++v) sm.q[v][slot][col] = q.v[v];
    if (TC) sm.t[TC ? slot : 0][TC ? col : 0] = t;
  }

  static FV2D_HD R Q(const Smem &sm, int v, int j, int col, int jb) { return sm.q[v][qslot(j, jb)][col]; }
  static FV2D_HD R T(const Smem &sm, int j, int col, int jb) {
    return sm.t[TC ? qslot(j, jb) : 0][TC ? col : 0];
  }

  // Load one strip row j into the ring: lane l -> strip column l; the two right halo
  // columns 64, 65 of `nextra` rows starting at j go to lanes 0 .. 2*nextra-1 of wave `xw`
  // (nextra <= 32).
  static FV2D_HD void load_row(int lane, int w, int j, int i0, int jb, Smem &sm, const KParams &kp,
                               const StageIO<S> &io) {
    St<R> q;
    R t = R(0);
    fetch_q<true>(i0 - 2 + lane, j, kp, io, q, t);
    put(sm, qslot(j, jb), lane, q, t);
    (void)w;
  }
  static FV2D_HD void load_extra(int lane, int j0, int nrows, int i0, int jb, Smem &sm, const KParams &kp,
                                 const StageIO<S> &io) {
    if (lane < 2 * nrows) {
      const int j = j0 + (lane >> 1), col = STRIP_W + (lane & 1);
      St<R> q;
      R t = R(0);
      fetch_q<false>(i0 - 2 + col, j, kp, io, q, t);
      put(sm, qslot(j, jb), col, q, t);
    }
  }

  // Own-cell U (and U^n) of output lanes, used by update: issued with the chunk's ring rows
  // (FP32 forms, which hide the latency with the row prefetch) or after the flux phase (FP64:
  // 16 fewer registers live through that phase; measured 0.4-1% faster, and 1.2% slower for
  // FP32). In RK2 stage 2, U^n is also the output array: each cell's U^n is read by the lane
  // that later writes that cell, and by no other lane.
  static constexpr bool kLateOwn = sizeof(R) == 8;
  static FV2D_HD void load_own(int lane, int j, int i0, const KParams &kp, const StageIO<S> &io, Regs &r) {
    const int gi = i0 - 2 + lane;
    if (lane >= 2 && gi < kp.Nx && j < kp.Ny) {
      const long long k = (long long)j * kp.pitch;
#pragma unroll
      for (int v = 0; v < 4; ++v) r.u[v] = URow<S>(io.src + (k + (long long)v * kp.plane))[gi];
      if (io.avg) {
#pragma unroll
        for (int v = 0; v < 4; ++v) r.u0[v] = URow<S>(io.avg + (k + (long long)v * kp.plane))[gi];
      }
    }
  }

  static FV2D_HD R slope(R a, R b, R c) { return minmod(b - a, c - b); }

  // y slope of (col, j) from rows j-1, j, j+1 of the ring.
  static FV2D_HD void yslope(const Smem &sm, int j, int col, int jb, R s[4]) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (RECON == PLM)
        s[v] = slope(Q(sm, v, j - 1, col, jb), Q(sm, v, j, col, jb), Q(sm, v, j + 1, col, jb));
      else
        s[v] = R(0);
    }
  }

  // y interface between rows j and j+1 at column col, given both rows' y slopes; returns
  // the flux in (rho, mx, my, E) order and the interface pressure.
  static FV2D_HD void yflux(const Smem &sm, int j, int col, int jb, const R slo[4], const R shi[4],
                            const KParams &kp, St<R> &f, R &pout) {
    St<R> a, b;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const R lo = Q(sm, v, j, col, jb), hi = Q(sm, v, j + 1, col, jb);
      a.v[v] = (RECON == PLM) ? plm_face(lo, slo[v], R(1)) : lo;
      b.v[v] = (RECON == PLM) ? plm_face(hi, shi[v], R(-1)) : hi;
    }
    St<R> fs;
    riemann<RIEM>(swap_uv(a), swap_uv(b), R(kp.gamma0), R(kp.gm1), R(kp.entho), R(kp.fslp_K), R(kp.gdx_y), fs,
                  pout);
    f = swap_uv(fs);
  }

  // ---------------------------------------------------------------- segment prologue
  // Rows jb-2 .. jb+1 into the ring (wave w < 4: row jb-2+w; wave 0 also the right halo).
  static FV2D_HD void pro_load(int lane, int w, int i0, int jb, Smem &sm, const KParams &kp, const StageIO<S> &io) {
    if (w < 4) load_row(lane, w, jb - 2 + w, i0, jb, sm, kp, io);
    if (w == 0) load_extra(lane, jb - 2, 4, i0, jb, sm, kp, io);
  }
  // y slopes of rows jb-1 (wave 0) and jb (wave 1).
  static FV2D_HD void pro_slopes(int lane, int w, int jb, Smem &sm) {
    if (w < 2) {
      R s[4];
      yslope(sm, jb - 1 + w, lane, jb, s);
#pragma unroll
      for (int v = 0; v < 4; ++v) sm.yr[v][yslot(jb - 1 + w, jb)][lane] = s[v];
    }
  }
  // The bottom interface of the segment (rows jb-1, jb) and its viscous terms, by wave 0.
  static FV2D_HD void pro_flux(int lane, int w, int jb, Smem &sm, const KParams &kp) {
    if (w != 0) return;
    R slo[4], shi[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      slo[v] = sm.yr[v][yslot(jb - 1, jb)][lane];
      shi[v] = sm.yr[v][yslot(jb, jb)][lane];
    }
    St<R> f;
    R pout;
    yflux(sm, jb - 1, lane, jb, slo, shi, kp, f, pout);
#pragma unroll
    for (int v = 0; v < 4; ++v) sm.yr[v][yslot(jb - 1, jb)][lane] = f.v[v];
    if (XG && jb - 1 == kp.Ny - 2) sm.pw[lane] = pout;
    if (VISC) visc_yface(lane, jb - 1, jb, sm, kp);
  }

  // Viscous terms of the y face between rows j and j+1 at column lane (Viscosity.h:87-106,
  // side 2 of row j = side 1 of row j+1), stored in the vy ring.
  static FV2D_HD void visc_yface(int lane, int j, int jb, Smem &sm, const KParams &kp) {
    const R mu = R(kp.mu), odx = R(kp.inv_dx), ody = R(kp.inv_dy);
    const int cm = (lane > 0) ? lane - 1 : 0, cp = lane + 1;  // lane 0's face is never used
    const R ub = Q(sm, IU, j + 1, lane, jb), ua = Q(sm, IU, j, lane, jb);
    const R vb = Q(sm, IV, j + 1, lane, jb), va = Q(sm, IV, j, lane, jb);
    const R qiu = (ub + ua) * R(0.5);
    const R qiv = (vb + va) * R(0.5);
    const R dudy = ody * (ub - ua);
    const R dvdy = ody * (vb - va);
    const R dudx = R(0.25) * odx *
                   (Q(sm, IU, j + 1, cp, jb) - Q(sm, IU, j + 1, cm, jb) + Q(sm, IU, j, cp, jb) - Q(sm, IU, j, cm, jb));
    const R dvdx = R(0.25) * odx *
                   (Q(sm, IV, j + 1, cp, jb) - Q(sm, IV, j + 1, cm, jb) + Q(sm, IV, j, cp, jb) - Q(sm, IV, j, cm, jb));
    const R tau_yy = R(4.0 / 3.0) * dvdy - R(2.0 / 3.0) * dudx;
    const R tau_xy = dvdx + dudy;
    const int sl = vslot(j, jb);
    sm.vy[VISC ? 0 : 0][VISC ? sl : 0][lane] = mu * tau_xy;
    sm.vy[VISC ? 1 : 0][VISC ? sl : 0][lane] = mu * tau_yy;
    sm.vy[VISC ? 2 : 0][VISC ? sl : 0][lane] = mu * (tau_xy * qiu + tau_yy * qiv);
  }

  // ---------------------------------------------------------------- chunk phases
  // Rows jc+2 .. jc+NW+1 (wave w: jc+2+w); the right halo columns by wave `xw`.
  static FV2D_HD void chunk_load(int lane, int w, int xw, int i0, int jb, int jc, Smem &sm, const KParams &kp,
                                 const StageIO<S> &io, Regs &r) {
    load_row(lane, w, jc + 2 + w, i0, jb, sm, kp, io);
    if (w == xw) load_extra(lane, jc + 2, NW, i0, jb, sm, kp, io);
  }

  // Prefetched form of chunk_load's ring rows: the loads of chunk jn's rows jn+2+w (and of
  // its halo columns by wave xwn) are issued one chunk ahead; chunk_commit converts them.
  static FV2D_HD void chunk_prefetch(int lane, int w, int xwn, int i0, int jn, const KParams &kp,
                                     const StageIO<S> &io, RowPrefetch<S> &p) {
    bool fx, fy;
    fetch_raw<true>(i0 - 2 + lane, jn + 2 + w, kp, io, p.row, fx, fy);
    p.flags = (fx ? 1 : 0) | (fy ? 2 : 0);
    if (kPFX && w == xwn && lane < 2 * NW) {
      fetch_raw<false>(i0 - 2 + STRIP_W + (lane & 1), jn + 2 + (lane >> 1), kp, io, p.extra, fx, fy);
      p.flags |= (fx ? 4 : 0) | (fy ? 8 : 0);
    }
  }
  static FV2D_HD void chunk_commit(int lane, int w, int xw, int i0, int jb, int jc, Smem &sm, const KParams &kp,
                                   const StageIO<S> &io, const RowPrefetch<S> &p, Regs &r) {
    St<R> q;
    R t = R(0);
    raw_to_q(p.row, (p.flags & 1) != 0, (p.flags & 2) != 0, kp, io, q, t);
    put(sm, qslot(jc + 2 + w, jb), lane, q, t);
    if (w == xw) {
      if (kPFX) {
        if (lane < 2 * NW) {
          raw_to_q(p.extra, (p.flags & 4) != 0, (p.flags & 8) != 0, kp, io, q, t);
          put(sm, qslot(jc + 2 + (lane >> 1), jb), STRIP_W + (lane & 1), q, t);
        }
      } else {
        load_extra(lane, jc + 2, NW, i0, jb, sm, kp, io);
      }
    }
  }

  // Pass 1a: x slope of column lane+1 of row jc+w, and the y slope of (lane, jc+w+1).
  static FV2D_HD void pass1_slopes(int lane, int w, int jb, int jc, Smem &sm, Regs &r) {
    const int j = jc + w;
#pragma unroll
    for (int v = 0; v < 4; ++v)
      r.sx[v] = (RECON == PLM) ? slope(Q(sm, v, j, lane, jb), Q(sm, v, j, lane + 1, jb), Q(sm, v, j, lane + 2, jb))
                               : R(0);
    yslope(sm, j + 1, lane, jb, r.syu);
#pragma unroll
    for (int v = 0; v < 4; ++v) sm.yr[v][yslot(j + 1, jb)][lane] = r.syu[v];
  }

  // Pass 1b: the x interface (lane, lane+1) of row jc+w (needs sxl = slope of column lane).
  static FV2D_HD void pass1_flux(int lane, int w, int jb, int jc, const Smem &sm, const KParams &kp,
                                 Regs &r) {
    const int j = jc + w;
    St<R> a, b;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const R lo = Q(sm, v, j, lane, jb), hi = Q(sm, v, j, lane + 1, jb);
      a.v[v] = (RECON == PLM) ? plm_face(lo, r.sxl[v], R(1)) : lo;
      b.v[v] = (RECON == PLM) ? plm_face(hi, r.sx[v], R(-1)) : hi;
    }
    St<R> f;
    R pout;
    riemann<RIEM>(a, b, R(kp.gamma0), R(kp.gm1), R(kp.entho), R(kp.fslp_K), R(kp.gdx_x), f, pout);
#pragma unroll
    for (int v = 0; v < 4; ++v) r.fx[v] = f.v[v];
  }

  // Pass 2: the y interface (jc+w, jc+w+1) at column lane, and its viscous terms.
  static FV2D_HD void pass2_flux(int lane, int w, int jb, int jc, Smem &sm, const KParams &kp, Regs &r) {
    const int j = jc + w;
    R slo[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) slo[v] = sm.yr[v][yslot(j, jb)][lane];
    St<R> f;
    yflux(sm, j, lane, jb, slo, r.syu, kp, f, r.py);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      r.fy[v] = f.v[v];
      sm.yr[v][yslot(j, jb)][lane] = f.v[v];  // the slope of row j is no longer needed
    }
    if (XG && j == kp.Ny - 2) sm.pw[lane] = r.py;
    if (VISC) visc_yface(lane, j, jb, sm, kp);
  }

  // Both cell sizes are powers of two (the FP64 divisions by dx and dy are exact multiplies):
  // the update phases are then compiled with that mode fixed, which drops the mode tests of
  // every division from the chunk loop. The FP32 form multiplies in every mode.
  static FV2D_HD bool exact_dxdy(const KParams &kp) {
    if (DIVF != DIVF_RUNTIME) return DIVF == DIVF_EXACT;
    return sizeof(R) == 8 && kp.divx == DIV_EXACT && kp.divy == DIV_EXACT;
  }

  // Update part 1: conduction and viscous terms of the x face (lane, lane+1) of row jc+w.
  static FV2D_HD void update_faces(int lane, int w, int jb, int jc, const Smem &sm, const KParams &kp,
                                   Regs &r) {
    if (exact_dxdy(kp))
      update_faces_m<true>(lane, w, jb, jc, sm, kp, r);
    else
      update_faces_m<false>(lane, w, jb, jc, sm, kp, r);
  }
  template <bool EXD>
  static FV2D_HD void update_faces_m(int lane, int w, int jb, int jc, const Smem &sm, const KParams &kp,
                                     Regs &r) {
    const int j = jc + w;
    if (TC) {
      const R kc = R(kp.kappa);
      const R kf = R(0.5) * (kc + kc);
      r.tcx = Ops<R>::div_c(kf * (T(sm, j, lane + 1, jb) - T(sm, j, lane, jb)), R(kp.dx), R(kp.inv_dx),
                            EXD ? int(DIV_EXACT) : kp.divx);
    }
    if (VISC) {
      const R mu = R(kp.mu), odx = R(kp.inv_dx), ody = R(kp.inv_dy);
      const int ca = lane, cb = lane + 1;
      const R ub = Q(sm, IU, j, cb, jb), ua = Q(sm, IU, j, ca, jb);
      const R vb = Q(sm, IV, j, cb, jb), va = Q(sm, IV, j, ca, jb);
      const R qiu = (ub + ua) * R(0.5);
      const R qiv = (vb + va) * R(0.5);
      const R dudx = odx * (ub - ua);
      const R dvdx = odx * (vb - va);
      const R dudy = R(0.25) * ody *
                     (Q(sm, IU, j + 1, cb, jb) - Q(sm, IU, j - 1, cb, jb) + Q(sm, IU, j + 1, ca, jb) -
                      Q(sm, IU, j - 1, ca, jb));
      const R dvdy = R(0.25) * ody *
                     (Q(sm, IV, j + 1, cb, jb) - Q(sm, IV, j - 1, cb, jb) + Q(sm, IV, j + 1, ca, jb) -
                      Q(sm, IV, j - 1, ca, jb));
      const R tau_xx = R(4.0 / 3.0) * dudx - R(2.0 / 3.0) * dvdy;
      const R tau_xy = dvdx + dudy;
      r.vx[0] = mu * tau_xx;
      r.vx[1] = mu * tau_xy;
      r.vx[2] = mu * (tau_xx * qiu + tau_xy * qiv);
    }
  }

  // Accumulator of the update. FP64: U itself, every term added in the reference's order
  // (Update.h:158-166, States.h:49-101). FP32 form: the increment dt*dF of the step in
  // FP32, added to U once at the end (in FP32 for the FP32 state, in FP64 for the FP64 one).
  typedef typename std::conditional<sizeof(R) == 8, double, float>::type A;

  // U += ((FL - FR) * dt) / d (Update.h:158-159); FP32 form: (FL - FR) * (dt / d).
  static FV2D_HD void add_flux(A acc[4], const R FL[4], const R FR[4], double dt, double dt_d, double d, double rd,
                               int mode) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (sizeof(R) == 8)
        acc[v] = acc[v] + A(Ops<double>::div_c((double)(FL[v] - FR[v]) * dt, d, rd, mode));
      else
        acc[v] = acc[v] + A(FL[v] - FR[v]) * A(dt_d);
    }
  }

  // Gravity source (Update.h:161-166): always into the y momentum; energy from the mass
  // fluxes of the sweep.
  static FV2D_HD void add_gravity(A acc[4], R rho, R FL0, R FR0, double dt, double g) {
    acc[IV] = acc[IV] + A(dt) * A(rho) * A(g);
    acc[IE] = acc[IE] + A(dt) * A(0.5) * (A(FL0) + A(FR0)) * A(g);
  }

  // Update part 2: the cell (lane, jc+w) (needs fxl, tcxl, vxl from lane-1).
  static FV2D_HD void update_cell(int lane, int w, int i0, int jb, int jc, int je, const Smem &sm,
                                  const KParams &kp, const StageIO<S> &io, double dt, double dt_dx, double dt_dy,
                                  Regs &r) {
    if (exact_dxdy(kp))
      update_cell_m<true>(lane, w, i0, jb, jc, je, sm, kp, io, dt, dt_dx, dt_dy, r);
    else
      update_cell_m<false>(lane, w, i0, jb, jc, je, sm, kp, io, dt, dt_dx, dt_dy, r);
  }
  template <bool EXD>
  static FV2D_HD void update_cell_m(int lane, int w, int i0, int jb, int jc, int je, const Smem &sm,
                                    const KParams &kp, const StageIO<S> &io, double dt, double dt_dx,
                                    double dt_dy, Regs &r) {
    const int divx = EXD ? int(DIV_EXACT) : kp.divx, divy = EXD ? int(DIV_EXACT) : kp.divy;
    const int j = jc + w, gi = i0 - 2 + lane;
    if (lane < 2 || gi >= kp.Nx || j >= je) return;
    const R rho = Q(sm, IR, j, lane, jb);

    A acc[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] = (sizeof(R) == 8) ? A(r.u[v]) : A(0);

    double gX = 0.0, gY = 0.0;
    if (XG) {
      if (kp.grav == GRAV_CONSTANT) {
        gX = kp.gxf;
        gY = kp.gyf;
      } else if (kp.grav == GRAV_ANALYTICAL) {
        gX = gY = (double)io.grav_row[j];
      }
    }

    // x sweep (Update.h:171)
    add_flux(acc, r.fxl, r.fx, dt, dt_dx, kp.dx, kp.inv_dx, divx);
    if (XG && kp.grav != GRAV_NONE) add_gravity(acc, rho, r.fxl[IR], r.fx[IR], dt, gX);

    // y sweep (Update.h:172)
    {
      R FL[4], FR[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        FL[v] = sm.yr[v][yslot(j - 1, jb)][lane];
        FR[v] = r.fy[v];
      }
      // well-balanced wall flux (Update.h:149-156), with the pout of the other face
      if (XG && kp.wb && (j == 0 || j == kp.Ny - 1)) {
        const R g = R(gY);
        if (j == 0) {
          FL[0] = R(0);
          FL[1] = R(0);
          FL[2] = r.py - rho * g * R(kp.dy);
          FL[3] = R(0);
        } else {
          FR[0] = R(0);
          FR[1] = R(0);
          FR[2] = sm.pw[lane] + rho * g * R(kp.dy);
          FR[3] = R(0);
        }
      }
      add_flux(acc, FL, FR, dt, dt_dy, kp.dy, kp.inv_dy, divy);
      if (XG && kp.grav != GRAV_NONE) add_gravity(acc, rho, FL[IR], FR[IR], dt, gY);
    }

    // thermal conduction into E (ThermalConduction.h:42-106)
    if (TC) {
      const R kc = R(kp.kappa);
      const R kf = R(0.5) * (kc + kc);
      const R dy = R(kp.dy);
      const R Tc = T(sm, j, lane, jb);
      R FL = r.tcxl;
      R FR = r.tcx;
      const R FU = Ops<R>::div_c(kf * (Tc - T(sm, j - 1, lane, jb)), dy, R(kp.inv_dy), divy);
      const R FD = Ops<R>::div_c(kf * (T(sm, j + 1, lane, jb) - Tc), dy, R(kp.inv_dy), divy);
      // wall modes replace the x fluxes on the first / last row (ThermalConduction.h:75-103)
      if (XG && j == 0 && kp.bctc_ymin != BCTC_NONE) {
        if (kp.bctc_ymin == BCTC_FIXED_TEMPERATURE)
          FL = Ops<R>::div_c(kf * R(2) * (Tc - R(kp.bctc_ymin_value)), dy, R(kp.inv_dy), divy);
        else if (kp.bctc_ymin == BCTC_FIXED_GRADIENT)
          FL = kf * R(kp.bctc_ymin_value);
      }
      if (XG && j == kp.Ny - 1 && kp.bctc_ymax != BCTC_NONE) {
        if (kp.bctc_ymax == BCTC_FIXED_TEMPERATURE)
          FR = Ops<R>::div_c(kf * R(2) * (R(kp.bctc_ymax_value) - Tc), dy, R(kp.inv_dy), divy);
        else if (kp.bctc_ymax == BCTC_FIXED_GRADIENT)
          FR = kf * R(kp.bctc_ymax_value);
      }
      acc[IE] = acc[IE] + (A(dt_dx) * A(FR - FL) + A(dt_dy) * A(FD - FU));
    }

    // viscosity (Viscosity.h:33-117): every face term is computed once (x faces by the lane
    // to the left of the face, y faces in pass 2); per cell and direction the flux is
    // 0 + (-mu)*tau(low face) + (+mu)*tau(high face), as the reference accumulates it
    if (VISC && sizeof(R) == 4) {
      // FP32 form: the same face terms, each difference taken once (no zero-based
      // accumulation, no zero density term)
      const int lo = vslot(j - 1, jb), hi = vslot(j, jb);
      const R fu = (r.vx[0] - r.vxl[0]) + (sm.vy[0][VISC ? hi : 0][lane] - sm.vy[0][VISC ? lo : 0][lane]);
      const R fv = (r.vx[1] - r.vxl[1]) +
                   (sm.vy[VISC ? 1 : 0][VISC ? hi : 0][lane] - sm.vy[VISC ? 1 : 0][VISC ? lo : 0][lane]);
      const R fe = (r.vx[2] - r.vxl[2]) +
                   (sm.vy[VISC ? 2 : 0][VISC ? hi : 0][lane] - sm.vy[VISC ? 2 : 0][VISC ? lo : 0][lane]);
      acc[IU] = acc[IU] + A(fu) * A(dt);
      acc[IV] = acc[IV] + A(fv) * A(dt);
      acc[IE] = acc[IE] + A(fe) * A(dt);
    } else if (VISC) {
      const int lo = vslot(j - 1, jb), hi = vslot(j, jb);
      const R fxu = (R(0) + (-r.vxl[0])) + r.vx[0];
      const R fxv = (R(0) + (-r.vxl[1])) + r.vx[1];
      const R fxe = (R(0) + (-r.vxl[2])) + r.vx[2];
      const R fyu = (R(0) + (-sm.vy[0][VISC ? lo : 0][lane])) + sm.vy[0][VISC ? hi : 0][lane];
      const R fyv = (R(0) + (-sm.vy[VISC ? 1 : 0][VISC ? lo : 0][lane])) + sm.vy[VISC ? 1 : 0][VISC ? hi : 0][lane];
      const R fye = (R(0) + (-sm.vy[VISC ? 2 : 0][VISC ? lo : 0][lane])) + sm.vy[VISC ? 2 : 0][VISC ? hi : 0][lane];
      // U += dt * (vf_x + vf_y) (Viscosity.h:116); the density components are zero
      acc[IR] = acc[IR] + A(R(0) + R(0)) * A(dt);
      acc[IU] = acc[IU] + A(fxu + fyu) * A(dt);
      acc[IV] = acc[IV] + A(fxv + fyv) * A(dt);
      acc[IE] = acc[IE] + A(fxe + fye) * A(dt);
    }

    S uns[4];
    if (sizeof(R) == 8 || sizeof(S) == 8) {
      double un64[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) un64[v] = (sizeof(R) == 8) ? double(acc[v]) : double(r.u[v]) + double(acc[v]);
      // SSP-RK2 average (Update.h:214-220)
      if (io.avg) {
#pragma unroll
        for (int v = 0; v < 4; ++v) un64[v] = 0.5 * (double(r.u0[v]) + un64[v]);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) uns[v] = S(un64[v]);
    } else {
      // FP32 state: U + dt*dF and the RK2 average in FP32, each rounded once, like the storage
#pragma unroll
      for (int v = 0; v < 4; ++v) uns[v] = S(r.u[v] + S(acc[v]));
      if (io.avg) {
#pragma unroll
        for (int v = 0; v < 4; ++v) uns[v] = S(0.5f) * (r.u0[v] + uns[v]);
      }
    }

    const long long k = (long long)j * kp.pitch;
    St<R> un;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      URow<S>(io.out + (k + (long long)v * kp.plane)).store(gi, uns[v]);
      un.v[v] = R(uns[v]);
    }

    // epilogue: consToPrim + checkNegatives (SimInfo.h:576-646) and the next dt's cell term
    if (io.epilogue) {
      const DivBy<R> byr(un.v[IR], un.v[IU]);
      St<R> q = cons_to_prim(un, R(kp.gm1), byr);
      const R eps = R(kp.eps_reset);
      if (q.v[IR] < R(0)) {
        q.v[IR] = eps;
        r.neg_rho += 1;
      }
      if (q.v[IP] < R(0)) {
        q.v[IP] = eps;
        r.neg_prs += 1;
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) r.nan_cnt += (q.v[v] != q.v[v]) ? 1u : 0u;
      // speedOfSound (States.h:45-46): (p * gamma) / rho, by the same divisor unless reset
      R c2;
      if (q.v[IR] == un.v[IR])
        c2 = byr(q.v[IP] * R(kp.gamma0));
      else
        c2 = Ops<R>::div(q.v[IP] * R(kp.gamma0), q.v[IR]);
      const double inv = inv_dt_hyp_cs(q, Ops<R>::sqrt_(c2), kp, divx, divy);
      if (inv > r.inv) r.inv = inv;  // NaN never wins: the reference's fmax-based Max ignores it
    }
  }
};

}  // namespace fv2d
