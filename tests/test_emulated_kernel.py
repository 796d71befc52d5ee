"""The stage kernel's phase code (fv2d_amd/csrc/fv2d_stage.h), run by the host emulation
(tests/emu), against the CPU oracle: bitwise-equal dt, reset counts and fields for every
combination of boundary types, reconstruction, Riemann solver, time stepping, conduction,
viscosity, gravity and the well-balanced wall flux, on an odd grid (partial tiles). The
sanitized build (ASan + UBSan, LDS image poisoned with NaN) checks every global and LDS
index of the same code."""
import itertools
import os

import numpy as np
import pytest

import emu as EMU
from fv2d_amd import problems as P
from fv2d_amd.capi import BC_PERIODIC, BC_REFLECTING, FP32, HLLC, PLM, Params
from oracle.oracle import Oracle

COMBOS = list(itertools.product([0, 1, 2], [0, 1, 2], [0, 1, 2], [0, 1, 2], [0, 1], [0, 1], [0, 1], [0, 1, 2], [0, 1]))


def make(n, bcx, bcy, recon, riem, ts, tc, vi, grav, wb, Nx=37, Ny=19):
    p = Params.defaults(Nx=Nx, Ny=Ny, xmin=0.1, xmax=1.3, ymin=-0.2, ymax=0.7, boundary_x=bcx, boundary_y=bcy,
                        reconstruction=recon, riemann_solver=riem, time_stepping=ts, tc_active=tc, kappa=3e-3,
                        visc_active=vi, mu=2e-3, gravity_mode=grav, gx=0.3, gy=-0.7, hot_bubble_g0=0.5,
                        well_balanced_flux_at_y_bc=wb, gamma0=1.4, CFL=0.4,
                        bctc_ymin=1 if (tc and bcy == 1) else 0, bctc_ymax=2 if (tc and bcy == 1) else 0,
                        bctc_ymin_value=1.5, bctc_ymax_value=0.3)
    return p, P.random_state(p, seed=n)


def compare(p, q0, lib, oracle_lib, steps=3, banded=0):
    o, e = Oracle(p, oracle_lib), EMU.Emu(p, lib)
    if banded:
        e.use_banded(banded)
    o.set_prim(q0)
    e.set_prim(q0)
    for s in range(steps):
        dto, _ = o.compute_dt()
        dte, _ = e.compute_dt()
        assert dte == dto, f"step {s}: dt {dte!r} != {dto!r}"
        co, ce = o.step(dto), e.step(dte)
        assert ce == co
        qo, qe = o.get_prim(), e.get_prim()
        assert np.array_equal(qo, qe, equal_nan=True), f"step {s}: max diff {np.nanmax(np.abs(qo - qe))}"


@pytest.fixture(scope="module")
def emu_lib():
    EMU.build()
    return EMU.load()


def test_all_switch_combinations_bitwise(emu_lib, oracle_lib):
    for n, c in enumerate(COMBOS):
        p, q0 = make(n, *c)
        compare(p, q0, emu_lib, oracle_lib)


@pytest.mark.parametrize("Nx,Ny", [(2, 2), (31, 9), (32, 8), (33, 7), (64, 3), (5, 70)])
def test_grid_shapes_bitwise(Nx, Ny, emu_lib, oracle_lib):
    for n, c in enumerate([(1, 1, PLM, 1, 1, 1, 1, 1, 1), (2, 0, PLM, 0, 0, 0, 1, 0, 0), (0, 2, 0, 2, 1, 1, 0, 2, 1)]):
        p, q0 = make(n, *c, Nx=Nx, Ny=Ny)
        compare(p, q0, emu_lib, oracle_lib, steps=2)


@pytest.mark.parametrize("Nx,Ny,seg", [(124, 66, 8), (125, 130, 12), (130, 7, 4), (5, 301, 0), (62, 4, 0)])
def test_segment_heights_bitwise(Nx, Ny, seg, emu_lib, oracle_lib):
    """Segments of 4, 8 and 12 rows (a last segment shorter than a chunk, several segments
    per strip), exactly one strip, a tall grid with the default segment height."""
    for n, c in enumerate([(1, 1, PLM, 1, 1, 1, 1, 1, 1), (2, 0, PLM, 0, 0, 0, 1, 0, 0)]):
        p, q0 = make(n, *c, Nx=Nx, Ny=Ny)
        compare(p.copy(seg_rows=seg), q0, emu_lib, oracle_lib, steps=2)


@pytest.mark.parametrize("nw", [8])
@pytest.mark.parametrize("Nx,Ny,seg", [(124, 66, 0), (125, 130, 12), (37, 19, 0), (130, 7, 4), (2, 2, 0),
                                       (63, 45, 20)])
def test_chunk_heights_bitwise(nw, Nx, Ny, seg, emu_lib, oracle_lib):
    """Chunks of 8 rows (workgroups of 8 waves): default segments, segments
    shorter than a chunk or not a multiple of it, grids shorter than one chunk."""
    for n, c in enumerate([(1, 1, PLM, 1, 1, 1, 1, 1, 1), (2, 0, PLM, 0, 0, 0, 1, 0, 0), (0, 2, 0, 2, 1, 1, 0, 2, 1)]):
        p, q0 = make(n, *c, Nx=Nx, Ny=Ny)
        compare(p.copy(seg_rows=seg, chunk_rows=nw), q0, emu_lib, oracle_lib, steps=2)
    if (Nx, Ny) == (124, 66):
        fp32_close(emu_lib, oracle_lib, N=48, steps=6, chunk_rows=nw)


@pytest.mark.parametrize("xr,yr", [(1.0, 1.0), (0.5, 2.0), (1.0, 0.7), (0.9, 1.0)])
def test_power_of_two_cell_sizes_bitwise(xr, yr, emu_lib, oracle_lib):
    """Cell sizes that are powers of two in both directions (the update phases are compiled
    with the divisions by dx and dy as exact multiplies) and in one direction only (the
    general form, one direction exact at run time); conduction with wall modes, viscosity
    and gravity on."""
    for n, c in enumerate([(1, 1, PLM, 1, 1, 1, 1, 1, 1), (0, 1, PLM, 2, 1, 1, 1, 2, 1), (2, 0, PLM, 0, 0, 1, 0, 0, 0)]):
        p, _ = make(n, *c, Nx=32, Ny=16)
        p = p.copy(xmin=0.0, xmax=xr, ymin=0.0, ymax=yr)
        compare(p, P.random_state(p, seed=n), emu_lib, oracle_lib, steps=3)


@pytest.mark.parametrize("Nx,Ny,seg,nw", [(124, 150, 8, 4), (63, 301, 12, 4), (130, 37, 4, 4), (37, 19, 8, 4),
                                          (124, 200, 16, 8), (5, 3, 8, 4), (70, 64, 32, 4)])
def test_banded_schedule_bitwise(Nx, Ny, seg, nw, emu_lib, oracle_lib):
    """The banded long-then-short block schedule of large single grids, forced on small ones:
    bands of zero to many chunks, long segments cut short by the tail, one-chunk and
    two-chunk short segments, padding blocks (dt partial zero), 8-row chunks."""
    for n, c in enumerate([(1, 1, PLM, 1, 1, 1, 1, 1, 1), (2, 0, PLM, 0, 0, 0, 1, 0, 0), (0, 2, 0, 2, 1, 1, 0, 2, 1)]):
        p, q0 = make(n, *c, Nx=Nx, Ny=Ny)
        compare(p.copy(chunk_rows=nw), q0, emu_lib, oracle_lib, steps=2, banded=seg)


@pytest.mark.parametrize("nw", [4, 8])
def test_reset_and_nan_counts(nw, emu_lib, oracle_lib):
    """Negative-pressure resets (a strong rarefaction) and NaN fields (a NaN patch spreading
    through the stencil) counted alike, 4- and 8-row chunks, several strips and segments."""
    p = Params.defaults(Nx=130, Ny=40, boundary_x=BC_REFLECTING, boundary_y=BC_PERIODIC, gamma0=1.4, CFL=0.9,
                        riemann_solver=HLLC, reconstruction=PLM, chunk_rows=nw, seg_rows=12)
    q0 = np.zeros((4, 40, 130))
    q0[0] = 1.0
    q0[3] = 1e-6
    q0[1, :, :65] = -0.5
    q0[1, :, 65:] = 0.5
    q0[2, 20:23, 100:103] = np.nan
    o, e = Oracle(p, oracle_lib), EMU.Emu(p, emu_lib)
    o.set_prim(q0)
    e.set_prim(q0)
    seen = [0, 0, 0]
    for n in range(4):
        dto, _ = o.compute_dt()
        dte, _ = e.compute_dt()
        assert dte == dto or (dte != dte and dto != dto)
        co, ce = o.step(dto), e.step(dte)
        assert ce == co, f"step {n}: counts {ce} != {co}"
        seen = [a + b for a, b in zip(seen, co)]
        assert np.array_equal(o.get_prim(), e.get_prim(), equal_nan=True)
    assert min(seen) > 0, f"not every kind of reset was triggered: {seen}"


def test_sanitized_bounds(oracle_lib):
    """ASan/UBSan build: every 9th combination and three strip / segment edge shapes, 2 steps;
    a bad index aborts the process, so this test runs the sanitized library in a child
    interpreter."""
    import subprocess
    import sys
    EMU.build()
    code = (
        "import sys; sys.path[:0] = %r\n"
        "import test_emulated_kernel as T, emu as EMU\n"
        "from oracle import oracle as OR\n"
        "lib = EMU.load(sanitized=True); olib = OR.load()\n"
        "for n, c in enumerate(T.COMBOS):\n"
        "    if n %% 9: continue\n"
        "    p, q0 = T.make(n, *c)\n"
        "    T.compare(p, q0, lib, olib, steps=2)\n"
        "for Nx, Ny, seg, nw in ((124, 66, 8, 4), (130, 7, 4, 4), (2, 2, 0, 4), (124, 66, 0, 8), (63, 45, 20, 8),\n"
        "                        (5, 3, 0, 8)):\n"
        "    p, q0 = T.make(0, 1, 1, 2, 1, 1, 1, 1, 1, 1, Nx=Nx, Ny=Ny)\n"
        "    T.compare(p.copy(seg_rows=seg, chunk_rows=nw), q0, lib, olib, steps=2)\n"
        "for Nx, Ny, seg, nw in ((124, 150, 8, 4), (5, 3, 8, 4), (124, 200, 16, 8)):\n"
        "    p, q0 = T.make(0, 1, 1, 2, 1, 1, 1, 1, 1, 1, Nx=Nx, Ny=Ny)\n"
        "    T.compare(p.copy(chunk_rows=nw), q0, lib, olib, steps=2, banded=seg)\n"
        "T.fp32_close(lib, olib, N=24, steps=4)\n"
        "print('ok')\n" % ([os.path.dirname(__file__), os.path.dirname(os.path.dirname(__file__)),
                           os.path.join(os.path.dirname(__file__), "emu")],))
    import subprocess as sp
    asan = sp.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    ubsan = sp.run(["g++", "-print-file-name=libubsan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=1200)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr[-3000:]


def test_fp32_form_close_to_fp64(emu_lib, oracle_lib):
    """FP32 form (emulated): KH 64^2, 20 steps, relative L1 of every field <= 1e-5."""
    fp32_close(emu_lib, oracle_lib)


def fp32_close(emu_lib, oracle_lib, N=64, steps=20, chunk_rows=0):
    p, q0, _ = P.workload("kh", N=N)
    o = Oracle(p, oracle_lib)
    e = EMU.Emu(p.copy(precision=FP32, chunk_rows=chunk_rows), emu_lib)
    o.set_prim(q0)
    e.set_prim(q0)
    for _ in range(steps):
        dto, _ = o.compute_dt()
        dte, _ = e.compute_dt()
        assert abs(dte - dto) <= 1e-5 * dto
        o.step(dto)
        e.step(dte)
    qo, qe = o.get_prim(), e.get_prim()
    for k in range(4):
        l1 = np.abs(qe[k] - qo[k]).mean() / np.abs(qo[k]).mean()
        assert l1 <= 1e-5, (k, l1)
