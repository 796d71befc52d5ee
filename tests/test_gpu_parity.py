"""FP64 parity of the HIP path with the CPU oracle, through the C ABI.

Bar: the same dt at every step and every primitive field equal to the oracle's. The
comparison asserts relative L-infinity <= 1e-12 (the stated tolerance) and records
whether the fields are bitwise equal (expected: yes, both sides keep the reference's
operation order with IEEE division / sqrt and no FMA contraction).
"""
import itertools

import numpy as np
import pytest

from fv2d_amd import problems as P
from fv2d_amd.capi import (BC_ABSORBING, BC_PERIODIC, BC_REFLECTING, FSLP, GRAV_ANALYTICAL,
                           GRAV_CONSTANT, GRAV_NONE, HLL, HLLC, PCM, PCM_WB, PLM, Params)
from fv2d_amd.solver import Solver
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def rel_linf(a, b):
    scale = np.maximum(np.abs(b).max(axis=(1, 2)), 1e-300)
    return (np.abs(a - b).max(axis=(1, 2)) / scale).max()


def run_pair(p, q0, nsteps, oracle_lib, gpu_lib, every=1):
    o = Oracle(p, oracle_lib)
    g = Solver(p, gpu_lib)
    o.set_prim(q0)
    g.set_prim(q0)
    exact = True
    for n in range(nsteps):
        dto, io = o.compute_dt()
        dtg, ig = g.compute_dt()
        assert dtg == dto, f"step {n}: dt {dtg!r} != oracle {dto!r} (inv {ig} vs {io})"
        co = o.step(dto)
        cg = g.step(dtg)
        assert cg == co, f"step {n}: reset counts {cg} != oracle {co}"
        if (n + 1) % every == 0 or n == nsteps - 1:
            qo, qg = o.get_prim(), g.get_prim()
            err = rel_linf(qg, qo)
            assert err <= RTOL, f"step {n}: rel Linf {err:.3e} > {RTOL}"
            exact &= bool(np.array_equal(qg, qo))
    g.close()
    o.close()
    return exact


def test_sod_full_run_bitwise(oracle_lib, gpu_lib):
    """Sod 512x8 to t=0.2 (PLM/HLLC/RK2, CFL 0.8, absorbing x, periodic y), every step."""
    p, q0, run = P.workload("sod")
    o = Oracle(p, oracle_lib)
    o.set_prim(q0)
    t_o, dts_o, _ = o.run(run["tend"])
    g = Solver(p, gpu_lib)
    g.set_prim(q0)
    exact = run_pair(p, q0, len(dts_o), oracle_lib, gpu_lib)
    assert exact, "Sod parity within 1e-12 but not bitwise"
    # the device-resident loop reproduces the step count and the dt sequence
    t_g, dts_g, counts, n = g.advance(tend=run["tend"], max_steps=10000)
    assert n == len(dts_o)
    np.testing.assert_array_equal(dts_g, dts_o)
    assert t_g == t_o
    np.testing.assert_array_equal(g.get_prim(), o.get_prim())


COMBOS = list(itertools.product([BC_ABSORBING, BC_REFLECTING, BC_PERIODIC],
                                [PCM, PLM], [HLL, HLLC, FSLP], [0, 1]))


@pytest.mark.parametrize("bc,recon,riem,ts", COMBOS)
def test_switch_combos_random_state(bc, recon, riem, ts, oracle_lib, gpu_lib):
    """Random states on an odd grid (partial tiles): every boundary, reconstruction,
    Riemann solver and time stepping, with viscosity, conduction and gravity on."""
    p = Params.defaults(Nx=67, Ny=41, xmin=0.1, xmax=1.3, ymin=-0.2, ymax=0.7, boundary_x=bc,
                        boundary_y=(bc + 1) % 3, reconstruction=recon, riemann_solver=riem,
                        time_stepping=ts, tc_active=1, kappa=3e-3, visc_active=1, mu=2e-3,
                        gravity_mode=GRAV_CONSTANT, gx=0.3, gy=-0.7, gamma0=1.4, CFL=0.4)
    q0 = P.random_state(p, seed=bc * 31 + recon * 7 + riem * 3 + ts)
    assert run_pair(p, q0, 4, oracle_lib, gpu_lib)


SHAPE_SWITCHES = [  # (bcx, bcy, recon, riem, ts, tc, visc, grav, wb)
    (BC_REFLECTING, BC_REFLECTING, PLM, HLLC, 1, 1, 1, GRAV_CONSTANT, 1),
    (BC_PERIODIC, BC_ABSORBING, PLM, HLL, 0, 0, 1, GRAV_NONE, 0),
    (BC_ABSORBING, BC_PERIODIC, PCM, FSLP, 1, 1, 0, GRAV_ANALYTICAL, 1),
]


@pytest.mark.parametrize("Nx,Ny,seg", [(2, 2, 0), (3, 5, 0), (62, 4, 0), (63, 9, 0), (124, 66, 8), (125, 130, 12),
                                       (5, 301, 0), (130, 7, 4), (300, 3, 0)])
def test_grid_shapes_bitwise(Nx, Ny, seg, oracle_lib, gpu_lib):
    """Grids at the edges of the strip and segment decomposition (62 output columns per
    strip, segments of `seg` rows, 4-row chunks): the smallest grid, one column over a strip,
    exactly one and two strips, a last segment of a few rows, tall and flat grids."""
    for n, (bcx, bcy, recon, riem, ts, tc, vi, grav, wb) in enumerate(SHAPE_SWITCHES):
        p = Params.defaults(Nx=Nx, Ny=Ny, xmin=0.1, xmax=1.3, ymin=-0.2, ymax=0.7, boundary_x=bcx, boundary_y=bcy,
                            reconstruction=recon, riemann_solver=riem, time_stepping=ts, tc_active=tc, kappa=3e-3,
                            visc_active=vi, mu=2e-3, gravity_mode=grav, gx=0.3, gy=-0.7, hot_bubble_g0=0.5,
                            well_balanced_flux_at_y_bc=wb, gamma0=1.4, CFL=0.4, seg_rows=seg)
        q0 = P.random_state(p, seed=1000 * n + Nx + Ny)
        assert run_pair(p, q0, 3, oracle_lib, gpu_lib), f"{Nx}x{Ny} seg {seg} switches {n}: not bitwise"


@pytest.mark.parametrize("grav,wb,bctc", [(GRAV_NONE, 0, 0), (GRAV_CONSTANT, 1, 0), (GRAV_ANALYTICAL, 1, 1),
                                          (GRAV_CONSTANT, 0, 2)])
def test_gravity_wb_wall_conduction(grav, wb, bctc, oracle_lib, gpu_lib):
    p = Params.defaults(Nx=48, Ny=72, boundary_x=BC_PERIODIC, boundary_y=BC_REFLECTING, reconstruction=PLM,
                        riemann_solver=HLLC, time_stepping=1, gravity_mode=grav, gx=0.0, gy=-0.1,
                        hot_bubble_g0=0.3, well_balanced_flux_at_y_bc=wb, tc_active=1, kappa=1e-3,
                        bctc_ymin=bctc, bctc_ymax=(3 - bctc) % 3, bctc_ymin_value=1.2, bctc_ymax_value=0.4,
                        gamma0=1.4, CFL=0.5)
    q0 = P.random_state(p, seed=11 + grav)
    assert run_pair(p, q0, 5, oracle_lib, gpu_lib)


def test_pcm_wb_is_pcm(oracle_lib, gpu_lib):
    """PCM_WB falls through to PCM in the reference (Update.h:27-33)."""
    p = Params.defaults(Nx=40, Ny=24, reconstruction=PCM_WB, gravity_mode=GRAV_CONSTANT, gy=-1.0, gamma0=1.4)
    q0 = P.random_state(p, seed=5)
    assert run_pair(p, q0, 3, oracle_lib, gpu_lib)


@pytest.mark.parametrize("name,size,steps", [("sedov", dict(N=128), 20), ("kh", dict(N=128), 20),
                                            ("rt", dict(Nx=64, Ny=192), 20)])
def test_workloads_small(name, size, steps, oracle_lib, gpu_lib):
    p, q0, _ = P.workload(name, **size)
    assert run_pair(p, q0, steps, oracle_lib, gpu_lib, every=5)


def test_negative_reset_counts(oracle_lib, gpu_lib):
    """A strong rarefaction drives pressure negative: both sides clamp and count alike."""
    p = Params.defaults(Nx=64, Ny=8, boundary_x=BC_REFLECTING, boundary_y=BC_PERIODIC, gamma0=1.4, CFL=0.9,
                        riemann_solver=HLLC, reconstruction=PLM)
    q0 = np.zeros((4, 8, 64))
    q0[0] = 1.0
    q0[3] = 1e-6
    q0[1, :, :32] = -20.0
    q0[1, :, 32:] = 20.0
    o = Oracle(p, oracle_lib)
    g = Solver(p, gpu_lib)
    o.set_prim(q0)
    g.set_prim(q0)
    seen = 0
    for n in range(6):
        dto, _ = o.compute_dt()
        dtg, _ = g.compute_dt()
        assert dtg == dto
        co = o.step(dto)
        cg = g.step(dtg)
        assert cg == co
        seen += sum(co)
        np.testing.assert_array_equal(g.get_prim(), o.get_prim())
    assert seen > 0, "the case did not trigger the reset"


def test_closed_form_setups_on_gpu(gpu_lib):
    """The closed-form setups that pin the oracle's FSLP, wall-flux and wall-conduction
    code, run through the HIP path: a lake at rest with FSLP and the well-balanced wall
    flux stays bit for bit at rest, and the wall conduction modes change E on the wall rows
    by exactly -(dt/dx) F_wall / +(dt/dx) F_wall (numbers exact in binary, derivations in
    tests/oracle_pins.py: pin_well_balanced_column, pin_bctc_wall_modes)."""
    from fv2d_amd.capi import BCTC_FIXED_GRADIENT, BCTC_FIXED_TEMPERATURE, TS_EULER

    ps = [2.0, 1.875, 1.75, 1.625]
    q = np.stack([np.ones((4, 2)), np.zeros((4, 2)), np.zeros((4, 2)), np.array(ps)[:, None] * np.ones((4, 2))])
    p = Params.defaults(Nx=2, Ny=4, xmax=1.0, ymax=1.0, boundary_x=BC_PERIODIC, boundary_y=BC_REFLECTING,
                        riemann_solver=FSLP, reconstruction=PCM, time_stepping=TS_EULER, gamma0=1.4,
                        gravity_mode=GRAV_CONSTANT, gx=0.0, gy=-0.5, well_balanced_flux_at_y_bc=1)
    with Solver(p, gpu_lib) as g:
        g.set_prim(q)
        U0 = g.get_cons()
        g.step(2.0 ** -10)
        assert np.all(g.get_cons() - U0 == 0.0)
    dt = 2.0 ** -10
    for mode, vmin, vmax, d0, d1 in ((BCTC_FIXED_TEMPERATURE, 1.5, 3.0, -2.0 ** -7, 2.0 ** -6),
                                     (BCTC_FIXED_GRADIENT, 4.0, -8.0, -2.0 ** -7, -2.0 ** -6)):
        p = Params.defaults(Nx=8, Ny=4, xmax=1.0, ymax=1.0, boundary_x=BC_PERIODIC, boundary_y=BC_REFLECTING,
                            tc_active=1, kappa=0.25, bctc_ymin=mode, bctc_ymax=mode, bctc_ymin_value=vmin,
                            bctc_ymax_value=vmax, gamma0=1.4, time_stepping=TS_EULER, reconstruction=PLM)
        q = np.stack([np.ones((4, 8)), np.zeros((4, 8)), np.zeros((4, 8)), np.full((4, 8), 2.0)])
        with Solver(p, gpu_lib) as g:
            g.set_prim(q)
            U0 = g.get_cons()
            g.step(dt)
            dU = g.get_cons() - U0
        assert np.all(dU[:3] == 0.0) and np.all(dU[3, 1:3] == 0.0), mode
        assert np.all(dU[3, 0] == d0) and np.all(dU[3, 3] == d1), (mode, dU[3, 0], dU[3, 3])


def test_hot_bubble_rows_on_gpu(oracle_lib, gpu_lib):
    """Analytical (hot-bubble) gravity on a grid with ymin != 0, against the oracle's
    closed-form-pinned gravity (pin_hot_bubble_gravity), through a whole RK2 run."""
    p = Params.defaults(Nx=20, Ny=36, xmin=0.0, xmax=2.0, ymin=-1.0, ymax=2.0, boundary_x=BC_PERIODIC,
                        boundary_y=BC_REFLECTING, reconstruction=PLM, riemann_solver=HLLC, time_stepping=1,
                        gravity_mode=GRAV_ANALYTICAL, hot_bubble_g0=0.3, well_balanced_flux_at_y_bc=1,
                        gamma0=1.4, CFL=0.5)
    q0 = P.random_state(p, seed=77)
    assert run_pair(p, q0, 6, oracle_lib, gpu_lib)


@pytest.mark.parametrize("nw", [8])
@pytest.mark.parametrize("Nx,Ny,seg", [(124, 66, 0), (125, 130, 12), (37, 19, 0), (130, 7, 4), (2, 2, 0),
                                       (63, 45, 20)])
def test_chunk_heights_bitwise(nw, Nx, Ny, seg, oracle_lib, gpu_lib):
    """Workgroups of 8 waves (8-row chunks): default segments, segments
    shorter than a chunk or not a multiple of it, grids shorter than one chunk."""
    for n, (bcx, bcy, recon, riem, ts, tc, vi, grav, wb) in enumerate(SHAPE_SWITCHES):
        p = Params.defaults(Nx=Nx, Ny=Ny, xmin=0.1, xmax=1.3, ymin=-0.2, ymax=0.7, boundary_x=bcx, boundary_y=bcy,
                            reconstruction=recon, riemann_solver=riem, time_stepping=ts, tc_active=tc, kappa=3e-3,
                            visc_active=vi, mu=2e-3, gravity_mode=grav, gx=0.3, gy=-0.7, hot_bubble_g0=0.5,
                            well_balanced_flux_at_y_bc=wb, gamma0=1.4, CFL=0.4, seg_rows=seg, chunk_rows=nw)
        q0 = P.random_state(p, seed=2000 * n + Nx + Ny + nw)
        assert run_pair(p, q0, 3, oracle_lib, gpu_lib), f"{Nx}x{Ny} seg {seg} chunk {nw} switches {n}: not bitwise"


@pytest.mark.parametrize("nw", [8])
def test_chunk_heights_workloads(nw, oracle_lib, gpu_lib):
    """The KH and RT workloads at reduced size with 8-row chunks, every fifth step."""
    for name, size in (("kh", dict(N=128)), ("rt", dict(Nx=64, Ny=192))):
        p, q0, _ = P.workload(name, **size)
        assert run_pair(p.copy(chunk_rows=nw), q0, 10, oracle_lib, gpu_lib, every=5)


@pytest.mark.parametrize("Nx,Ny,wall", [(4096, 1100, 0), (2100, 2049, 1)])
def test_banded_schedule_general_division(Nx, Ny, wall, oracle_lib, gpu_lib):
    """Grids large enough for the banded long-then-short block schedule, with cell sizes that
    are not powers of two (the general division path of the update phases): PLM / HLLC /
    RK2 with conduction and viscosity, periodic or reflecting walls with gravity, the
    well-balanced wall flux and both wall conduction modes; 3 steps, every step compared."""
    p = Params.defaults(Nx=Nx, Ny=Ny, xmin=0.1, xmax=1.3, ymin=-0.2, ymax=0.7, gamma0=1.4, CFL=0.4,
                        reconstruction=PLM, riemann_solver=HLLC, tc_active=1, kappa=3e-4, visc_active=1,
                        mu=2e-4, boundary_x=BC_PERIODIC, boundary_y=BC_REFLECTING if wall else BC_PERIODIC,
                        gravity_mode=GRAV_CONSTANT if wall else GRAV_NONE, gy=-0.7,
                        well_balanced_flux_at_y_bc=wall, bctc_ymin=1 if wall else 0, bctc_ymax=2 if wall else 0,
                        bctc_ymin_value=1.5, bctc_ymax_value=0.3)
    assert run_pair(p, P.random_state(p, seed=Nx + Ny), 3, oracle_lib, gpu_lib)


@pytest.mark.parametrize("Nx,Ny,recon,riem,ts", [(2048, 1100, PCM_WB, HLL, 0), (1100, 2300, PLM, FSLP, 1)])
def test_banded_schedule_analytical_gravity(Nx, Ny, recon, riem, ts, oracle_lib, gpu_lib):
    """Banded-schedule grids with the branches the test above leaves out: analytical
    (hot-bubble) gravity on a domain with ymin != 0, the well-balanced PCM reconstruction
    or the FSLP solver, HLL, absorbing x walls, Euler or RK2 stepping, and the reversed wall
    conduction modes; 3 steps, every step compared."""
    p = Params.defaults(Nx=Nx, Ny=Ny, xmin=0.0, xmax=2.0, ymin=-1.0, ymax=2.0, gamma0=1.4, CFL=0.4,
                        reconstruction=recon, riemann_solver=riem, time_stepping=ts, tc_active=1, kappa=3e-4,
                        visc_active=1, mu=2e-4, boundary_x=BC_ABSORBING, boundary_y=BC_REFLECTING,
                        gravity_mode=GRAV_ANALYTICAL, hot_bubble_g0=0.3, well_balanced_flux_at_y_bc=1,
                        bctc_ymin=2, bctc_ymax=1, bctc_ymin_value=0.3, bctc_ymax_value=1.5)
    assert run_pair(p, P.random_state(p, seed=Nx + 3 * Ny), 3, oracle_lib, gpu_lib)


def extreme_state(p, seed):
    """A smooth random state with patches whose values leave the exact ranges of the fast
    division and square-root sequences: cold gas (p = 1e-250: squared sound speeds below
    2^-767), p = 0, tiny densities (1e-300), subnormal velocities, and very large density
    and pressure (1e150)."""
    q = P.random_state(p, seed=seed)
    Ny, Nx = q.shape[1:]
    w = max(Nx // 8, 2)
    patches = [(1.0, 0.1, 0.0, 1e-250), (1.0, 0.0, 0.0, 0.0), (1e-300, 1.0, 1e-10, 1e-300),
               (1.0, 1e-315, 1e-320, 1.0), (1e150, 1e-70, 0.0, 1e150)]
    for k, vals in enumerate(patches):
        i0 = (k * Nx) // len(patches)
        for v, x in enumerate(vals):
            q[v, Ny // 4:Ny // 4 + Ny // 2, i0:i0 + w] = x
    return q


@pytest.mark.parametrize("riem,domain", [(HLLC, 1.0), (HLLC, 1.3), (HLL, 1.0), (FSLP, 1.3)])
def test_extreme_states_take_exact_fallbacks(riem, domain, oracle_lib, gpu_lib):
    """Values outside the exact ranges of the fast FP64 sequences, inside a stage: every lane
    that leaves them must get the IEEE quotient and square root, so dt, reset and NaN counts
    and every field (NaN positions included) equal the oracle's, bitwise; power-of-two and
    general cell sizes."""
    p = Params.defaults(Nx=64, Ny=40, xmin=0.0, xmax=domain, ymin=0.0, ymax=domain * 40 / 64, gamma0=1.4,
                        CFL=0.4, reconstruction=PLM, riemann_solver=riem, tc_active=1, kappa=1e-3, visc_active=1,
                        mu=1e-3, boundary_x=BC_PERIODIC, boundary_y=BC_REFLECTING,
                        gravity_mode=GRAV_CONSTANT if riem == FSLP else GRAV_NONE, gy=-0.5)
    q0 = extreme_state(p, seed=7 + riem)
    o, g = Oracle(p, oracle_lib), Solver(p, gpu_lib)
    o.set_prim(q0)
    g.set_prim(q0)
    for n in range(3):
        dto, _ = o.compute_dt()
        dtg, _ = g.compute_dt()
        assert dtg == dto or (dtg != dtg and dto != dto), f"step {n}: dt {dtg!r} != {dto!r}"
        co, cg = o.step(dto), g.step(dtg)
        assert cg == co, f"step {n}: counts {cg} != {co}"
        qo, qg = o.get_prim(), g.get_prim()
        assert np.array_equal(qg, qo, equal_nan=True), f"step {n}: fields differ"
    g.close()
    o.close()


def test_banded_schedule_chunk_rows_8(oracle_lib, gpu_lib):
    """8-row chunks on a grid large enough for the banded schedule (KH 2048^2: long segments
    of 4 chunks, short ones of 2), 4 steps, every field compared."""
    p, q0, _ = P.workload("kh", N=2048)
    assert run_pair(p.copy(chunk_rows=8), q0, 4, oracle_lib, gpu_lib, every=2)

