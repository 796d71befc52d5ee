"""The FP32 form (state stored in FP32, and in FP64 for every switch combination) on the GPU
against the FP64 form (which is bitwise the oracle's arithmetic)
for every boundary type, reconstruction, Riemann solver and time stepping, with conduction,
viscosity and constant gravity on, and with the gravity / wall-flux / wall-conduction
variants: relative L1 of density and pressure <= 1e-5 (the stated bar) after 10 steps from
a smooth state, the same step count, and dt within 1e-5."""
import itertools

import numpy as np
import pytest

from fv2d_amd.capi import (BC_ABSORBING, BC_PERIODIC, BC_REFLECTING, FP32, FP32_STATE64, FSLP, GRAV_ANALYTICAL, GRAV_CONSTANT, HLL,
                           HLLC, PCM, PLM, Params)
from fv2d_amd.solver import Solver

pytestmark = pytest.mark.gpu

STEPS = 10


def smooth_state(p):
    x = p.xmin + (np.arange(p.Nx) + 0.5) * (p.xmax - p.xmin) / p.Nx
    y = p.ymin + (np.arange(p.Ny) + 0.5) * (p.ymax - p.ymin) / p.Ny
    X, Y = np.meshgrid(x, y)
    q = np.empty((4, p.Ny, p.Nx))
    q[0] = 1.0 + 0.2 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    q[1] = 0.3 * np.sin(2 * np.pi * Y)
    q[2] = 0.2 * np.cos(2 * np.pi * X)
    q[3] = 1.0 + 0.1 * np.cos(2 * np.pi * (X + Y))
    return q


def run(p, q0):
    with Solver(p) as s:
        s.set_prim(q0)
        _, dts, _, n = s.advance(max_steps=STEPS)
        return s.get_prim(), dts, n


def compare(p, form=FP32):
    q0 = smooth_state(p)
    q64, dts64, n64 = run(p, q0)
    q32, dts32, n32 = run(p.copy(precision=form), q0)
    assert n32 == n64 == STEPS
    assert np.all(np.abs(dts32 - dts64) <= 1e-5 * dts64), (dts32, dts64)
    for k, name in ((0, "rho"), (3, "prs")):
        l1 = np.abs(q32[k] - q64[k]).mean() / np.abs(q64[k]).mean()
        assert l1 <= 1e-5, f"{name}: relative L1 {l1:.2e}"


COMBOS = list(itertools.product([BC_ABSORBING, BC_REFLECTING, BC_PERIODIC], [PCM, PLM], [HLL, HLLC, FSLP], [0, 1]))


@pytest.mark.parametrize("form", [FP32, FP32_STATE64])
@pytest.mark.parametrize("bc,recon,riem,ts", COMBOS)
def test_fp32_form_every_switch(bc, recon, riem, ts, form):
    compare(Params.defaults(Nx=67, Ny=41, boundary_x=bc, boundary_y=(bc + 1) % 3, reconstruction=recon,
                            riemann_solver=riem, time_stepping=ts, tc_active=1, kappa=3e-3, visc_active=1, mu=2e-3,
                            gravity_mode=GRAV_CONSTANT, gx=0.3, gy=-0.7, gamma0=1.4, CFL=0.4), form)


@pytest.mark.parametrize("grav,wb,bctc", [(GRAV_CONSTANT, 1, 0), (GRAV_ANALYTICAL, 1, 1), (GRAV_CONSTANT, 0, 2)])
def test_fp32_form_walls_and_gravity(grav, wb, bctc):
    compare(Params.defaults(Nx=48, Ny=72, boundary_x=BC_PERIODIC, boundary_y=BC_REFLECTING, reconstruction=PLM,
                            riemann_solver=HLLC, time_stepping=1, gravity_mode=grav, gx=0.0, gy=-0.1,
                            hot_bubble_g0=0.3, well_balanced_flux_at_y_bc=wb, tc_active=1, kappa=1e-3,
                            bctc_ymin=bctc, bctc_ymax=(3 - bctc) % 3, bctc_ymin_value=1.2, bctc_ymax_value=0.4,
                            gamma0=1.4, CFL=0.5))
