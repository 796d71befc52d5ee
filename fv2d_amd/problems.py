"""Initial conditions (host side, numpy) and the solver settings of each workload.

Every generator returns primitives in the snapshot order of the reference's IOManager
(IOManager.h:155-175): an array of shape (4, Ny, Nx) holding [rho, u, v, prs], j-outer,
i-inner, domain cells only. Cell centres follow getPos (SimInfo.h:495-499):
x_i = xmin + (i + 0.5) * dx.

Workloads (sizes as named):
  sod_x 512x8      -- Sod tube along x, PLM / HLLC / RK2, CFL 0.8, to t = 0.2
  sedov 2048^2     -- point blast E = 1 in r < 0.02, PLM / HLL / RK2, CFL 0.7, 500 steps
  kh 4096^2        -- shear band with viscosity and conduction, PLM / HLLC / RK2, 300 steps
  rt 1024x3072     -- Rayleigh-Taylor, constant gravity, PLM / HLLC / RK2, 1000 steps
The reference's own problems (Init.h:20-265) live in the C++ host; the ones needed by
tests are restated here too (sod_x / sod_y / blast / diffusion).

Randomness is counter based (SplitMix64 keyed by seed and cell index), so any cell's
noise is reproducible on its own.
"""
from __future__ import annotations

import math

import numpy as np

from .capi import (BC_ABSORBING, BC_PERIODIC, BC_REFLECTING, GRAV_CONSTANT, HLL, HLLC, PLM,
                   TS_RK2, Params)

_GOLD = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on uint64 arrays."""
    with np.errstate(over="ignore"):
        z = (x + _GOLD).astype(np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, idx: np.ndarray) -> np.ndarray:
    """U[0,1) from the top 53 bits of splitmix64(seed * golden + idx)."""
    with np.errstate(over="ignore"):
        key = np.uint64(seed) * _GOLD + idx.astype(np.uint64)
    z = splitmix64(key)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def centres(p: Params):
    dx = (p.xmax - p.xmin) / p.Nx
    dy = (p.ymax - p.ymin) / p.Ny
    x = p.xmin + (np.arange(p.Nx) + 0.5) * dx
    y = p.ymin + (np.arange(p.Ny) + 0.5) * dy
    return np.meshgrid(x, y)  # each (Ny, Nx)


# ------------------------------------------------------------------ workloads

def sod_params(Nx=512, Ny=8, **kw) -> Params:
    """Sod tube along x on [0,1]x[0,0.02] (absorbing x, periodic y)."""
    base = dict(Nx=Nx, Ny=Ny, xmin=0.0, xmax=1.0, ymin=0.0, ymax=0.02 * Ny / 8.0, gamma0=1.4,
                reconstruction=PLM, riemann_solver=HLLC, time_stepping=TS_RK2, CFL=0.8,
                boundary_x=BC_ABSORBING, boundary_y=BC_PERIODIC)
    base.update(kw)
    return Params.defaults(**base)


def sod_ic(p: Params) -> np.ndarray:
    x, _ = centres(p)
    q = np.zeros((4, p.Ny, p.Nx))
    left = x < 0.5
    q[0] = np.where(left, 1.0, 0.125)
    q[3] = np.where(left, 1.0, 0.1)
    return q


def sedov_params(N=2048, **kw) -> Params:
    base = dict(Nx=N, Ny=N, gamma0=1.4, reconstruction=PLM, riemann_solver=HLL, time_stepping=TS_RK2,
                CFL=0.7, boundary_x=BC_REFLECTING, boundary_y=BC_REFLECTING)
    base.update(kw)
    return Params.defaults(**base)


def sedov_ic(p: Params, E=1.0, r0=0.02, rho0=1.0, p0=1.0e-5) -> np.ndarray:
    """Point blast: E deposited uniformly in cells within r < r0 of the centre."""
    x, y = centres(p)
    xm = 0.5 * (p.xmin + p.xmax)
    ym = 0.5 * (p.ymin + p.ymax)
    r = np.sqrt((x - xm) ** 2 + (y - ym) ** 2)
    q = np.zeros((4, p.Ny, p.Nx))
    q[0] = rho0
    pin = (p.gamma0 - 1.0) * E / (math.pi * r0 * r0)
    q[3] = np.where(r < r0, pin, p0)
    return q


def kh_params(N=4096, **kw) -> Params:
    base = dict(Nx=N, Ny=N, gamma0=5.0 / 3.0, reconstruction=PLM, riemann_solver=HLLC, time_stepping=TS_RK2,
                CFL=0.8, boundary_x=BC_PERIODIC, boundary_y=BC_PERIODIC, visc_active=1, mu=1.0e-4,
                tc_active=1, kappa=1.0e-4)
    base.update(kw)
    return Params.defaults(**base)


def kh_ic(p: Params, seed=12345, amp=0.01, noise=1.0e-3, prs=2.5) -> np.ndarray:
    """Shear band 0.25 < y < 0.75 (rho 2, u +0.5) in rho 1, u -0.5; v = amp sin(4 pi x) + noise."""
    x, y = centres(p)
    band = (y > 0.25) & (y < 0.75)
    q = np.zeros((4, p.Ny, p.Nx))
    q[0] = np.where(band, 2.0, 1.0)
    q[1] = np.where(band, 0.5, -0.5)
    idx = np.arange(p.Nx * p.Ny, dtype=np.uint64).reshape(p.Ny, p.Nx)
    q[2] = amp * np.sin(4.0 * math.pi * x) + noise * (2.0 * uniform01(seed, idx) - 1.0)
    q[3] = prs
    return q


def rt_params(Nx=1024, Ny=3072, **kw) -> Params:
    base = dict(Nx=Nx, Ny=Ny, xmin=0.0, xmax=0.5, ymin=0.0, ymax=1.5, gamma0=1.4, reconstruction=PLM,
                riemann_solver=HLLC, time_stepping=TS_RK2, CFL=0.6, boundary_x=BC_PERIODIC,
                boundary_y=BC_REFLECTING, gravity_mode=GRAV_CONSTANT, gx=0.0, gy=-0.1,
                well_balanced_flux_at_y_bc=1)
    base.update(kw)
    return Params.defaults(**base)


def rt_ic(p: Params, y0=0.75, p0=2.5, amp=0.01) -> np.ndarray:
    """Heavy (rho 2) over light (rho 1) at y0, hydrostatic p with p0 at y0, seeded v."""
    x, y = centres(p)
    g = p.gy
    q = np.zeros((4, p.Ny, p.Nx))
    rho = np.where(y > y0, 2.0, 1.0)
    q[0] = rho
    q[3] = p0 + rho * g * (y - y0)
    pert = amp * (1.0 + np.cos(4.0 * math.pi * x)) * (1.0 + np.cos(3.0 * math.pi * (y - y0))) / 4.0
    q[2] = np.where(np.abs(y - y0) < 1.0 / 3.0, pert, 0.0)
    return q


WORKLOADS = {
    "sod": (sod_params, sod_ic, dict(tend=0.2, steps=None)),
    "sedov": (sedov_params, sedov_ic, dict(tend=1e30, steps=500)),
    "kh": (kh_params, kh_ic, dict(tend=1e30, steps=300)),
    "rt": (rt_params, rt_ic, dict(tend=1e30, steps=1000)),
}


def workload(name: str, **size):
    """(params, q0, run) for a named workload; `size` overrides the grid (e.g. N=64)."""
    mk_params, mk_ic, run = WORKLOADS[name]
    p = mk_params(**size)
    return p, mk_ic(p), dict(run)


# ------------------------------------------------------------------ reference problems (Init.h)

def blast_ic(p: Params) -> np.ndarray:
    """initBlast (Init.h:103-131)."""
    x, y = centres(p)
    xm = 0.5 * (p.xmin + p.xmax)
    ym = 0.5 * (p.ymin + p.ymax)
    xr, yr = xm - x, ym - y
    r = np.sqrt(xr * xr + yr * yr)
    q = np.zeros((4, p.Ny, p.Nx))
    inside = r < 0.2
    q[0] = np.where(inside, 1.0, 1.2)
    q[3] = np.where(inside, 10.0, 0.1)
    return q


def sod_x_ic(p: Params) -> np.ndarray:
    """initSodX (Init.h:20-35): x <= 0.5 is the high state."""
    x, _ = centres(p)
    q = np.zeros((4, p.Ny, p.Nx))
    q[0] = np.where(x <= 0.5, 1.0, 0.125)
    q[3] = np.where(x <= 0.5, 1.0, 0.1)
    return q


def sod_y_ic(p: Params) -> np.ndarray:
    """initSodY (Init.h:83-98)."""
    _, y = centres(p)
    q = np.zeros((4, p.Ny, p.Nx))
    q[0] = np.where(y <= 0.5, 1.0, 0.125)
    q[3] = np.where(y <= 0.5, 1.0, 0.1)
    return q


def diffusion_ic(p: Params) -> np.ndarray:
    """initDiffusion (Init.h:185-206)."""
    x, y = centres(p)
    xm = 0.5 * (p.xmin + p.xmax)
    ym = 0.5 * (p.ymin + p.ymax)
    r = np.sqrt((x - xm) ** 2 + (y - ym) ** 2)
    q = np.ones((4, p.Ny, p.Nx))
    q[0] = np.where(r < 0.2, 1.0, 0.1)
    return q


def random_state(p: Params, seed=7, rho=(0.5, 2.0), vel=(-0.5, 0.5), prs=(0.5, 2.0)) -> np.ndarray:
    """Smooth-free random primitives (every cell independent) for kernel parity tests."""
    idx = np.arange(p.Nx * p.Ny, dtype=np.uint64).reshape(p.Ny, p.Nx)
    n = p.Nx * p.Ny
    q = np.empty((4, p.Ny, p.Nx))
    lo_hi = [rho, vel, vel, prs]
    for k in range(4):
        u = uniform01(seed * 4 + k, idx + np.uint64(k * n))
        q[k] = lo_hi[k][0] + (lo_hi[k][1] - lo_hi[k][0]) * u
    return q
