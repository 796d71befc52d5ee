"""Exact solution of the 1D Riemann problem for the ideal-gas Euler equations.

Independent of the oracle: the standard two-rarefaction / two-shock pressure function with
Newton iteration for the star pressure, then sampling of the self-similar solution
(E. F. Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics", ch. 4).
Used to pin the oracle's scheme (Sod convergence) and to sanity-check its Riemann fluxes.
"""
from __future__ import annotations

import math

import numpy as np


def _fk(p, rk, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1.0) * rk)
        B = (g - 1.0) / (g + 1.0) * pk
        f = (p - pk) * math.sqrt(A / (p + B))
        df = math.sqrt(A / (B + p)) * (1.0 - (p - pk) / (2.0 * (B + p)))
    else:  # rarefaction
        f = 2.0 * ck / (g - 1.0) * ((p / pk) ** ((g - 1.0) / (2.0 * g)) - 1.0)
        df = 1.0 / (rk * ck) * (p / pk) ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rL, uL, pL, rR, uR, pR, g):
    cL, cR = math.sqrt(g * pL / rL), math.sqrt(g * pR / rR)
    du = uR - uL
    # two-rarefaction guess, then Newton on f_L(p) + f_R(p) + du = 0
    p = ((cL + cR - 0.5 * (g - 1.0) * du) / (cL / pL ** ((g - 1) / (2 * g)) + cR / pR ** ((g - 1) / (2 * g)))) ** (
        2 * g / (g - 1))
    p = max(p, 1e-12)
    for _ in range(100):
        fL, dfL = _fk(p, rL, pL, cL, g)
        fR, dfR = _fk(p, rR, pR, cR, g)
        pn = p - (fL + fR + du) / (dfL + dfR)
        pn = max(pn, 1e-14)
        if abs(pn - p) / (0.5 * (pn + p)) < 1e-14:
            p = pn
            break
        p = pn
    fL, _ = _fk(p, rL, pL, cL, g)
    fR, _ = _fk(p, rR, pR, cR, g)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    return p, u


def sample(rL, uL, pL, rR, uR, pR, g, s):
    """Primitive state (rho, u, p) at similarity coordinate s = x / t."""
    ps, us = star_state(rL, uL, pL, rR, uR, pR, g)
    cL, cR = math.sqrt(g * pL / rL), math.sqrt(g * pR / rR)
    gm1, gp1 = g - 1.0, g + 1.0
    if s <= us:  # left of contact
        if ps > pL:  # left shock
            SL = uL - cL * math.sqrt(gp1 / (2 * g) * ps / pL + gm1 / (2 * g))
            if s <= SL:
                return rL, uL, pL
            r = rL * (ps / pL + gm1 / gp1) / (gm1 / gp1 * ps / pL + 1.0)
            return r, us, ps
        cs = cL * (ps / pL) ** (gm1 / (2 * g))
        if s <= uL - cL:
            return rL, uL, pL
        if s >= us - cs:
            return rL * (ps / pL) ** (1.0 / g), us, ps
        c = 2.0 / gp1 * (cL + gm1 / 2.0 * (uL - s))
        u = 2.0 / gp1 * (cL + gm1 / 2.0 * uL + s)
        r = rL * (c / cL) ** (2.0 / gm1)
        return r, u, pL * (c / cL) ** (2.0 * g / gm1)
    # right of contact
    if ps > pR:  # right shock
        SR = uR + cR * math.sqrt(gp1 / (2 * g) * ps / pR + gm1 / (2 * g))
        if s >= SR:
            return rR, uR, pR
        r = rR * (ps / pR + gm1 / gp1) / (gm1 / gp1 * ps / pR + 1.0)
        return r, us, ps
    cs = cR * (ps / pR) ** (gm1 / (2 * g))
    if s >= uR + cR:
        return rR, uR, pR
    if s <= us + cs:
        return rR * (ps / pR) ** (1.0 / g), us, ps
    c = 2.0 / gp1 * (cR - gm1 / 2.0 * (uR - s))
    u = 2.0 / gp1 * (-cR + gm1 / 2.0 * uR + s)
    r = rR * (c / cR) ** (2.0 / gm1)
    return r, u, pR * (c / cR) ** (2.0 * g / gm1)


def profile(x, t, x0, left, right, g):
    """rho, u, p arrays of the exact solution at positions x, time t, interface x0."""
    out = np.array([sample(*left, *right, g, (xi - x0) / t) for xi in np.asarray(x)])
    return out[:, 0], out[:, 1], out[:, 2]


def godunov_flux(left, right, g):
    """Exact Godunov flux (rho u, rho u^2 + p, (E + p) u) at x/t = 0, 1D."""
    r, u, p = sample(*left, *right, g, 0.0)
    E = p / (g - 1.0) + 0.5 * r * u * u
    return np.array([r * u, r * u * u + p, (E + p) * u])
