"""The .ini surface of the host side (no GPU): inih's parsing rules, the reference's
parameter table read through single precision, unknown-key warnings (checkValidityIni),
the last.ini dump, and the initial-condition library.

The reference settings files under tests/golden/settings/ are inputs of the reference's
runs (settings/*.ini), used here as fixtures. Expected values are fixed by inih's documented
rules and the reference's defaults (SimInfo.h:356-568), not by running the reference.
"""
import math
import os

import numpy as np
import pytest

from fv2d_amd import host
from fv2d_amd import problems as P
from fv2d_amd.capi import (BC_ABSORBING, BC_PERIODIC, BC_REFLECTING, FSLP, GRAV_CONSTANT, HLL, HLLC, PCM, PCM_WB,
                           PLM, TS_EULER, TS_RK2, Params)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SETTINGS = os.path.join(ROOT, "tests", "golden", "settings")


def f32(x):
    return float(np.float32(x))


def ini(text):
    return host.read_ini(text, text=True)


# ------------------------------------------------------------------ parser rules

def test_keys_case_insensitive_values_case_sensitive():
    d = ini("[Solvers]\nRiemann_Solver = hllc\n[PHYSICS]\nProblem = Sod_X\n")
    assert d["riemann_solver"] == HLLC
    assert d["problem"] == "Sod_X"  # values keep their case


def test_comments_separators_and_whitespace():
    d = ini("; comment\n# another\n[mesh]\n  Nx : 48   ; inline comment\nNy=24;not a comment\n"
            "[run]\ntend = 0.5 # not an inline comment prefix\n")
    assert d["Nx"] == 48
    # ';' not preceded by whitespace is part of the value: strtol stops at ';'
    assert d["Ny"] == 24
    # '#' is a comment only at the start of a line; strtof stops at ' #'
    assert d["tend"] == f32(0.5)


def test_duplicate_key_is_appended_and_parsed_prefix():
    # inih joins repeated keys with "\n"; strtol reads the first line
    d = ini("[mesh]\nNx=40\nNx=80\n")
    assert d["Nx"] == 40


def test_duplicate_enum_key_is_an_error():
    with pytest.raises(host.HostError, match="bad parameter for riemann_solver"):
        host.read_ini("[solvers]\nriemann_solver=hll\nriemann_solver=hllc\n", text=True)


def test_continuation_line():
    d = ini("[physics]\nproblem = sod_x\n   continued\n")
    # the value becomes "sod_x\ncontinued": the problem is then unknown to InitFunctor
    assert d["problem"] == "sod_x\ncontinued"
    with pytest.raises(host.HostError, match="unknown problem"):
        host.init_problem("[physics]\nproblem = sod_x\n   continued\n", text=True)


def test_parse_errors_are_ignored_like_the_reference():
    d = ini("[mesh\nNx=40\njunk line\n[run]\ntend=0.3\n")
    assert d["parse_error"] == 1  # first bad line
    # "[mesh" is no section: Nx=40 sits in the empty section and is not mesh's Nx
    assert d["Nx"] == 32
    assert d["tend"] == f32(0.3)


def test_bom_and_crlf():
    d = ini("﻿[mesh]\r\nNx = 20\r\nNy=10\r\n")
    assert (d["Nx"], d["Ny"]) == (20, 10)


def test_integer_base_prefixes_and_garbage():
    d = ini("[mesh]\nNx=0x20\nNy=abc\n[misc]\nseed=017\n")
    assert d["Nx"] == 32  # strtol base 0: hex
    assert d["Ny"] == 32  # no digits: default
    assert d["seed"] == 15  # strtol base 0: octal


def test_booleans():
    for v, want in (("true", 1), ("YES", 1), ("on", 1), ("1", 1), ("false", 0), ("No", 0), ("OFF", 0), ("0", 0),
                    ("maybe", 0)):
        assert ini(f"[viscosity]\nactive={v}\n")["visc_active"] == want, v


def test_floats_go_through_single_precision():
    d = ini("[solvers]\nCFL=0.8\n[physics]\ngamma0=1.4\n[mesh]\nxmax=0.3\n")
    assert d["CFL"] == f32(0.8) != 0.8
    assert d["gamma0"] == f32(1.4)
    assert d["xmax"] == f32(0.3)
    assert d["dx"] == f32(0.3) / 32
    # defaults too: 5/3, 1e-6, 1.1 pass through float
    d = ini("")
    assert d["gamma0"] == f32(5.0 / 3.0)
    assert d["epsilon"] == f32(1e-6)
    assert d["fslp_K"] == f32(1.1)
    assert d["epsilon_reset_negative"] == f32(1e-8)


def test_reference_defaults():
    d = ini("")
    assert (d["Nx"], d["Ny"], d["Ng"]) == (32, 32, 2)
    assert (d["boundary_x"], d["boundary_y"]) == (BC_REFLECTING, BC_REFLECTING)
    assert (d["reconstruction"], d["riemann_solver"], d["time_stepping"]) == (PCM, HLLC, TS_EULER)
    assert d["problem"] == "blast" and d["seed"] == 12345 and d["log_frequency"] == 10
    assert d["tend"] == 1.0 and d["save_freq"] == f32(0.1) and d["CFL"] == f32(0.8)
    assert d["output_filename"] == "run" and d["output_path"] == "./" and d["restart_file"] == ""
    assert d["kh_y1"] == 0.5 and d["kh_y2"] == 1.5 and d["kh_uflow"] == 1.0


@pytest.mark.parametrize("key,val,frag", [
    ("[run]\nboundaries_x", "open", "bad parameter for boundaries_x"),
    ("[solvers]\nreconstruction", "weno", "bad parameter for reconstruction"),
    ("[solvers]\ntime_stepping", "rk2", "bad parameter for time_stepping"),  # the map key is "RK2"
    ("[gravity]\nmode", "radial", "bad parameter for mode"),
])
def test_bad_enum_values(key, val, frag):
    with pytest.raises(host.HostError, match=frag):
        ini(f"{key}={val}\n")


def test_enum_maps():
    d = ini("[solvers]\ntime_stepping=RK2\nreconstruction=pcm_wb\nriemann_solver=fslp\n"
            "[run]\nboundaries_x=absorbing\nboundaries_y=periodic\n[gravity]\nmode=constant\n")
    assert (d["time_stepping"], d["reconstruction"], d["riemann_solver"]) == (TS_RK2, PCM_WB, FSLP)
    assert (d["boundary_x"], d["boundary_y"], d["gravity_mode"]) == (BC_ABSORBING, BC_PERIODIC, GRAV_CONSTANT)


# ------------------------------------------------------------------ warnings and dump

def test_unknown_sections_and_keys_warn():
    d = ini("[mesh]\nNx=16\nNz=4\n[foo]\nbar=1\n[Mesh]\nNy=8\n")
    w = d["warnings"].splitlines()
    assert "WARNING: parameter `nz` in section [mesh] is unknown." in w
    assert "WARNING: section [foo] is unknown." in w
    # sections are matched as written: [Mesh] is read (keys are case-insensitive) but warned about
    assert "WARNING: section [Mesh] is unknown." in w
    assert d["Ny"] == 8


def test_settings_files_warnings():
    kh = host.read_ini(os.path.join(SETTINGS, "kelvin_helmholtz.ini"))
    w = set(kh["warnings"].splitlines())
    assert w == {f"WARNING: parameter `{k}` in section [kelvin_helmholtz] is unknown." for k in ("uflow", "z1", "z2")}
    # the reference reads uflow / y1 / y2 from a section spelled "kelvin_helmholts": defaults stay
    assert (kh["kh_uflow"], kh["kh_y1"], kh["kh_y2"]) == (1.0, 0.5, 1.5)
    assert kh["kh_amp"] == f32(0.01) and kh["kh_P0"] == 10.0 and kh["kh_rho_fac"] == 1.0
    h84 = host.read_ini(os.path.join(SETTINGS, "H84_chi1.5.ini"))
    assert "WARNING: section [H84] is unknown." in h84["warnings"]
    assert h84["h84_pert"] == f32(1e-4)  # read all the same (keys are case-insensitive)
    rt = host.read_ini(os.path.join(SETTINGS, "rayleigh_taylor.ini"))
    assert rt["warnings"].strip() == "WARNING: parameter `gravity` in section [physics] is unknown."
    for f in ("sod_x", "sod_y", "blast", "diffusion", "gresho_vortex"):
        assert host.read_ini(os.path.join(SETTINGS, f + ".ini"))["warnings"] == "", f


def test_last_ini_dump_format():
    d = host.read_ini(os.path.join(SETTINGS, "sod_x.ini"))
    lines = d["last_ini"].splitlines()
    assert lines[0] == "; Parameters used for the problem: sod_x"
    # sections present in the file only, in name order, keys lower case in name order
    sections = [ln for ln in lines if ln.startswith("[")]
    assert sections == ["[mesh]", "[physics]", "[run]", "[solvers]"]
    assert "nghosts                    = 2                    ; default " in lines
    assert "nx                         = 64                  " in lines
    assert "cfl                        = 1.000000014901e-01  " in lines
    assert "gamma0                     = 1.666666626930e+00  " in lines
    assert "problem                    = sod_x               " in lines


def test_settings_values_and_translation():
    p = host.fv2d_params(os.path.join(SETTINGS, "rayleigh_taylor.ini"))
    assert (p.Nx, p.Ny) == (50, 150)
    assert p.xmin == f32(-0.25) and p.ymax == f32(0.75)
    assert p.reconstruction == PLM and p.riemann_solver == HLLC and p.gravity_mode == GRAV_CONSTANT
    assert p.gy == f32(-0.1) and p.CFL == f32(0.1) and p.gamma0 == f32(1.4)
    h = host.fv2d_params(os.path.join(SETTINGS, "H84_chi1.5.ini"))
    assert h.tc_active == 1 and h.kappa == f32(0.0071) and h.boundary_x == BC_PERIODIC


# ------------------------------------------------------------------ initial conditions

def _params(d):
    return Params.defaults(Nx=d["Nx"], Ny=d["Ny"], xmin=d["xmin"], xmax=d["xmax"], ymin=d["ymin"], ymax=d["ymax"])


@pytest.mark.parametrize("name,ref", [("sod_x", P.sod_x_ic), ("sod_y", P.sod_y_ic), ("blast", P.blast_ic),
                                      ("diffusion", P.diffusion_ic)])
def test_ic_matches_numpy_generators(name, ref):
    path = os.path.join(SETTINGS, name + ".ini")
    q = host.init_problem(path)
    assert np.array_equal(q, ref(_params(host.read_ini(path))))


def test_ic_gresho_profile():
    path = os.path.join(SETTINGS, "gresho_vortex.ini")
    d = host.read_ini(path)
    q = host.init_problem(path)
    x, y = np.meshgrid(d["xmin"] + (np.arange(d["Nx"]) + 0.5) * d["dx"], d["ymin"] + (np.arange(d["Ny"]) + 0.5) * d["dy"])
    xm, ym = 0.5 * (d["xmin"] + d["xmax"]), 0.5 * (d["ymin"] + d["ymax"])
    r = np.hypot(x - xm, y - ym)
    speed = np.hypot(q[1], q[2])
    uphi = np.where(r < 0.2, 5 * r, np.where(r < 0.4, 2 - 5 * r, 0.0))
    np.testing.assert_allclose(speed, uphi, atol=1e-14)
    # velocity is azimuthal: u*x + v*y = 0
    np.testing.assert_allclose(q[1] * (x - xm) + q[2] * (y - ym), 0.0, atol=1e-14)
    p0 = d["gresho_density"] / (d["gamma0"] * d["gresho_Mach"] ** 2)
    # the piecewise pressure is continuous at r = 0.2 and r = 0.4 (Miczek et al. 2015)
    assert abs((p0 + 12.5 * 0.04) - (p0 + 12.5 * 0.04 + 4 * (1 - 1 + math.log(1.0)))) < 1e-12
    assert abs((p0 + 2.0 + 4 * (1 - 2 + math.log(2.0))) - (p0 - 2 + 4 * math.log(2.0))) < 1e-12
    assert np.allclose(q[3][r >= 0.4], p0 - 2 + 4 * math.log(2.0))
    assert np.all(q[0] == d["gresho_density"])


def test_ic_kelvin_helmholtz_lecoanet():
    path = os.path.join(SETTINGS, "kelvin_helmholtz.ini")
    d = host.read_ini(path)
    q = host.init_problem(path)
    y = d["ymin"] + (np.arange(d["Ny"]) + 0.5) * d["dy"]
    # far from both interfaces (y1 = 0.5, y2 = 1.5, a = 0.05): u = uflow (inside), -uflow (outside)
    inside = np.abs(y - 1.0) < 0.15
    outside = y < 0.15
    tol = 2 * (1 - math.tanh(0.35 / 0.05))
    assert np.allclose(q[1][inside], d["kh_uflow"], rtol=0, atol=tol)
    assert np.allclose(q[1][outside], -d["kh_uflow"], rtol=0, atol=tol)
    assert np.allclose(q[0][inside], 1 + d["kh_rho_fac"], rtol=0, atol=tol)
    assert np.allclose(q[0][outside], 1.0, rtol=0, atol=tol)
    assert np.all(q[3] == d["kh_P0"])
    # v is odd about x = 0.5 within a period of sin(2 pi x) ... and bounded by 2 amp
    assert np.abs(q[2]).max() <= 2 * d["kh_amp"] + 1e-15


def test_ic_rayleigh_taylor_and_stratified():
    path = os.path.join(SETTINGS, "rayleigh_taylor.ini")
    d = host.read_ini(path)
    q = host.init_problem(path)
    y = (d["ymin"] + (np.arange(d["Ny"]) + 0.5) * d["dy"])[:, None] * np.ones((1, d["Nx"]))
    assert np.array_equal(q[0], np.where(y < 0.5 * (d["ymin"] + d["ymax"]), 1.0, 2.0))
    assert np.array_equal(q[3], 2.5 + 0.1 * d["gy"] * y)
    band = (y > -1 / 3) & (y < 1 / 3)
    assert np.all(q[2][~band] == 0.0) and np.all(q[2][band] >= 0.0) and q[2].max() <= 0.01 + 1e-15
    for name, key in (("H84_chi1.5", "h84_pert"), ("C91", "c91_pert")):
        path = os.path.join(SETTINGS, name + ".ini")
        d = host.read_ini(path)
        q = host.init_problem(path)
        y = (d["ymin"] + (np.arange(d["Ny"]) + 0.5) * d["dy"])[:, None] * np.ones((1, d["Nx"]))
        if name.startswith("H84"):
            assert np.array_equal(q[0], y ** d["m1"]) and np.array_equal(q[3], y ** (d["m1"] + 1.0))
            pert = q[2]
        else:
            T = 1.0 + d["theta1"] * y
            assert np.array_equal(q[0], T ** d["m1"])
            pert = q[3] / T ** (d["m1"] + 1.0) - 1.0
        # uniform in [-pert/2, pert/2): bounds, mean, seeded reproducibility
        assert np.abs(pert).max() <= 0.5 * d[key] * (1 + 1e-12)
        assert abs(pert.mean()) < 0.05 * d[key]
        assert np.array_equal(q, host.init_problem(path))
