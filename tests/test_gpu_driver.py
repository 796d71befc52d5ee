"""The fv2d executable on the GPU: a .ini-driven run writes the snapshots of the reference's
main loop (main.cpp:62-88), bitwise equal to the CPU oracle driven by the same loop rule,
in both loop modes; restart continues from a snapshot; the FP32 form runs the same file.

The oracle side uses the host library's own parameter translation and initial state of the
.ini file, so both sides start from the same bits; the loop rule (save when t + epsilon >
next_save, at the top of an iteration, before the step) is restated below on the oracle.
"""
import os
import shutil

import numpy as np
import pytest

from fv2d_amd import host
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SETTINGS = os.path.join(ROOT, "tests", "golden", "settings")


def make_ini(tmp_path, base="sod_x", out=None, **run):
    text = open(os.path.join(SETTINGS, base + ".ini")).read()
    extra = "".join(f"{k}={v}\n" for k, v in run.items())
    text = text.replace("[run]\n", "[run]\n" + extra + f"output_path={out or tmp_path / 'out'}\n", 1)
    path = tmp_path / f"{base}.ini"
    path.write_text(text)
    return str(path)


def oracle_saves(ini, oracle_lib, q0=None, t0=0.0, ite0=0):
    """(iteration, t, q) of every snapshot the reference loop writes, and the final one."""
    d = host.read_ini(ini)
    p = host.fv2d_params(ini)
    o = Oracle(p, oracle_lib)
    o.set_prim(host.init_problem(ini) if q0 is None else q0)
    t, ite = t0, ite0
    next_save = 0.0 if q0 is None else t0 + d["save_freq"]
    saves = []
    while t + d["epsilon"] < d["tend"]:
        dt, _ = o.compute_dt()
        if t + d["epsilon"] > next_save:
            saves.append((ite, t, o.get_prim()))
            ite += 1
            next_save += d["save_freq"]
        o.step(dt)
        t += dt
    saves.append((ite, t, o.get_prim()))
    o.close()
    return saves


def check_file(path, saves):
    names = [n for n in host.h5_list(path) if n.startswith("ite_")]
    assert names == [f"ite_{k:04d}" for k, _, _ in saves]
    for k, t, q in saves:
        g = f"ite_{k:04d}"
        assert host.h5_attr(path, g, "time", "double") == t, g
        assert host.h5_attr(path, g, "iteration") == k
        assert np.array_equal(host.snapshot(path, g), q), f"{g}: snapshot differs from the oracle"


@pytest.mark.parametrize("mode", [[], ["--per-step"]])
def test_sod_run_matches_oracle(tmp_path, oracle_lib, mode):
    ini = make_ini(tmp_path, tend=0.06, save_freq=0.02)
    r = host.run(ini, *mode, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert "Time at end is" in r.stdout and r.stdout.count(" - Saving at time") == 3
    check_file(str(tmp_path / "out" / "run.h5"), oracle_saves(ini, oracle_lib))


def test_loop_modes_same_output(tmp_path):
    a = tmp_path / "a"
    b = tmp_path / "b"
    a.mkdir()
    b.mkdir()
    ia = make_ini(a, base="blast", out="out", tend=0.01, save_freq=0.004)
    ib = make_ini(b, base="blast", out="out", tend=0.01, save_freq=0.004)
    ra = host.run(ia, cwd=str(a))
    rb = host.run(ib, "--per-step", cwd=str(b))
    assert ra.returncode == 0 and rb.returncode == 0, ra.stderr + rb.stderr
    assert ra.stdout == rb.stdout
    # the first hyperbolic dt of settings/blast.ini, as printed in the reference's README.md:86
    assert "Computing dts at (t=0) : dt_hyp=0.000521995\n" in ra.stdout
    fa, fb = str(a / "out" / "run.h5"), str(b / "out" / "run.h5")
    assert host.h5_list(fa) == host.h5_list(fb)
    for g in [n for n in host.h5_list(fa) if n.startswith("ite_")]:
        assert np.array_equal(host.snapshot(fa, g), host.snapshot(fb, g))


def test_restart_continues_from_snapshot(tmp_path, oracle_lib):
    ini = make_ini(tmp_path, tend=0.04, save_freq=0.01)
    assert host.run(ini, cwd=str(tmp_path)).returncode == 0
    saved = tmp_path / "saved.h5"
    shutil.copy(tmp_path / "out" / "run.h5", saved)
    k, t, q = 2, host.h5_attr(str(saved), "ite_0002", "time", "double"), host.snapshot(str(saved), "ite_0002")
    ini2 = make_ini(tmp_path, tend=0.07, save_freq=0.01, restart_file=f"{saved}:/ite_0002")
    r = host.run(ini2, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert f"Restart at iteration {k} and time" in r.stdout
    saves = oracle_saves(ini2, oracle_lib, q0=q, t0=t, ite0=k + 1)
    check_file(str(tmp_path / "out" / "run.h5"), [(k, t, q)] + saves)


def test_fp32_form_runs_the_same_file(tmp_path, oracle_lib):
    ini = make_ini(tmp_path, base="kelvin_helmholtz", tend=0.05, save_freq=1.0)
    r = host.run(ini, "--fp32", cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    # settings/kelvin_helmholtz.ini writes one file per snapshot: t = 0, then the end
    got = host.snapshot(str(tmp_path / "out" / "run_0001.h5"))
    ref = oracle_saves(ini, oracle_lib)[-1][2]
    err = np.abs(got[0] - ref[0]).mean() / np.abs(ref[0]).mean()
    assert err <= 1e-5, f"FP32 form rel L1(rho) {err:.2e}"


def test_errors_exit_with_message(tmp_path):
    bad = tmp_path / "bad.ini"
    bad.write_text(f"[physics]\nproblem=tornado\n[run]\noutput_path={tmp_path}\n")
    r = host.run(str(bad), cwd=str(tmp_path))
    assert r.returncode == 1 and "ERROR: Error unknown problem tornado" in r.stderr
    b02 = tmp_path / "b02.ini"
    b02.write_text(f"[thermal_conduction]\nactive=true\nconductivity_mode=B02\n[run]\noutput_path={tmp_path}\n")
    r = host.run(str(b02), cwd=str(tmp_path))
    assert r.returncode == 1 and "B02" in r.stderr


def last_snapshot(out):
    """The last snapshot a run wrote: the last group of run.h5 (unique mode) or the last
    run_NNNN.h5 file (multiple mode)."""
    unique = os.path.join(out, "run.h5")
    if os.path.exists(unique):
        return host.snapshot(unique, [n for n in host.h5_list(unique) if n.startswith("ite_")][-1])
    files = sorted(f for f in os.listdir(out) if f.startswith("run_") and f.endswith(".h5"))
    return host.snapshot(os.path.join(out, files[-1]))


SETTINGS_FILES = sorted(os.path.splitext(f)[0] for f in os.listdir(SETTINGS) if f.endswith(".ini"))


@pytest.mark.parametrize("base", SETTINGS_FILES)
def test_every_reference_settings_file(tmp_path, oracle_lib, base):
    """Each of the reference's settings files (its problem, solvers, boundaries, gravity,
    conduction, viscosity as written) through the fv2d executable for ~8 steps: the final
    snapshot equals the oracle driven by the same loop, bit for bit."""
    ini0 = make_ini(tmp_path, base=base)
    o = Oracle(host.fv2d_params(ini0), oracle_lib)
    o.set_prim(host.init_problem(ini0))
    dt0, _ = o.compute_dt()
    o.close()
    tend = 8 * dt0
    ini = make_ini(tmp_path, base=base, tend=repr(tend), save_freq=repr(tend))
    r = host.run(ini, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    saves = oracle_saves(ini, oracle_lib)
    assert len(saves) >= 2
    assert np.array_equal(last_snapshot(str(tmp_path / "out")), saves[-1][2]), f"{base}: final snapshot differs"
