"""FP64 parity with the CPU oracle at the full benchmark sizes, the time loop of
main.cpp:62-84 with the dt of ComputeDt.h:18-65 on both sides:

  * Sod 512x8 to t = 0.2, Sedov 2048^2 for 500 steps, Rayleigh-Taylor 1024x3072 for 1000
    steps, Kelvin-Helmholtz 4096^2 for 300 steps, from the same seeded initial states;
  * the same number of steps, the same dt at every step (==), the same end time;
  * every primitive field within relative L-infinity 1e-12 of the oracle's (the stated
    tolerance; bitwise equality is recorded as a test property and expected);
  * the FP32 form's error per field against the oracle's fields: relative L1 <= 1e-5 for
    density and pressure (the stated bar for density), same step count; L1 and L-infinity
    of every field recorded.

The oracle runs on the host cores (OpenMP, at most 16 threads) beside the GPU; the four
configurations take about four minutes of oracle time together."""
import os

import numpy as np
import pytest

from fv2d_amd import problems as P
from fv2d_amd.capi import FP32
from fv2d_amd.solver import Solver
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-12
FIELDS = ("rho", "u", "v", "prs")


def rel_errors(a, b):
    """Relative L1 and L-infinity of a against b, per field."""
    out = {}
    for k, f in enumerate(FIELDS):
        d = np.abs(a[k] - b[k])
        out[f] = (float(d.mean() / max(np.abs(b[k]).mean(), 1e-300)),
                  float(d.max() / max(np.abs(b[k]).max(), 1e-300)))
    return out


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name", ["sod", "sedov", "rt", "kh"])
def test_fullsize_parity_with_oracle(name, oracle_lib, gpu_lib, record_property):
    p, q0, run = P.workload(name)
    steps = run["steps"] or 0
    oracle_lib.orc_set_threads(min(os.cpu_count() or 1, 16))
    with Oracle(p, oracle_lib) as o:
        o.set_prim(q0)
        t_o, dts_o, counts_o = o.run(run["tend"], steps)
        qo = o.get_prim()
    n_o = len(dts_o)
    if steps:
        assert n_o == steps
    with Solver(p, gpu_lib) as g:
        g.set_prim(q0)
        t_g, dts_g, counts_g, n_g = g.advance(tend=run["tend"], max_steps=steps or 100000)
        qg = g.get_prim()
    assert n_g == n_o, f"{name}: {n_g} steps on the GPU, {n_o} in the oracle"
    np.testing.assert_array_equal(dts_g, dts_o, err_msg=f"{name}: dt sequence")
    assert t_g == t_o and counts_g == counts_o, (name, t_g, t_o, counts_g, counts_o)
    e = rel_errors(qg, qo)
    worst = max(v[1] for v in e.values())
    bitwise = bool(np.array_equal(qg, qo))
    record_property("fp64_rel_linf", worst)
    record_property("fp64_bitwise", bitwise)
    assert worst <= RTOL, f"{name}: rel Linf {worst:.3e} > {RTOL}"

    # the FP32 form against the oracle's fields
    with Solver(p.copy(precision=FP32), gpu_lib) as s:
        s.set_prim(q0)
        _, _, _, n32 = s.advance(tend=run["tend"], max_steps=steps or 100000)
        q32 = s.get_prim()
    e32 = rel_errors(q32, qo)
    for f in FIELDS:
        record_property(f"fp32_{f}_rel_L1", e32[f][0])
        record_property(f"fp32_{f}_rel_Linf", e32[f][1])
    assert n32 == n_o, f"{name}: FP32 form took {n32} steps, the oracle {n_o}"
    assert e32["rho"][0] <= 1e-5 and e32["prs"][0] <= 1e-5, (name, e32)
