"""The largest grid the C ABI accepts (2^30 cells: 32768^2, 4.3e9 field values, 96 GiB of
FP64 state), Kelvin-Helmholtz set up on the device: every element offset is 64-bit, so the
run conserves mass and both momenta to round-off (relative 1e-11 at 1e9 cells), has no
negative or NaN values, repeats bit for bit, and the FP32 form's sums agree with FP64's."""
import numpy as np
import pytest

from fv2d_amd import problems as P
from fv2d_amd.capi import FP32, PROB_KH
from fv2d_amd.solver import Solver

pytestmark = pytest.mark.gpu

N = 32768
STEPS = 3


def run(prec=0):
    p = P.kh_params(N=N, precision=prec)
    assert p.Nx * p.Ny == 1 << 30
    with Solver(p) as s:
        s.set_problem(PROB_KH, seed=12345, amplitude=0.01)
        s0 = s.sums()
        t, dts, counts, n = s.advance(max_steps=STEPS)
        return s0, s.sums(), dts, counts, n


def test_max_grid_conservation_determinism_fp32():
    s0, s1, dts, counts, n = run()
    assert n == STEPS and counts == (0, 0, 0) and np.all(dts > 0)
    scale = abs(s0[0])
    for k, name in ((0, "mass"), (1, "x momentum"), (2, "y momentum")):
        assert abs(s1[k] - s0[k]) <= 1e-11 * scale, f"{name}: {s0[k]!r} -> {s1[k]!r}"
    assert abs(s1[3] - s0[3]) <= 1e-11 * abs(s0[3]), "energy"
    r0, r1, rdts, _, _ = run()
    assert np.array_equal(r0, s0) and np.array_equal(r1, s1) and np.array_equal(rdts, dts)
    f0, f1, fdts, fcounts, fn = run(FP32)
    assert fn == STEPS and fcounts == (0, 0, 0)
    assert np.all(np.abs(fdts - dts) <= 1e-5 * dts)
    for k in (0, 3):
        assert abs(f1[k] - s1[k]) <= 1e-6 * abs(s1[k])
