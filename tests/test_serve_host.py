"""Host logic of the serving layer (no GPU): the seeded job stream, the per-setup solver
settings, the FIFO latency model of one-at-a-time serving, and the result statistics."""
import numpy as np

from fv2d_amd import serve as S
from fv2d_amd.capi import (BC_ABSORBING, BC_PERIODIC, BC_REFLECTING, FP32, HLL, HLLC, PLM, PROB_BLAST, PROB_KH,
                           PROB_SOD_X, TS_RK2)


def test_stream_is_seeded_and_in_range():
    a = S.make_stream(500, rate=100.0, seed=3)
    b = S.make_stream(500, rate=100.0, seed=3)
    assert [vars(x) for x in a] == [vars(x) for x in b]
    assert all((s.Nx, s.Ny) in S.SERVE_SIZES for s in a)
    assert {s.kind for s in a} == {PROB_SOD_X, PROB_BLAST, PROB_KH}
    assert all(1.4 <= s.gamma <= 5 / 3 and 1e-3 <= s.amplitude <= 1e-2 and s.steps == 200 for s in a)
    arr = np.array([s.arrival for s in a])
    assert np.all(np.diff(arr) >= 0)
    assert abs(np.diff(arr).mean() - 0.01) < 0.002  # mean inter-arrival 1 / rate
    # the same jobs whatever the rate; rate 0: all at t = 0
    c = S.make_stream(500, rate=0, seed=3)
    assert all(x.arrival == 0.0 for x in c)
    assert [(x.kind, x.Nx, x.Ny, x.gamma, x.seed) for x in a] == [(x.kind, x.Nx, x.Ny, x.gamma, x.seed) for x in c]


def test_with_arrivals_keeps_the_jobs():
    a = S.make_stream(50, rate=0, seed=1)
    b = S.with_arrivals(a, 10.0, seed=2)
    assert [(x.idx, x.kind, x.Nx, x.seed) for x in a] == [(x.idx, x.kind, x.Nx, x.seed) for x in b]
    assert b[0].arrival > 0 and np.all(np.diff([x.arrival for x in b]) > 0)


def test_job_params_per_setup():
    base = dict(idx=0, arrival=0.0, Nx=64, Ny=32, gamma=1.5, amplitude=5e-3, seed=1, steps=10)
    sod = S.job_params(S.JobSpec(kind=PROB_SOD_X, **base), FP32, 16)
    assert (sod.riemann_solver, sod.boundary_x, sod.boundary_y, sod.CFL) == (HLLC, BC_ABSORBING, BC_PERIODIC, 0.8)
    blast = S.job_params(S.JobSpec(kind=PROB_BLAST, **base))
    assert (blast.riemann_solver, blast.boundary_x, blast.boundary_y, blast.CFL) == (HLL, BC_REFLECTING,
                                                                                     BC_REFLECTING, 0.7)
    kh = S.job_params(S.JobSpec(kind=PROB_KH, **base))
    assert (kh.tc_active, kh.visc_active, kh.kappa, kh.mu) == (1, 1, 1e-4, 1e-4)
    for p in (sod, blast, kh):
        assert (p.reconstruction, p.time_stepping, p.gamma0) == (PLM, TS_RK2, 1.5)
        assert (p.xmin, p.xmax, p.ymin, p.ymax, p.Nx, p.Ny) == (0.0, 1.0, 0.0, 1.0, 64, 32)
    assert sod.precision == FP32 and sod.seg_rows == 16


def test_fifo_latencies():
    specs = [S.JobSpec(i, a, PROB_SOD_X, 8, 8, 1.4, 1e-3, 1, 1) for i, a in enumerate((0.0, 1.0, 1.5, 10.0))]
    service = {0: 2.0, 1: 1.0, 2: 0.5, 3: 1.0}
    # finishes: 2.0, 3.0, 3.5, 11.0
    np.testing.assert_allclose(S.fifo_latencies(specs, service), [2.0, 2.0, 2.0, 1.0])


def test_result_statistics():
    r = S.ServeResult("x", makespan=2.0, cell_steps=4e9)
    r.records = [S.JobRecord(i, 0.0, 0.0, float(i + 1), 1, 0.0, []) for i in range(100)]
    assert r.throughput == 2e9
    p = r.latency_percentiles()
    assert abs(p["p50"] - 50.5) < 1e-9 and abs(p["p99"] - 99.01) < 1e-9


def test_field_errors():
    ref = np.ones((4, 3, 3))
    q = ref.copy()
    q[0, 0, 0] = 1.1
    e = S.field_errors(q, ref)
    assert abs(e["rho"][0] - 0.1 / 9) < 1e-12 and abs(e["rho"][1] - 0.1) < 1e-12
    assert e["u"] == (0.0, 0.0)
