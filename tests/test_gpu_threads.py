"""Distinct handles are independent (include/fv2d_amd.h): two host threads, each driving its
own simulation handle, and a third driving a batch, all at once on one GPU, give bitwise the
results of the same work run alone, one after the other."""
import threading

import numpy as np
import pytest

from fv2d_amd import problems as P
from fv2d_amd import serve as S
from fv2d_amd.capi import FP32, FP64, PROB_BLAST, PROB_KH, BatchConfig
from fv2d_amd.solver import Batch, Solver

pytestmark = pytest.mark.gpu


def single(name, size, steps, prec):
    p, q0, _ = P.workload(name, **size)
    with Solver(p.copy(precision=prec)) as s:
        s.set_prim(q0)
        _, dts, _, n = s.advance(max_steps=steps)
        # a host-driven stretch too: dt read back, one step per call
        for _ in range(5):
            dt, _ = s.compute_dt()
            s.step(dt)
        return s.get_prim(), dts, n


def batched():
    specs = [S.JobSpec(k, 0.0, kind, n, n, 1.4 + 0.05 * k, 5e-3, 11 + k, 12)
             for k, (kind, n) in enumerate([(PROB_KH, 128), (PROB_BLAST, 96), (PROB_KH, 64)])]
    out = {}
    with Batch(BatchConfig.defaults(precision=FP32, max_jobs=4, steps_per_launch=4, arena_bytes=256 << 20)) as b:
        ids = {b.submit(S.job_params(sp, FP32), problem=S.job_problem(sp), max_steps=sp.steps): sp.idx
               for sp in specs}
        while ids:
            b.run(1)
            for jid in b.poll():
                q, _, inf = b.collect(jid)
                out[ids.pop(jid)] = (q, inf.nsteps, inf.t)
    return out


WORK = [("sod", dict(Nx=512, Ny=8), 60, FP64), ("kh", dict(N=256), 30, FP64), ("sedov", dict(N=256), 40, FP32)]


def test_concurrent_handles_match_sequential():
    alone = [single(*w) for w in WORK]
    alone_b = batched()

    results = [None] * (len(WORK) + 1)
    errors = []

    def work(i):
        try:
            results[i] = single(*WORK[i]) if i < len(WORK) else batched()
        except Exception as e:  # reported below, in the main thread
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(WORK) + 1)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not errors, errors
    for (qa, dta, na), (qc, dtc, nc) in zip(alone, results[:len(WORK)]):
        assert na == nc and np.array_equal(dta, dtc) and np.array_equal(qa, qc)
    got_b = results[-1]
    assert sorted(got_b) == sorted(alone_b)
    for k in alone_b:
        assert alone_b[k][1:] == got_b[k][1:] and np.array_equal(alone_b[k][0], got_b[k][0]), f"job {k}"
