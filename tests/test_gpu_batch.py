"""Batched serving on the GPU: every job's result equals the same job run alone.

"Alone" is the single-simulation path (fv2d_create / fv2d_set_problem or fv2d_set_prim /
fv2d_advance), whose FP64 form is bitwise equal to the oracle (test_gpu_parity.py). The
batch must reproduce it bit for bit, in FP64 and in the FP32 form, whatever else runs in
the batch, whatever step the other jobs are at, with or without graph replay. Tolerance:
zero (bitwise), stated in each assertion.
"""
import numpy as np
import pytest

from fv2d_amd import problems as P
from fv2d_amd import serve as S
from fv2d_amd.capi import (FP32, FP64, PROB_BLAST, PROB_KH, PROB_SOD_X, TS_EULER, BatchConfig, Params,
                           Problem)
from fv2d_amd.solver import Batch, BatchBusy, Solver
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def spec(kind, nx, ny, gamma=1.5, amp=4e-3, seed=11, steps=12, idx=0):
    return S.JobSpec(idx, 0.0, kind, nx, ny, gamma, amp, seed, steps)


def alone(sp, prec, lib, seg=0, **kw):
    p = S.job_params(sp, prec, seg).copy(**kw)
    with Solver(p, lib) as s:
        s.set_problem(sp.kind, sp.seed, sp.amplitude)
        t, _, counts, n = s.advance(max_steps=sp.steps, record_dts=False)
        return s.get_prim(), s.sums(), t, n


def run_batch(jobs, prec, lib, cfg=None, stagger=0):
    """jobs: list of (params, problem, q, steps); the last `stagger` jobs are submitted after
    two launches. Returns {index: (q, sums, info)}."""
    cfg = cfg or BatchConfig.defaults(precision=prec, steps_per_launch=4, max_jobs=16, arena_bytes=1 << 30)
    out = {}
    with Batch(cfg, lib=lib) as b:
        ids = {}
        first = jobs[: len(jobs) - stagger]
        for k, (p, pr, q, steps) in enumerate(first):
            ids[b.submit(p, problem=pr, q=q, max_steps=steps)] = k
        if stagger:
            b.run(2)
            for k, (p, pr, q, steps) in enumerate(jobs[len(first):], start=len(first)):
                ids[b.submit(p, problem=pr, q=q, max_steps=steps)] = k
        for _ in range(1000):
            running = b.run(1)
            for jid in b.poll():
                out[ids[jid]] = b.collect(jid)
            if running == 0 and not b.poll():
                break
        assert len(out) == len(jobs)
    return out


MIX = [spec(PROB_SOD_X, 130, 70, 1.4, steps=9), spec(PROB_BLAST, 96, 96, 1.6, steps=14),
       spec(PROB_KH, 200, 64, 5 / 3, 7e-3, 5, steps=11), spec(PROB_KH, 64, 130, 1.45, 2e-3, 9, steps=6),
       spec(PROB_SOD_X, 256, 8, 1.55, steps=17), spec(PROB_BLAST, 62, 124, 1.5, steps=5)]


def rel_linf(a, b):
    scale = np.maximum(np.abs(b).max(axis=(1, 2)), 1e-300)
    return float((np.abs(a - b).max(axis=(1, 2)) / scale).max())


def test_mix_equals_alone_fp64(gpu_lib):
    """Six jobs of three setups and six grids, three of them admitted two launches after the
    others: every job's primitives, sums, step count and time equal its alone run through
    the single-simulation path, bitwise (FP64; the batch also uses other segment heights)."""
    jobs = [(S.job_params(sp, FP64), S.job_problem(sp), None, sp.steps) for sp in MIX]
    got = run_batch(jobs, FP64, gpu_lib, stagger=3)
    for k, sp in enumerate(MIX):
        qa, sa, ta, na = alone(sp, FP64, gpu_lib)
        q, sums, inf = got[k]
        assert inf.nsteps == na == sp.steps
        assert inf.t == ta, f"job {k}: t {inf.t!r} != alone {ta!r}"
        assert np.array_equal(q, qa), f"job {k} ({S.KIND_NAMES[sp.kind]} {sp.Nx}x{sp.Ny}): not bitwise equal alone"
        assert np.array_equal(sums, sa)


def test_mix_equals_alone_fp32(gpu_lib):
    """FP32 form: each job in the mix equals the same job alone in a batch, bitwise, and the
    single-simulation kernel within relative Linf 1e-5 (the two kernels are compiled
    separately and may contract a*b+c into FMA at different places)."""
    jobs = [(S.job_params(sp, FP32, 16), S.job_problem(sp), None, sp.steps) for sp in MIX]
    got = run_batch(jobs, FP32, gpu_lib, stagger=3)
    for k, sp in enumerate(MIX):
        one = run_batch([jobs[k]], FP32, gpu_lib)[0]
        q, sums, inf = got[k]
        assert inf.nsteps == one[2].nsteps == sp.steps and inf.t == one[2].t
        assert np.array_equal(q, one[0]), f"job {k}: mix differs from alone (FP32, bitwise)"
        qa, sa, ta, na = alone(sp, FP32, gpu_lib, seg=16)
        assert na == sp.steps
        err = rel_linf(q, qa)
        assert err <= 1e-5, f"job {k}: FP32 batch vs single kernel rel Linf {err:.2e}"


def test_single_job_batch_equals_mix(gpu_lib):
    """A job alone in a batch equals the same job in a mix (FP32 form, bitwise)."""
    sp = MIX[2]
    job = (S.job_params(sp, FP32), S.job_problem(sp), None, sp.steps)
    one = run_batch([job], FP32, gpu_lib)[0]
    others = [(S.job_params(s, FP32), S.job_problem(s), None, s.steps) for s in MIX if s is not sp]
    mixed = run_batch(others + [job], FP32, gpu_lib, stagger=1)[len(others)]
    assert np.array_equal(one[0], mixed[0]) and one[2].t == mixed[2].t


def test_graph_and_direct_launches_agree(gpu_lib):
    jobs = [(S.job_params(sp, FP32), S.job_problem(sp), None, sp.steps) for sp in MIX[:4]]
    a = run_batch(jobs, FP32, gpu_lib, BatchConfig.defaults(precision=FP32, steps_per_launch=3, use_graph=1))
    b = run_batch(jobs, FP32, gpu_lib, BatchConfig.defaults(precision=FP32, steps_per_launch=5, use_graph=0))
    for k in a:
        assert np.array_equal(a[k][0], b[k][0])


@pytest.mark.parametrize("steps", [4, 7])
def test_euler_ping_pong(steps, gpu_lib):
    """Euler jobs alternate U and Us; an odd or even step count is collected right."""
    sp = spec(PROB_BLAST, 70, 66, steps=steps)
    p = S.job_params(sp, FP64).copy(time_stepping=TS_EULER)
    qa, sa, ta, na = alone(sp, FP64, gpu_lib, time_stepping=TS_EULER)
    got = run_batch([(p, S.job_problem(sp), None, steps)], FP64, gpu_lib)[0]
    assert got[2].nsteps == steps and np.array_equal(got[0], qa) and np.array_equal(got[1], sa)


def test_tend_loop_rule(gpu_lib):
    """A job run to a time: same step count and end time as fv2d_advance (main.cpp:62-84)."""
    p, q0, run = P.workload("sod")
    with Solver(p, gpu_lib) as s:
        s.set_prim(q0)
        t, _, _, n = s.advance(tend=run["tend"], max_steps=0, record_dts=False)
        qa = s.get_prim()
    with Batch(BatchConfig.defaults(precision=FP64, steps_per_launch=16), lib=gpu_lib) as b:
        jid = b.submit(p, q=q0, tend=run["tend"])
        while b.run(1):
            pass
        q, sums, inf = b.collect(b.poll()[0])
    assert inf.nsteps == n and inf.t == t and np.array_equal(q, qa)


def test_batch_fp64_equals_oracle(oracle_lib, gpu_lib):
    """A KH job (viscosity and conduction on) in a mix: bitwise equal to the CPU oracle."""
    sp = spec(PROB_KH, 72, 40, 5 / 3, 8e-3, 3, steps=6)
    p = S.job_params(sp, FP64)
    q0 = P.kh_ic(p, seed=sp.seed, amp=sp.amplitude)
    o = Oracle(p, oracle_lib)
    o.set_prim(q0)
    for _ in range(sp.steps):
        dt, _ = o.compute_dt()
        o.step(dt)
    qo = o.get_prim()
    o.close()
    others = [(S.job_params(s, FP64), S.job_problem(s), None, s.steps) for s in MIX[:3]]
    got = run_batch(others + [(p, None, q0, sp.steps)], FP64, gpu_lib, stagger=1)[3]
    assert np.array_equal(got[0], qo), "batched FP64 job differs from the oracle"


def test_set_problem_matches_host_generators(gpu_lib):
    """Device initial conditions equal the numpy generators (Sod and blast bitwise, KH within
    1e-16 absolute: sin() may round differently by one unit)."""
    for kind, mk in ((PROB_SOD_X, lambda p: P.sod_ic(p)), (PROB_BLAST, lambda p: P.sedov_ic(p)),
                     (PROB_KH, lambda p: P.kh_ic(p, seed=77, amp=6e-3))):
        sp = spec(kind, 90, 50, 1.45, 6e-3, 77)
        p = S.job_params(sp, FP64)
        with Solver(p, gpu_lib) as s:
            s.set_problem(kind, 77, 6e-3)
            q = s.get_prim()
        ref = mk(p)
        if kind == PROB_KH:
            assert np.array_equal(q[[0, 1, 3]], ref[[0, 1, 3]])
            assert np.abs(q[2] - ref[2]).max() <= 1e-16
        else:
            assert np.array_equal(q, ref)


def test_admission_by_slots_and_arena(gpu_lib):
    sp = spec(PROB_SOD_X, 128, 128, steps=4)
    p = S.job_params(sp, FP32)
    from fv2d_amd.solver import job_bytes
    nbytes = job_bytes(p, gpu_lib)
    small = S.job_params(spec(PROB_SOD_X, 16, 8), FP32)
    cfg = BatchConfig.defaults(precision=FP32, max_jobs=3, arena_bytes=2 * nbytes + job_bytes(small, gpu_lib),
                               steps_per_launch=8)
    with Batch(cfg, lib=gpu_lib) as b:
        a = b.submit(p, problem=S.job_problem(sp), max_steps=4)
        b.submit(p, problem=S.job_problem(sp), max_steps=4)
        with pytest.raises(BatchBusy):  # arena full
            b.submit(p, problem=S.job_problem(sp), max_steps=4)
        b.submit(small, problem=S.job_problem(sp), max_steps=4)
        with pytest.raises(BatchBusy):  # slots full
            b.submit(small, problem=S.job_problem(sp), max_steps=4)
        assert b.run(1) == 0
        done = b.poll()
        assert len(done) == 3
        b.collect(a)
        b.submit(p, problem=S.job_problem(sp), max_steps=4)  # room again
        b.run(1)
        assert len(b.poll()) == 3


def test_cancel_and_reuse(gpu_lib):
    sp = spec(PROB_KH, 100, 100, steps=40)
    cfg = BatchConfig.defaults(precision=FP64, max_jobs=2, steps_per_launch=4)
    with Batch(cfg, lib=gpu_lib) as b:
        j = b.submit(S.job_params(sp, FP64), problem=S.job_problem(sp), max_steps=40)
        b.run(1)
        assert b.info(j).nsteps == 4
        b.cancel(j)
        sp2 = spec(PROB_BLAST, 80, 48, steps=6)
        j2 = b.submit(S.job_params(sp2, FP64), problem=S.job_problem(sp2), max_steps=6)
        while b.run(1):
            pass
        q, _, inf = b.collect(j2)
    qa, _, _, _ = alone(sp2, FP64, gpu_lib)
    assert inf.nsteps == 6 and np.array_equal(q, qa)


def test_serve_loop_small_stream(gpu_lib):
    """The request loop on a small stream (FP64): every job's step count, time and sums equal
    the one-at-a-time run, bitwise."""
    specs = S.make_stream(24, rate=0, seed=5, steps=10, sizes=((128, 128), (256, 64), (64, 192)))
    rb, _ = S.serve_batched(specs, FP64, BatchConfig.defaults(precision=FP64, max_jobs=8, steps_per_launch=4),
                            realtime=False, lib=gpu_lib)
    rs, _ = S.serve_sequential(specs, FP64, lib=gpu_lib)
    sb = {r.idx: (r.nsteps, r.t, r.sums) for r in rb.records}
    ss = {r.idx: (r.nsteps, r.t, r.sums) for r in rs.records}
    assert sb == ss


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_host_step_loop_equals_device_loop(prec, gpu_lib):
    """One job at a time with the host-driven loop (dt read back, one blocking step per RK2
    step) gives the device loop's step count, time, sums and primitives, bitwise."""
    specs = S.make_stream(6, rate=0, seed=11, steps=7, sizes=((128, 64), (96, 96)))
    keep = {s.idx for s in specs}
    rd, qd = S.serve_sequential(specs, prec, keep_prim=keep, lib=gpu_lib)
    rh, qh = S.serve_sequential(specs, prec, keep_prim=keep, lib=gpu_lib, loop="host")
    assert {r.idx: (r.nsteps, r.t, r.sums) for r in rd.records} == {r.idx: (r.nsteps, r.t, r.sums) for r in rh.records}
    for i in keep:
        assert np.array_equal(qd[i], qh[i]), f"job {i}"
    with pytest.raises(ValueError):
        S.serve_sequential(specs[:1], prec, lib=gpu_lib, loop="other")


def _random_jobs(n, seed, prec):
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n):
        kind = int(rng.integers(0, 3))
        nx, ny = int(rng.integers(16, 300)), int(rng.integers(8, 200))
        sp = spec(kind, nx, ny, float(rng.uniform(1.4, 5 / 3)), float(rng.uniform(1e-3, 1e-2)),
                  int(rng.integers(1, 1 << 40)), steps=int(rng.integers(3, 25)), idx=k)
        jobs.append(sp)
    return jobs


def _run_random_stream(jobs, prec, lib, seed):
    """Admit the jobs in a random order at random launch boundaries into a small batch
    (few slots: admission waits for finished jobs); returns {idx: (q, info)}."""
    rng = np.random.default_rng(seed)
    order = list(rng.permutation(len(jobs)))
    cfg = BatchConfig.defaults(precision=prec, max_jobs=7, steps_per_launch=3, arena_bytes=256 << 20)
    out, ids = {}, {}
    with Batch(cfg, lib=lib) as b:
        while order or ids:
            for _ in range(int(rng.integers(0, 4))):
                if not order:
                    break
                sp = jobs[order[0]]
                try:
                    jid = b.submit(S.job_params(sp, prec), problem=S.job_problem(sp), max_steps=sp.steps)
                except BatchBusy:
                    break
                ids[jid] = order.pop(0)
            if ids:
                b.run(1)
            for jid in b.poll():
                q, _, inf = b.collect(jid)
                out[ids.pop(jid)] = (q, inf)
    return out


def test_random_stream_fp32_each_job_equals_alone(gpu_lib):
    """40 jobs of random setup, grid (16..300 x 8..200), gamma, amplitude, seed and step
    count, admitted in random order at random launches into a 7-slot batch: every result is
    bitwise the result of the same job alone in a batch (FP32 form)."""
    jobs = _random_jobs(40, 123, FP32)
    got = _run_random_stream(jobs, FP32, gpu_lib, 9)
    assert sorted(got) == list(range(40))
    for k, sp in enumerate(jobs):
        one = run_batch([(S.job_params(sp, FP32), S.job_problem(sp), None, sp.steps)], FP32, gpu_lib)[0]
        q, inf = got[k]
        assert inf.nsteps == sp.steps and inf.t == one[2].t
        assert np.array_equal(q, one[0]), f"job {k} ({S.KIND_NAMES[sp.kind]} {sp.Nx}x{sp.Ny}) differs alone"


def test_random_stream_fp64_each_job_equals_single_path(gpu_lib):
    """The same with FP64 jobs, against the single-simulation path (bitwise)."""
    jobs = _random_jobs(16, 321, FP64)
    got = _run_random_stream(jobs, FP64, gpu_lib, 4)
    for k, sp in enumerate(jobs):
        qa, _, ta, na = alone(sp, FP64, gpu_lib)
        q, inf = got[k]
        assert inf.nsteps == na and inf.t == ta
        assert np.array_equal(q, qa), f"job {k} ({S.KIND_NAMES[sp.kind]} {sp.Nx}x{sp.Ny}) differs from the single path"


# The served stream's grid sizes (1024^2, 2048x512, 1024x256, 512^2) mixed with 128^2 jobs.
STREAM_MIX = [spec(PROB_KH, 1024, 1024, 5 / 3, 6e-3, 21, steps=40), spec(PROB_BLAST, 2048, 512, 1.45, steps=35),
              spec(PROB_SOD_X, 1024, 256, 1.6, steps=30), spec(PROB_KH, 512, 512, 1.5, 9e-3, 4, steps=45),
              spec(PROB_KH, 128, 128, 1.55, 3e-3, 8, steps=50), spec(PROB_BLAST, 128, 128, 1.62, steps=25),
              spec(PROB_SOD_X, 128, 128, 1.41, steps=33), spec(PROB_KH, 512, 512, 1.4, 2e-3, 17, steps=20)]


def test_stream_sizes_mix_equals_alone_fp32(gpu_lib):
    """Jobs at the served stream's real sizes (1024^2, 2048x512, 1024x256, 512^2) in one
    batch with 128^2 jobs, half of them admitted two launches after the others (the FP32
    form, as the stream runs): every job bitwise equal to the same job run alone through the
    single-simulation path. Tolerance: zero."""
    cfg = BatchConfig.defaults(precision=FP32, steps_per_launch=8, max_jobs=16, arena_bytes=2 << 30)
    jobs = [(S.job_params(sp, FP32), S.job_problem(sp), None, sp.steps) for sp in STREAM_MIX]
    got = run_batch(jobs, FP32, gpu_lib, cfg=cfg, stagger=4)
    for k, sp in enumerate(STREAM_MIX):
        qa, sa, ta, na = alone(sp, FP32, gpu_lib)
        q, sums, inf = got[k]
        assert inf.nsteps == na == sp.steps, (k, inf.nsteps, na)
        assert inf.t == ta, f"job {k}: t {inf.t!r} != alone {ta!r}"
        assert np.array_equal(q, qa), f"job {k} ({S.KIND_NAMES[sp.kind]} {sp.Nx}x{sp.Ny}): not bitwise equal alone"
        assert np.array_equal(sums, sa)


def test_stream_size_fp64_job_equals_oracle(oracle_lib, gpu_lib):
    """One FP64 512^2 Kelvin-Helmholtz job of the stream in a batch with FP64 jobs of the
    other stream sizes: bitwise equal to the oracle from the same initial state (the device
    initial condition read back before the run), step by step dt included."""
    sp = spec(PROB_KH, 512, 512, 1.5, 9e-3, 4, steps=30)
    p = S.job_params(sp, FP64)
    with Solver(p, gpu_lib) as s:  # the device-generated initial state, as the batch job gets it
        s.set_problem(sp.kind, sp.seed, sp.amplitude)
        q0 = s.get_prim()
    others = [spec(PROB_BLAST, 1024, 256, 1.45, steps=12), spec(PROB_SOD_X, 128, 128, 1.6, steps=40),
              spec(PROB_KH, 1024, 1024, 1.4, 2e-3, 3, steps=8)]
    jobs = [(p, S.job_problem(sp), None, sp.steps)] + [(S.job_params(o, FP64), S.job_problem(o), None, o.steps)
                                                        for o in others]
    cfg = BatchConfig.defaults(precision=FP64, steps_per_launch=8, max_jobs=8, arena_bytes=2 << 30)
    got = run_batch(jobs, FP64, gpu_lib, cfg=cfg, stagger=2)
    o = Oracle(p, oracle_lib)
    o.set_prim(q0)
    t_o, dts_o, _ = o.run(1e300, sp.steps)
    qo = o.get_prim()
    o.close()
    q, _, inf = got[0]
    assert inf.nsteps == len(dts_o) == sp.steps
    assert inf.t == t_o, (inf.t, t_o)
    assert np.array_equal(q, qo), f"stream-size FP64 job differs from the oracle: rel Linf {rel_linf(q, qo):.3e}"


def test_chunk_heights_in_one_batch(gpu_lib):
    """Jobs with 4- and 8-row chunks (two kernel classes of one setup: the autotuner's
    waves-per-workgroup candidates) in one batch, FP64 and FP32: each bitwise equal to the
    same job alone with the default tile."""
    for prec in (FP64, FP32):
        specs = [spec(PROB_KH, 200, 64, 5 / 3, 7e-3, 5, steps=11), spec(PROB_KH, 130, 90, 1.45, 2e-3, 9, steps=7),
                 spec(PROB_BLAST, 96, 96, 1.6, steps=9), spec(PROB_SOD_X, 256, 24, 1.55, steps=13)]
        tiles = [(32, 8), (0, 4), (16, 8), (24, 8)]
        jobs = [(S.job_params(sp, prec, seg, nw), S.job_problem(sp), None, sp.steps) for sp, (seg, nw) in
                zip(specs, tiles)]
        got = run_batch(jobs, prec, gpu_lib, stagger=2)
        for k, sp in enumerate(specs):
            qa, sa, ta, na = alone(sp, prec, gpu_lib)
            q, sums, inf = got[k]
            assert inf.nsteps == na == sp.steps and inf.t == ta, (prec, k)
            assert np.array_equal(q, qa), f"{prec} job {k} tile {tiles[k]}: not bitwise equal alone"
