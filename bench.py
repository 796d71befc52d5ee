#!/usr/bin/env python3
"""Benchmark of the explicit 2D hydro step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--cpu-seconds S]
                    [--serve-jobs J] [--serve-steps S] [--serve-load L] [--no-serve]

Workload: the Kelvin-Helmholtz shear layer at N x N (default 4096^2), PLM-minmod / HLLC /
SSP-RK2, CFL 0.8, periodic, viscosity mu = 1e-4 and thermal conduction kappa = 1e-4 on,
synthetic seeded initial state, all state resident in HBM. One step = one full RK2 step
(dt reduction + two fused stage launches). The FP64 parity form gives `value`
(Gcell-updates/s = Nx*Ny*K / time of K steps, HIP events on the solver stream, device
synchronised on both sides); the FP32 form is timed the same way and reported beside it.
The K steps run as one device loop (fv2d_advance), after W untimed warm-up steps; before
those, each form runs its untimed 3000-step KH drift measurement (the GPU clocks rise over
the first ~50-100 ms of sustained load, so the timed loop starts at the running clock).
`repetitions` gives the median, minimum and maximum rate of --reps further loops of K steps
each (when K >= 16). `pcie_inclusive`
adds the upload of the host primitives before and their read-back after (host clock).

roofline: the dominant kernel is the fused stage kernel (two launches per step). Its
algorithmic HBM bytes per cell are 160 B per FP64 RK2 step (stage 1 reads U^n and writes
U*, stage 2 reads U* and U^n and writes U^n+1: 5 states x 4 fields x 8 B), i.e. 80 B per
launch on average (80 B / 40 B per launch in the FP32 form, whose U is stored in FP32).
achieved = bytes per launch x cells / mean launch time (HIP events around every stage
launch in the timed region). traffic = measured HBM bytes per launch (rocprofv3
FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel calibration) from the committed summary
under profiles/, when present.

serving: the served job stream (2000 jobs of seven grid sizes from 128^2 to 2048x512,
Sod / point blast / Kelvin-Helmholtz, seeded gamma and amplitude, 200 RK2 steps each, FP32
form) through the batch engine with continuous batching, against the same jobs one at a
time through the single-simulation path: aggregate Gcell-updates/s of each (all jobs
queued at t = 0), p50/p99 latency (arrival to result) under seeded Poisson arrivals at
--serve-load x the batched capacity, and a 5% seeded subsample re-run in FP64 with the
FP32 form's error against it. A seeded 10% subsample also runs one job at a time with the
reference's loop structure (dt read back to the host, one blocking step per RK2 step).

cpu_baseline: the CPU oracle (a plain-C FP64 restatement of the reference path with the
reference's pass structure, OpenMP) on every host core it may use (OMP_NUM_THREADS, else
nproc), median of 3 bounded samples of the same workload (as many steps as fit in
--cpu-seconds each), with nproc, the CPU model, OMP_NUM_THREADS and the host's STREAM-triad
bandwidth. The FP32 form's error at every configuration is measured against the oracle's
fields at full size (the oracle runs on the host cores; a few minutes).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s & % HBM peak at 4096² FP64 vs FP32; served-job p50/p99 latency"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_STEP = {"fp64": 160.0, "fp32": 80.0, "fp32s64": 160.0}
STAGE_LAUNCHES_PER_STEP = 2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(prec: str, size: int):
    """Measured HBM bytes per stage launch from profiles/ (None when not measured)."""
    path = os.path.join(ROOT, "profiles", f"stage_kh{size}_{prec}_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def run_gpu(prec: str, size: int, steps: int, warmup: int, q0: np.ndarray, reps: int = 5, before=None):
    """One form on KH size^2: upload q0, run `before` (untimed GPU work of the bench's own, see
    below), `warmup` untimed steps, then the K = `steps` timed steps as one device loop, then
    `reps` further loops of K steps for the spread."""
    from fv2d_amd import hiprt
    from fv2d_amd import problems as P
    from fv2d_amd.capi import FP32, FP32_STATE64, FP64
    from fv2d_amd.solver import Solver

    p = P.kh_params(N=size, precision={"fp64": FP64, "fp32": FP32, "fp32s64": FP32_STATE64}[prec])
    s = Solver(p)
    hiprt.synchronize()
    t_up = time.perf_counter()
    s.set_prim(q0)  # host primitives over PCIe, then U = primToCons on the device
    hiprt.synchronize()
    t_up = time.perf_counter() - t_up
    # The GPU raises its clocks over the first ~50-100 ms of sustained load (a K = 20 loop
    # right after 5 warm-up steps reads ~7% slow in FP64, ~20% in FP32, a second loop close
    # to the steady rate; profiling on or off makes no difference). The accuracy run of this
    # form goes here, after the upload, so the warm-up and timed steps start at the running clock.
    before_out = before() if before is not None else None
    if warmup:
        s.advance(max_steps=warmup, record_dts=False)
    stream = s.lib.fv2d_get_stream(s.h)
    hiprt.synchronize()
    # per-launch events for the K steps' stage launches, created here, before the timed loop
    assert s.lib.fv2d_profile(s.h, STAGE_LAUNCHES_PER_STEP * steps) == 0
    # the K timed steps: one device loop, HIP events on the solver stream around it
    ev = [hiprt.Event() for _ in range(2)]
    ev[0].record(stream)
    t, _, counts, n = s.advance(max_steps=steps, t0=0.0, record_dts=False)
    ev[1].record(stream)
    hiprt.synchronize()
    ms = ev[0].elapsed_ms(ev[1])
    import ctypes as C

    st_ms = C.c_double()
    st_n = C.c_int64()
    s.lib.fv2d_profile_read(s.h, C.byref(st_ms), C.byref(st_n))
    s.lib.fv2d_profile(s.h, 0)
    assert n == steps, f"advance took {n} steps, expected {steps}"
    cells = p.Nx * p.Ny
    # repetitions (spread): `reps` more device loops of the same K steps, each event-timed;
    # only when a loop is long enough (>= 16 steps) for its rate to mean something
    rep_rates = []
    if steps >= 16:
        for _ in range(reps):
            ev[0].record(stream)
            t, _, _, nr = s.advance(max_steps=steps, t0=t, record_dts=False)
            ev[1].record(stream)
            hiprt.synchronize()
            rep_rates.append(cells * nr / (ev[0].elapsed_ms(ev[1]) * 1e-3) / 1e9)
    gcells = cells * steps / (ms * 1e-3) / 1e9
    mean_launch_s = st_ms.value * 1e-3 / max(st_n.value, 1)
    bpl = BYTES_PER_CELL_STEP[prec] / STAGE_LAUNCHES_PER_STEP
    achieved = bpl * cells / mean_launch_s / 1e9
    t_down = time.perf_counter()
    q = s.get_prim()  # primitives back to the host
    t_down = time.perf_counter() - t_down
    sums = s.sums()
    s.close()
    reps_out = {"n": len(rep_rates), "steps_each": steps}
    if rep_rates:
        reps_out.update(gcells_median=round(float(np.median(rep_rates)), 4), gcells_min=round(min(rep_rates), 4),
                        gcells_max=round(max(rep_rates), 4))
    return dict(before=before_out, gcells=gcells, ms_per_step=ms / steps, mean_launch_ms=mean_launch_s * 1e3, launches=int(st_n.value),
                achieved=achieved, bytes_per_launch=bpl * cells, counts=counts, q=q, sums=sums,
                pcie={"upload_ms": round(t_up * 1e3, 2), "download_ms": round(t_down * 1e3, 2),
                      "gcells": round(cells * steps / (ms * 1e-3 + t_up + t_down) / 1e9, 4)},
                blocks=reps_out)


def host_info():
    """nproc, the CPU model and OMP_NUM_THREADS of the host the baseline runs on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    return {"nproc": os.cpu_count(), "cpu_model": model, "OMP_NUM_THREADS": omp}


def run_cpu(size: int, seconds: float, q0: np.ndarray, samples: int = 3):
    """The CPU baseline: the oracle (the reference path restated in C, its pass structure,
    OpenMP) on every host core it may use (OMP_NUM_THREADS when set, else nproc), on a
    bounded sample of the KH workload: `samples` runs of as many RK2 steps as fit in
    `seconds` each (at least one step), median reported; plus the host's STREAM-triad
    bandwidth in the same process for context."""
    from fv2d_amd import problems as P
    from oracle import oracle as OR

    info = host_info()
    threads = int(info["OMP_NUM_THREADS"]) if (info["OMP_NUM_THREADS"] or "").isdigit() else (os.cpu_count() or 1)
    lib = OR.load()
    lib.orc_set_threads(threads)
    triad = lib.orc_stream_triad(1 << 26, 5)  # 3 x 512 MiB arrays
    rates, steps = [], []
    o = OR.Oracle(P.kh_params(N=size), lib)
    o.set_prim(q0)
    dt, _ = o.compute_dt()
    o.step(dt)  # warm-up: first touch of every array
    for _ in range(samples):
        n = 0
        t0 = time.perf_counter()
        while True:
            dt, _ = o.compute_dt()
            o.step(dt)
            n += 1
            el = time.perf_counter() - t0
            if el >= seconds or n >= 200:
                break
        rates.append(size * size * n / el / 1e9)
        steps.append(n)
    o.close()
    return dict(value=float(np.median(rates)), samples=[round(r, 5) for r in rates], steps=steps,
                seconds=seconds, threads=lib.orc_get_threads(), stream_triad_gbs=round(triad, 1), **info)


def run_serving(n_jobs: int, steps: int, load: float, check_frac: float, seed: int = 2024, use_graph: bool = True):
    """The served job stream: autotune, batched and one-at-a-time capacity (all jobs queued
    at t = 0), batched latency under Poisson arrivals at `load` x the batched capacity, and
    the FP64 re-run of a seeded subsample as the accuracy check of the FP32 form."""
    from fv2d_amd import serve as S
    from fv2d_amd.capi import FP32, FP64, BatchConfig

    t0 = time.perf_counter()
    tune = S.autotune(FP32, log=log, use_graph=int(use_graph))
    t_tune = time.perf_counter() - t0
    specs = S.make_stream(n_jobs, rate=0, seed=seed, steps=steps)
    rng = np.random.default_rng(seed + 1)
    check = set(int(i) for i in rng.choice(n_jobs, max(1, int(round(check_frac * n_jobs))), replace=False))
    cfg = BatchConfig.defaults(precision=FP32, max_jobs=256, use_graph=int(use_graph))
    log(f"bench: serving {n_jobs} jobs x {steps} RK2 steps, batched (all queued at t=0)")
    rb, _ = S.serve_batched(specs, FP32, cfg, tuning=tune, realtime=False)
    log(f"bench: batched {rb.throughput / 1e9:.2f} Gcell/s, makespan {rb.makespan:.2f} s")
    log("bench: serving one at a time (single-simulation path)")
    rs, _ = S.serve_sequential(specs, FP32)
    log(f"bench: one-at-a-time {rs.throughput / 1e9:.2f} Gcell/s, makespan {rs.makespan:.2f} s")
    by_size = {}
    svc = {r.idx: r.finished - r.admitted for r in rs.records}
    for sp in specs:
        k = f"{sp.Nx}x{sp.Ny}"
        c, t = by_size.get(k, (0.0, 0.0))
        by_size[k] = (c + float(sp.cells) * sp.steps, t + svc[sp.idx])
    seq_by_size = {k: round(c / t / 1e9, 3) for k, (c, t) in sorted(by_size.items(), key=lambda kv: kv[1][0])}
    # the reference's loop structure (dt to the host, one blocking step per RK2 step), one
    # job at a time, on a seeded 10% subsample of the same stream
    sub_host = sorted(int(i) for i in np.random.default_rng(seed + 3).choice(n_jobs, max(1, n_jobs // 10),
                                                                             replace=False))
    log(f"bench: {len(sub_host)} of the jobs one at a time through a host-driven step loop")
    rh, _ = S.serve_sequential([specs[i] for i in sub_host], FP32, loop="host")
    dev_sub = sum(svc[i] for i in sub_host)
    log(f"bench: host loop {rh.throughput / 1e9:.2f} Gcell/s (device loop on the same jobs "
        f"{rh.cell_steps / dev_sub / 1e9:.2f})")
    cap_jobs = n_jobs / rb.makespan
    rate = load * cap_jobs
    specs_rt = S.with_arrivals(specs, rate, seed=seed + 2)
    log(f"bench: batched under Poisson arrivals at {rate:.1f} jobs/s ({load:.2f} x capacity)")
    rr, _ = S.serve_batched(specs_rt, FP32, cfg, tuning=tune, realtime=True)
    lat_b = rr.latency_percentiles()
    service = {r.idx: r.finished - r.admitted for r in rs.records}
    lat_s = S.fifo_latencies(specs_rt, service)
    log(f"bench: FP64 re-run of {len(check)} jobs")
    sub = [s for s in specs if s.idx in check]
    # the FP32 results of the subset come from a batch of their own (a job's results do not
    # depend on the jobs it shares a batch with), so that the timed stream copies nothing back
    r32, prim32 = S.serve_batched(sub, FP32, cfg, tuning=tune, keep_prim=check, realtime=False)
    r64, prim64 = S.serve_batched(sub, FP64, BatchConfig.defaults(precision=FP64, max_jobs=64,
                                                                  use_graph=int(use_graph)),
                                  keep_prim=check, realtime=False)
    steps32 = {r.idx: r.nsteps for r in r32.records}
    steps64 = {r.idx: r.nsteps for r in r64.records}
    errs = {name: [] for name in ("rho", "u", "v", "prs")}
    for i in sorted(check):
        e = S.field_errors(prim32[i], prim64[i])
        for name in errs:
            errs[name].append(e[name])
    fp64_check = {
        "jobs": len(check),
        "same_step_count": all(steps32[i] == steps64[i] for i in check),
        "rel_L1_max": {k: float(max(v[0] for v in errs[k])) for k in errs},
        "rel_L1_mean": {k: float(np.mean([v[0] for v in errs[k]])) for k in errs},
        "rel_Linf_max": {k: float(max(v[1] for v in errs[k])) for k in errs},
    }
    return {
        "jobs": n_jobs,
        "steps_per_job": steps,
        "cell_updates": rb.cell_steps,
        "precision": "fp32 form",
        "value": round(rb.throughput / 1e9, 4),
        "unit": "Gcell-updates/s",
        "one_at_a_time": round(rs.throughput / 1e9, 4),
        "one_at_a_time_by_grid": seq_by_size,
        "one_at_a_time_host_loop": {
            "jobs": len(sub_host),
            "value": round(rh.throughput / 1e9, 4),
            "device_loop_same_jobs": round(rh.cell_steps / dev_sub / 1e9, 4),
            "speedup_of_batched": round(rb.throughput / rh.throughput, 3),
            "note": "seeded 10% subsample, each job alone with the reference's loop structure: dt read back "
                    "to the host and one blocking step per RK2 step (fv2d_compute_dt + fv2d_step)",
        },
        "speedup_vs_one_at_a_time": round(rb.throughput / rs.throughput, 3),
        "makespan_s": {"batched": round(rb.makespan, 3), "one_at_a_time": round(rs.makespan, 3)},
        "gpu_busy_frac": round(rb.gpu_ms / 1e3 / rb.makespan, 3),
        "arrival_rate_jobs_per_s": round(rate, 2),
        "load": load,
        "latency_ms": {"p50": round(lat_b["p50"] * 1e3, 2), "p99": round(lat_b["p99"] * 1e3, 2)},
        "one_at_a_time_latency_ms": {"p50": round(float(np.percentile(lat_s, 50)) * 1e3, 2),
                                     "p99": round(float(np.percentile(lat_s, 99)) * 1e3, 2),
                                     "note": "single FIFO server with the measured one-at-a-time service times "
                                             "under the same arrivals"},
        "latency_all_queued_ms": {k: round(v * 1e3, 2) for k, v in rb.latency_percentiles().items()},
        "fp64_check": fp64_check,
        "autotune": {"seconds": round(t_tune, 1), "tile_choice": tune.get("tile_choice"),
                     "blocks_per_cu": tune.get("blocks_per_cu") or "occupancy",
                     "winner_per_class": {k: {"seg_rows": v[0], "chunk_rows": v[1]} for k, v in tune["tile"].items()},
                     "candidates_gcells": {k: round(v / 1e9, 3) for k, v in tune["rates"].items()}},
    }


def errors_vs_oracle_all_configs():
    """Every benchmark configuration at full size against the CPU oracle (FP64, the
    reference path restated in C, on the host cores): the FP64 form's parity (same step
    count and dt sequence, worst relative L-infinity, bitwise equality) and the FP32 form's
    relative L1 and L-infinity error per primitive field."""
    from fv2d_amd import problems as P
    from fv2d_amd.capi import FP32, FP32_STATE64, FP64
    from fv2d_amd.solver import Solver
    from oracle import oracle as OR

    lib = OR.load()
    info = host_info()
    lib.orc_set_threads(int(info["OMP_NUM_THREADS"]) if (info["OMP_NUM_THREADS"] or "").isdigit()
                        else (os.cpu_count() or 1))
    out = {}
    for name in ("sod", "sedov", "kh", "rt"):
        p, q0, r = P.workload(name)
        steps = r["steps"] or 0
        t0 = time.perf_counter()
        with OR.Oracle(p, lib) as o:
            o.set_prim(q0)
            _, dts_o, _ = o.run(r["tend"], steps)
            qo = o.get_prim()
        t_oracle = time.perf_counter() - t0
        res = {}
        for prec in (FP64, FP32, FP32_STATE64):
            with Solver(p.copy(precision=prec)) as s:
                s.set_prim(q0)
                _, dts, _, n = s.advance(tend=r["tend"], max_steps=steps or 100000)
                res[prec] = (s.get_prim(), n, dts)
        def field_err(q):
            e = {}
            for k, f in enumerate(("rho", "u", "v", "prs")):
                d = np.abs(q[k] - qo[k])
                e[f] = {"L1": float(d.mean() / max(np.abs(qo[k]).mean(), 1e-300)),
                        "Linf": float(d.max() / max(np.abs(qo[k]).max(), 1e-300))}
            return e

        e = field_err(res[FP32][0])
        e64s = field_err(res[FP32_STATE64][0])
        q64 = res[FP64][0]
        linf64 = max(float(np.abs(q64[k] - qo[k]).max() / max(np.abs(qo[k]).max(), 1e-300)) for k in range(4))
        out[name] = {"grid": [p.Nx, p.Ny], "steps_oracle": len(dts_o), "steps_fp64": res[FP64][1],
                     "steps_fp32": res[FP32][1], "oracle_seconds": round(t_oracle, 1),
                     "fp64_parity": {"same_dt_sequence": bool(np.array_equal(res[FP64][2], dts_o)),
                                     "rel_Linf": linf64, "bitwise": bool(np.array_equal(q64, qo))},
                     "fp32_rel_error": e, "fp32_state64_steps": res[FP32_STATE64][1],
                     "fp32_state64_rel_error": e64s}
        log(f"bench: {name} vs oracle ({t_oracle:.0f} s): FP64 rel Linf {linf64:.1e}, "
            f"FP32 L1(rho) {e['rho']['L1']:.2e}")
    return out


KH_DRIFT_STEPS = 3000


def kh_drift(prec: str, size: int = 4096, steps: int = KH_DRIFT_STEPS):
    """Relative drift of the conserved totals (mass, x / y momentum, energy; momenta relative
    to the mass) over `steps` KH steps of one form ("fp64", "fp32", "fp32s64")."""
    from fv2d_amd.capi import FP32, FP32_STATE64, FP64, PROB_KH
    from fv2d_amd import problems as P
    from fv2d_amd.solver import Solver

    with Solver(P.kh_params(N=size, precision={"fp64": FP64, "fp32": FP32, "fp32s64": FP32_STATE64}[prec])) as s:
        s.set_problem(PROB_KH, seed=12345, amplitude=0.01)
        s0 = s.sums()
        s.advance(max_steps=steps, record_dts=False)
        s1 = s.sums()
    return {k: float(abs(s1[i] - s0[i]) / max(abs(s0[i]), abs(s0[0])))
            for i, k in enumerate(("mass", "x_momentum", "y_momentum", "energy"))}


def calibrate_copy(mib: int, reps: int = 4) -> int:
    """hipMemcpy device-to-device of `mib` MiB, `reps` times: a copy kernel that reads and
    writes exactly mib MiB per dispatch, the yardstick for the HBM byte counters."""
    import ctypes as C

    from fv2d_amd import hiprt

    rt = hiprt.rt()
    n = mib << 20
    a_, b_ = C.c_void_p(), C.c_void_p()
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipFree.argtypes = [C.c_void_p]
    assert rt.hipMalloc(C.byref(a_), n) == 0 and rt.hipMalloc(C.byref(b_), n) == 0
    assert rt.hipMemset(a_, 1, n) == 0
    hiprt.synchronize()
    for _ in range(reps):
        assert rt.hipMemcpy(b_, a_, n, 3) == 0  # hipMemcpyDeviceToDevice
    hiprt.synchronize()
    rt.hipFree(a_)
    rt.hipFree(b_)
    print(json.dumps({"calibrate_copy_bytes": n, "reps": reps}))
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--reps", type=int, default=5,
                    help="further event-timed device loops of K steps for the spread (median reported; K >= 16)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="length of each of the 3 CPU-baseline samples")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-errors", action="store_true",
                    help="skip the FP32-vs-FP64 error runs of every configuration (for profiling the KH loops)")
    ap.add_argument("--serve-only", action="store_true", help="only the serving measurement (for profiling)")
    ap.add_argument("--serve-no-graph", action="store_true",
                    help="batched launches without graph replay (same kernels; for tracing tools)")
    ap.add_argument("--serve-jobs", type=int, default=2000)
    ap.add_argument("--serve-steps", type=int, default=200)
    ap.add_argument("--serve-load", type=float, default=0.7,
                    help="Poisson arrival rate as a fraction of the measured batched capacity")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (calibrates rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE against a copy of known size) and exit")
    a = ap.parse_args(argv)

    if a.calibrate_copy:
        return calibrate_copy(a.calibrate_copy)

    if a.steps < 1 or a.warmup < 0 or a.serve_steps < 1 or a.serve_jobs < 1:
        print("bench.py: --steps, --serve-steps and --serve-jobs must be >= 1 and --warmup >= 0", file=sys.stderr)
        return 2
    if a.gpus != 1:
        print(f"bench.py: --gpus {a.gpus} is not supported: this build measures one GPU only; the "
              "multi-GPU decomposition is a design (y-slab halo exchange), not a measured path", file=sys.stderr)
        return 2

    from fv2d_amd import capi
    from fv2d_amd import problems as P

    capi.load_library()  # fail loudly when the HIP library is missing
    if a.serve_only:
        print(json.dumps({"serving": run_serving(a.serve_jobs, a.serve_steps, a.serve_load, 0.05,
                                                 use_graph=not a.serve_no_graph)}), flush=True)
        return 0
    log(f"bench: building KH {a.size}^2 initial state")
    q0 = P.kh_ic(P.kh_params(N=a.size))

    fp32_err = drift = None
    accuracy = not a.no_fp32 and not a.no_errors and a.size == 4096

    def drift_run(prec):
        if not accuracy:
            return None
        return lambda: kh_drift(prec, a.size)

    log(f"bench: FP64 form, {a.warmup} warmup + {a.steps} timed steps")
    r64 = run_gpu("fp64", a.size, a.steps, a.warmup, q0, a.reps, before=drift_run("fp64"))
    log(f"bench: FP64 {r64['gcells']:.3f} Gcell/s, {r64['ms_per_step']:.3f} ms/step")
    r32 = None
    if not a.no_fp32:
        log(f"bench: FP32 form, {a.warmup} warmup + {a.steps} timed steps")
        r32 = run_gpu("fp32", a.size, a.steps, a.warmup, q0, a.reps, before=drift_run("fp32"))
        log(f"bench: FP32 {r32['gcells']:.3f} Gcell/s, {r32['ms_per_step']:.3f} ms/step")
        log(f"bench: FP32 form with the state in FP64, {a.warmup} warmup + {a.steps} timed steps")
        r32s = run_gpu("fp32s64", a.size, a.steps, a.warmup, q0, a.reps, before=drift_run("fp32s64"))
        log(f"bench: FP32 (state FP64) {r32s['gcells']:.3f} Gcell/s, {r32s['ms_per_step']:.3f} ms/step")
    if accuracy:
        drift = {"steps": KH_DRIFT_STEPS, "grid": [a.size, a.size],
                 "rel_drift": {"fp64": r64["before"], "fp32": r32["before"], "fp32s64": r32s["before"]}}
        log(f"bench: KH {KH_DRIFT_STEPS}-step drift: {drift['rel_drift']}")
        fp32_err = errors_vs_oracle_all_configs()
    serving = None
    if not a.no_serve:
        serving = run_serving(a.serve_jobs, a.serve_steps, a.serve_load, 0.05)
    cpu = None
    if not a.no_cpu:
        log(f"bench: CPU oracle baseline, 3 x ~{a.cpu_seconds:.0f} s samples")
        cpu = run_cpu(a.size, a.cpu_seconds, q0)
        log(f"bench: CPU {cpu['value']:.4f} Gcell/s (median of {cpu['samples']}) on {cpu['threads']} threads, "
            f"STREAM triad {cpu['stream_triad_gbs']} GB/s")

    cells = a.size * a.size
    out = {
        "metric": METRIC,
        "value": round(r64["gcells"], 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r64["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"kelvin-helmholtz {a.size}x{a.size} fp64",
            "grid": [a.size, a.size],
            "reconstruction": "plm-minmod",
            "riemann": "hllc",
            "time_stepping": "ssp-rk2",
            "cfl": 0.8,
            "viscosity_mu": 1e-4,
            "conduction_kappa": 1e-4,
            "boundaries": "periodic",
        },
        "repetitions": r64["blocks"],
        "before_warmup": ("the form's untimed KH 3000-step drift run (the GPU at its running clock when the timed "
                          "loop starts)" if accuracy else "nothing (the timed loop may start at a ramping GPU clock)"),
        "pcie_inclusive": dict(r64["pcie"], note="the same K steps with the host primitives uploaded before "
                               "(fv2d_set_prim) and read back after (fv2d_get_prim), host clock; not `value`"),
        "hbm_frac_step": round(BYTES_PER_CELL_STEP["fp64"] * cells / (r64["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {
            "bound": "hbm",
            "kernel": "fused stage (2 launches per RK2 step)",
            "achieved": round(r64["achieved"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(r64["achieved"] / HBM_PEAK_GBS, 4),
            "traffic": read_traffic("fp64", a.size),
            "algorithmic_bytes_per_launch": r64["bytes_per_launch"],
            "mean_launch_ms": round(r64["mean_launch_ms"], 4),
        },
    }
    if r32 is not None:
        q64, q32 = r64["q"], r32["q"]
        l1 = {}
        for k, name in enumerate(("rho", "u", "v", "prs")):
            l1[name] = float(np.abs(q32[k] - q64[k]).mean() / max(np.abs(q64[k]).mean(), 1e-300))
        out["fp32_form"] = {
            "value": round(r32["gcells"], 4),
            "ms_per_step": round(r32["ms_per_step"], 4),
            "speedup_vs_fp64": round(r32["gcells"] / r64["gcells"], 3),
            "repetitions": r32["blocks"],
            "pcie_inclusive": r32["pcie"],
            "hbm_frac_step": round(BYTES_PER_CELL_STEP["fp32"] * cells / (r32["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "roofline": {
                "bound": "hbm",
                "achieved": round(r32["achieved"], 1),
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": round(r32["achieved"] / HBM_PEAK_GBS, 4),
                "traffic": read_traffic("fp32", a.size),
                "algorithmic_bytes_per_launch": r32["bytes_per_launch"],
                "mean_launch_ms": round(r32["mean_launch_ms"], 4),
            },
            "rel_L1_vs_fp64_after_steps": l1,
        }
        out["fp32_form"]["state_fp64"] = {
            "value": round(r32s["gcells"], 4),
            "ms_per_step": round(r32s["ms_per_step"], 4),
            "speedup_vs_fp64": round(r32s["gcells"] / r64["gcells"], 3),
            "roofline": {"bound": "hbm", "achieved": round(r32s["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(r32s["achieved"] / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes_per_launch": r32s["bytes_per_launch"],
                         "mean_launch_ms": round(r32s["mean_launch_ms"], 4)},
            "note": "the FP32 form's arithmetic with the state U stored in FP64 (160 B per cell-step)",
        }
        if drift is not None:
            out["fp32_form"]["kh_long_run"] = drift
        if fp32_err is not None:
            out["fp32_form"]["error_vs_oracle_every_config"] = fp32_err
    if serving is not None:
        out["serving"] = serving
    if cpu is not None:
        out["cpu_baseline"] = {
            "value": round(cpu["value"], 5),
            "unit": "Gcell-updates/s",
            "cores": cpu["threads"],
            "kind": "port",
            "sample": f"CPU oracle (FP64 restatement of the reference path with its pass structure, OpenMP), "
                      f"KH {a.size}^2: median of {len(cpu['samples'])} runs of ~{cpu['seconds']:.0f} s "
                      f"({cpu['steps']} RK2 steps)",
            "samples": cpu["samples"],
            "nproc": cpu["nproc"],
            "cpu_model": cpu["cpu_model"],
            "OMP_NUM_THREADS": cpu["OMP_NUM_THREADS"],
            "stream_triad_gbs": cpu["stream_triad_gbs"],
        }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
