"""Serving: a stream of independent simulations on one GPU, batched (continuous batching
over fv2d_batch_*) or one at a time (the single-simulation entry points), with an
autotuner for the batch's launch shape.

The job stream: jobs arrive by a seeded Poisson process; each is a seeded instance of the
Sod tube, the point blast or the Kelvin-Helmholtz layer on one of the grids of
SERVE_SIZES, with gamma ~ U[1.4, 5/3], perturbation amplitude ~ U[1e-3, 1e-2], run for a
fixed number of RK2 steps (200 by default). Every job's domain is [0,1]^2.

Batched serving (`serve_batched`) is continuous: the request loop admits arrived jobs
whenever a slot and room in the batch's HBM arena are free, queues a launch (every running
job advances by steps_per_launch steps, each at its own step and dt), collects the jobs
that the previous launch finished while the next launch runs, and repeats. A job's latency
is the time from its arrival to the end of its collection.

One-at-a-time serving (`serve_sequential`) runs each job alone through the
single-simulation path: fv2d_create, fv2d_set_problem, fv2d_advance, fv2d_sums,
fv2d_destroy.
"""
from __future__ import annotations

import json
import math
import os
import time
from collections import deque
from dataclasses import asdict, dataclass, field

import numpy as np

from . import capi
from .capi import (BC_ABSORBING, BC_PERIODIC, BC_REFLECTING, FP32, FP64, HLL, HLLC, PLM, PROB_BLAST, PROB_KH,
                   PROB_SOD_X, TS_RK2, BatchConfig, Params, Problem)
from .solver import Batch, BatchBusy, Solver

SERVE_SIZES = ((128, 128), (256, 256), (512, 128), (512, 512), (1024, 256), (1024, 1024), (2048, 512))
KINDS = (PROB_SOD_X, PROB_BLAST, PROB_KH)
KIND_NAMES = {PROB_SOD_X: "sod", PROB_BLAST: "blast", PROB_KH: "kh"}


@dataclass
class JobSpec:
    idx: int
    arrival: float  # seconds after the start of the stream
    kind: int
    Nx: int
    Ny: int
    gamma: float
    amplitude: float
    seed: int
    steps: int

    @property
    def cells(self) -> int:
        return self.Nx * self.Ny


def job_params(spec: JobSpec, precision: int = FP32, seg_rows: int = 0, chunk_rows: int = 0) -> Params:
    """Solver settings of a served job: the numerics of its setup on [0,1]^2; seg_rows and
    chunk_rows (the tile: rows per work item, rows per chunk = waves per workgroup) as tuned."""
    base = dict(Nx=spec.Nx, Ny=spec.Ny, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, gamma0=spec.gamma,
                reconstruction=PLM, time_stepping=TS_RK2, precision=precision, seg_rows=seg_rows,
                chunk_rows=chunk_rows)
    if spec.kind == PROB_SOD_X:
        base.update(riemann_solver=HLLC, CFL=0.8, boundary_x=BC_ABSORBING, boundary_y=BC_PERIODIC)
    elif spec.kind == PROB_BLAST:
        base.update(riemann_solver=HLL, CFL=0.7, boundary_x=BC_REFLECTING, boundary_y=BC_REFLECTING)
    else:
        base.update(riemann_solver=HLLC, CFL=0.8, boundary_x=BC_PERIODIC, boundary_y=BC_PERIODIC,
                    visc_active=1, mu=1.0e-4, tc_active=1, kappa=1.0e-4)
    return Params.defaults(**base)


def job_problem(spec: JobSpec) -> Problem:
    return Problem(kind=spec.kind, seed=spec.seed, amplitude=spec.amplitude)


def make_stream(n_jobs: int = 2000, rate: float = 200.0, seed: int = 2024, steps: int = 200,
                sizes=SERVE_SIZES) -> list[JobSpec]:
    """Seeded Poisson arrivals at `rate` jobs/s (rate <= 0 or inf: all at t = 0)."""
    rng = np.random.default_rng(seed)
    if rate and math.isfinite(rate) and rate > 0:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    else:
        rng.exponential(1.0, n_jobs)  # keep the other draws identical
        arrivals = np.zeros(n_jobs)
    size_ix = rng.integers(0, len(sizes), n_jobs)
    kinds = rng.integers(0, len(KINDS), n_jobs)
    gammas = rng.uniform(1.4, 5.0 / 3.0, n_jobs)
    amps = rng.uniform(1.0e-3, 1.0e-2, n_jobs)
    seeds = rng.integers(1, 2**62, n_jobs)
    out = []
    for i in range(n_jobs):
        nx, ny = sizes[size_ix[i]]
        out.append(JobSpec(i, float(arrivals[i]), int(KINDS[kinds[i]]), nx, ny, float(gammas[i]), float(amps[i]),
                           int(seeds[i]), steps))
    return out


def with_arrivals(specs: list[JobSpec], rate: float, seed: int = 7) -> list[JobSpec]:
    """The same jobs with new Poisson arrival times at `rate` jobs/s."""
    rng = np.random.default_rng(seed)
    arr = np.cumsum(rng.exponential(1.0 / rate, len(specs)))
    return [JobSpec(**{**asdict(s), "arrival": float(a)}) for s, a in zip(specs, arr)]


@dataclass
class JobRecord:
    idx: int
    arrival: float
    admitted: float
    finished: float
    nsteps: int
    t: float
    sums: list

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class ServeResult:
    mode: str
    records: list = field(default_factory=list)
    makespan: float = 0.0  # first arrival to last completion (s)
    cell_steps: float = 0.0
    gpu_ms: float = 0.0  # summed launch durations (batched) or 0
    launches: int = 0

    @property
    def throughput(self) -> float:
        """Cell updates per second over the makespan."""
        return self.cell_steps / self.makespan if self.makespan > 0 else 0.0

    def latency_percentiles(self, ps=(50, 99)) -> dict:
        lat = np.array([r.latency for r in self.records])
        return {f"p{p}": float(np.percentile(lat, p)) for p in ps} if len(lat) else {}


def _tuned_tile(tuning: dict | None, spec: JobSpec) -> tuple[int, int]:
    """(seg_rows, chunk_rows) the autotuner chose for the job's setup and grid (0, 0: defaults)."""
    if not tuning:
        return 0, 0
    t = tuning.get("tile", {}).get(f"{KIND_NAMES[spec.kind]}:{spec.Nx}x{spec.Ny}", (0, 0))
    return int(t[0]), int(t[1])


def serve_batched(specs: list[JobSpec], precision: int = FP32, config: BatchConfig | None = None,
                  tuning: dict | None = None, keep_prim: set | None = None, realtime: bool = True,
                  lib=None) -> tuple[ServeResult, dict]:
    """Continuous batching of the stream. Returns (result, {idx: final primitives}) where the
    primitives are kept only for the job indices in keep_prim."""
    cfg = config or BatchConfig.defaults(precision=precision)
    if tuning and "blocks_per_cu" in tuning and cfg.blocks_per_cu == 0:
        cfg.blocks_per_cu = int(tuning["blocks_per_cu"])
    if tuning and "steps_per_launch" in tuning:
        cfg.steps_per_launch = int(tuning["steps_per_launch"])
    cfg.precision = precision
    keep_prim = keep_prim or set()
    res = ServeResult("batched")
    prims = {}
    order = sorted(specs, key=lambda s: (s.arrival, s.idx))
    with Batch(cfg, lib=lib) as b:
        pending = deque(order)
        queue: deque = deque()
        running: dict[int, tuple[JobSpec, float]] = {}
        to_collect: list[int] = []
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))
        while pending or queue or running or to_collect:
            now = clock()
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            while queue:
                s = queue[0]
                try:
                    jid = b.submit(job_params(s, precision, *_tuned_tile(tuning, s)), problem=job_problem(s),
                                   max_steps=s.steps)
                except BatchBusy as e:
                    if not running and not to_collect:
                        raise RuntimeError(f"job {s.idx} ({s.Nx}x{s.Ny}) does not fit an empty batch: {e}") from e
                    break
                queue.popleft()
                running[jid] = (s, time.perf_counter() - t0)
            launched = False
            if len(running) > len(to_collect):
                b.launch(1)
                launched = True
            for jid in to_collect:  # overlaps the launch just queued
                s, adm = running.pop(jid)
                q, sums, inf = b.collect(jid, prim=s.idx in keep_prim)
                if q is not None:
                    prims[s.idx] = q
                res.records.append(JobRecord(s.idx, s.arrival, adm, time.perf_counter() - t0, inf.nsteps, inf.t,
                                             sums.tolist()))
                res.cell_steps += float(s.cells) * inf.nsteps
            to_collect = []
            if launched:
                b.wait()
                res.gpu_ms += b.last_launch_ms()
                res.launches += 1
                to_collect = b.poll()
            elif not running and pending:
                # idle: wait for the next arrival
                dt = pending[0].arrival - clock()
                if dt > 0:
                    time.sleep(min(dt, 2e-4))
        t_end = time.perf_counter() - t0
    first = min(s.arrival for s in specs) if realtime else 0.0
    res.makespan = t_end - first
    return res, prims


def serve_sequential(specs: list[JobSpec], precision: int = FP32, realtime: bool = False,
                     keep_prim: set | None = None, lib=None, loop: str = "device") -> tuple[ServeResult, dict]:
    """Each job alone through the single-simulation path, in arrival order. With
    realtime=False all jobs are taken as arrived at t = 0 (service times only).

    loop="device": the time loop runs on the device (fv2d_advance). loop="host": the
    reference's loop structure (main.cpp:62-84), dt read back to the host and one blocking
    step per RK2 step (fv2d_compute_dt + fv2d_step)."""
    if loop not in ("device", "host"):
        raise ValueError(f"loop must be 'device' or 'host', not {loop!r}")
    keep_prim = keep_prim or set()
    res = ServeResult("sequential" if loop == "device" else "sequential-host-loop")
    prims = {}
    order = sorted(specs, key=lambda s: (s.arrival, s.idx))
    t0 = time.perf_counter()
    for s in order:
        if realtime:
            dt = s.arrival - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
        adm = time.perf_counter() - t0
        with Solver(job_params(s, precision), lib=lib) as sv:
            sv.set_problem(s.kind, s.seed, s.amplitude)
            if loop == "device":
                t, _, _, n = sv.advance(max_steps=s.steps, record_dts=False)
            else:
                t, n = 0.0, 0
                while n < s.steps:
                    dt, _ = sv.compute_dt()
                    sv.step(dt)
                    t += dt
                    n += 1
            sums = sv.sums()
            if s.idx in keep_prim:
                prims[s.idx] = sv.get_prim()
        fin = time.perf_counter() - t0
        res.records.append(JobRecord(s.idx, s.arrival if realtime else 0.0, adm, fin, n, t, sums.tolist()))
        res.cell_steps += float(s.cells) * n
    res.makespan = time.perf_counter() - t0
    return res, prims


def fifo_latencies(specs: list[JobSpec], service: dict) -> np.ndarray:
    """Latencies of a single FIFO server with the given per-job service times (s) under the
    specs' arrival times: finish_i = max(arrival_i, finish_{i-1}) + service_i."""
    order = sorted(specs, key=lambda s: (s.arrival, s.idx))
    fin = 0.0
    lat = []
    for s in order:
        fin = max(s.arrival, fin) + service[s.idx]
        lat.append(fin - s.arrival)
    return np.array(lat)


# ------------------------------------------------------------------ autotuner

# Tile candidates of the batched stage kernel: (segment height = rows per work item, chunk
# rows = waves per workgroup).
TILE_CANDIDATES = ((32, 4), (64, 4), (32, 8), (64, 8))


def _time_batch(specs_params, config: BatchConfig, launches: int = 3, warm: int = 1, lib=None) -> float:
    """Cell updates per second of a batch holding the given (params, problem) jobs."""
    with Batch(config, lib=lib) as b:
        steps = config.steps_per_launch * (warm + launches) + 1
        cells = 0
        for p, pr in specs_params:
            b.submit(p, problem=pr, max_steps=steps)
            cells += p.Nx * p.Ny
        b.run(warm)
        t = 0.0
        for _ in range(launches):
            b.run(1)
            t += b.last_launch_ms() * 1e-3
        return cells * config.steps_per_launch * launches / t


def autotune(precision: int = FP32, sizes=SERVE_SIZES, kinds=KINDS, target_cells: int = 12 << 20,
             tiles=TILE_CANDIDATES, blocks_per_cu=(0, 3, 2), mix_tiles=((64, 4), (64, 8)), cache: str | None = None,
             lib=None, log=None, use_graph: int = 1) -> dict:
    """Choose the tile of the batched stage kernel per kernel class (setup) and grid: the
    segment height (rows per work item) and the chunk height (rows per chunk = waves per
    workgroup, a compiled variant), then the persistent workgroups per CU of the batch.
    First by timing batches of identical jobs (enough copies to fill the GPU) for each
    candidate tile; then on a mixed sample of the stream, where the per-size table competes
    with uniform tiles (`mix_tiles`), then the workgroup counts. Every candidate's rate is
    kept in "rates", the choice per class and grid in "tile". Cached in `cache` (JSON)."""
    lib = lib or capi.load_library()
    key = f"prec{precision}"
    if cache and os.path.exists(cache):
        with open(cache) as f:
            data = json.load(f)
        if key in data:
            return data[key]
    out = {"tile": {}, "rates": {}}
    for kind in kinds:
        for nx, ny in sizes:
            spec = JobSpec(0, 0.0, kind, nx, ny, 1.4, 5e-3, 1, 0)
            copies = max(1, min(256, target_cells // (nx * ny)))
            best, best_rate = tiles[0], -1.0
            for seg, nw in tiles:
                if seg > (ny + nw - 1) // nw * nw and (seg, nw) != tiles[0]:
                    continue
                p = job_params(spec, precision, seg, nw)
                cfg = BatchConfig.defaults(precision=precision, max_jobs=max(copies, 1), use_graph=use_graph)
                rate = _time_batch([(p, job_problem(spec))] * copies, cfg, lib=lib)
                out["rates"][f"{KIND_NAMES[kind]}:{nx}x{ny}:seg{seg}:nw{nw}"] = rate
                if rate > best_rate:
                    best, best_rate = (seg, nw), rate
            out["tile"][f"{KIND_NAMES[kind]}:{nx}x{ny}"] = list(best)
            if log:
                log(f"autotune {KIND_NAMES[kind]} {nx}x{ny}: seg {best[0]} x {best[1]} waves "
                    f"({best_rate / 1e9:.2f} Gcell/s)")
    # On a mixed sample of the stream the classes run concurrently, which favours other
    # tiles than batches of one class: the per-size table and uniform tiles compete there,
    # then the persistent workgroups per CU. Heights above 64 rows were measured to add ~2%
    # throughput but ~50% p50 latency (every job of a class waits for the slowest work item
    # of each step), so they are not candidates by default.
    sample = make_stream(96, rate=0, seed=99, steps=0, sizes=sizes)

    def mix_rate(table, bpc):
        mix = [(job_params(s, precision, *table.get(f"{KIND_NAMES[s.kind]}:{s.Nx}x{s.Ny}", (0, 0))),
                job_problem(s)) for s in sample]
        cfg = BatchConfig.defaults(precision=precision, max_jobs=len(mix), blocks_per_cu=bpc, use_graph=use_graph)
        return _time_batch(mix, cfg, lib=lib)

    tables = {"per-size": dict(out["tile"])}
    for seg, nw in mix_tiles:
        tables[f"uniform-seg{seg}-nw{nw}"] = {k: [seg, nw] for k in out["tile"]}
    best_t, best_rate = "per-size", -1.0
    for name, table in tables.items():
        rate = mix_rate(table, 0)
        out["rates"][f"mix:{name}"] = rate
        if rate > best_rate:
            best_t, best_rate = name, rate
    out["tile"] = tables[best_t]
    out["tile_choice"] = best_t
    if log:
        log(f"autotune mix: tile table {best_t} ({best_rate / 1e9:.2f} Gcell/s)")
    best_b = 0
    for bpc in blocks_per_cu:
        if bpc == 0:
            continue
        rate = mix_rate(out["tile"], bpc)
        out["rates"][f"mix:bpc{bpc}"] = rate
        if rate > best_rate:
            best_b, best_rate = bpc, rate
    out["blocks_per_cu"] = best_b
    if log:
        log(f"autotune mix: blocks_per_cu {best_b or 'occupancy'} ({best_rate / 1e9:.2f} Gcell/s)")
    if cache:
        data = {}
        if os.path.exists(cache):
            with open(cache) as f:
                data = json.load(f)
        data[key] = out
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(data, f, indent=1)
    return out


def field_errors(q: np.ndarray, ref: np.ndarray) -> dict:
    """Relative L1 and Linf error per primitive field (rho, u, v, prs)."""
    out = {}
    for k, name in enumerate(("rho", "u", "v", "prs")):
        d = np.abs(q[k] - ref[k])
        s1 = max(float(np.abs(ref[k]).mean()), 1e-300)
        si = max(float(np.abs(ref[k]).max()), 1e-300)
        out[name] = (float(d.mean()) / s1, float(d.max()) / si)
    return out
