"""Python mirror of the C ABI (include/fv2d_amd.h) over libfv2d_amd.so.

`Solver` is one simulation on the GPU. Its methods map one to one onto the C entry
points, which replace the reference calls of main.cpp:

    set_prim     InitFunctor::init / IOManager::loadSnapshot + primToCons (main.cpp:47-58)
    compute_dt   ComputeDtFunctor::computeDt (ComputeDt.h:18-65)
    step         UpdateFunctor::update + consToPrim + checkNegatives (main.cpp:79-81)
    advance      the time loop of main.cpp:62-84, dt kept on the device
    get_prim     the Q read-back of a snapshot (IOManager.h:151-170)

Errors raise Fv2dError carrying the library's message; there is no fallback path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import BatchConfig, JobInfo, Params, Problem, StepStats


class Fv2dError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"fv2d error {code}: {msg}")
        self.code = code


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def check_params(p: Params, lib=None) -> None:
    lib = lib or capi.load_library()
    rc = lib.fv2d_params_check(C.byref(p))
    if rc:
        raise Fv2dError(rc, (lib.fv2d_last_error(None) or b"").decode())


class Solver:
    def __init__(self, params: Params, lib=None):
        self.lib = lib or capi.load_library()
        self.params = params.copy()
        h = C.c_void_p()
        rc = self.lib.fv2d_create(C.byref(self.params), C.byref(h))
        if rc:
            raise Fv2dError(rc, (self.lib.fv2d_last_error(None) or b"").decode())
        self.h = h
        self.Nx, self.Ny = params.Nx, params.Ny

    # lifecycle ----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.fv2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ok(self, rc: int) -> None:
        if rc:
            raise Fv2dError(rc, (self.lib.fv2d_last_error(self.h) or b"").decode())

    # path ---------------------------------------------------------------
    def set_prim(self, q: np.ndarray) -> None:
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.size != 4 * self.Nx * self.Ny:
            raise ValueError(f"q has {q.size} values, expected 4*{self.Nx}*{self.Ny}")
        self._ok(self.lib.fv2d_set_prim(self.h, _dp(q)))

    def set_problem(self, kind: int, seed: int = 0, amplitude: float = 0.0) -> None:
        """Initial condition generated on the device (fv2d_set_problem)."""
        pr = Problem(kind=kind, seed=seed, amplitude=amplitude)
        self._ok(self.lib.fv2d_set_problem(self.h, C.byref(pr)))

    def compute_dt(self):
        dt = C.c_double()
        inv = np.zeros(3)
        self._ok(self.lib.fv2d_compute_dt(self.h, C.byref(dt), _dp(inv)))
        return dt.value, inv

    def step(self, dt: float):
        s = StepStats()
        self._ok(self.lib.fv2d_step(self.h, float(dt), C.byref(s)))
        return s.as_tuple()

    def advance(self, tend: float = 1e300, max_steps: int = 0, t0: float = 0.0, record_dts: bool = True):
        """Run the loop on the device until t + epsilon >= tend or max_steps steps; returns
        (t, dts, counts, steps taken). One of the two ends must be given."""
        if max_steps <= 0 and not tend < 1e300:
            raise ValueError("advance: give tend or max_steps > 0 (the loop would not end)")
        t = C.c_double(t0)
        n = C.c_int32(0)
        s = StepStats()
        dts = np.zeros(max(max_steps, 1)) if (record_dts and max_steps > 0) else None
        self._ok(self.lib.fv2d_advance(self.h, C.byref(t), float(tend), int(max_steps),
                                       _dp(dts) if dts is not None else None, C.byref(n), C.byref(s)))
        return t.value, (dts[: n.value].copy() if dts is not None else np.zeros(0)), s.as_tuple(), n.value

    def get_prim(self) -> np.ndarray:
        out = np.empty((4, self.Ny, self.Nx))
        self._ok(self.lib.fv2d_get_prim(self.h, _dp(out)))
        return out

    def get_cons(self) -> np.ndarray:
        out = np.empty((4, self.Ny, self.Nx))
        self._ok(self.lib.fv2d_get_cons(self.h, _dp(out)))
        return out

    def sums(self) -> np.ndarray:
        out = np.zeros(4)
        self._ok(self.lib.fv2d_sums(self.h, _dp(out)))
        return out

    def synchronize(self) -> None:
        self._ok(self.lib.fv2d_synchronize(self.h))


class Batch:
    """Many independent simulations on one GPU (the fv2d_batch_* entry points).

    submit() admits a job (raises BatchBusy when no slot or arena range is free), launch()
    queues launches of every running job, wait() blocks for them and returns the number of
    jobs still running, poll() lists finished jobs and collect() returns a finished job's
    primitives, conservation sums and info and frees its slot.
    """

    def __init__(self, config: BatchConfig | None = None, lib=None, **kw):
        self.lib = lib or capi.load_library()
        self.config = config if config is not None else BatchConfig.defaults(**kw)
        b = C.c_void_p()
        rc = self.lib.fv2d_batch_create(C.byref(self.config), C.byref(b))
        if rc:
            raise Fv2dError(rc, (self.lib.fv2d_batch_last_error(None) or b"").decode())
        self.b = b
        self.shapes: dict[int, tuple[int, int]] = {}

    def close(self) -> None:
        if getattr(self, "b", None):
            self.lib.fv2d_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ok(self, rc: int) -> None:
        if rc == capi.FV2D_E_BUSY:
            raise BatchBusy(rc, (self.lib.fv2d_batch_last_error(self.b) or b"").decode())
        if rc:
            raise Fv2dError(rc, (self.lib.fv2d_batch_last_error(self.b) or b"").decode())

    def submit(self, params: Params, problem: Problem | None = None, q: np.ndarray | None = None,
               tend: float = 1e300, max_steps: int = 0) -> int:
        if max_steps <= 0 and not tend < 1e300:
            raise ValueError("submit: give tend or max_steps > 0 (the job would not end)")
        jid = C.c_int64()
        qa = None
        if q is not None:
            qa = np.ascontiguousarray(q, dtype=np.float64)
            if qa.size != 4 * params.Nx * params.Ny:
                raise ValueError("q has the wrong size")
        self._ok(self.lib.fv2d_batch_submit(self.b, C.byref(params), C.byref(problem) if problem is not None else None,
                                            _dp(qa) if qa is not None else None, float(tend), int(max_steps),
                                            C.byref(jid)))
        self.shapes[jid.value] = (params.Nx, params.Ny)
        return jid.value

    def launch(self, launches: int = 1) -> None:
        self._ok(self.lib.fv2d_batch_launch(self.b, int(launches)))

    def wait(self) -> int:
        r = C.c_int32()
        self._ok(self.lib.fv2d_batch_wait(self.b, C.byref(r)))
        return r.value

    def run(self, launches: int = 1) -> int:
        self.launch(launches)
        return self.wait()

    def poll(self, cap: int = 4096) -> list[int]:
        ids = (C.c_int64 * cap)()
        n = C.c_int32()
        self._ok(self.lib.fv2d_batch_poll(self.b, ids, cap, C.byref(n)))
        return [ids[i] for i in range(n.value)]

    def info(self, jid: int) -> JobInfo:
        inf = JobInfo()
        self._ok(self.lib.fv2d_batch_job_info(self.b, int(jid), C.byref(inf)))
        return inf

    def collect(self, jid: int, prim: bool = True):
        """(q or None, sums, info) of a finished job; frees its slot."""
        if jid not in self.shapes:
            raise Fv2dError(capi.FV2D_E_INVALID, f"no job {jid} in this batch")
        nx, ny = self.shapes[jid]
        q = np.empty((4, ny, nx)) if prim else None
        sums = np.zeros(4)
        inf = JobInfo()
        self._ok(self.lib.fv2d_batch_collect(self.b, int(jid), _dp(q) if q is not None else None, _dp(sums),
                                             C.byref(inf)))
        del self.shapes[jid]
        return q, sums, inf

    def cancel(self, jid: int) -> None:
        self._ok(self.lib.fv2d_batch_cancel(self.b, int(jid)))
        self.shapes.pop(jid, None)

    def last_launch_ms(self) -> float:
        ms = C.c_double()
        self._ok(self.lib.fv2d_batch_last_launch_ms(self.b, C.byref(ms)))
        return ms.value


class BatchBusy(Fv2dError):
    """No slot, work-list room or arena range is free for the job right now."""


def job_bytes(params: Params, lib=None) -> int:
    lib = lib or capi.load_library()
    return int(lib.fv2d_job_bytes(C.byref(params)))
