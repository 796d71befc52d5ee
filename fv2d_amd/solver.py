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
from .capi import Params, StepStats


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
        """Run the loop on the device; returns (t, dts, counts)."""
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
