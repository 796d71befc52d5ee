"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle (liboracle_fv2d.so).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may use
this module, and only as the parity reference / the timed CPU baseline. The product path never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from fv2d_amd.capi import Params

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_fv2d.so")


def build(plant: int = 0) -> str:
    """Compile the oracle (or planted-mistake variant `plant`) with its Makefile."""
    if plant:
        subprocess.run(["make", "-s", "-C", HERE, "plant", f"N={plant}"], check=True)
        return os.path.join(HERE, "_plant", f"liboracle_fv2d_p{plant}.so")
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _declare(lib):
    P = C.POINTER
    dp = P(C.c_double)
    u64p = P(C.c_uint64)
    H = C.c_void_p
    sig = {
        "orc_create": (H, [P(Params), C.c_char_p]),
        "orc_destroy": (None, [H]),
        "orc_set_threads": (None, [C.c_int]),
        "orc_get_threads": (C.c_int, []),
        "orc_set_prim": (None, [H, dp]),
        "orc_get_prim": (None, [H, dp]),
        "orc_get_cons": (None, [H, dp]),
        "orc_get_prim_full": (None, [H, dp]),
        "orc_get_cons_full": (None, [H, dp]),
        "orc_poke_prim": (None, [H, dp]),
        "orc_compute_dt": (C.c_double, [H, dp]),
        "orc_step": (None, [H, C.c_double, u64p]),
        "orc_check_negatives": (None, [H, u64p]),
        "orc_run": (C.c_int, [H, dp, C.c_double, C.c_int, dp, u64p]),
        "orc_fill_boundaries": (None, [H]),
        "orc_sums": (None, [H, dp]),
        "orc_prim_to_cons": (None, [dp, C.c_double, dp]),
        "orc_cons_to_prim": (None, [dp, C.c_double, dp]),
        "orc_speed_of_sound": (C.c_double, [dp, C.c_double]),
        "orc_hll": (None, [dp, dp, C.c_double, dp, dp]),
        "orc_hllc": (None, [dp, dp, C.c_double, dp, dp]),
        "orc_fslp": (None, [dp, dp, C.c_double, C.c_double, C.c_double, dp, dp]),
        "orc_minmod": (C.c_double, [C.c_double, C.c_double]),
        "orc_gravity": (C.c_double, [H, C.c_int, C.c_int, C.c_int]),
        "orc_mesh": (None, [H, P(C.c_int32), dp]),
        "orc_stream_triad": (C.c_double, [C.c_long, C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIBS: dict = {}


def load(path: str | None = None):
    p = path or LIB_PATH
    if p not in _LIBS:
        if not os.path.exists(p):
            raise RuntimeError(f"{p} is missing: build the oracle first (make -C oracle)")
        _LIBS[p] = _declare(C.CDLL(p))
    return _LIBS[p]


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _v4(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(4))


class Oracle:
    """One oracle simulation (the reference's Q/U arrays with ghosts, on the CPU)."""

    def __init__(self, params: Params, lib=None):
        self.lib = lib or load()
        self.params = params.copy()
        err = C.create_string_buffer(256)
        self.h = self.lib.orc_create(C.byref(self.params), err)
        if not self.h:
            raise ValueError(err.value.decode())
        oi = (C.c_int32 * 8)()
        orl = (C.c_double * 2)()
        self.lib.orc_mesh(self.h, oi, orl)
        self.Ntx, self.Nty, self.ibeg, self.iend, self.jbeg, self.jend, self.Nx, self.Ny = list(oi)
        self.dx, self.dy = orl[0], orl[1]

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
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

    # state -------------------------------------------------------------
    def set_prim(self, q: np.ndarray) -> None:
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(4, self.Ny, self.Nx)
        self.lib.orc_set_prim(self.h, _dp(q))

    def poke_prim(self, q: np.ndarray) -> None:
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(4, self.Ny, self.Nx)
        self.lib.orc_poke_prim(self.h, _dp(q))

    def get_prim(self) -> np.ndarray:
        out = np.empty((4, self.Ny, self.Nx))
        self.lib.orc_get_prim(self.h, _dp(out))
        return out

    def get_cons(self) -> np.ndarray:
        out = np.empty((4, self.Ny, self.Nx))
        self.lib.orc_get_cons(self.h, _dp(out))
        return out

    def get_prim_full(self) -> np.ndarray:
        out = np.empty((4, self.Nty, self.Ntx))
        self.lib.orc_get_prim_full(self.h, _dp(out))
        return out

    def get_cons_full(self) -> np.ndarray:
        out = np.empty((4, self.Nty, self.Ntx))
        self.lib.orc_get_cons_full(self.h, _dp(out))
        return out

    def fill_boundaries(self) -> None:
        self.lib.orc_fill_boundaries(self.h)

    # path ----------------------------------------------------------------
    def compute_dt(self):
        inv = np.zeros(3)
        dt = self.lib.orc_compute_dt(self.h, _dp(inv))
        return dt, inv

    def step(self, dt: float):
        c = (C.c_uint64 * 3)()
        self.lib.orc_step(self.h, dt, c)
        return tuple(int(x) for x in c)

    def check_negatives(self):
        c = (C.c_uint64 * 3)()
        self.lib.orc_check_negatives(self.h, c)
        return tuple(int(x) for x in c)

    def run(self, tend: float, max_steps: int = 0, t0: float = 0.0):
        cap = max_steps if max_steps > 0 else 1_000_000
        dts = np.zeros(cap)
        t = C.c_double(t0)
        c = (C.c_uint64 * 3)()
        n = self.lib.orc_run(self.h, C.byref(t), tend, max_steps, _dp(dts), c)
        return t.value, dts[:n].copy(), tuple(int(x) for x in c)

    def sums(self) -> np.ndarray:
        out = np.zeros(4)
        self.lib.orc_sums(self.h, _dp(out))
        return out

    def gravity(self, i: int, j: int, d: int) -> float:
        return self.lib.orc_gravity(self.h, i, j, d)


# point functions ------------------------------------------------------------

def prim_to_cons(q, gamma0, lib=None):
    lib = lib or load()
    out = np.zeros(4)
    lib.orc_prim_to_cons(_dp(_v4(q)), gamma0, _dp(out))
    return out


def cons_to_prim(u, gamma0, lib=None):
    lib = lib or load()
    out = np.zeros(4)
    lib.orc_cons_to_prim(_dp(_v4(u)), gamma0, _dp(out))
    return out


def speed_of_sound(q, gamma0, lib=None):
    lib = lib or load()
    return lib.orc_speed_of_sound(_dp(_v4(q)), gamma0)


def riemann(kind: str, qL, qR, gamma0, lib=None, fslp_K=1.1, gdx=0.0):
    lib = lib or load()
    f = np.zeros(4)
    pout = C.c_double(0.0)
    a, b = _v4(qL), _v4(qR)
    if kind == "hll":
        lib.orc_hll(_dp(a), _dp(b), gamma0, _dp(f), C.byref(pout))
    elif kind == "hllc":
        lib.orc_hllc(_dp(a), _dp(b), gamma0, _dp(f), C.byref(pout))
    elif kind == "fslp":
        lib.orc_fslp(_dp(a), _dp(b), gamma0, fslp_K, gdx, _dp(f), C.byref(pout))
    else:
        raise ValueError(kind)
    return f, pout.value


def minmod(a: float, b: float, lib=None) -> float:
    lib = lib or load()
    return lib.orc_minmod(a, b)
