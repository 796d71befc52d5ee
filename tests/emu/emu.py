"""TEST INFRASTRUCTURE: ctypes wrapper of the host emulation of the stage kernel."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from fv2d_amd.capi import Params

HERE = os.path.dirname(os.path.abspath(__file__))


def build() -> None:
    subprocess.run(["make", "-s", "-j8", "-C", HERE], check=True)


def load(sanitized: bool = False):
    name = "libtile_emu_san.so" if sanitized else "libtile_emu.so"
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    u64p = C.POINTER(C.c_uint64)
    lib.emu_create.restype = C.c_void_p
    lib.emu_create.argtypes = [C.POINTER(Params), C.c_char_p]
    lib.emu_destroy.argtypes = [C.c_void_p]
    lib.emu_set_prim.argtypes = [C.c_void_p, dp]
    lib.emu_compute_dt.restype = C.c_double
    lib.emu_compute_dt.argtypes = [C.c_void_p, dp]
    lib.emu_step.argtypes = [C.c_void_p, C.c_double, u64p]
    lib.emu_get_prim.argtypes = [C.c_void_p, dp]
    lib.emu_get_cons.argtypes = [C.c_void_p, dp]
    lib.emu_markstein_mismatches.restype = C.c_longlong
    lib.emu_markstein_mismatches.argtypes = [dp, C.c_longlong, C.c_double, C.c_int]
    lib.emu_div_mode.restype = C.c_int
    lib.emu_div_mode.argtypes = [C.c_double]
    lib.emu_use_banded.argtypes = [C.c_void_p, C.c_int]
    lib.emu_block_rows.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]
    lib.emu_block_rows.restype = C.c_int
    lib.emu_sim_geom.argtypes = [C.POINTER(Params), C.POINTER(C.c_int)]
    return lib


def block_rows(Nx: int, Ny: int, seg: int, nw: int, banded: bool, lib=None) -> np.ndarray:
    """(sx, jb, je) of every block of the stage kernel's geometry, shape (nblocks, 3)."""
    lib = lib or load()
    n = lib.emu_block_rows(Nx, Ny, seg, nw, int(banded), None)
    out = np.zeros(3 * n, dtype=np.int32)
    lib.emu_block_rows(Nx, Ny, seg, nw, int(banded), out.ctypes.data_as(C.POINTER(C.c_int)))
    return out.reshape(n, 3)


def sim_geom(params: Params, lib=None) -> dict:
    """Geometry the single-grid kernel picks for params."""
    lib = lib or load()
    out = (C.c_int * 6)()
    p = params.copy()
    lib.emu_sim_geom(C.byref(p), out)
    return dict(zip(("nblocks", "seg", "sched", "lseg", "sseg", "tail"), list(out)))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Emu:
    def __init__(self, params: Params, lib=None):
        self.lib = lib or load()
        self.params = params.copy()
        err = C.create_string_buffer(256)
        self.h = self.lib.emu_create(C.byref(self.params), err)
        if not self.h:
            raise ValueError(err.value.decode())
        self.Nx, self.Ny = params.Nx, params.Ny

    def close(self):
        if self.h:
            self.lib.emu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def use_banded(self, seg: int):
        """Run the stage with the banded long-then-short schedule (long segments of seg rows)."""
        self.lib.emu_use_banded(self.h, int(seg))

    def set_prim(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        self.lib.emu_set_prim(self.h, _dp(q))

    def compute_dt(self):
        inv = np.zeros(3)
        return self.lib.emu_compute_dt(self.h, _dp(inv)), inv

    def step(self, dt):
        c = (C.c_uint64 * 3)()
        self.lib.emu_step(self.h, dt, c)
        return tuple(int(x) for x in c)

    def get_prim(self):
        out = np.empty((4, self.Ny, self.Nx))
        self.lib.emu_get_prim(self.h, _dp(out))
        return out

    def get_cons(self):
        out = np.empty((4, self.Ny, self.Nx))
        self.lib.emu_get_cons(self.h, _dp(out))
        return out
