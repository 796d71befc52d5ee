"""Python access to the host side (libfv2d_host.so and the fv2d executable).

The host side keeps the reference's run surface: the .ini reader and parameter set
(SimInfo.h:100-568), the initial-condition library (Init.h) and the HDF5/XDMF snapshot
layout with restart (IOManager.h). This module wraps its C entry points with ctypes for
tests and tools; `run` starts the fv2d executable.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess

import numpy as np

from .capi import Params

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(HERE, "libfv2d_host.so")
FV2D_EXE = os.path.join(HERE, "fv2d")

_LIB = None


class HostError(RuntimeError):
    pass


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is missing (python -m fv2d_amd.build)")
        L = C.CDLL(HOST_LIB)
        L.fv2dh_last_error.restype = C.c_char_p
        L.fv2dh_read_ini.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int]
        L.fv2dh_fv2d_params.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(Params)]
        L.fv2dh_init_problem.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_double)]
        L.fv2dh_save.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_double), C.c_int, C.c_double, C.c_int]
        L.fv2dh_load.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_int)]
        L.fv2dh_h5_read.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_double), C.c_long]
        L.fv2dh_h5_read.restype = C.c_long
        L.fv2dh_h5_attr.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_char_p, C.c_int]
        L.fv2dh_h5_list.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
        _LIB = L
    return _LIB


def _err() -> str:
    return (lib().fv2dh_last_error() or b"").decode()


def _src(ini: str, text: bool):
    return ini.encode(), 1 if text else 0


def read_ini(ini: str, text: bool = False) -> dict:
    """Parameters of a .ini file (or its text), with "warnings" and "last_ini"."""
    s, t = _src(ini, text)
    n = lib().fv2dh_read_ini(s, t, None, 0)
    if n < 0:
        raise HostError(_err())
    buf = C.create_string_buffer(n + 1)
    lib().fv2dh_read_ini(s, t, buf, n + 1)
    return json.loads(buf.value.decode())


def fv2d_params(ini: str, text: bool = False, precision: int = 0) -> Params:
    s, t = _src(ini, text)
    p = Params()
    if lib().fv2dh_fv2d_params(s, t, precision, C.byref(p)):
        raise HostError(_err())
    return p


def init_problem(ini: str, text: bool = False) -> np.ndarray:
    d = read_ini(ini, text)
    q = np.zeros((4, d["Ny"], d["Nx"]))
    s, t = _src(ini, text)
    if lib().fv2dh_init_problem(s, t, q.ctypes.data_as(C.POINTER(C.c_double))):
        raise HostError(_err())
    return q


def save(ini: str, q: np.ndarray, iteration: int, time: float, text: bool = False,
         force_truncate: bool = False) -> None:
    qa = np.ascontiguousarray(q, dtype=np.float64)
    s, t = _src(ini, text)
    if lib().fv2dh_save(s, t, qa.ctypes.data_as(C.POINTER(C.c_double)), iteration, time, int(force_truncate)):
        raise HostError(_err())


def load(ini: str, text: bool = False):
    """(q, time, iteration) of the run's restart_file."""
    d = read_ini(ini, text)
    q = np.zeros((4, d["Ny"], d["Nx"]))
    tm = C.c_double()
    it = C.c_int()
    s, t = _src(ini, text)
    if lib().fv2dh_load(s, t, q.ctypes.data_as(C.POINTER(C.c_double)), C.byref(tm), C.byref(it)):
        raise HostError(_err())
    return q, tm.value, it.value


def h5_read(file: str, name: str) -> np.ndarray:
    n = lib().fv2dh_h5_read(file.encode(), name.encode(), None, 0)
    if n < 0:
        raise HostError(_err())
    out = np.zeros(n)
    lib().fv2dh_h5_read(file.encode(), name.encode(), out.ctypes.data_as(C.POINTER(C.c_double)), n)
    return out


def h5_attr(file: str, obj: str, name: str, kind: str = "int"):
    k = {"int": 0, "double": 1, "string": 2}[kind]
    buf = C.create_string_buffer(4096)
    n = lib().fv2dh_h5_attr(file.encode(), obj.encode(), name.encode(), k, buf, 4096)
    if n < 0:
        raise HostError(_err())
    v = buf.value.decode()
    return int(v) if k == 0 else float(v) if k == 1 else v


def h5_list(file: str, group: str = "/") -> list[str]:
    buf = C.create_string_buffer(1 << 16)
    n = lib().fv2dh_h5_list(file.encode(), group.encode(), buf, 1 << 16)
    if n < 0:
        raise HostError(_err())
    return [x for x in buf.value.decode().split("\n") if x]


def snapshot(file: str, group: str = "") -> np.ndarray:
    """rho, u, v, prs of a snapshot as (4, Ny, Nx)."""
    nx = h5_attr(file, "/", "Nx")
    ny = h5_attr(file, "/", "Ny")
    pre = group.rstrip("/") + "/" if group else ""
    return np.stack([h5_read(file, pre + f).reshape(ny, nx) for f in ("rho", "u", "v", "prs")])


def run(ini: str, *args: str, cwd: str | None = None, timeout: float = 600) -> subprocess.CompletedProcess:
    """Run the fv2d executable on a .ini file."""
    if not os.path.exists(FV2D_EXE):
        raise RuntimeError(f"{FV2D_EXE} is missing (python -m fv2d_amd.build)")
    return subprocess.run([FV2D_EXE, ini, *args], cwd=cwd, capture_output=True, text=True, timeout=timeout)
