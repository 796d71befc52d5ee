"""Minimal ctypes access to the HIP runtime: events and device synchronisation.

Used to time work on a solver's own stream (fv2d_get_stream) with HIP events, and to
bracket timed regions with a device-wide synchronisation.
"""
from __future__ import annotations

import ctypes as C
import os

_RT = None


def rt():
    global _RT
    if _RT is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _RT = C.CDLL(name)
                break
            except OSError:
                continue
        if _RT is None:
            raise RuntimeError("libamdhip64.so not found")
        _RT.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _RT.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _RT.hipEventSynchronize.argtypes = [C.c_void_p]
        _RT.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _RT.hipEventDestroy.argtypes = [C.c_void_p]
        _RT.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _RT.hipGetErrorString.restype = C.c_char_p
        _RT.hipGetErrorString.argtypes = [C.c_int]
    return _RT


def _ok(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what}: {rt().hipGetErrorString(rc).decode()}")


def device_count() -> int:
    n = C.c_int(0)
    rc = rt().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def synchronize() -> None:
    _ok(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Event:
    def __init__(self):
        self.e = C.c_void_p()
        _ok(rt().hipEventCreate(C.byref(self.e)), "hipEventCreate")

    def record(self, stream) -> None:
        _ok(rt().hipEventRecord(self.e, stream), "hipEventRecord")

    def elapsed_ms(self, end: "Event") -> float:
        _ok(rt().hipEventSynchronize(end.e), "hipEventSynchronize")
        ms = C.c_float()
        _ok(rt().hipEventElapsedTime(C.byref(ms), self.e, end.e), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            rt().hipEventDestroy(self.e)
        except Exception:
            pass
