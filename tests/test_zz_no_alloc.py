"""fv2d_run_noalloc, the entry point that allocates nothing, against the handle path it is a
variant of (fv2d_create + fv2d_set_prim + fv2d_advance + fv2d_get_prim): same step count,
same dt sequence, same end time, same reset counts and bitwise the same primitives, on
calls in a row whose later grids are smaller (the scratch array keeps the previous call's
contents), for both arithmetic forms and analytical gravity; and its limits and argument
errors return FV2D_E_INVALID with a message.

The test allocates the device buffers q_in / q_out itself (the caller's job); the entry
point takes only their addresses."""
import ctypes as C

import numpy as np
import pytest

from fv2d_amd import capi, hiprt
from fv2d_amd import problems as P
from fv2d_amd.capi import FP32, GRAV_ANALYTICAL, Params, StepStats
from fv2d_amd.solver import Solver

pytestmark = pytest.mark.gpu


class DeviceArray:
    """A device buffer of float64 (hipMalloc'd by the test, not by the library)."""

    def __init__(self, n):
        rt = hiprt.rt()
        rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        rt.hipFree.argtypes = [C.c_void_p]
        self.n = n
        self.ptr = C.c_void_p()
        assert rt.hipMalloc(C.byref(self.ptr), 8 * n) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        assert a.size == self.n
        assert hiprt.rt().hipMemcpy(self.ptr, a.ctypes.data, 8 * self.n, 1) == 0  # host to device

    def download(self):
        out = np.empty(self.n)
        assert hiprt.rt().hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, 2) == 0  # device to host
        return out

    def free(self):
        hiprt.rt().hipFree(self.ptr)


def dp(ptr):
    return C.cast(ptr, C.POINTER(C.c_double))


def run_noalloc(lib, p, q0, tend, max_steps):
    n = 4 * p.Nx * p.Ny
    qi, qo = DeviceArray(n), DeviceArray(n)
    try:
        qi.upload(q0)
        t = C.c_double(0.0)
        ns = C.c_int32(0)
        st = StepStats()
        dts = np.zeros(max(max_steps, 1))
        rc = lib.fv2d_run_noalloc(C.byref(p), dp(qi.ptr), dp(qo.ptr), C.byref(t), float(tend), int(max_steps),
                                  dts.ctypes.data_as(C.POINTER(C.c_double)) if max_steps > 0 else None,
                                  C.byref(ns), C.byref(st), None)
        assert rc == 0, lib.fv2d_last_error(None)
        q = qo.download().reshape(4, p.Ny, p.Nx)
    finally:
        qi.free()
        qo.free()
    return q, t.value, dts[: ns.value], ns.value, st.as_tuple()


def run_handle(lib, p, q0, tend, max_steps):
    with Solver(p, lib) as s:
        s.set_prim(q0)
        t, dts, counts, n = s.advance(tend=tend, max_steps=max_steps)
        return s.get_prim(), t, dts, n, counts


CASES = [  # (name, params, steps, tend): each later grid is smaller than an earlier one
    ("kh 256x192 fp64", P.workload("kh", N=256)[0].copy(Ny=192), 24, 1e300),
    ("rt 64x192 fp64 analytical gravity", P.workload("rt", Nx=64, Ny=192)[0].copy(gravity_mode=GRAV_ANALYTICAL,
                                                                                  hot_bubble_g0=-0.1), 12, 1e300),
    ("sod 512x8 fp64 to t=0.05", P.workload("sod")[0], 1000, 0.05),
    ("kh 128x96 fp32", P.workload("kh", N=128)[0].copy(Ny=96, precision=FP32), 16, 1e300),
    ("sedov 64x64 fp32", P.workload("sedov", N=64)[0].copy(precision=FP32), 10, 1e300),
]


def _ic(p):
    return P.random_state(p, seed=p.Nx * 7 + p.Ny) if p.gravity_mode == GRAV_ANALYTICAL else None


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_equals_handle_path(order, gpu_lib):
    for name, p, steps, tend in CASES:
        q0 = _ic(p)
        if q0 is None:
            q0 = {"kh": P.kh_ic, "sod": P.sod_ic, "sedov": P.sedov_ic, "rt": P.rt_ic}[name.split()[0]](p)
        print(f"noalloc case: {name}", flush=True)
        qa, ta, da, na, ca = run_noalloc(gpu_lib, p, q0, tend, steps)
        qb, tb, db, nb, cb = run_handle(gpu_lib, p, q0, tend, steps)
        assert na == nb and na > 0, (name, na, nb)
        np.testing.assert_array_equal(da, db, err_msg=name)
        assert ta == tb and ca == cb, (name, ta, tb, ca, cb)
        assert np.array_equal(qa, qb), (name, float(np.abs(qa - qb).max()))


def test_noalloc_twice_second_smaller(gpu_lib):
    """Two calls in a row, the second on a smaller grid, each against the handle path."""
    for N, steps in ((192, 20), (48, 9)):
        p, q0, _ = P.workload("kh", N=N)
        qa, ta, da, na, _ = run_noalloc(gpu_lib, p, q0, 1e300, steps)
        qb, tb, db, nb, _ = run_handle(gpu_lib, p, q0, 1e300, steps)
        assert na == nb == steps and ta == tb
        np.testing.assert_array_equal(da, db)
        assert np.array_equal(qa, qb), N


def test_noalloc_limits_and_errors(gpu_lib):
    lib = gpu_lib
    q = DeviceArray(4 * 64 * 64)
    try:
        t = C.c_double(0.0)
        p = Params.defaults(Nx=64, Ny=64)
        # no end
        assert lib.fv2d_run_noalloc(C.byref(p), dp(q.ptr), dp(q.ptr), C.byref(t), 1e300, 0, None, None, None,
                                    None) == capi.FV2D_E_INVALID
        assert b"no end" in lib.fv2d_last_error(None)
        # more steps than the dt log holds
        assert lib.fv2d_run_noalloc(C.byref(p), dp(q.ptr), dp(q.ptr), C.byref(t), 1e300, (1 << 14) + 1, None, None,
                                    None, None) == capi.FV2D_E_INVALID
        assert b"max_steps" in lib.fv2d_last_error(None)
        # a grid beyond the scratch size (pitch 1088 x 1024 rows > 2^20 cells); nothing runs
        big = Params.defaults(Nx=1030, Ny=1024)
        assert lib.fv2d_run_noalloc(C.byref(big), dp(q.ptr), dp(q.ptr), C.byref(t), 1e300, 1, None, None, None,
                                    None) == capi.FV2D_E_INVALID
        assert b"scratch limit" in lib.fv2d_last_error(None)
        # NULL buffers
        assert lib.fv2d_run_noalloc(C.byref(p), None, dp(q.ptr), C.byref(t), 1e300, 1, None, None, None,
                                    None) == capi.FV2D_E_INVALID
    finally:
        q.free()
