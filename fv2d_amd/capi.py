"""ctypes mirror of the C ABI declared in include/fv2d_amd.h.

The structures and enums here follow the header field by field. The library itself is
``fv2d_amd/libfv2d_amd.so`` (hand-written HIP kernels for gfx950 plus the C-ABI layer);
``load_library`` fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfv2d_amd.so")

# Option values: same members and order as the reference enums (SimInfo.h:41-97).
HLL, HLLC, FSLP = 0, 1, 2
BC_ABSORBING, BC_REFLECTING, BC_PERIODIC = 0, 1, 2
TS_EULER, TS_RK2 = 0, 1
PCM, PCM_WB, PLM = 0, 1, 2
GRAV_NONE, GRAV_CONSTANT, GRAV_ANALYTICAL = 0, 1, 2
TCM_CONSTANT, TCM_B02 = 0, 1
BCTC_NONE, BCTC_FIXED_TEMPERATURE, BCTC_FIXED_GRADIENT = 0, 1, 2
FP64, FP32, FP32_STATE64 = 0, 1, 2

FV2D_OK = 0
FV2D_E_INVALID = -1
FV2D_E_UNSUPPORTED = -2
FV2D_E_HIP = -3
FV2D_E_NOMEM = -4
FV2D_E_STATE = -5
FV2D_E_BUSY = -6

PROB_SOD_X, PROB_BLAST, PROB_KH = 0, 1, 2
JOB_FREE, JOB_RUNNING, JOB_DONE = 0, 1, 2


class Params(C.Structure):
    """fv2d_params (include/fv2d_amd.h); the hot-path subset of DeviceParams/Params."""

    _fields_ = [
        ("Nx", C.c_int32),
        ("Ny", C.c_int32),
        ("Ng", C.c_int32),
        ("xmin", C.c_double),
        ("xmax", C.c_double),
        ("ymin", C.c_double),
        ("ymax", C.c_double),
        ("CFL", C.c_double),
        ("reconstruction", C.c_int32),
        ("riemann_solver", C.c_int32),
        ("time_stepping", C.c_int32),
        ("boundary_x", C.c_int32),
        ("boundary_y", C.c_int32),
        ("gamma0", C.c_double),
        ("well_balanced_flux_at_y_bc", C.c_int32),
        ("fslp_K", C.c_double),
        ("gravity_mode", C.c_int32),
        ("gx", C.c_double),
        ("gy", C.c_double),
        ("hot_bubble_g0", C.c_double),
        ("tc_active", C.c_int32),
        ("tc_mode", C.c_int32),
        ("kappa", C.c_double),
        ("bctc_ymin", C.c_int32),
        ("bctc_ymax", C.c_int32),
        ("bctc_ymin_value", C.c_double),
        ("bctc_ymax_value", C.c_double),
        ("visc_active", C.c_int32),
        ("mu", C.c_double),
        ("epsilon", C.c_double),
        ("epsilon_reset_negative", C.c_double),
        ("precision", C.c_int32),
        ("seg_rows", C.c_int32),
        ("chunk_rows", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]

    @classmethod
    def defaults(cls, **kw) -> "Params":
        """Reference .ini defaults (SimInfo.h:360-459, 536-551), then the given overrides.

        Computed in Python so that it needs no native library; the C function
        fv2d_params_default returns the same values (checked by the tests).
        """
        p = cls()
        p.Nx, p.Ny, p.Ng = 32, 32, 2
        p.xmin, p.xmax, p.ymin, p.ymax = 0.0, 1.0, 0.0, 1.0
        p.CFL = 0.8
        p.reconstruction = PCM
        p.riemann_solver = HLLC
        p.time_stepping = TS_EULER
        p.boundary_x = p.boundary_y = BC_REFLECTING
        p.gamma0 = 5.0 / 3.0
        p.well_balanced_flux_at_y_bc = 0
        p.fslp_K = 1.1
        p.gravity_mode = GRAV_NONE
        p.gx = p.gy = 0.0
        p.hot_bubble_g0 = 0.0
        p.tc_active = 0
        p.tc_mode = TCM_CONSTANT
        p.kappa = 0.0
        p.bctc_ymin = p.bctc_ymax = BCTC_NONE
        p.bctc_ymin_value = p.bctc_ymax_value = 1.0
        p.visc_active = 0
        p.mu = 0.0
        p.epsilon = 1.0e-6
        p.epsilon_reset_negative = 1.0e-8
        p.precision = FP64
        p.seg_rows = 0
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(f"fv2d_params has no field {k!r}")
            setattr(p, k, v)
        return p

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}

    def copy(self, **kw) -> "Params":
        q = Params()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(Params))
        for k, v in kw.items():
            setattr(q, k, v)
        return q


class StepStats(C.Structure):
    _fields_ = [
        ("negative_density", C.c_uint64),
        ("negative_pressure", C.c_uint64),
        ("nan_count", C.c_uint64),
    ]

    def as_tuple(self):
        return (int(self.negative_density), int(self.negative_pressure), int(self.nan_count))


class Problem(C.Structure):
    """fv2d_problem: an initial condition generated on the device."""

    _fields_ = [
        ("kind", C.c_int32),
        ("reserved0", C.c_int32),
        ("seed", C.c_uint64),
        ("amplitude", C.c_double),
        ("reserved", C.c_double * 5),
    ]


class BatchConfig(C.Structure):
    """fv2d_batch_config."""

    _fields_ = [
        ("precision", C.c_int32),
        ("max_jobs", C.c_int32),
        ("arena_bytes", C.c_int64),
        ("steps_per_launch", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("max_items", C.c_int32),
        ("use_graph", C.c_int32),
        ("reserved", C.c_int32 * 8),
    ]

    @classmethod
    def defaults(cls, **kw) -> "BatchConfig":
        """The values of fv2d_batch_config_default, then the given overrides."""
        c = cls()
        c.precision = FP32
        c.max_jobs = 256
        c.arena_bytes = 8 << 30
        c.steps_per_launch = 8
        c.blocks_per_cu = 0
        c.max_items = 1 << 18
        c.use_graph = 1
        for k, v in kw.items():
            if not hasattr(c, k):
                raise AttributeError(f"fv2d_batch_config has no field {k!r}")
            setattr(c, k, v)
        return c


class JobInfo(C.Structure):
    """fv2d_job_info."""

    _fields_ = [
        ("id", C.c_int64),
        ("status", C.c_int32),
        ("nsteps", C.c_int32),
        ("t", C.c_double),
        ("dt_last", C.c_double),
        ("negative_density", C.c_uint64),
        ("negative_pressure", C.c_uint64),
        ("nan_count", C.c_uint64),
    ]


_LIB = None


def load_library(path: str | None = None) -> C.CDLL:
    """Load libfv2d_amd.so and declare its signatures. Raises if it is not built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: the HIP library is not built (run __graft_entry__.build() "
            "or python -m fv2d_amd.build); there is no CPU fallback"
        )
    lib = C.CDLL(p)
    declare(lib)
    if path is None:
        _LIB = lib
    return lib


def declare(lib: C.CDLL) -> None:
    P, H = C.POINTER, C.c_void_p
    dp = P(C.c_double)
    sig = {
        "fv2d_abi_version": (C.c_int, []),
        "fv2d_params_default": (None, [P(Params)]),
        "fv2d_params_check": (C.c_int, [P(Params)]),
        "fv2d_create": (C.c_int, [P(Params), P(H)]),
        "fv2d_destroy": (None, [H]),
        "fv2d_last_error": (C.c_char_p, [H]),
        "fv2d_set_prim": (C.c_int, [H, dp]),
        "fv2d_compute_dt": (C.c_int, [H, dp, dp]),
        "fv2d_step": (C.c_int, [H, C.c_double, P(StepStats)]),
        "fv2d_advance": (C.c_int, [H, dp, C.c_double, C.c_int32, dp, P(C.c_int32), P(StepStats)]),
        "fv2d_get_prim": (C.c_int, [H, dp]),
        "fv2d_get_cons": (C.c_int, [H, dp]),
        "fv2d_sums": (C.c_int, [H, dp]),
        "fv2d_synchronize": (C.c_int, [H]),
        "fv2d_get_stream": (C.c_void_p, [H]),
        "fv2d_profile": (C.c_int, [H, C.c_int]),
        "fv2d_profile_read": (C.c_int, [H, dp, P(C.c_int64)]),
        "fv2d_set_problem": (C.c_int, [H, P(Problem)]),
        "fv2d_batch_config_default": (None, [P(BatchConfig)]),
        "fv2d_job_bytes": (C.c_int64, [P(Params)]),
        "fv2d_batch_create": (C.c_int, [P(BatchConfig), P(H)]),
        "fv2d_batch_destroy": (None, [H]),
        "fv2d_batch_last_error": (C.c_char_p, [H]),
        "fv2d_batch_submit": (C.c_int, [H, P(Params), P(Problem), dp, C.c_double, C.c_int32, P(C.c_int64)]),
        "fv2d_batch_launch": (C.c_int, [H, C.c_int32]),
        "fv2d_batch_wait": (C.c_int, [H, P(C.c_int32)]),
        "fv2d_batch_poll": (C.c_int, [H, P(C.c_int64), C.c_int32, P(C.c_int32)]),
        "fv2d_batch_job_info": (C.c_int, [H, C.c_int64, P(JobInfo)]),
        "fv2d_batch_collect": (C.c_int, [H, C.c_int64, dp, dp, P(JobInfo)]),
        "fv2d_batch_cancel": (C.c_int, [H, C.c_int64]),
        "fv2d_batch_last_launch_ms": (C.c_int, [H, dp]),
        "fv2d_run_noalloc": (C.c_int, [P(Params), dp, dp, dp, C.c_double, C.c_int32, dp, P(C.c_int32),
                                       P(StepStats), C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


# Every symbol include/fv2d_amd.h declares (checked against the built library by the tests).
EXPORTED_SYMBOLS = (
    "fv2d_abi_version",
    "fv2d_params_default",
    "fv2d_params_check",
    "fv2d_create",
    "fv2d_destroy",
    "fv2d_last_error",
    "fv2d_set_prim",
    "fv2d_compute_dt",
    "fv2d_step",
    "fv2d_advance",
    "fv2d_get_prim",
    "fv2d_get_cons",
    "fv2d_sums",
    "fv2d_synchronize",
    "fv2d_get_stream",
    "fv2d_profile",
    "fv2d_profile_read",
    "fv2d_set_problem",
    "fv2d_batch_config_default",
    "fv2d_job_bytes",
    "fv2d_batch_create",
    "fv2d_batch_destroy",
    "fv2d_batch_last_error",
    "fv2d_batch_submit",
    "fv2d_batch_launch",
    "fv2d_batch_wait",
    "fv2d_batch_poll",
    "fv2d_batch_job_info",
    "fv2d_batch_collect",
    "fv2d_batch_cancel",
    "fv2d_batch_last_launch_ms",
    "fv2d_run_noalloc",
)
