"""Error behaviour of the C ABI on a live device: calls in the wrong order, bad arguments
and busy batches return the documented codes with a message, and leave the handle usable."""
import ctypes as C

import numpy as np
import pytest

from fv2d_amd import capi
from fv2d_amd import problems as P
from fv2d_amd import serve as S
from fv2d_amd.capi import FP32, FP64, BatchConfig, JobInfo, Params, Problem
from fv2d_amd.solver import Batch, Fv2dError, Solver

pytestmark = pytest.mark.gpu


def test_single_handle_order_and_arguments(gpu_lib):
    p = Params.defaults(Nx=16, Ny=8)
    with Solver(p, gpu_lib) as s:
        for call in (lambda: s.step(1e-3), lambda: s.compute_dt(), lambda: s.get_prim(), lambda: s.sums(),
                     lambda: s.advance(max_steps=2)):
            with pytest.raises(Fv2dError) as e:
                call()
            assert e.value.code == capi.FV2D_E_STATE
        assert gpu_lib.fv2d_set_prim(s.h, None) == capi.FV2D_E_INVALID
        assert b"NULL" in gpu_lib.fv2d_last_error(s.h)
        assert gpu_lib.fv2d_set_problem(s.h, C.byref(Problem(kind=9))) == capi.FV2D_E_INVALID
        t = C.c_double(0.0)
        assert gpu_lib.fv2d_advance(s.h, None, 1.0, 1, None, None, None) == capi.FV2D_E_INVALID
        # a loop with neither an end time nor a step limit is refused (it would not end)
        s.set_problem(capi.PROB_SOD_X)
        for tend in (float("inf"), 1e300, float("nan")):
            assert gpu_lib.fv2d_advance(s.h, C.byref(t), tend, 0, None, None, None) == capi.FV2D_E_INVALID
            assert b"no end" in gpu_lib.fv2d_last_error(s.h)
        with pytest.raises(ValueError):
            s.advance()
        # still usable
        s.set_problem(capi.PROB_SOD_X)
        s.advance(max_steps=3)
        assert np.isfinite(s.get_prim()).all()
        del t


def test_batch_submit_errors(gpu_lib):
    sp = S.JobSpec(0, 0.0, capi.PROB_KH, 32, 16, 1.5, 5e-3, 1, 4)
    with Batch(BatchConfig.defaults(precision=FP32, max_jobs=2, arena_bytes=64 << 20), lib=gpu_lib) as b:
        with pytest.raises(Fv2dError) as e:  # precision differs from the batch's
            b.submit(S.job_params(sp, FP64), problem=S.job_problem(sp), max_steps=4)
        assert e.value.code == capi.FV2D_E_INVALID and "precision" in str(e.value)
        jid = C.c_int64()
        p = S.job_params(sp, FP32)
        q = np.zeros((4, 16, 32))
        rc = gpu_lib.fv2d_batch_submit(b.b, C.byref(p), C.byref(S.job_problem(sp)),
                                       q.ctypes.data_as(C.POINTER(C.c_double)), 1e300, 4, C.byref(jid))
        assert rc == capi.FV2D_E_INVALID  # both prob and q
        for tend in (float("inf"), 1e300):  # neither an end time nor a step limit
            rc = gpu_lib.fv2d_batch_submit(b.b, C.byref(p), C.byref(S.job_problem(sp)), None, tend, 0, C.byref(jid))
            assert rc == capi.FV2D_E_INVALID and b"no end" in gpu_lib.fv2d_batch_last_error(b.b)
        with pytest.raises(ValueError):
            b.submit(p, problem=S.job_problem(sp))
        with pytest.raises(Fv2dError):
            b.submit(p.copy(Nx=1), problem=S.job_problem(sp), max_steps=4)  # invalid parameters
        j = b.submit(p, problem=S.job_problem(sp), max_steps=4)
        with pytest.raises(Fv2dError) as e:  # not finished yet
            b.collect(j)
        assert e.value.code == capi.FV2D_E_STATE
        with pytest.raises(Fv2dError):
            b.info(12345)  # no such job
        b.run(1)
        q, sums, inf = b.collect(j)
        assert inf.nsteps == 4 and np.isfinite(q).all()
        with pytest.raises(Fv2dError):
            b.collect(j)  # collected already


def test_batch_job_that_never_fits(gpu_lib):
    sp = S.JobSpec(0, 0.0, capi.PROB_SOD_X, 512, 512, 1.4, 1e-3, 1, 2)
    with Batch(BatchConfig.defaults(precision=FP32, max_jobs=4, arena_bytes=1 << 20), lib=gpu_lib) as b:
        with pytest.raises(Fv2dError) as e:
            b.submit(S.job_params(sp, FP32), problem=S.job_problem(sp), max_steps=2)
        assert e.value.code == capi.FV2D_E_BUSY and "arena" in str(e.value)
    with pytest.raises(RuntimeError, match="does not fit an empty batch"):
        S.serve_batched([sp], FP32, BatchConfig.defaults(precision=FP32, arena_bytes=1 << 20), realtime=False,
                        lib=gpu_lib)


def test_unsupported_option_is_rejected_at_create(gpu_lib):
    with pytest.raises(Fv2dError) as e:
        Solver(Params.defaults(tc_active=1, tc_mode=capi.TCM_B02), gpu_lib)
    assert e.value.code == capi.FV2D_E_UNSUPPORTED and "B02" in str(e.value)


@pytest.mark.parametrize("enable", [1, 6, 40])
def test_stage_profiling_counts(enable, gpu_lib):
    """fv2d_profile: every stage launch of an RK2 step (2 per step) is timed, whether the events
    were created up front (enable = launch count) or some are created as the launches need
    them (enable = 1 or fewer than the launches); a read resets the count; enable < 0 is
    refused and leaves profiling as it was."""
    p = Params.defaults(Nx=64, Ny=32, time_stepping=1)
    with Solver(p, gpu_lib) as s:
        s.set_problem(capi.PROB_SOD_X)
        assert gpu_lib.fv2d_profile(s.h, -1) == capi.FV2D_E_INVALID
        assert b"enable" in gpu_lib.fv2d_last_error(s.h)
        assert gpu_lib.fv2d_profile(s.h, enable) == 0
        ms, n = C.c_double(), C.c_int64()
        for steps in (10, 3):
            s.advance(max_steps=steps, record_dts=False)
            assert gpu_lib.fv2d_profile_read(s.h, C.byref(ms), C.byref(n)) == 0
            assert n.value == 2 * steps and ms.value > 0.0
        assert gpu_lib.fv2d_profile(s.h, 0) == 0
        s.advance(max_steps=2, record_dts=False)
        assert gpu_lib.fv2d_profile_read(s.h, C.byref(ms), C.byref(n)) == 0
        assert n.value == 0 and ms.value == 0.0
