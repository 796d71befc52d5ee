"""C-ABI library checks that need no GPU: it loads, exports every function include/*.h
declares, its defaults are the reference's .ini defaults, rejected options fail with the
right code and message, and a create without a GPU fails cleanly."""
import ctypes as C
import os
import re

import pytest

from fv2d_amd import capi
from fv2d_amd.capi import Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(fv2d_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = capi.load_library()
    declared = header_functions()
    assert len(declared) >= 17
    for name in declared:
        assert hasattr(lib, name), f"{name} declared in include/ but not exported"
    assert sorted(declared) == sorted(capi.EXPORTED_SYMBOLS)


def test_abi_version_and_defaults():
    lib = capi.load_library()
    assert lib.fv2d_abi_version() == 1
    p = Params()
    lib.fv2d_params_default(C.byref(p))
    assert p.as_dict() == Params.defaults().as_dict()


@pytest.mark.parametrize("kw,code,frag", [
    (dict(tc_active=1, tc_mode=1), capi.FV2D_E_UNSUPPORTED, "B02"),
    (dict(Nx=1), capi.FV2D_E_INVALID, "Nx and Ny"),
    (dict(Nx=(1 << 26) + 1, Ny=2), capi.FV2D_E_INVALID, "wider than 2^26"),
    (dict(Nx=1 << 16, Ny=(1 << 14) + 1), capi.FV2D_E_INVALID, "2^30 cells"),
    (dict(reconstruction=2, Ng=1), capi.FV2D_E_INVALID, "Nghosts"),
    (dict(xmax=-1.0), capi.FV2D_E_INVALID, "empty domain"),
    (dict(riemann_solver=7), capi.FV2D_E_INVALID, "out of range"),
])
def test_rejected_options(kw, code, frag):
    lib = capi.load_library()
    p = Params.defaults(**kw)
    assert lib.fv2d_params_check(C.byref(p)) == code
    assert frag in lib.fv2d_last_error(None).decode()
    h = C.c_void_p()
    assert lib.fv2d_create(C.byref(p), C.byref(h)) == code and not h.value


def test_accepted_options():
    lib = capi.load_library()
    for kw in (dict(), dict(reconstruction=1, riemann_solver=2, gravity_mode=2, tc_active=1, bctc_ymin=1,
                            visc_active=1, time_stepping=1, precision=1)):
        assert lib.fv2d_params_check(C.byref(Params.defaults(**kw))) == 0


def test_create_without_gpu_fails_cleanly():
    from fv2d_amd import hiprt
    if hiprt.device_count() > 0:
        pytest.skip("a GPU is present")
    lib = capi.load_library()
    h = C.c_void_p()
    rc = lib.fv2d_create(C.byref(Params.defaults()), C.byref(h))
    assert rc in (capi.FV2D_E_HIP, capi.FV2D_E_NOMEM) and not h.value
    assert lib.fv2d_last_error(None)


def test_batch_config_default_and_job_bytes():
    lib = capi.load_library()
    c = capi.BatchConfig()
    lib.fv2d_batch_config_default(C.byref(c))
    d = capi.BatchConfig.defaults()
    for name, _ in capi.BatchConfig._fields_:
        if name != "reserved":
            assert getattr(c, name) == getattr(d, name), name
    # U, Us, Q0 (4 planes of pitch-64 rows) + one dt partial per work item, 256-byte aligned
    p = Params.defaults(Nx=100, Ny=30, precision=capi.FP32)
    state = -(-(4 * 128 * 30 * 4) // 256) * 256
    assert lib.fv2d_job_bytes(C.byref(p)) == 3 * state + 256  # 2 strips x 1 segment: 16 B of partials
    p64 = Params.defaults(Nx=4096, Ny=4096)
    nblk = (-(-4096 // 62)) * (4096 // 64)
    assert lib.fv2d_job_bytes(C.byref(p64)) == 3 * 4 * 4096 * 4096 * 8 + -(-(8 * nblk) // 256) * 256
    # invalid parameters have no footprint
    assert lib.fv2d_job_bytes(C.byref(Params.defaults(Nx=1))) == -1


@pytest.mark.parametrize("kw,frag", [
    (dict(precision=5), "precision"),
    (dict(max_jobs=0), "max_jobs"),
    (dict(max_jobs=2000), "max_jobs"),
    (dict(steps_per_launch=0), "steps_per_launch"),
    (dict(arena_bytes=16), "arena_bytes"),
])
def test_batch_config_rejected(kw, frag):
    lib = capi.load_library()
    c = capi.BatchConfig.defaults(**kw)
    b = C.c_void_p()
    assert lib.fv2d_batch_create(C.byref(c), C.byref(b)) == capi.FV2D_E_INVALID and not b.value
    assert frag in lib.fv2d_batch_last_error(None).decode()


def test_batch_null_handles():
    lib = capi.load_library()
    assert lib.fv2d_batch_launch(None, 1) == capi.FV2D_E_INVALID
    assert lib.fv2d_batch_wait(None, None) == capi.FV2D_E_INVALID
    assert lib.fv2d_batch_collect(None, 1, None, None, None) == capi.FV2D_E_INVALID
    lib.fv2d_batch_destroy(None)
    assert lib.fv2d_set_problem(None, None) == capi.FV2D_E_INVALID


def test_batch_create_without_gpu_fails_cleanly():
    from fv2d_amd import hiprt
    if hiprt.device_count() > 0:
        pytest.skip("a GPU is present")
    lib = capi.load_library()
    b = C.c_void_p()
    rc = lib.fv2d_batch_create(C.byref(capi.BatchConfig.defaults()), C.byref(b))
    assert rc in (capi.FV2D_E_HIP, capi.FV2D_E_NOMEM) and not b.value
    assert lib.fv2d_batch_last_error(None)
