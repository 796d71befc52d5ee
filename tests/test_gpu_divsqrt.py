"""The FP64 quotient and square-root helpers of the stage kernel on the GPU
(fv2d_amd/csrc/fv2d_math.h: DivBy, which reuses one refined reciprocal for several
numerators, sqrt_ieee, which skips the scaling steps for in-range inputs, and the
constant-divisor markstein_div, whose waves share the fast path) against the
compiler's own a / b and sqrt in the same kernel, bit for bit (NaNs as NaNs): operands
log-uniform over the whole exponent range with both signs, subnormals, zeros, infinities and
NaNs, numerator pairs of very different magnitude (the reuse check fails and a / b is taken),
and the magnitudes the hydro step divides."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "gpu_probe", "libdivsqrt_probe.so")


@pytest.fixture(scope="module")
def probe():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built (make -C tests/gpu_probe)")
    lib = C.CDLL(LIB)
    lib.divsqrt_probe.argtypes = [C.POINTER(C.c_double)] * 3 + [C.c_int, C.POINTER(C.c_double)]
    lib.divsqrt_probe.restype = C.c_int
    lib.markstein_probe.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_double, C.c_double, C.c_int,
                                    C.POINTER(C.c_double)]
    lib.markstein_probe.restype = C.c_int
    return lib


def run(lib, a0, a1, b):
    a0, a1, b = (np.ascontiguousarray(x, dtype=np.float64) for x in (a0, a1, b))
    out = np.zeros(6 * a0.size)
    dp = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    assert lib.divsqrt_probe(dp(a0), dp(a1), dp(b), a0.size, dp(out)) == 0
    return out.reshape(-1, 6)


def same_bits(x, y):
    return (x.view(np.uint64) == y.view(np.uint64)) | (np.isnan(x) & np.isnan(y))


def loguniform(rng, n, lo=-1074, hi=1023):
    e = rng.integers(lo, hi + 1, n)
    return np.ldexp(rng.uniform(1.0, 2.0, n), e) * rng.choice([-1.0, 1.0], n)


SPECIALS = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, -5e-324, 2.2250738585072014e-308,
                     1.7976931348623157e308, -1.7976931348623157e308, 2.0 ** -767, np.nextafter(2.0 ** -767, 0),
                     2.0 ** -969, 2.0 ** 969, 1.0, -1.0, 3.0, 0.1])


def check(out):
    bad_q0 = ~same_bits(out[:, 0], out[:, 2])
    bad_q1 = ~same_bits(out[:, 1], out[:, 3])
    bad_s = ~same_bits(out[:, 4], out[:, 5])
    assert not bad_q0.any(), f"first quotient differs at {np.flatnonzero(bad_q0)[:5]}"
    assert not bad_q1.any(), f"shared quotient differs at {np.flatnonzero(bad_q1)[:5]}"
    assert not bad_s.any(), f"sqrt differs at {np.flatnonzero(bad_s)[:5]}"


def test_random_full_range(probe):
    rng = np.random.default_rng(7)
    n = 1 << 20
    check(run(probe, loguniform(rng, n), loguniform(rng, n), loguniform(rng, n)))


def test_specials_all_pairs(probe):
    s = SPECIALS
    a0, a1, b = np.meshgrid(s, s, s, indexing="ij")
    check(run(probe, a0.ravel(), a1.ravel(), b.ravel()))


def test_numerators_far_apart(probe):
    """The second numerator far from the first (the scaled divisor differs near the ends of
    the exponent range), and quotients that overflow or become subnormal."""
    rng = np.random.default_rng(8)
    n = 1 << 18
    b = loguniform(rng, n, -60, 60)
    a0 = loguniform(rng, n, -20, 20)
    a1 = np.concatenate([loguniform(rng, n // 2, -1074, -900), loguniform(rng, n - n // 2, 900, 1023)])
    check(run(probe, a0, a1, b))
    b2 = np.concatenate([loguniform(rng, n // 2, -1074, -1000), loguniform(rng, n - n // 2, 1000, 1023)])
    check(run(probe, a0, a1, b2))


def test_hydro_magnitudes(probe):
    """Densities, momenta, pressures and wave-speed differences of the workloads."""
    rng = np.random.default_rng(9)
    n = 1 << 20
    rho = np.exp(rng.uniform(np.log(1e-6), np.log(1e3), n))
    mom = rng.normal(0.0, 1.0, n) * rho
    prs = np.exp(rng.uniform(np.log(1e-8), np.log(1e4), n))
    check(run(probe, mom, prs, rho))
    check(run(probe, prs * 1.4, mom * 0.5, rho))
    check(run(probe, rng.normal(0.0, 1e-3, n), rng.normal(0.0, 1e-9, n), rng.normal(0.0, 1.0, n)))


def div_mode(d):
    """The mode the solver picks for a cell size d (div_mode in fv2d_math.h)."""
    m, _ = np.frexp(d)
    if 2.0 ** -1022 <= d <= 2.0 ** 1022 and m == 0.5:
        return 2
    return 1 if 2.0 ** -60 <= d <= 2.0 ** 60 else 0


def test_markstein_on_device(probe):
    """The constant-divisor division (markstein_div) on the device, where lanes of one wave
    share the fast path or fall back together: cell sizes and random divisors, numerators
    over the whole range with zeros of both signs and special values mixed into every wave."""
    rng = np.random.default_rng(10)
    n = 1 << 18
    for d in (1.0 / 4096, 1.0 / 2048, 0.02 / 512, 1.5 / 3072, 0.5 / 1024) + tuple(
            float(np.ldexp(rng.uniform(1, 2), int(rng.integers(-60, 61)))) for _ in range(12)):
        a = np.concatenate([loguniform(rng, n), rng.uniform(-1e3, 1e3, n), np.zeros(64), -np.zeros(64),
                            np.tile(SPECIALS, 8)])
        rng.shuffle(a)
        a = np.ascontiguousarray(a)
        dp = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        # the mode the solver picks (one multiply for a power of two) and the Markstein mode
        for fast in sorted({div_mode(d), 1}):
            out = np.zeros(2 * a.size)
            assert probe.markstein_probe(dp(a), a.size, d, 1.0 / d, fast, dp(out)) == 0
            out = out.reshape(-1, 2)
            bad = ~same_bits(out[:, 0], out[:, 1])
            assert not bad.any(), f"d={d!r} mode {fast}: {np.count_nonzero(bad)} mismatches, e.g. " \
                                  f"a={a[np.flatnonzero(bad)[:3]]}"
            # whole waves of zeros and of ordinary values take the fast path
            z = np.ascontiguousarray(np.concatenate([np.zeros(256), -np.zeros(256), rng.uniform(1, 2, 256)]))
            out = np.zeros(2 * z.size)
            assert probe.markstein_probe(dp(z), z.size, d, 1.0 / d, fast, dp(out)) == 0
            assert same_bits(out[0::2], out[1::2]).all()
