"""The constant-divisor division of the FP64 path (markstein_div in fv2d_math.h) against
IEEE division, bit for bit: the cell sizes of every workload and random divisors, with
numerators spread over the whole exponent range (log-uniform), both signs, zeros, and
the special values that take the IEEE fallback; each divisor in the mode the solver picks
for it (div_mode: one multiply for powers of two, Markstein's correction otherwise) and
in the Markstein mode as well."""
import numpy as np
import pytest

import emu as EMU
from fv2d_amd import problems as P


@pytest.fixture(scope="module")
def lib():
    EMU.build()
    return EMU.load()


def numerators(rng, n):
    e = rng.uniform(-1074, 1023, n)
    m = rng.uniform(1.0, 2.0, n)
    a = np.ldexp(m, e.astype(int)) * rng.choice([-1.0, 1.0], n)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, -5e-324, 2.0 ** -960, 2.0 ** 960,
                        np.nextafter(2.0 ** -960, 0), np.nextafter(2.0 ** 960, np.inf), 1.0, -1.0, 3.0])
    near = rng.uniform(-1e3, 1e3, n // 4)  # the magnitudes a flux difference times dt takes
    return np.concatenate([a, special, near, near * 1e-9])


def divisors():
    ds = set()
    for name, size in (("sod", {}), ("sedov", dict(N=2048)), ("kh", dict(N=4096)), ("rt", {})):
        p = P.WORKLOADS[name][0](**size)
        ds.add((p.xmax - p.xmin) / p.Nx)
        ds.add((p.ymax - p.ymin) / p.Ny)
    for N in (37, 64, 100, 128, 256, 512, 1000, 1024, 3072):
        ds.add(1.0 / N)
        ds.add(1.2 / N)
        ds.add(0.9 / N)
    return sorted(ds)


def test_cell_sizes(lib):
    rng = np.random.default_rng(0)
    modes = set()
    for d in divisors():
        a = np.ascontiguousarray(numerators(rng, 200_000))
        assert lib.emu_markstein_mismatches(EMU._dp(a), a.size, d, -1) == 0, d
        assert lib.emu_markstein_mismatches(EMU._dp(a), a.size, d, 1) == 0, d
        modes.add(lib.emu_div_mode(d))
    assert modes == {1, 2}  # KH 4096^2 (dx = 2^-12) takes the exact multiply, Sod's dy Markstein


def test_power_of_two_divisors(lib):
    """Powers of two over the whole range where the reciprocal is normal: one multiply."""
    rng = np.random.default_rng(2)
    for e in list(range(-1022, 1023, 37)) + [-1022, -60, -12, 0, 60, 1022]:
        d = 2.0 ** e
        assert lib.emu_div_mode(d) == 2, e
        a = np.ascontiguousarray(numerators(rng, 20_000))
        assert lib.emu_markstein_mismatches(EMU._dp(a), a.size, d, 2) == 0, e
    for d in (2.0 ** -1023, 2.0 ** 1023, 3.0 * 2.0 ** -12, 0.1):
        assert lib.emu_div_mode(d) != 2, d


def test_random_divisors(lib):
    rng = np.random.default_rng(1)
    for _ in range(200):
        d = float(np.ldexp(rng.uniform(1.0, 2.0), int(rng.integers(-60, 61))))
        a = np.ascontiguousarray(numerators(rng, 20_000))
        assert lib.emu_markstein_mismatches(EMU._dp(a), a.size, d, -1) == 0, d
