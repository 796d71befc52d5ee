"""The constant-divisor division of the FP64 path (markstein_div in fv2d_math.h) against
IEEE division, bit for bit: the cell sizes of every workload and random divisors, with
numerators spread over the whole exponent range (log-uniform), both signs, zeros, and
the special values that take the IEEE fallback."""
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
    for d in divisors():
        a = np.ascontiguousarray(numerators(rng, 200_000))
        assert lib.emu_markstein_mismatches(EMU._dp(a), a.size, d) == 0, d


def test_random_divisors(lib):
    rng = np.random.default_rng(1)
    for _ in range(200):
        d = float(np.ldexp(rng.uniform(1.0, 2.0), int(rng.integers(-60, 61))))
        a = np.ascontiguousarray(numerators(rng, 20_000))
        assert lib.emu_markstein_mismatches(EMU._dp(a), a.size, d) == 0, d
