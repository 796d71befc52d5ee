import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests", "emu")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as OR
    if not os.path.exists(OR.LIB_PATH):
        OR.build()
    return OR.load()


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library; a GPU test fails loudly (no fallback) when it is missing."""
    from fv2d_amd import capi
    return capi.load_library()
