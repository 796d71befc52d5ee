"""The CPU oracle against closed forms, invariants, the exact Riemann solver and the one
value the reference publishes (README.md:86). See tests/oracle_pins.py."""
import pytest

import oracle_pins as PN


@pytest.mark.parametrize("pin", PN.ALL_PINS, ids=[f.__name__ for f in PN.ALL_PINS])
def test_pin(pin, oracle_lib):
    pin(oracle_lib)
