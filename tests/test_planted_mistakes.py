"""Every planted mistake in the oracle (built with -DORACLE_PLANT=n, see PLANT(n) in
oracle/fv2d_oracle.c) must make at least one pin fail: the pins are not a restatement of
the oracle's formulas but constraints a plausible mistake violates."""
import pytest

import oracle_pins as PN
from oracle import oracle as OR

PLANTS = {
    1: "HLL star flux: wrong sign of the SL*SR*(uR-uL) term",
    2: "HLLC right star state takes the left tangential velocity",
    3: "reflecting x boundary does not negate the normal velocity",
    4: "minmod picks the larger magnitude",
    5: "gravity not passed through float",
    6: "gravity energy source uses FL - FR",
    7: "thermal conduction y flux difference with the wrong sign",
    8: "consToPrim uses gamma instead of gamma - 1",
    9: "viscous update divided by dx",
    10: "x and y slopes swapped",
    11: "RK2 averages with U* instead of U^n",
    12: "hyperbolic dt uses dx for the y term",
    13: "FSLP theta takes the smaller of |u|/c over the two states",
    14: "FSLP a_i takes the smaller acoustic impedance",
    15: "FSLP gravity term of u* with the wrong sign",
    16: "hot-bubble gravity evaluated at x instead of y",
    17: "hot-bubble gravity not passed through float",
    18: "hot-bubble gravity divides by the domain length instead of ymax",
    19: "fixed-temperature wall flux without the factor 2",
    20: "wall conduction modes replace the y fluxes FU / FD instead of FL / FR",
    21: "well-balanced wall flux with the wrong sign of rho g dy",
    22: "well-balanced bottom wall flux takes the pout of its own face",
}


@pytest.mark.parametrize("n", sorted(PLANTS), ids=[f"plant{n}" for n in sorted(PLANTS)])
def test_planted_mistake_is_caught(n):
    lib = OR.load(OR.build(n))
    failed = []
    for pin in PN.ALL_PINS:
        try:
            pin(lib)
        except AssertionError:
            failed.append(pin.__name__)
    assert failed, f"planted mistake {n} ({PLANTS[n]}) passes every pin"
