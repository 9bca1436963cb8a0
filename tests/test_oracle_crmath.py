"""The correctly rounded powers of the contact models (the oracle's copy; the
engine's device copy performs the same IEEE operations in the same order, and
the GPU contact tests compare forces and trajectories bit for bit).

Each routine is checked against the exact value rounded to the nearest
double, computed with 60-digit decimal arithmetic, over the magnitudes the
Hertz and JKR models produce, and against pow's special values.  libm's pow
is within an ulp of the same references (not always equal to them)."""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

from oracle import binding as orc

getcontext().prec = 60
THIRD = Decimal(1.0 / 3.0)  # the double nearest 1/3, exactly


def exact(which, x):
    d = Decimal(float(x))
    if which == "1.5":
        return d * d.sqrt()
    if which == "3":
        return d * d * d
    if which == "0.5":
        return d.sqrt()
    return (d.ln() * THIRD).exp()


@pytest.mark.parametrize("which,lo,hi,signed", [("1.5", -30, 10, False), ("1/3", -300, 300, False),
                                                ("3", -100, 100, True), ("0.5", -300, 300, False)])
def test_correctly_rounded(which, lo, hi, signed):
    rng = np.random.default_rng(20261203)
    xs = 10.0 ** rng.uniform(lo, hi, 3000)
    if signed:
        xs = xs * rng.choice([-1.0, 1.0], xs.size)
    got = orc.cr_pow(which, xs)
    ref = np.array([float(exact(which, x)) for x in xs])
    assert np.array_equal(got, ref)
    y = {"1.5": 1.5, "1/3": 1.0 / 3.0, "3": 3.0, "0.5": 0.5}[which]
    lib = np.array([math.pow(x, y) for x in xs])
    ulp = np.abs(lib - ref) / np.spacing(np.abs(ref))
    assert np.all(ulp <= 1.0)  # libm agrees to an ulp


def test_exact_and_special_values():
    x = np.array([0.0, -0.0, 1.0, 8.0, 27.0, 64.0, 5e-324, 1e300, math.inf])
    # x^(1/3 - 2^-54/3): perfect cubes sit just below the integer root
    assert [repr(float(v)) for v in orc.cr_pow("1/3", x)] == [repr(float(v)) for v in (
        0.0, 0.0, 1.0, 2.0, 3.0, 3.9999999999999996, 1.7031839360032837e-108, 9.999999999999872e+99, math.inf)]
    for which, y in (("1.5", 1.5), ("1/3", 1.0 / 3.0), ("3", 3.0), ("0.5", 0.5)):
        for v in (0.0, -0.0, math.inf, -2.0, math.nan):
            try:
                want = math.pow(v, y)
            except ValueError:
                want = math.nan
            got = float(orc.cr_pow(which, np.array([v]))[0])
            assert (math.isnan(want) and math.isnan(got)) or repr(got) == repr(want), (which, v, got, want)
    assert orc.cr_pow("3", np.array([-3.0]))[0] == -27.0
