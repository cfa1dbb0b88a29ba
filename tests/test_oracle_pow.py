"""pow(norm, -0.2) of the step control (dp45.cu:120) is computed correctly
rounded (exponent = the double nearest -0.2), checked against 50-digit
Decimal arithmetic.  glibc's pow is within 1 ulp of it."""
import math
from decimal import Decimal, getcontext

import numpy as np


def exact(x):
    getcontext().prec = 50
    return float(Decimal(x) ** Decimal(-0.2))


def test_pow_correctly_rounded(olib, rng):
    xs = np.concatenate([np.exp(rng.uniform(-15, 15, 3000)), rng.uniform(1e-3, 10, 1000),
                         10.0 ** rng.uniform(-300, 300, 300), [5e-324, 2.2250738585072014e-308,
                                                              1.7976931348623157e308, 1.0, 2.0,
                                                              32.0, 1 / 32.0]])
    for x in map(float, xs):
        assert olib.oracle_pow_m02(x) == exact(x), x


def test_pow_special_values(olib):
    assert olib.oracle_pow_m02(0.0) == math.inf
    assert olib.oracle_pow_m02(math.inf) == 0.0
    assert math.isnan(olib.oracle_pow_m02(math.nan))
    assert math.isnan(olib.oracle_pow_m02(-1.0))


def test_glibc_pow_within_one_ulp(olib, rng):
    xs = np.exp(rng.uniform(-10, 10, 2000))
    for x in map(float, xs):
        a, b = olib.oracle_pow_m02(x), math.pow(x, -0.2)
        assert abs(a - b) <= math.ulp(a)
