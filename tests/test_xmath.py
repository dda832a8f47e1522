"""The shared numerics (medyan_amd/csrc/mdy_xmath.h) that the HIP kernels and
the oracle's exact mode both use: atan and exp within one ulp of the C
library's (correctly rounded in almost all cases), and an exact
order-independent double sum equal to math.fsum (exactly rounded)."""
import math

import numpy as np
import pytest

RNG = np.random.default_rng(7)


def ulp_diff(a, b):
    a = np.asarray(a, np.float64).view(np.int64)
    b = np.asarray(b, np.float64).view(np.int64)
    return np.abs(a - b)


@pytest.mark.parametrize("scale", [1e-6, 1e-2, 0.5, 1.0, 3.0, 1e3, 1e12])
def test_atan_within_one_ulp(oracle, scale):
    x = RNG.standard_normal(200000) * scale
    y = oracle.xmath_atan(x)
    assert ulp_diff(y, np.arctan(x)).max() <= 1
    # odd symmetry holds exactly
    np.testing.assert_array_equal(oracle.xmath_atan(-x), -y)


def test_atan_special_values(oracle):
    x = np.array([0.0, -0.0, np.inf, -np.inf, 1.0, -1.0, 1e300, 5e-324])
    y = oracle.xmath_atan(x)
    assert y[0] == 0.0 and y[2] == pytest.approx(math.pi / 2, rel=1e-16) and y[3] == -y[2]
    assert y[4] == pytest.approx(math.pi / 4, rel=1e-16)
    assert y[7] == 5e-324
    assert np.isnan(oracle.xmath_atan(np.array([np.nan]))[0])


@pytest.mark.parametrize("lo,hi", [(-700, -100), (-100, -1), (-1, 1), (-1e-8, 1e-8), (1, 100), (100, 709)])
def test_exp_within_one_ulp(oracle, lo, hi):
    x = RNG.uniform(lo, hi, 200000)
    assert ulp_diff(oracle.xmath_exp(x), np.exp(x)).max() <= 1


def test_exp_special_values(oracle):
    y = oracle.xmath_exp(np.array([0.0, -np.inf, np.inf, 710.0, -746.0, 1.0]))
    assert y[0] == 1.0 and y[1] == 0.0 and y[2] == np.inf and y[3] == np.inf and y[4] == 0.0
    assert y[5] == pytest.approx(math.e, rel=1e-16)


@pytest.mark.parametrize("kind", ["normal", "cancel", "wide", "large_n"])
def test_xsum_equals_fsum(oracle, kind):
    if kind == "normal":
        x = RNG.standard_normal(100000) * 1e3
    elif kind == "cancel":
        a = RNG.standard_normal(50000) * 1e12
        x = np.concatenate([a, -a, RNG.standard_normal(1000) * 1e-3])
        RNG.shuffle(x)
    elif kind == "wide":
        x = RNG.standard_normal(20000) * 10.0 ** RNG.integers(-30, 30, 20000)
    else:
        x = RNG.uniform(0, 2 ** 40, 3 * 10 ** 6)   # beyond 2^15 additions between renormalizations
    ref = math.fsum(x)
    assert oracle.xsum(x) == ref
    assert oracle.xsum(x, stride=13) == ref
    assert oracle.xsum(x[::-1].copy()) == ref


def test_xsum_truncates_below_grid(oracle):
    """Terms are placed on a grid of unit 2^-144, truncated toward zero."""
    for scale in (1e-40, 1e-310):
        x = RNG.standard_normal(10000) * scale
        on_grid = np.trunc(np.ldexp(x, 144))
        assert oracle.xsum(x) == math.ldexp(math.fsum(on_grid), -144)


def test_xsum_nonfinite(oracle):
    assert oracle.xsum(np.array([1.0, np.inf, 2.0])) == np.inf
    assert oracle.xsum(np.array([1.0, -np.inf])) == -np.inf
    assert np.isnan(oracle.xsum(np.array([np.inf, -np.inf])))
    assert np.isnan(oracle.xsum(np.array([1.0, np.nan])))
    assert oracle.xsum(np.array([1e57, 1e57, -1e57])) == 1e57        # no intermediate overflow
    assert oracle.xsum(np.array([1e300])) == np.inf                  # beyond 2^191: treated as overflow
