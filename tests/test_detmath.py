"""Pins of frt_detmath.h against glibc (Python's math module calls libm).

The deterministic routines replace libm on both the GPU and the CPU checker;
these tests bound how far they are from the library the reference links.
"""
import math
import struct

import numpy as np
import pytest

from oracle import pyoracle as P


def ulp_diff(a, b):
    if a == b:
        return 0
    ia = struct.unpack("<q", struct.pack("<d", a))[0]
    ib = struct.unpack("<q", struct.pack("<d", b))[0]
    if ia < 0:
        ia = -(1 << 63) - ia
    if ib < 0:
        ib = -(1 << 63) - ib
    return abs(ia - ib)


@pytest.fixture(scope="module")
def rng():
    return np.random.default_rng(12345)


def test_exp_within_2ulp(rng):
    xs = np.concatenate([rng.uniform(-745, 709.7, 20000), rng.uniform(-1, 1, 20000),
                         rng.uniform(-60, 0, 20000), [0.0, -0.0, 1e-300, -1e-300, 709.78, -745.0]])
    worst = max(ulp_diff(P.dm("exp", float(x)), math.exp(float(x))) for x in xs)
    assert worst <= 2, worst


def test_exp_special_values():
    assert P.dm("exp", 0.0) == 1.0
    assert P.dm("exp", -0.0) == 1.0
    assert P.dm("exp", -746.0) == 0.0
    assert math.isinf(P.dm("exp", 710.0))
    assert math.isnan(P.dm("exp", float("nan")))
    # subnormal range stays within 1 ulp of a subnormal
    for x in (-740.0, -744.5, -720.3):
        assert abs(P.dm("exp", x) - math.exp(x)) <= 5e-324 * 2


def test_log_within_2ulp(rng):
    xs = np.concatenate([10.0 ** rng.uniform(-300, 300, 20000), rng.uniform(0.5, 2.0, 20000),
                         [1.0, 2.0, 4.0, 1e-310, 5e-324, 1.7e308]])
    worst = max(ulp_diff(P.dm("log", float(x)), math.log(float(x))) for x in xs)
    assert worst <= 2, worst
    assert P.dm("log", 1.0) == 0.0
    assert math.isinf(P.dm("log", 0.0)) and P.dm("log", 0.0) < 0
    assert math.isnan(P.dm("log", -1.0))


def test_pow_accuracy(rng):
    # bases and exponents of the rate fits: lambda(T) = 2 E_th / kT over 1e1..1e9 K
    xs = 10.0 ** rng.uniform(-3, 7, 5000)
    ys = rng.uniform(-6, 6, 5000)
    for x, y in zip(xs, ys):
        ref = math.pow(float(x), float(y))
        got = P.dm("pow", float(x), float(y))
        bound = 2.0 ** -52 * (2.0 + abs(y * math.log(x)) * 2.0)
        assert abs(got - ref) <= bound * abs(ref), (x, y, got, ref)


def test_pow_exact_cases():
    for x in (0.3, 1.7, 12345.678, 1e-20):
        assert P.dm("pow", x, 2.0) == x * x
        assert P.dm("pow", x, 0.5) == math.sqrt(x)
        assert P.dm("pow", x, 1.0) == x
    for k in range(-30, 31):
        assert P.dm("pow", 2.0, float(k)) == math.ldexp(1.0, k)
    assert P.dm("pow", -2.0, 3.0) == pytest.approx(-8.0, rel=1e-15)
    assert P.dm("pow", -2.0, 2.0) == 4.0
    assert math.isnan(P.dm("pow", -2.0, 0.5))
    assert P.dm("pow", float("nan"), 0.0) == 1.0


def test_sin_cos_within_2ulp(rng):
    xs = np.concatenate([rng.uniform(0, 2 * math.pi, 20000), rng.uniform(-7, 7, 5000),
                         [0.0, math.pi / 4, math.pi / 2, math.pi, 1.5 * math.pi]])
    for x in xs:
        s, c = P.dm("sin", float(x)), P.dm("cos", float(x))
        # absolute error relative to 1 (the function values are bounded by 1)
        assert abs(s - math.sin(x)) <= 2 * 2.0 ** -53 * max(abs(math.sin(x)), 2.0 ** -3)
        assert abs(c - math.cos(x)) <= 2 * 2.0 ** -53 * max(abs(math.cos(x)), 2.0 ** -3)


def test_atan2_within_2ulp(rng):
    ys = rng.standard_normal(20000)
    xs = rng.standard_normal(20000)
    worst = max(ulp_diff(P.dm("atan2", float(y), float(x)), math.atan2(y, x)) for y, x in zip(ys, xs))
    assert worst <= 2, worst
    for y, x in [(0.0, 1.0), (0.0, -1.0), (1.0, 0.0), (-1.0, 0.0), (-0.0, -1.0), (1.0, 1.0)]:
        assert P.dm("atan2", y, x) == pytest.approx(math.atan2(y, x), abs=1e-16)
    assert math.copysign(1.0, P.dm("atan2", -0.0, 1.0)) == -1.0
