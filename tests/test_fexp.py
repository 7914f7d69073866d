"""The mixed form's FP32 exp(-d)/expm1(-d) helper (csrc/frt_fexp.h), compiled
for the host, against double-precision exp and expm1 (numpy, i.e. libm)."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle


def fexp(d):
    lib = pyoracle.load()
    d = np.ascontiguousarray(d, dtype=np.float32)
    e = np.empty_like(d)
    m = np.empty_like(d)
    fp = C.POINTER(C.c_float)
    lib.orc_fexp_neg(d.ctypes.data_as(fp), e.ctypes.data_as(fp), m.ctypes.data_as(fp), d.size)
    return e, m


def ulps(got, want):
    """Error in units of the float32 spacing at the exact value."""
    want32 = np.abs(want).astype(np.float32)
    spacing = np.spacing(np.maximum(want32, np.float32(np.finfo(np.float32).tiny))).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / spacing


@pytest.fixture
def samples():
    rng = np.random.default_rng(7)
    d = np.concatenate([
        10.0 ** rng.uniform(-12, 1.9, 200000),  # optical depths 1e-12 .. ~80
        rng.uniform(0, 1, 50000),
        np.array([0.0, 1e-30, 1e-8, 0.0312499, 0.03125, 0.0312501, 0.34657, 0.34658, 0.5, 0.6931472, 1.0, 10.0, 50.0, 87.0]),
    ]).astype(np.float32)
    return d


def test_exp_within_2ulp(samples):
    e, _ = fexp(samples)
    want = np.exp(-samples.astype(np.float64))
    assert ulps(e, want).max() <= 2.0


def test_expm1_within_2ulp(samples):
    _, m = fexp(samples)
    want = np.expm1(-samples.astype(np.float64))
    ok = want != 0
    assert ulps(m[ok], want[ok]).max() <= 2.0


def test_exact_and_limit_values():
    e, m = fexp(np.array([0.0, 88.0, 1000.0, np.inf], dtype=np.float32))
    assert e[0] == 1.0 and m[0] == 0.0
    assert np.all(e[1:] == 0.0) and np.all(m[1:] == -1.0)
    e, m = fexp(np.array([np.nan], dtype=np.float32))
    assert np.isnan(e[0]) and np.isnan(m[0])


def test_absorbed_fraction_identity(samples):
    """1 - exp(-(a+b)) = oa + ob - oa*ob with oa = 1 - exp(-a): the combined
    absorbed fraction the deposit forms from the two per-species ones."""
    rng = np.random.default_rng(8)
    a = samples
    b = rng.permutation(samples)
    _, ma = fexp(a)
    _, mb = fexp(b)
    oa, ob = -ma, -mb
    comb = (oa + ob - oa * ob).astype(np.float64)
    want = -np.expm1(-(a.astype(np.float64) + b.astype(np.float64)))
    ok = want > 0
    assert (np.abs(comb[ok] - want[ok]) / want[ok]).max() < 5e-7
