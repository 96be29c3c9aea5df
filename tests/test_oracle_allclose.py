"""Pins of the oracle's allclose restatement (o_allclose: allclose_all of
feature_rayon/auto_impl/reduction.rs:725-750 with isclose of rstsr-dtype-traits/src/isclose.rs:92-106).

Checked against: NumPy's isclose on finite f64 data (the same formula |a - b| <= atol + rtol |b|),
Python's exact integer arithmetic for the wrapping signed-integer difference, math.hypot for
complex norms, and the reference's special cases: inf - inf is NaN and never close (where NumPy
says close), NaN only with equal_nan, zero size an error."""
import math

import numpy as np
import pytest

from oracle_ffi import View


def close(oracle, a, b, **kw):
    a, b = np.broadcast_arrays(np.asarray(a), np.asarray(b))
    return oracle.allclose(View.of(a), View.of(b), **kw)


def test_matches_numpy_on_finite_f64(oracle):
    rng = np.random.default_rng(41)
    for _ in range(300):
        a = rng.normal(size=7) * 10.0 ** rng.integers(-9, 3)
        b = a + rng.normal(size=7) * 10.0 ** rng.integers(-12, 0)
        rtol, atol = 10.0 ** rng.integers(-9, -2), 10.0 ** rng.integers(-12, -5)
        assert close(oracle, a, b, rtol=rtol, atol=atol) == bool(np.all(np.abs(a - b) <= atol + rtol * np.abs(b)))


def test_special_values(oracle):
    inf, nan = np.inf, np.nan
    assert close(oracle, [inf], [inf]) is False  # |inf - inf| = NaN: not close
    assert bool(np.isclose(inf, inf))  # NumPy differs here, so the pin tells them apart
    # b = +-inf makes the bound infinite: any a whose difference is not NaN is "close"
    assert close(oracle, [-inf], [inf]) is True
    assert close(oracle, [1.0], [inf]) is True
    assert close(oracle, [inf], [1.0]) is False
    assert close(oracle, [nan], [nan]) is False
    assert close(oracle, [nan], [nan], equal_nan=True) is True
    assert close(oracle, [nan], [1.0], equal_nan=True) is False
    assert close(oracle, [1.0, 2.0], [1.0, 2.0 + 1.5e-5]) is True  # 1.5e-5 <= 1e-8 + 1e-5 * 2.000015
    assert close(oracle, [1.0, 2.0], [1.0, 2.0 + 2.5e-5]) is False  # 2.5e-5 > 1e-8 + 1e-5 * 2.000025
    assert close(oracle, np.zeros(0), np.zeros(0)) is None  # zero size: an error in the reference


def test_asymmetry_exact(oracle):
    """The bound uses |b| only: with rtol 0.00995, |100 - 101| = 1 is within 0.00995 * 101 but
    not within 0.00995 * 100."""
    assert close(oracle, [100.0], [101.0], rtol=0.00995, atol=0.0) is True
    assert close(oracle, [101.0], [100.0], rtol=0.00995, atol=0.0) is False


def test_integers_wrap(oracle):
    def wrap(v, bits):
        m = 1 << bits
        v %= m
        return v - m if v >= m // 2 else v

    rng = np.random.default_rng(43)
    for dt, bits in ((np.int32, 32), (np.int64, 64)):
        info = np.iinfo(dt)
        xs = rng.integers(info.min, info.max, 200, endpoint=True).tolist() + [info.max, info.min, 0, -5]
        ys = rng.integers(info.min, info.max, 200, endpoint=True).tolist() + [info.min, info.max, info.min, 5]
        for x, y in zip(xs, ys):
            d = wrap(x - y, bits) if x >= y else wrap(y - x, bits)
            m = wrap(-y, bits) if y < 0 else y
            exp = float(d) <= 1e-8 + 1e-5 * float(m)
            assert close(oracle, np.array([x], dtype=dt), np.array([y], dtype=dt)) == exp, (x, y)


def test_complex_norm(oracle):
    rng = np.random.default_rng(47)
    for _ in range(200):
        a = complex(*rng.normal(size=2))
        b = a + complex(*(rng.normal(size=2) * 10.0 ** rng.integers(-8, -4)))
        exp = math.hypot(a.real - b.real, a.imag - b.imag) <= 1e-8 + 1e-5 * math.hypot(b.real, b.imag)
        assert close(oracle, np.array([a]), np.array([b])) == exp


def test_broadcast_shapes(oracle):
    a = np.arange(12.0).reshape(3, 4)
    assert close(oracle, a, a[1:2, :] * 0 + a) is True
    assert close(oracle, a, np.arange(4.0)) is False
    assert close(oracle, np.tile(np.arange(4.0), (3, 1)), np.arange(4.0)) is True
