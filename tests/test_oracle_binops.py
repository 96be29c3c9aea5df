"""Pins of the oracle's extended binary operators (o_binary with Div/Rem/bit ops/shifts), the
operators of rstsr-core/src/feature_rayon/auto_impl/op_ternary_arithmetic.rs:3-14 over the std and
num-complex 0.4 impls.

Each is checked against an independent computation: Python's exact integer arithmetic (truncating
division, masked shifts, two's-complement wrap), math.fmod, NumPy's bitwise ops, and for complex
division and remainder exact rational arithmetic on Gaussian-integer inputs (where the floating
formula's result is exact) plus a relative bound against Python's own complex division.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle_ffi import View, c_view

DIV, REM, AND, OR, XOR, SHL, SHR = 3, 4, 5, 6, 7, 8, 9


def o_bin(oracle, op, a, b):
    a, b = np.broadcast_arrays(a, b)
    vc = c_view(a.shape, a.dtype)
    oracle.binary(op, vc, View.of(a), View.of(b))
    return vc.array()


def wrap(v, bits):
    m = 1 << bits
    v %= m
    return v - m if v >= m // 2 else v


def trunc_div(x, y):
    q = abs(x) // abs(y)
    return q if (x < 0) == (y < 0) else -q


@pytest.mark.parametrize("dt,bits", [(np.int32, 32), (np.int64, 64)])
def test_integer_div_rem_truncate(oracle, dt, bits):
    rng = np.random.default_rng(11 + bits)
    lo, hi = np.iinfo(dt).min, np.iinfo(dt).max
    a = np.concatenate([rng.integers(-1000, 1000, 200), rng.integers(lo, hi, 200, endpoint=True),
                        [7, -7, 7, -7, lo, lo, hi, 0]]).astype(dt)
    b = np.concatenate([rng.integers(-30, 30, 200), rng.integers(lo, hi, 200, endpoint=True),
                        [2, 2, -2, -2, -1, 1, -1, 5]]).astype(dt)
    q = o_bin(oracle, DIV, a, b)
    r = o_bin(oracle, REM, a, b)
    for x, y, gq, gr in zip(a.tolist(), b.tolist(), q.tolist(), r.tolist()):
        if y == 0:  # a panic in Rust: the documented substitute is 0
            assert gq == 0 and gr == 0
        elif y == -1:  # x / -1 wraps (MIN stays MIN), x % -1 is 0
            assert gq == wrap(-x, bits) and gr == 0
        else:
            eq = trunc_div(x, y)
            assert gq == eq and gr == x - eq * y, (x, y)
            assert gr == 0 or (gr < 0) == (x < 0)  # sign of the dividend (not NumPy's floored %)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_float_rem_is_fmod(oracle, dt):
    rng = np.random.default_rng(5)
    a = (rng.normal(size=300) * 100).astype(dt)
    b = (rng.normal(size=300) * 3).astype(dt)
    a[:4] = [5.5, -5.5, 5.5, -5.5]
    b[:4] = [2.0, 2.0, -2.0, -2.0]
    got = o_bin(oracle, REM, a, b)
    ref = np.array([math.fmod(float(x), float(y)) for x, y in zip(a, b)]).astype(dt)
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
    assert got[:4].tolist() == [1.5, -1.5, 1.5, -1.5]
    # NumPy's remainder is floored: it differs here, so the pin can tell them apart
    assert not np.array_equal(got[:4], np.remainder(a[:4], b[:4]))


@pytest.mark.parametrize("dt,bits", [(np.int32, 32), (np.int64, 64)])
def test_bit_ops_and_shifts(oracle, dt, bits):
    rng = np.random.default_rng(17 + bits)
    lo, hi = np.iinfo(dt).min, np.iinfo(dt).max
    a = rng.integers(lo, hi, 500, endpoint=True).astype(dt)
    b = rng.integers(lo, hi, 500, endpoint=True).astype(dt)
    s = np.concatenate([rng.integers(-3 * bits, 3 * bits, 496), [0, bits - 1, bits, -1]]).astype(dt)
    assert np.array_equal(o_bin(oracle, AND, a, b), a & b)
    assert np.array_equal(o_bin(oracle, OR, a, b), a | b)
    assert np.array_equal(o_bin(oracle, XOR, a, b), a ^ b)
    shl = o_bin(oracle, SHL, a, s)
    shr = o_bin(oracle, SHR, a, s)
    for x, k, gl, gr in zip(a.tolist(), s.tolist(), shl.tolist(), shr.tolist()):
        m = k & (bits - 1)  # Rust release: the amount modulo the width (two's complement of k)
        assert gl == wrap(x << m, bits)
        assert gr == x >> m  # arithmetic


def test_bool_bit_ops(oracle):
    a = np.array([False, False, True, True] * 5)
    b = np.array([False, True, False, True] * 5)
    assert np.array_equal(o_bin(oracle, AND, a, b), a & b)
    assert np.array_equal(o_bin(oracle, OR, a, b), a | b)
    assert np.array_equal(o_bin(oracle, XOR, a, b), a ^ b)


def _gauss(rng, n, lim):
    return rng.integers(-lim, lim + 1, n) + 1j * rng.integers(-lim, lim + 1, n)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_complex_div_exact_on_gaussian_multiples(oracle, dt):
    """(q b) / b with Gaussian integers q, b whose products stay exact: num-complex's formula
    gives q exactly (the numerators and n are exact integers and the quotients integers)."""
    rng = np.random.default_rng(23)
    b = _gauss(rng, 400, 40)
    b[b == 0] = 3 + 4j
    q = _gauss(rng, 400, 40)
    a = (q * b).astype(dt)
    got = o_bin(oracle, DIV, a, b.astype(dt))
    assert np.array_equal(got, q.astype(dt))


def test_complex_div_close_to_python(oracle):
    rng = np.random.default_rng(29)
    a = rng.normal(size=300) + 1j * rng.normal(size=300)
    b = rng.normal(size=300) + 1j * rng.normal(size=300)
    got = o_bin(oracle, DIV, a, b)
    ref = np.array([complex(x) / complex(y) for x, y in zip(a, b)])
    assert np.all(np.abs(got - ref) <= 8 * 2.0 ** -53 * np.abs(ref))


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_complex_rem_exact_on_gaussian_integers(oracle, dt):
    """On Gaussian integers a, b the remainder is a - b * trunc(a / b) with a / b taken exactly
    (the floating quotient truncates to the same integers: a non-integral exact quotient is at
    least 1/|b|^2 from an integer, far more than its rounding), every product and difference
    exact."""
    rng = np.random.default_rng(31)
    a = _gauss(rng, 500, 200)
    b = _gauss(rng, 500, 30)
    b[b == 0] = 2 - 1j
    got = o_bin(oracle, REM, a.astype(dt), b.astype(dt))
    for x, y, g in zip(a, b, got):
        n = int(y.real) ** 2 + int(y.imag) ** 2
        qr = Fraction(int(x.real) * int(y.real) + int(x.imag) * int(y.imag), n)
        qi = Fraction(int(x.imag) * int(y.real) - int(x.real) * int(y.imag), n)
        tr, ti = int(qr), int(qi)  # int() of a Fraction truncates toward zero
        rr = int(x.real) - (int(y.real) * tr - int(y.imag) * ti)
        ri = int(x.imag) - (int(y.real) * ti + int(y.imag) * tr)
        assert (g.real, g.imag) == (rr, ri), (x, y, g)


def test_planted_mistakes_fail(oracle):
    """A floored remainder, a logical right shift and an unmasked shift each differ from the
    oracle on these inputs, so the pins above would catch them."""
    a = np.array([-7, -7, -8], dtype=np.int32)
    b = np.array([2, 3, 33], dtype=np.int32)
    r = o_bin(oracle, REM, a, b)
    assert not np.array_equal(r, np.remainder(a, b))
    sr = o_bin(oracle, SHR, a, np.array([1, 1, 1], dtype=np.int32))
    assert not np.array_equal(sr, (a.view(np.uint32) >> 1).view(np.int32))
    sl = o_bin(oracle, SHL, np.array([1], dtype=np.int32), np.array([33], dtype=np.int32))
    assert sl[0] == 2  # 33 mod 32 = 1, not 0 (an unmasked 1 << 33 overflows to 0)


# ---------------- mixed complex / real operands (num-complex Complex<T> op T, T op Complex<T>) ----------------

ADD, SUB, MUL = 0, 1, 2


def o_mixed(oracle, op, a, b, tc):
    a, b = np.broadcast_arrays(a, b)
    vc = c_view(a.shape, tc)
    oracle.binary_mixed(op, vc, View.of(a), View.of(b))
    return vc.array()


@pytest.mark.parametrize("tc,tr", [(np.complex128, np.float64), (np.complex64, np.float32)])
def test_mixed_add_sub_mul_match_python(oracle, tc, tr):
    """Python's complex-with-float operators treat the float as (t, 0) and the results are exact
    componentwise for finite values: (re + t, im), (t - re, 0 - im), (re t, im t) ... here in
    float64, cast to the pair's precision for complex64 inputs (one rounding each)."""
    rng = np.random.default_rng(53)
    z = (rng.normal(size=200) + 1j * rng.normal(size=200)).astype(tc)
    t = rng.normal(size=200).astype(tr)
    for op, f in ((ADD, lambda x, y: x + y), (SUB, lambda x, y: x - y), (MUL, lambda x, y: x * y)):
        got_cr = o_mixed(oracle, op, z, t, tc)
        got_rc = o_mixed(oracle, op, t, z, tc)
        for zi, ti, g1, g2 in zip(z, t, got_cr, got_rc):
            r1 = f(complex(zi), float(ti))
            r2 = f(float(ti), complex(zi))
            assert g1 == np.array(r1).astype(tc) and g2 == np.array(r2).astype(tc)


def test_mixed_div_exact_cases(oracle):
    """t / z on exact cases: 25 / (3 + 4i) = 3 - 4i, 10 / (1 + 3i) = 1 - 3i; z / t = (re/t, im/t)."""
    z = np.array([3 + 4j, 1 + 3j, -6 + 8j])
    t = np.array([25.0, 10.0, 50.0])
    got = o_mixed(oracle, DIV, t, z, np.complex128)
    assert got.tolist() == [3 - 4j, 1 - 3j, -3 - 4j]
    got2 = o_mixed(oracle, DIV, z, np.array([2.0, 0.5, -4.0]), np.complex128)
    assert got2.tolist() == [1.5 + 2j, 2 + 6j, 1.5 - 2j]


def test_mixed_div_close_to_python(oracle):
    rng = np.random.default_rng(59)
    z = rng.normal(size=300) + 1j * rng.normal(size=300)
    t = rng.normal(size=300)
    got = o_mixed(oracle, DIV, t, z, np.complex128)
    ref = np.array([float(x) / complex(y) for x, y in zip(t, z)])
    assert np.all(np.abs(got - ref) <= 8 * 2.0 ** -53 * np.abs(ref))


def test_mixed_sub_negative_zero(oracle):
    """t - z keeps (0 - im): for im = +0 the imaginary part is +0, not -0 (a plain negation would
    give -0), so a planted -im fails here."""
    got = o_mixed(oracle, SUB, np.array([1.0]), np.array([2.0 + 0.0j]), np.complex128)
    assert np.signbit(got[0].imag) == False  # noqa: E712
