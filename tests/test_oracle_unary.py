"""Pin the unary part of the CPU oracle (oracle/unary.c) before trusting it as the parity checker.

Checks, per function:
  * the reference's own known answers (tests/golden/reference_cases.json "unary_math":
    rstsr-core/tests/core_func/math/test_unary_math.rs, doc_draft/math/test_math_doc.rs,
    tensor/operators/op_unary_common.rs test_sign);
  * Python's math module on the same f64 inputs (the C library the reference's f64 methods
    call, so bit-exact: pins the op -> function mapping) and NumPy within 4 ulp (an
    independent implementation: pins the mathematics);
  * complex functions against NumPy / cmath principal values away from branch cuts and
    from |z| -> 0, where num-complex's log-based formulas cancel;
  * exact identities: round half away from zero, sign/signbit on signed zeros and NaN,
    wrapping integer neg/abs/square, the output-dtype table of the traits;
  * planted mistakes: the comparators reject a perturbed result and the pins above
    fail for a mapped-to-the-wrong-function or conjugated result.
No GPU.
"""
import cmath
import json
import math
import os

import numpy as np
import pytest

from oracle_ffi import UNARY_OPS

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_cases.json")))["unary_math"]


def _f(v):
    if isinstance(v, str):
        return {"nan": math.nan, "inf": math.inf, "-inf": -math.inf, "-0.0": -0.0, "0.0": 0.0,
                "pi": math.pi, "pi/2": math.pi / 2}[v]
    return float(v)


def ulp_dist(x, y):
    """Distance in units in the last place between two float arrays of one dtype (NaN == NaN)."""
    x = np.ascontiguousarray(x)
    y = np.ascontiguousarray(y)
    assert x.dtype == y.dtype and x.dtype.kind == "f"
    it, lo = (np.int64, -(2 ** 63)) if x.dtype == np.float64 else (np.int32, -(2 ** 31))
    xb = x.view(it).astype(np.int64)
    yb = y.view(it).astype(np.int64)
    xo = np.where(xb < 0, lo - xb, xb)
    yo = np.where(yb < 0, lo - yb, yb)
    same = (xo < 0) == (yo < 0)
    # exact int64 difference when both lie on one side of zero; otherwise |x| + |y| steps
    d = np.where(same, np.abs(np.where(same, xo - yo, 0)).astype(np.float64),
                 np.abs(xo).astype(np.float64) + np.abs(yo).astype(np.float64))
    nx, ny = np.isnan(x), np.isnan(y)
    d = np.where(nx & ny, 0.0, np.where(nx | ny, np.inf, d))
    return d


# (op, domain) for real functions: uniform samples over the domain
DOMAINS = {
    "acos": (-1.0, 1.0), "asin": (-1.0, 1.0), "atan": (-20.0, 20.0), "acosh": (1.0, 50.0),
    "asinh": (-50.0, 50.0), "atanh": (-0.999, 0.999), "cos": (-30.0, 30.0), "sin": (-30.0, 30.0),
    "tan": (-10.0, 10.0), "cosh": (-20.0, 20.0), "sinh": (-20.0, 20.0), "tanh": (-10.0, 10.0),
    "exp": (-50.0, 50.0), "expm1": (-5.0, 5.0), "log": (1e-3, 1e3), "log1p": (-0.999, 10.0),
    "log2": (1e-3, 1e3), "log10": (1e-3, 1e3), "sqrt": (0.0, 1e3),
}
NP_NAME = {"acos": "arccos", "asin": "arcsin", "atan": "arctan", "acosh": "arccosh", "asinh": "arcsinh",
           "atanh": "arctanh"}


@pytest.mark.parametrize("op", sorted(DOMAINS))
def test_real_f64_matches_libm_and_numpy(oracle, op):
    lo, hi = DOMAINS[op]
    x = np.random.default_rng(7).uniform(lo, hi, size=4000)
    got = oracle.unary_array(op, x)
    mfn = getattr(math, op)
    ref = np.array([mfn(v) for v in x])
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), op  # same C library
    npv = getattr(np, NP_NAME.get(op, op))(x)
    assert ulp_dist(got, npv).max() <= 4, op


@pytest.mark.parametrize("op", sorted(DOMAINS))
def test_real_f32_within_2ulp_of_rounded_f64(oracle, op):
    lo, hi = DOMAINS[op]
    if op in ("exp", "cosh", "sinh"):
        lo, hi = lo / 2, hi / 2  # stay inside f32 range
    x = np.random.default_rng(8).uniform(lo, hi, size=4000).astype(np.float32)
    got = oracle.unary_array(op, x)
    assert got.dtype == np.float32
    mfn = getattr(math, op)
    ref = np.array([mfn(float(v)) for v in x]).astype(np.float32)
    assert ulp_dist(got, ref).max() <= 2, op


def test_reciprocal_inv_sqrt_exact(oracle):
    x = np.random.default_rng(9).normal(size=3000) * 10.0 ** np.random.default_rng(10).integers(-30, 30, 3000)
    for op in ("inv", "reciprocal"):
        assert np.array_equal(oracle.unary_array(op, x), 1.0 / x)
        x32 = x.astype(np.float32)
        assert np.array_equal(oracle.unary_array(op, x32), np.float32(1) / x32)
    ax = np.abs(x)
    assert np.array_equal(oracle.unary_array("sqrt", ax), np.sqrt(ax))
    assert np.array_equal(oracle.unary_array("sqrt", ax.astype(np.float32)), np.sqrt(ax.astype(np.float32)))


def test_rounding_family(oracle):
    halves = np.array([-3.5, -2.5, -1.5, -0.5, 0.5, 1.5, 2.5, 3.5, -0.0, 0.0, 1e300, -7.25])
    got = oracle.unary_array("round", halves)
    want = np.array([math.copysign(math.floor(abs(v) + 0.5), v) for v in halves])  # half away from zero
    assert np.array_equal(got, want)
    assert not np.array_equal(got, np.rint(halves))  # NumPy's half-to-even differs: the pin is sharp
    for op, f in (("floor", np.floor), ("ceil", np.ceil), ("trunc", np.trunc)):
        assert np.array_equal(oracle.unary_array(op, halves).view(np.uint64), f(halves).view(np.uint64))
    r = REF
    x = np.array(r["rounding_in"])
    for op in ("floor", "ceil", "trunc"):
        assert oracle.unary_array(op, x).tolist() == r[op]


def test_reference_known_answers(oracle):
    r = REF
    assert oracle.unary_array("abs", np.array(r["abs_i64"]["in"], np.int64)).tolist() == r["abs_i64"]["out"]
    assert oracle.unary_array("sqrt", np.array(r["sqrt_f64"]["in"])).tolist() == r["sqrt_f64"]["out"]
    assert oracle.unary_array("sign", np.array(r["sign_f64"]["in"])).tolist() == r["sign_f64"]["out"]
    assert oracle.unary_array("sign", np.array(r["sign_i64"]["in"], np.int64)).tolist() == r["sign_i64"]["out"]
    s = oracle.unary_array("sign", np.array([_f(v) for v in r["sign_special"]["in"]]))
    assert math.isnan(s[0]) and s[1] == 1.0 and s[2] == -1.0 and s[3] == 0.0 and not math.copysign(1, s[3]) < 0
    assert s[4] == 0.0
    for op in ("exp", "sin", "cos", "tanh"):
        x = np.array([_f(v) for v in r[op]["in"]])
        assert np.allclose(oracle.unary_array(op, x), np.array(r[op]["out"]), rtol=1e-5, atol=1e-8), op
    ex = oracle.unary_array("exp", np.array([0.0, 1.0, 2.0]))
    assert np.allclose(oracle.unary_array("log", ex), [0.0, 1.0, 2.0], rtol=1e-5, atol=1e-8)
    sv = oracle.unary_array("sin", np.array([0.0, math.pi / 2]))
    assert abs(sv[0]) < 1e-12 and abs(sv[1] - 1.0) < 1e-12
    c = np.array([_f(v) for v in r["classify_in"]])
    for op in ("is_nan", "is_finite", "is_inf"):
        got = oracle.unary_array(op, c)
        assert got.dtype == np.bool_ and got.tolist() == r[op], op
    lc = r["sign_c128_linspace"]
    z = np.linspace(complex(*lc["start"]), complex(*lc["stop"]), lc["num"])
    sg = oracle.unary_array("sign", z)
    assert abs(np.abs(sg).sum() - lc["sum_abs"]) < lc["tol"]


def test_sign_signbit_abs_special_values(oracle):
    x = np.array([0.0, -0.0, np.nan, -np.nan, np.inf, -np.inf, 2.0, -3.0])
    sb = oracle.unary_array("signbit", x)
    # Signed::is_positive: true for +0.0, +NaN and +inf (the complement of C's signbit)
    assert sb.tolist() == [not math.copysign(1.0, v) < 0 for v in x]
    assert sb.tolist() == [True, False, True, False, True, False, True, False]
    xi = np.array([0, 1, -1, 2 ** 31 - 1, -(2 ** 31)], np.int32)
    assert oracle.unary_array("signbit", xi).tolist() == [False, True, False, True, False]
    sg = oracle.unary_array("sign", x)
    assert np.isnan(sg[2]) and np.isnan(sg[3])
    assert np.signbit(sg[3])  # NaN passes through with its payload and sign
    assert sg[[0, 1, 4, 5, 6, 7]].tolist() == [0.0, 0.0, 1.0, -1.0, 1.0, -1.0]
    assert not np.signbit(sg[1])
    ab = oracle.unary_array("abs", x)
    assert np.array_equal(ab.view(np.uint64), np.abs(x).view(np.uint64))
    z = np.array([0j, complex(3, 4), complex(-5, 12), complex(np.inf, 1)])
    assert oracle.unary_array("abs", z).tolist() == [0.0, 5.0, 13.0, np.inf]
    assert oracle.unary_array("sign", z)[:3].tolist() == [0j, complex(0.6, 0.8), complex(-5 / 13, 12 / 13)]


def test_integer_wrapping(oracle):
    m32 = np.array([-(2 ** 31), 2 ** 31 - 1, -7, 0, 46341], np.int32)
    assert oracle.unary_array("neg", m32).tolist() == [-(2 ** 31), -(2 ** 31) + 1, 7, 0, -46341]
    assert oracle.unary_array("abs", m32).tolist() == [-(2 ** 31), 2 ** 31 - 1, 7, 0, 46341]
    sq = oracle.unary_array("square", m32)
    want = [((int(v) * int(v) + 2 ** 31) % 2 ** 32) - 2 ** 31 for v in m32]
    assert sq.tolist() == want
    m64 = np.array([-(2 ** 63), 3037000500, -5], np.int64)
    assert oracle.unary_array("neg", m64).tolist() == [-(2 ** 63), -3037000500, 5]
    assert oracle.unary_array("square", m64).tolist() == [
        ((int(v) * int(v) + 2 ** 63) % 2 ** 64) - 2 ** 63 for v in m64]
    assert oracle.unary_array("not_", m32).tolist() == [~int(v) for v in m32]
    assert oracle.unary_array("not_", np.array([True, False])).tolist() == [False, True]
    assert oracle.unary_array("imag", m32).tolist() == [0] * 5
    assert oracle.unary_array("real", m32).tolist() == m32.tolist()


def test_integer_inputs_convert_to_f64(oracle):
    xi = np.array([-(2 ** 53) - 1, -3, 0, 5, 2 ** 62 + 1], np.int64)
    for op in ("exp", "sqrt", "floor", "conj", "inv", "atan"):
        got = oracle.unary_array(op, xi)
        assert got.dtype == np.float64
        ref = oracle.unary_array(op, xi.astype(np.float64))  # `as f64`, round to nearest
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), op


def test_out_dtype_table(oracle):
    real_only = {"ceil", "expm1", "floor", "log1p", "round", "trunc"}
    for op in UNARY_OPS:
        for dt in (np.float32, np.float64, np.int32, np.int64, np.complex64, np.complex128, np.bool_):
            got = oracle.unary_out_dtype(op, dt)
            d = np.dtype(dt)
            if op == "not_":
                want = d if d.kind in "ib" else None
            elif d.kind == "b":
                want = None
            elif op in ("is_finite", "is_inf", "is_nan"):
                want = None if d.kind == "i" else np.dtype(np.bool_)
            elif op == "signbit":
                want = None if d.kind == "c" else np.dtype(np.bool_)
            elif op in ("square", "sign", "neg"):
                want = d
            elif op in ("abs", "real", "imag"):
                want = {np.dtype(np.complex64): np.dtype(np.float32),
                        np.dtype(np.complex128): np.dtype(np.float64)}.get(d, d)
            elif op in real_only and d.kind == "c":
                want = None
            else:
                want = np.dtype(np.float64) if d.kind == "i" else d
            assert got == want, (op, d, got, want)


CPLX_NP = {
    "acos": np.arccos, "acosh": np.arccosh, "asin": np.arcsin, "asinh": np.arcsinh, "atan": np.arctan,
    "atanh": np.arctanh, "cos": np.cos, "cosh": np.cosh, "exp": np.exp, "log": np.log, "log2": np.log2,
    "log10": np.log10, "sin": np.sin, "sinh": np.sinh, "sqrt": np.sqrt, "tan": np.tan, "tanh": np.tanh,
    "inv": lambda z: 1 / z, "reciprocal": lambda z: 1 / z, "square": lambda z: z * z, "conj": np.conj,
    "neg": np.negative,
}


def cplx_samples(n, seed, dtype=np.complex128):
    """|z| in [0.25, 2.5], both parts at least 0.1 in magnitude: off every branch cut and away
    from z -> 0 and z -> +-1, +-i."""
    rng = np.random.default_rng(seed)
    re = rng.uniform(0.1, 1.7, n) * rng.choice([-1, 1], n)
    im = rng.uniform(0.1, 1.7, n) * rng.choice([-1, 1], n)
    return (re + 1j * im).astype(dtype)


def rel_err(got, ref):
    return np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)


@pytest.mark.parametrize("op", sorted(CPLX_NP))
def test_complex_principal_values(oracle, op):
    z = cplx_samples(3000, 11)
    got = oracle.unary_array(op, z)
    assert got.dtype == np.complex128
    ref = CPLX_NP[op](z)
    assert rel_err(got, ref).max() < 1e-13, op
    # cmath on a few points (another implementation of the same principal branches)
    cm = {"inv": lambda v: 1 / v, "reciprocal": lambda v: 1 / v, "square": lambda v: v * v,
          "conj": lambda v: v.conjugate(), "neg": lambda v: -v,
          "log2": lambda v: cmath.log(v) / math.log(2), "log10": cmath.log10}.get(op, getattr(cmath, op, None))
    for v, g in zip(z[:50], got[:50]):
        assert abs(g - cm(complex(v))) <= 1e-13 * abs(cm(complex(v)))
    z32 = z.astype(np.complex64)
    g32 = oracle.unary_array(op, z32)
    assert g32.dtype == np.complex64
    assert rel_err(g32.astype(np.complex128), CPLX_NP[op](z32.astype(np.complex128))).max() < 2e-5, op


def test_complex_special_cases(oracle):
    i = 1j
    at = oracle.unary_array("atan", np.array([i, -i]))
    assert at[0] == complex(0, np.inf) and at[1] == complex(0, -np.inf)
    ath = oracle.unary_array("atanh", np.array([1 + 0j, -1 + 0j]))
    assert ath[0] == complex(np.inf, 0) and ath[1] == complex(-np.inf, 0)
    sq = oracle.unary_array("sqrt", np.array([complex(-4, 0), complex(-4, -0.0), complex(0, 8), complex(0, -8),
                                              complex(9, 0)]))
    assert sq.tolist() == [2j, -2j, complex(2, 2), complex(2, -2), 3 + 0j]
    ex = oracle.unary_array("exp", np.array([complex(-np.inf, np.inf), complex(np.inf, 0), complex(np.nan, 0)]))
    assert ex[0] == 0 and ex[1] == complex(np.inf, 0) and np.isnan(ex[2].real) and ex[2].imag == 0
    cls = np.array([complex(1, 2), complex(np.inf, np.nan), complex(np.inf, 1), complex(1, np.nan)])
    assert oracle.unary_array("is_inf", cls).tolist() == [False, False, True, False]
    assert oracle.unary_array("is_nan", cls).tolist() == [False, True, False, True]
    assert oracle.unary_array("is_finite", cls).tolist() == [True, False, False, False]


def test_pins_catch_planted_mistakes(oracle):
    x = np.random.default_rng(12).uniform(-1, 1, 500)
    got = oracle.unary_array("asin", x)
    # a result perturbed by 8 ulp fails the 4-ulp comparator
    bad = (got.view(np.int64) + 8).view(np.float64)
    assert ulp_dist(bad, np.arcsin(x)).max() > 4
    # mapping an op to a neighbouring function fails the libm pin
    assert not np.array_equal(oracle.unary_array("acos", x), np.array([math.asin(v) for v in x]))
    # a conjugated complex result fails the principal-value pin
    z = cplx_samples(200, 13)
    g = oracle.unary_array("asin", z)
    assert rel_err(np.conj(g), np.arcsin(z)).max() > 1e-3
    # a dropped minus sign in acos (i instead of -i in front of the log) fails too
    assert rel_err(-oracle.unary_array("acos", z), np.arccos(z)).max() > 1e-3
