"""GPU parity of the unary elementwise ops (rstsr_hip_unary / rstsr_hip_unary_inplace) against
the CPU oracle (oracle/unary.c) on the same strided inputs.

Bars, written per op class:
  * bit-exact: rounding family, conj, neg, not, square, abs/real/imag/sign of real and integer
    inputs, signbit and the classifiers, real reciprocal and sqrt (single IEEE operations),
    and the integer -> f64 conversion;
  * real transcendental functions: within 4 ulp of the C library (the device's math library
    and the host's are separate implementations of the same functions);
  * complex functions (num-complex 0.4 formulas) and complex abs/sign/reciprocal: relative
    error to |oracle| below 5e-14 (c128) / 5e-6 (c64), on inputs off the branch cuts.
"""
import zlib

import numpy as np
import pytest

import rstsr_amd as rt
from gpu_util import bits_equal, from_np
from oracle_ffi import UNARY_OPS, View, c_view
from test_oracle_unary import DOMAINS, cplx_samples, ulp_dist

pytestmark = pytest.mark.gpu

EXACT_REAL = {"ceil", "floor", "trunc", "round", "conj", "square", "abs", "neg", "real", "imag", "sign",
              "signbit", "is_finite", "is_inf", "is_nan", "inv", "reciprocal", "sqrt", "not_"}
EXACT_CPLX = {"conj", "neg", "real", "imag", "square", "is_finite", "is_inf", "is_nan"}
ULP_TOL = 4
REL_TOL = {np.dtype(np.complex128): 5e-14, np.dtype(np.complex64): 5e-6}
DTYPES = [np.float64, np.float32, np.int32, np.int64, np.complex128, np.complex64]


def oracle_out(oracle, op, x):
    return oracle.unary_array(op, x)


def sample(rng, op, dtype, n):
    d = np.dtype(dtype)
    if d.kind == "c":
        return cplx_samples(n, int(rng.integers(1 << 30)), d)
    if d.kind == "i":
        if op in ("acosh",):
            return rng.integers(1, 60, n).astype(d)
        if op in ("log", "log2", "log10", "sqrt"):
            return rng.integers(1, 10 ** 6, n).astype(d)
        if op in ("asin", "acos", "atanh", "log1p"):
            return rng.integers(0, 2, n).astype(d)
        if op in ("exp", "expm1", "sinh", "cosh"):
            return rng.integers(-40, 40, n).astype(d)
        return rng.integers(-(2 ** 20), 2 ** 20, n).astype(d)
    if d.kind == "b":
        return rng.integers(0, 2, n).astype(bool)
    lo, hi = DOMAINS.get(op, (-50.0, 50.0))
    if d == np.float32 and op in ("exp", "cosh", "sinh"):
        lo, hi = lo / 2, hi / 2
    return rng.uniform(lo, hi, n).astype(d)


def check(op, dtype, got, ref):
    assert got.dtype == ref.dtype, (op, dtype, got.dtype, ref.dtype)
    assert got.shape == ref.shape
    d = np.dtype(dtype)
    if d.kind == "c":
        if op in EXACT_CPLX:
            assert bits_equal(got, ref), (op, dtype)
            return 0.0
        g = got.astype(np.complex128) if got.dtype.kind == "c" else got.astype(np.float64)
        r = ref.astype(np.complex128) if ref.dtype.kind == "c" else ref.astype(np.float64)
        err = np.abs(g - r) / np.maximum(np.abs(r), 1e-300)
        tol = REL_TOL[d]
        assert err.max() < tol, (op, dtype, float(err.max()))
        return float(err.max())
    if op in EXACT_REAL or ref.dtype.kind != "f":
        assert bits_equal(got, ref), (op, dtype)
        return 0.0
    u = ulp_dist(got, ref)
    assert u.max() <= ULP_TOL, (op, dtype, float(u.max()))
    return float(u.max())


def supported(oracle, op, dtype):
    return oracle.unary_out_dtype(op, dtype) is not None


CASES = [(op, dt) for op in UNARY_OPS for dt in DTYPES + [np.bool_]]


@pytest.mark.parametrize("op,dtype", CASES, ids=[f"{o}-{np.dtype(d).name}" for o, d in CASES])
def test_unary_strided(dev, oracle, op, dtype):
    if not supported(oracle, op, dtype):
        x = rt.asarray(np.zeros(4, dtype=dtype), device=dev)
        with pytest.raises(rt.RSTSRError):
            x.unary(op)
        return
    rng = np.random.default_rng(zlib.crc32(f"{op}-{np.dtype(dtype).name}".encode()))  # stable across processes
    # a permuted, sliced 3-d view (generic kernel) and a long contiguous row view (row kernel)
    base = sample(rng, op, dtype, 9 * 20 * 30)
    v1 = base.reshape(9, 20, 30).transpose(2, 0, 1)[1:-2:2, :, 3:]
    base2 = sample(rng, op, dtype, 2 * 5000)
    v2 = base2.reshape(2, 5000)[:, 7:4990]
    for v in (v1, v2):
        t, _ = from_np(dev, v)
        out = t.unary(op)
        assert out.layout == rt._rstsr.layout_for_array_copy(t.layout, "K")
        check(op, dtype, out.to_numpy(), oracle_out(oracle, op, v))


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.complex64])
def test_unary_special_values(dev, oracle, dtype):
    vals = [0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 0.5, -0.5, 2.5, -2.5, 1e-310, 710.0, -745.0, 1e300,
            3.0, 100.0, -1e-30]
    d = np.dtype(dtype)
    if d.kind == "c":
        re = np.array(vals)
        x = (re[:, None] + 1j * np.array([0.0, -0.0, 1.0, np.inf, np.nan])[None, :]).reshape(-1).astype(d)
    else:
        x = np.array(vals).astype(d)
    t = rt.asarray(x, device=dev)
    for op in UNARY_OPS:
        if not supported(oracle, op, dtype):
            continue
        got = t.unary(op).to_numpy()
        ref = oracle_out(oracle, op, x)
        if d.kind == "c" and op not in EXACT_CPLX:
            # special values: identical classification (NaN / inf / finite) per component
            gc = np.stack([got.real, got.imag]) if got.dtype.kind == "c" else got[None]
            rc = np.stack([ref.real, ref.imag]) if ref.dtype.kind == "c" else ref[None]
            fin = np.isfinite(rc) & np.isfinite(gc)
            assert np.array_equal(np.isnan(gc), np.isnan(rc)), op
            assert np.array_equal(np.isinf(gc), np.isinf(rc)), op
            assert np.array_equal(gc[np.isinf(gc)], rc[np.isinf(rc)]), op
            big = np.maximum(np.abs(rc), 1e-300)
            assert (np.abs(gc - rc)[fin] <= 64 * np.finfo(gc.dtype).eps * big[fin]
                    + 1e-300).all() or op in ("acos", "acosh", "asin", "asinh", "atan", "atanh", "tan", "tanh"), op
        else:
            check(op, dtype, got, ref)


def test_unary_integer_two_pass_and_wrapping(dev, oracle):
    x = np.array([-(2 ** 31), 2 ** 31 - 1, -7, 0, 46341, 3], np.int32)
    t = rt.asarray(x, device=dev)
    for op in ("neg", "abs", "square", "sign", "not_", "signbit", "real", "imag", "exp", "sqrt", "floor", "inv"):
        assert bits_equal(t.unary(op).to_numpy(), oracle_out(oracle, op, x)), op
    y = np.array([-(2 ** 63), 2 ** 53 + 1, -(2 ** 62) - 3, 5], np.int64)
    ty = rt.asarray(y, device=dev)
    for op in ("neg", "abs", "square", "conj", "round", "atan"):
        r = ty.unary(op).to_numpy()
        assert check(op, np.int64, r, oracle_out(oracle, op, y)) <= ULP_TOL, op


def test_unary_bool_not(dev, oracle):
    x = np.array([True, False, False, True])
    assert rt.asarray(x, device=dev).unary("not_").to_numpy().tolist() == [False, True, True, False]


@pytest.mark.parametrize("op", ["sin", "exp", "sqrt", "neg", "abs", "sign", "round", "square", "conj"])
def test_unary_inplace_matches(dev, oracle, op):
    rng = np.random.default_rng(3)
    x = sample(rng, op, np.float64, 40 * 37).reshape(40, 37)
    t = rt.asarray(x, device=dev)
    ref = t.unary(op).to_numpy()
    v = t[:, 3:30:2]
    before = t.to_numpy()
    v.unary_(op)
    after = t.to_numpy()
    assert bits_equal(after[:, 3:30:2], ref[:, 3:30:2])
    mask = np.ones_like(after, dtype=bool)
    mask[:, 3:30:2] = False
    assert bits_equal(after[mask], before[mask])  # elements outside the view untouched


def test_unary_inplace_rejects_dtype_change(dev):
    t = rt.asarray(np.arange(6, dtype=np.int64), device=dev)
    with pytest.raises(rt.RSTSRError):
        t.unary_("exp")  # i64 -> f64 has no in-place form
    z = rt.asarray(np.ones(3, dtype=np.complex128), device=dev)
    with pytest.raises(rt.RSTSRError):
        z.unary_("abs")  # complex -> real
    with pytest.raises(rt.RSTSRError):
        z.unary_("is_nan")


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32])
def test_unary_to_transposed_output(dev, oracle, dtype):
    """Output laid out differently from the input (a transpose): the generic kernel."""
    rng = np.random.default_rng(5)
    x = sample(rng, "exp", dtype, 96 * 80).reshape(96, 80)
    t = rt.asarray(x, device=dev)
    out_t = rt.empty((80, 96), dtype=dtype, device=dev)
    out = out_t.transpose([1, 0])  # shape (96, 80), column-major in memory
    t.unary_to("exp", out)
    check("exp", dtype, out.to_numpy(), oracle_out(oracle, "exp", x))


def test_unary_rejects_wrong_output_dtype(dev):
    t = rt.asarray(np.ones(4), device=dev)
    out = rt.empty((4,), dtype=np.float32, device=dev)
    with pytest.raises(rt.RSTSRError):
        t.unary_to("exp", out)
    out2 = rt.empty((5,), dtype=np.float64, device=dev)
    with pytest.raises(rt.RSTSRError):
        t.unary_to("exp", out2)


def test_unary_reference_known_answers(dev):
    import json
    import os

    r = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_cases.json")))["unary_math"]
    from test_oracle_unary import _f

    A = lambda v, dt=np.float64: rt.asarray(np.array(v, dtype=dt), device=dev)  # noqa: E731
    assert rt.abs(A(r["abs_i64"]["in"], np.int64)).to_vec() == r["abs_i64"]["out"]
    assert rt.sqrt(A(r["sqrt_f64"]["in"])).to_vec() == r["sqrt_f64"]["out"]
    assert rt.sign(A(r["sign_f64"]["in"])).to_vec() == r["sign_f64"]["out"]
    assert rt.sign(A(r["sign_i64"]["in"], np.int64)).to_vec() == r["sign_i64"]["out"]
    for op in ("floor", "ceil", "trunc"):
        assert getattr(rt, op)(A(r["rounding_in"])).to_vec() == r[op]
    s = rt.sign(A([_f(v) for v in r["sign_special"]["in"]])).to_numpy()
    assert np.isnan(s[0]) and s[1:].tolist() == [1.0, -1.0, 0.0, 0.0] and not np.signbit(s[3])
    for op in ("exp", "sin", "cos", "tanh"):
        x = A([_f(v) for v in r[op]["in"]])
        assert rt.allclose(getattr(rt, op)(x), A(r[op]["out"])), op
    ex = rt.exp(A([0.0, 1.0, 2.0]))
    assert rt.allclose(rt.log(ex), A([0.0, 1.0, 2.0]))
    c = A([_f(v) for v in r["classify_in"]])
    for op in ("is_nan", "is_finite", "is_inf"):
        assert getattr(rt, op)(c).to_vec() == r[op], op
    lc = r["sign_c128_linspace"]
    z = A(np.linspace(complex(*lc["start"]), complex(*lc["stop"]), lc["num"]), np.complex128)
    assert abs(rt.abs(rt.sign(z)).sum_all() - lc["sum_abs"]) < lc["tol"]
    neg = -A([1.0, -2.0])
    assert neg.to_vec() == [-1.0, 2.0]


def test_unary_large_contiguous(dev, oracle):
    """2^22 elements: the row kernel over one long axis, full-size check against the oracle."""
    rng = np.random.default_rng(17)
    x = rng.uniform(-30, 30, 1 << 22)
    t = rt.asarray(x, device=dev)
    for op in ("exp", "sin", "sqrt", "abs"):
        xx = np.abs(x) if op == "sqrt" else x
        tt = rt.asarray(xx, device=dev) if op == "sqrt" else t
        check(op, np.float64, tt.unary(op).to_numpy(), oracle_out(oracle, op, xx))
