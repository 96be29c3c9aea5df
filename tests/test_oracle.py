"""Pin the CPU oracle (oracle/oracle.c) before trusting it as the parity checker.

Each oracle function is checked against
  * the reference's own known answers (tests/golden/reference_cases.json, with file:line), and
  * results the mathematics fixes: NumPy's IEEE add/sub/mul/div on the same strided
    views (bit-exact), math.fsum (correctly rounded sums), first-occurrence argmax,
    Python complex products, exact integer matmul, an extended-precision matmul.
The last block plants mistakes (a dropped element, a transposed operand, a wrong tie
rule) into oracle CALLS and shows each pin catches them. No GPU.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle_ffi import View, c_view

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_cases.json")))

ADD, SUB, MUL, DIV = 0, 1, 2, 3
SUM, MAX, MIN, ARGMAX, ARGMIN = 0, 1, 2, 3, 4
PROD, MEAN, VAR, STD, L2 = 5, 6, 7, 8, 9


def view_of_numpy_slice(base, build):
    """base: flat array; build(ndarray) -> a strided numpy view of base. Returns (View, ndarray view)."""
    arr = build(base)
    it = base.itemsize
    off = (arr.__array_interface__["data"][0] - base.__array_interface__["data"][0]) // it
    return View(base, arr.shape, [s // it for s in arr.strides], off), arr


def o_binary(oracle, op, a_np, b_np):
    """Run the oracle on numpy views a_np, b_np (already broadcast); output c-contig."""
    a_np, b_np = np.broadcast_arrays(a_np, b_np)
    va = View.of(a_np) if a_np.flags.owndata is False else View.of(a_np)
    vb = View.of(b_np)
    vc = c_view(a_np.shape, a_np.dtype)
    oracle.binary(op, vc, va, vb)
    return vc.array()


def o_reduce(oracle, op, a_np, axes):
    va = View.of(a_np)
    kept = [i for i in range(a_np.ndim) if i not in [x % a_np.ndim for x in axes]]
    shape = [a_np.shape[i] for i in kept]
    odt = np.uint64 if op in (ARGMAX, ARGMIN) else a_np.dtype
    if op in (VAR, STD, L2) and np.issubdtype(a_np.dtype, np.complexfloating):
        odt = np.float32 if a_np.dtype == np.complex64 else np.float64
    vo = c_view(shape, odt)
    assert oracle.reduce_axes(op, va, list(axes), vo) == 0
    return vo.array()


# ---------------- reference known answers ----------------

def test_known_sum_all(oracle):
    base = np.arange(3240, dtype=np.int64)
    a = base.reshape(12, 15, 18).swapaxes(-1, -2)[2:-3, 1:-4:2, -1:3:-2]
    assert oracle.reduce_all(SUM, View.of(a)) == REF["sum_all_sliced_swapaxes"]["value"]


def test_known_sum_axes(oracle):
    a = np.arange(3240, dtype=np.int64).reshape(4, 6, 15, 9).transpose(2, 0, 3, 1)
    s = o_reduce(oracle, SUM, a, [0, -2])
    for (i, j), v in REF["sum_axes_transposed_4d"]["entries"]:
        assert s[i, j] == v


def test_known_min_and_mean(oracle):
    r = REF["min_4x3"]
    a = np.array(r["data"], dtype=np.int32).reshape(r["shape"])
    assert list(o_reduce(oracle, MIN, a, [0])) == r["min_axes_0"]
    assert list(o_reduce(oracle, MIN, a, [1])) == r["min_axes_1"]
    assert oracle.reduce_all(MIN, View.of(a)) == r["min_all"]
    r = REF["mean_arange24"]
    a = np.arange(24.0).reshape(2, 3, 4)
    assert list(o_reduce(oracle, SUM, a, [0, 2]) / 8) == r["mean_axes_0_2"]
    b = a[::-1, :, ::-2]
    assert list(o_reduce(oracle, SUM, b, [-1, 1]) / 6) == r["flipped_mean"]


def test_known_argmin(oracle):
    r = REF["argmin_4x3"]
    a = np.array(r["data"], dtype=np.int32).reshape(r["shape"])
    assert oracle.reduce_all(ARGMIN, View.of(a)) == r["argmin_all"]
    assert list(o_reduce(oracle, ARGMIN, a, [-1])) == r["argmin_axes_-1"]
    assert list(o_reduce(oracle, ARGMIN, a, [0])) == r["argmin_axes_0"]


def test_known_argmax_sum_max(oracle):
    r = REF["argmax_core"]
    for data, idx in r["all_cases"]:
        assert oracle.reduce_all(ARGMAX, View.of(np.array(data, dtype=np.int32))) == idx
    m = np.array(r["matrix"], dtype=np.int32)
    assert list(o_reduce(oracle, ARGMAX, m, [0])) == r["axes_0"]
    assert list(o_reduce(oracle, ARGMAX, m, [1])) == r["axes_1"]
    r = REF["sum_core"]
    m = np.array(r["matrix"], dtype=np.int32)
    assert list(o_reduce(oracle, SUM, m, [1])) == r["sum_axes_1"]
    assert oracle.reduce_all(SUM, View.of(m)) == r["sum_all"]
    assert list(o_reduce(oracle, SUM, np.arange(24, dtype=np.int32).reshape(2, 3, 4), [-2, -1])) == r[
        "arange24_sum_axes_-2_-1"]
    r = REF["max_core"]
    assert oracle.reduce_all(MAX, View.of(np.array(r["vector"]))) == r["max_all"]
    m = np.array(r["matrix"])
    assert list(o_reduce(oracle, MAX, m, [0])) == r["max_axes_0"]
    assert list(o_reduce(oracle, MAX, m, [1])) == r["max_axes_1"]


def _lin(spec):
    x = np.linspace(spec[0], spec[1], spec[2])
    return x.reshape(spec[3]) if len(spec) > 3 else x


def test_known_add(oracle):
    r = REF["add_row_major"]
    assert list(o_binary(oracle, ADD, _lin(r["contiguous"]["a"]), _lin(r["contiguous"]["b"])).ravel()) == r["contiguous"]["c"]
    k = r["broadcast_2x3_3"]
    assert list(o_binary(oracle, ADD, _lin(k["a"]), _lin(k["b"])).ravel()) == k["c"]
    k = r["broadcast_123_5121"]
    assert list(o_binary(oracle, ADD, _lin(k["a"]), _lin(k["b"])).ravel()) == k["c"]
    k = r["transposed"]
    assert list(o_binary(oracle, ADD, _lin(k["a"]), _lin(k["b_reversed"]).T).ravel()) == k["c"]
    k = r["flip_a"]
    assert list(o_binary(oracle, ADD, _lin(k["a"])[::-1], _lin(k["b"])).ravel()) == k["c"]
    k = r["flip_b"]
    assert list(o_binary(oracle, ADD, _lin(k["a"]), _lin(k["b"])[::-1]).ravel()) == k["c"]


def test_known_matmul(oracle):
    r = REF["matmul_core"]
    m1 = np.array(r["m1"], dtype=np.int32)
    m2 = np.array(r["m2"], dtype=np.int32)
    assert oracle.gemm_naive(m1, m2).tolist() == r["m1_m2"]
    assert oracle.gemm_naive(m2, m1).tolist() == r["m2_m1"]
    r = REF["faer_c32_minimal"]
    a = np.array([complex(*x) for x in r["a"]], dtype=np.complex64).reshape(r["a_shape"])
    b = np.array([complex(*x) for x in r["b"]], dtype=np.complex64).reshape(r["b_shape"])
    s = oracle.gemm_naive(a, b).sum()
    assert abs(s.real - r["sum_c"][0]) < 1e-5 and abs(s.imag - r["sum_c"][1]) < 1e-5


def test_known_large_matmul_stats(oracle):
    r = REF["large_matmul_stats"]
    a = np.linspace(0, 1, 1048576).reshape(16, 256, 256)
    b = np.linspace(1, 2, 1048576).reshape(16, 256, 256)
    c = np.stack([oracle.gemm_naive(a[i], b[i]) for i in range(16)])
    assert abs(c.mean() - r["mean"]) < 1e-6
    assert abs(c.std() - r["std"]) < 1e-6
    s01 = c.std(axis=(0, 1))
    for i, v in r["std_axes_0_1"]:
        assert abs(s01[i] - v) < 1e-6
    s12 = c.std(axis=(1, 2))
    for i, v in r["std_axes_1_2"]:
        assert abs(s12[i] - v) < 1e-6


# ---------------- mathematically fixed results ----------------

@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.int32, np.int64])
@pytest.mark.parametrize("op", [ADD, SUB, MUL, DIV])
def test_binary_matches_numpy_ieee(oracle, dtype, op):
    if op == DIV and np.dtype(dtype).kind == "i":
        pytest.skip("integer division is not an op of this path")
    rng = np.random.default_rng(7 + op)
    if np.dtype(dtype).kind == "f":
        base = rng.normal(size=6 * 10 * 12).astype(dtype)
        bb = rng.uniform(0.5, 2.0, size=10).astype(dtype)
    else:
        base = rng.integers(-(2 ** 20), 2 ** 20, size=6 * 10 * 12).astype(dtype)
        bb = rng.integers(-(2 ** 20), 2 ** 20, size=10).astype(dtype)
    x = base.reshape(6, 10, 12).transpose(2, 0, 1)[:, :, ::2]  # permuted + stepped view
    y = bb.reshape(1, 10, 1).transpose(2, 0, 1)[:, :, ::2]
    got = o_binary(oracle, op, x, y)
    with np.errstate(all="ignore"):
        ref = {ADD: np.add, SUB: np.subtract, MUL: np.multiply, DIV: np.divide}[op](x, y)
    assert got.dtype == ref.dtype
    assert np.array_equal(got.view(np.uint8), np.ascontiguousarray(ref).view(np.uint8))


def test_complex_mul_matches_python(oracle):
    rng = np.random.default_rng(3)
    a = (rng.normal(size=64) + 1j * rng.normal(size=64)).astype(np.complex128)
    b = (rng.normal(size=64) + 1j * rng.normal(size=64)).astype(np.complex128)
    got = o_binary(oracle, MUL, a, b)
    for x, y, g in zip(a, b, got):
        # (ac - bd, ad + bc), each product and sum rounded once
        re = x.real * y.real - x.imag * y.imag
        im = x.real * y.imag + x.imag * y.real
        assert g.real == re and g.imag == im


@pytest.mark.parametrize("op", [ADD, SUB])
def test_complex_add_sub_componentwise_ieee(oracle, op):
    """num-complex add/sub are componentwise IEEE operations: c128 against Python's complex
    arithmetic, c64 against NumPy's componentwise float32 add/sub, bit for bit, on strided views
    with a broadcast operand and signed zeros, infinities and NaN."""
    rng = np.random.default_rng(40 + op)
    re = rng.normal(size=(7, 9))
    im = rng.normal(size=(7, 9))
    re[0, :4] = [0.0, -0.0, np.inf, -np.inf]
    im[1, :4] = [-0.0, 0.0, np.nan, np.inf]
    z = np.empty(re.shape, np.complex128)
    z.real, z.imag = re, im                         # no 0 * inf products
    a = z.T[::2]                                    # transposed, stepped
    b = (rng.normal(size=7) - 1j * rng.normal(size=7))[None, :]  # broadcast row
    got = o_binary(oracle, op, a, b)
    f = (lambda x, y: x + y) if op == ADD else (lambda x, y: x - y)
    A, B = np.broadcast_arrays(a, b)
    ref = np.array([[f(complex(x), complex(y)) for x, y in zip(ra, rb)] for ra, rb in zip(A, B)])
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    a32, b32 = a.astype(np.complex64), b.astype(np.complex64)
    got32 = o_binary(oracle, op, a32, b32)
    A32, B32 = np.broadcast_arrays(a32, b32)
    g = np.add if op == ADD else np.subtract
    ref32 = np.empty(A32.shape, np.complex64)
    ref32.real = g(A32.real, B32.real)
    ref32.imag = g(A32.imag, B32.imag)
    assert np.array_equal(got32.view(np.uint32), ref32.view(np.uint32))


def test_complex_mul_python_builtin(oracle):
    """c128 multiply against CPython's own complex product (an independent C implementation of
    (ac - bd, ad + bc) with no fused multiply-add), bit for bit."""
    rng = np.random.default_rng(43)
    a = (rng.normal(size=500) + 1j * rng.normal(size=500)) * np.exp2(rng.integers(-30, 30, size=500))
    b = (rng.normal(size=500) + 1j * rng.normal(size=500)) * np.exp2(rng.integers(-30, 30, size=500))
    got = o_binary(oracle, MUL, a, b)
    ref = np.array([complex(x) * complex(y) for x, y in zip(a, b)])
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_complex_mul_exact_identities(oracle, dtype):
    """Products the mathematics fixes exactly: Gaussian integers (exact in both precisions, checked
    against Python integers), multiplication by i, by a real number, and z * conj(z) = |z|^2 + 0i
    (imaginary part a(-b) + ba = 0 exactly when nothing is fused)."""
    rng = np.random.default_rng(44)
    p = rng.integers(-2000, 2000, size=(4, 200))
    a = (p[0] + 1j * p[1]).astype(dtype)
    b = (p[2] + 1j * p[3]).astype(dtype)
    got = o_binary(oracle, MUL, a, b)
    assert np.array_equal(got.real, (p[0] * p[2] - p[1] * p[3]).astype(got.real.dtype))
    assert np.array_equal(got.imag, (p[0] * p[3] + p[1] * p[2]).astype(got.real.dtype))
    z = (rng.normal(size=300) + 1j * rng.normal(size=300)).astype(dtype)
    gi = o_binary(oracle, MUL, z, np.full_like(z, 1j))
    assert np.array_equal(gi.real, -z.imag) and np.array_equal(gi.imag, z.real)
    r = rng.normal(size=300).astype(z.real.dtype)
    gr = o_binary(oracle, MUL, z, r.astype(dtype))
    assert np.array_equal(gr.real, z.real * r) and np.array_equal(gr.imag, z.imag * r)
    gc = o_binary(oracle, MUL, z, np.conj(z))
    assert np.all(gc.imag == 0)
    rt_ = z.real.dtype.type
    assert np.array_equal(gc.real, np.array([rt_(x.real) * rt_(x.real) + rt_(x.imag) * rt_(x.imag) for x in z]))


def test_complex_mul_pin_bites(oracle):
    """The identities above fail for the plausible mistakes: a conjugated operand, swapped
    components, or the sign of bd flipped."""
    rng = np.random.default_rng(45)
    p = rng.integers(-50, 50, size=(4, 64))
    a = p[0] + 1j * p[1]
    b = p[2] + 1j * p[3]
    good = o_binary(oracle, MUL, a, b)
    for wrong in (a * np.conj(b), (a * b).imag + 1j * (a * b).real, (p[0] * p[2] + p[1] * p[3]) + 1j * (a * b).imag):
        assert not np.array_equal(good, wrong)


def test_assign_matches_numpy_copy(oracle):
    rng = np.random.default_rng(11)
    base = rng.normal(size=4 * 5 * 6)
    x = base.reshape(4, 5, 6).transpose(2, 0, 1)
    vc = c_view(x.shape, np.float64)
    oracle.assign(vc, View.of(x))
    assert np.array_equal(vc.array(), np.ascontiguousarray(x))
    # different shapes, same count: row-major logical order on both sides
    vc2 = c_view((12, 10), np.float64)
    oracle.assign(vc2, View.of(x))
    assert np.array_equal(vc2.array(), np.ascontiguousarray(x).reshape(12, 10))


@pytest.mark.parametrize("n", [1, 7, 1000, 65537])
def test_sum_is_correctly_rounded(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) * np.exp2(rng.integers(-30, 30, size=n))
    got = oracle.reduce_all(SUM, View.of(x))
    ref = math.fsum(x.tolist())
    assert got == ref or abs(got - ref) <= np.spacing(abs(ref))


def test_sum_axes_vs_fsum(oracle):
    rng = np.random.default_rng(5)
    x = rng.normal(size=(7, 9, 11)).transpose(1, 2, 0)
    got = o_reduce(oracle, SUM, x, [0, 2])
    for j in range(x.shape[1]):
        ref = math.fsum(x[:, j, :].ravel().tolist())
        assert abs(got[j] - ref) <= np.spacing(abs(ref))


def test_max_min_nan_rules(oracle):
    # maxNum: NaN ignored; start from f64::MIN so all -inf gives f64::MIN (ext_real.rs:70-87)
    assert oracle.reduce_all(MAX, View.of(np.array([np.nan, 1.0, 2.0]))) == 2.0
    assert oracle.reduce_all(MIN, View.of(np.array([3.0, np.nan, -1.0]))) == -1.0
    assert oracle.reduce_all(MAX, View.of(np.array([-np.inf, -np.inf]))) == -np.finfo(np.float64).max
    rng = np.random.default_rng(2)
    x = rng.normal(size=(8, 9, 10)).transpose(2, 0, 1)
    assert np.array_equal(o_reduce(oracle, MAX, x, [1]), x.max(axis=1))
    assert np.array_equal(o_reduce(oracle, MIN, x, [0, 2]), x.min(axis=(0, 2)))


def test_argmax_first_occurrence(oracle):
    rng = np.random.default_rng(9)
    x = rng.integers(-5, 5, size=(6, 7, 8)).astype(np.int32).transpose(1, 2, 0)  # many ties
    for ax in range(3):
        assert np.array_equal(o_reduce(oracle, ARGMAX, x, [ax]), np.argmax(x, axis=ax))
        assert np.array_equal(o_reduce(oracle, ARGMIN, x, [ax]), np.argmin(x, axis=ax))
    # multi-axis: ravel over the reduced axes in the order given
    for axes in ([1, 2], [2, 1], [0, 2]):
        moved = np.moveaxis(x, axes, [-2, -1])
        ref = np.argmax(moved.reshape(moved.shape[:-2] + (-1,)), axis=-1)
        assert np.array_equal(o_reduce(oracle, ARGMAX, x, axes), ref)
    assert oracle.reduce_all(ARGMAX, View.of(x)) == np.argmax(x)


def test_argmax_nan_rules(oracle):
    assert oracle.reduce_all(ARGMAX, View.of(np.array([np.nan, 5.0, 3.0]))) == 0
    assert oracle.reduce_all(ARGMAX, View.of(np.array([1.0, np.nan, 5.0]))) == 2
    assert oracle.reduce_all(ARGMIN, View.of(np.array([1.0, np.nan, -5.0, -5.0]))) == 2


def test_gemm_integer_exact(oracle):
    rng = np.random.default_rng(4)
    a = rng.integers(-(2 ** 20), 2 ** 20, size=(9, 13)).astype(np.int32)
    b = rng.integers(-(2 ** 20), 2 ** 20, size=(13, 7)).astype(np.int32)
    got = oracle.gemm_naive(a, b)
    ref = (a.astype(object) @ b.astype(object))
    ref = np.array([[int(v) % 2 ** 32 for v in row] for row in ref], dtype=np.uint64).astype(np.uint32).view(np.int32)
    assert np.array_equal(got, ref)


def test_gemm_f64_vs_extended(oracle):
    rng = np.random.default_rng(8)
    a = rng.normal(size=(33, 65)) * np.exp2(rng.integers(-8, 9, size=(33, 65)))
    b = rng.normal(size=(65, 17)) * np.exp2(rng.integers(-8, 9, size=(65, 17)))
    ref = a.astype(np.longdouble) @ b.astype(np.longdouble)
    naive = oracle.gemm_naive(a, b)
    dot2 = oracle.gemm_dot2(a, b)
    scale = np.abs(a) @ np.abs(b)
    assert np.max(np.abs(naive - ref) / scale) < 65 * 2.0 ** -52
    # Dot2 is accurate to ~u|c| + u^2 sum|a||b|: far tighter than the naive order
    assert np.all(np.abs(dot2 - ref) <= 2.0 ** -53 * np.abs(ref) + 2.0 ** -58 * scale)
    assert np.max(np.abs(naive - ref) / scale) > np.max(np.abs(dot2 - ref) / scale)


def test_gemm_strided_operands(oracle):
    rng = np.random.default_rng(12)
    A = rng.normal(size=(20, 30)).T  # column-major view
    B = rng.normal(size=(40, 20))[::2, :].T[::1, :]  # stepped
    got = oracle.gemm_naive(A, B[:, :17] if B.shape[1] > 17 else B)
    Bv = B[:, :17] if B.shape[1] > 17 else B
    ref = A.astype(np.longdouble) @ Bv.astype(np.longdouble)
    assert np.max(np.abs(got - ref)) < 1e-12


# ---------------- planted mistakes are caught ----------------

def test_pins_catch_planted_mistakes(oracle):
    # a dropped element changes the known sum
    base = np.arange(3240, dtype=np.int64)
    a = base.reshape(12, 15, 18).swapaxes(-1, -2)[2:-3, 1:-4:2, -1:3:-2]
    assert oracle.reduce_all(SUM, View.of(a[:, :, 1:])) != REF["sum_all_sliced_swapaxes"]["value"]
    # a transposed operand fails the matmul known answer
    r = REF["matmul_core"]
    m1 = np.array(r["m1"], dtype=np.int32)
    m2 = np.array(r["m2"], dtype=np.int32)
    assert oracle.gemm_naive(m1.T, m2).tolist() != r["m1_m2"]
    # a wrong sign (sub for add) fails the IEEE add pin
    x = np.linspace(1, 5, 5)
    y = np.linspace(2, 10, 5)
    assert not np.array_equal(o_binary(oracle, SUB, x, y), x + y)
    # last-occurrence tie rule would fail the argmax known answer
    d = np.array([3, 3, 3, 3, 2, 2, 2, 2], dtype=np.int32)
    assert len(d) - 1 - int(np.argmax(d[::-1])) != REF["argmax_core"]["all_cases"][1][1]


# ---------------- prod / mean / var / std / l2_norm ----------------


def _allclose_f64(a, b):
    """dev_utilities/mod.rs:11-36: |a - b| <= 1e-8 + 1e-5 |b| elementwise."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and bool(np.all(np.abs(a - b) <= 1e-8 + 1e-5 * np.abs(b)))


def test_known_moments(oracle):
    r = REF["moments_4x3"]
    a = np.array(r["data"], dtype=np.float64).reshape(r["shape"])
    assert abs(oracle.reduce_all(VAR, View.of(a)) - r["var_all"]) < 1e-10
    assert _allclose_f64(o_reduce(oracle, VAR, a, [0]), r["var_axes_0"])
    assert _allclose_f64(o_reduce(oracle, VAR, a, [1]), r["var_axes_1"])
    assert abs(oracle.reduce_all(STD, View.of(a)) - r["std_all"]) < 1e-10
    assert _allclose_f64(o_reduce(oracle, STD, a, [0]), r["std_axes_0"])
    assert _allclose_f64(o_reduce(oracle, STD, a, [1]), r["std_axes_1"])
    z = (np.array(r["data"], dtype=np.float64) + 1j * np.array(r["imag"], dtype=np.float64)).reshape(r["shape"])
    assert abs(oracle.reduce_all(STD, View.of(z)) - r["std_all_complex"]) < 1e-10
    assert abs(oracle.reduce_all(L2, View.of(z)) - r["l2_norm_all_complex"]) < 1e-10


def test_known_mean_direct(oracle):
    r = REF["mean_arange24"]
    a = np.arange(24.0).reshape(2, 3, 4)
    assert oracle.reduce_all(MEAN, View.of(a)) == r["mean_all"]
    assert list(o_reduce(oracle, MEAN, a, [0, 2])) == r["mean_axes_0_2"]
    assert list(o_reduce(oracle, MEAN, a[::-1, :, ::-2], [-1, 1])) == r["flipped_mean"]


def test_var_std_are_population_statistics(oracle):
    """statistics.pvariance works in exact rational arithmetic, so on data with a small mean the
    reference formula (E|x|^2 - |E x|^2 with compensated sums) must agree to a few ulp."""
    import statistics

    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 100, 1001):
        x = rng.normal(size=n)
        pv = statistics.pvariance(x.tolist())
        assert abs(oracle.reduce_all(VAR, View.of(x)) - pv) <= 1e-13 * max(pv, 1e-300) + 1e-300
        assert abs(oracle.reduce_all(STD, View.of(x)) - math.sqrt(pv)) <= 1e-13 * math.sqrt(pv) + 1e-300
    # complex: population variance of x is var(re) + var(im)
    z = rng.normal(size=257) + 1j * rng.normal(size=257)
    pv = statistics.pvariance(z.real.tolist()) + statistics.pvariance(z.imag.tolist())
    assert abs(oracle.reduce_all(VAR, View.of(z)) - pv) <= 1e-13 * pv


def test_l2_mean_prod_closed_forms(oracle):
    rng = np.random.default_rng(4)
    x = rng.normal(size=513)
    assert abs(oracle.reduce_all(L2, View.of(x)) - math.sqrt(math.fsum(x * x))) <= 2 ** -52 * math.sqrt(math.fsum(x * x))
    assert abs(oracle.reduce_all(MEAN, View.of(x)) - math.fsum(x) / 513) <= 2 ** -52 * abs(math.fsum(x) / 513)
    # prod: the sequential fold in the element type, bit for bit (Python floats are f64)
    y = rng.uniform(0.5, 1.5, size=200)
    acc = 1.0
    for v in y:
        acc *= float(v)
    assert oracle.reduce_all(PROD, View.of(y)) == acc
    # integers wrap (i32): exact modular product
    k = rng.integers(-1000, 1000, size=50).astype(np.int32)
    ref = 1
    for v in k.tolist():
        ref = (ref * v) % (1 << 32)
    ref = ref - (1 << 32) if ref >= (1 << 31) else ref
    assert int(oracle.reduce_all(PROD, View.of(k))) == ref
    # complex: Python's complex product on the same order
    z = rng.uniform(0.5, 1.2, size=40) + 1j * rng.uniform(-0.3, 0.3, size=40)
    accz = complex(1, 0)
    for v in z:
        accz = complex(accz.real * v.real - accz.imag * v.imag, accz.real * v.imag + accz.imag * v.real)
    assert oracle.reduce_all(PROD, View.of(z)) == accz


def test_empty_reduction_values(oracle):
    e = np.zeros((0,))
    assert oracle.reduce_all(PROD, View.of(e)) == 1.0
    assert math.isnan(oracle.reduce_all(MEAN, View.of(e)))
    assert math.isnan(oracle.reduce_all(VAR, View.of(e)))
    assert oracle.reduce_all(L2, View.of(e)) == 0.0


def test_planted_mistakes_in_moments_are_caught(oracle):
    """A sample variance (divide by n-1), a mean over the wrong count and a missing square root
    each fail the pins above."""
    import statistics

    x = np.random.default_rng(5).normal(size=100)
    pv = statistics.pvariance(x.tolist())
    sample = oracle.reduce_all(VAR, View.of(x)) * 100 / 99
    assert abs(sample - pv) > 1e-13 * pv
    wrong_mean = oracle.reduce_all(SUM, View.of(x)) / 99
    assert abs(wrong_mean - math.fsum(x) / 100) > 2 ** -52 * abs(math.fsum(x) / 100)
    r = REF["moments_4x3"]
    a = np.array(r["data"], dtype=np.float64).reshape(r["shape"])
    assert abs(oracle.reduce_all(VAR, View.of(a)) - r["std_all"]) > 1e-10
