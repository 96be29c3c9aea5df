"""HIP reductions vs the oracle.

max/min/argmax/argmin and integer sums: bit-exact. Floating sums: the device accumulates
in FP64 with TwoSum compensation, the oracle in double-double; both are the exact sum
rounded once up to O(n u^2) terms, so the test bound is 2 ulp of the result plus
1e-30 * sum|x| (stated tolerance; the reference's own sums are only ~n*u accurate).
"""
import zlib

import numpy as np
import pytest

import rstsr_amd as rt
from gpu_util import bits_equal, from_np, oracle_reduce, rand
from oracle_ffi import View

pytestmark = pytest.mark.gpu

SUM, MAX, MIN, ARGMAX, ARGMIN = 0, 1, 2, 3, 4
METH = {SUM: "sum_axes", MAX: "max_axes", MIN: "min_axes", ARGMAX: "argmax_axes", ARGMIN: "argmin_axes"}
ALL = {SUM: "sum_all", MAX: "max_all", MIN: "min_all", ARGMAX: "argmax_all", ARGMIN: "argmin_all"}


def check_sum(got, ref, absx_sum, dtype):
    got = np.asarray(got, dtype=np.float64 if np.dtype(dtype).kind != "c" else np.complex128)
    ref = np.asarray(ref, dtype=got.dtype)
    eps = np.finfo(np.dtype(dtype).type(0).real.dtype if np.dtype(dtype).kind == "c" else dtype).eps
    tol = 2 * eps * np.abs(ref) + 1e-30 * absx_sum
    assert np.all(np.abs(got - ref) <= tol), float(np.max(np.abs(got - ref) / np.maximum(tol, 1e-300)))


CASES = [
    # (X, perm, step, axes)  view = base.reshape(X).transpose(perm)[..., ::step]
    ((16, 32, 64), (2, 0, 1), 2, [0]),          # unit-stride reduced axis -> row kernel
    ((16, 32, 64), (2, 0, 1), 2, [1, 2]),       # unit-stride kept axis -> column kernel + split
    ((16, 32, 64), (2, 0, 1), 2, [2, 1]),
    ((64, 128, 256), (0, 1, 2), 1, [1]),        # the small config's sum over axis 1
    ((64, 128, 256), (0, 1, 2), 1, [2]),        # ... and argmax over axis 2
    ((9, 300, 7), (1, 0, 2), 1, [0, 2]),
    ((4, 6, 15, 9), (2, 0, 3, 1), 1, [0, -2]),
    ((5, 6, 7), (0, 1, 2), 1, [0, 1, 2]),
    ((2000, 3), (0, 1), 1, [0]),                # long strided reduced axis, few outputs
    ((3, 100000), (0, 1), 1, [1]),              # long unit reduced axis, few outputs (split)
    ((1000, 70), (0, 1), 1, [0]),               # column kernel, 15 splits of 67 rows: ragged unroll tails
    ((999, 33), (0, 1), 1, [0]),                # odd kept extent: one output per thread
    ((37, 5, 96), (1, 0, 2), 1, [0, 1]),        # two reduced dims that do not chain: runs of 37
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.int32, np.int64])
@pytest.mark.parametrize("op", [SUM, MAX, MIN, ARGMAX, ARGMIN])
@pytest.mark.parametrize("ci", range(len(CASES)))
def test_reduce_axes(dev, oracle, dtype, op, ci):
    X, perm, step, axes = CASES[ci]
    rng = np.random.default_rng(zlib.crc32(repr((ci, op, str(dtype))).encode()))
    n = int(np.prod(X))
    base = rand(rng, n, dtype)
    if op in (ARGMAX, ARGMIN) and np.dtype(dtype).kind == "i":
        base = rng.integers(-8, 8, size=n).astype(dtype)  # many ties
    x = base.reshape(X).transpose(perm)[..., ::step]
    tx, vx = from_np(dev, x)
    out = getattr(tx, METH[op])(axes)
    ref = oracle_reduce(oracle, op, vx, axes, dtype)
    got = out.to_numpy()
    assert got.shape == ref.shape
    if op == SUM and np.dtype(dtype).kind == "f":
        check_sum(got, ref, np.abs(x).sum(), dtype)
    else:
        assert bits_equal(got, ref)
    # output layout: layout_for_array_copy(rest, K)
    ls, lm = rt._rstsr.dim_split_axes(tx.layout, axes)
    assert out.layout == rt._rstsr.layout_for_array_copy(lm, "K")


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.int32, np.int64])
@pytest.mark.parametrize("op", [SUM, MAX, MIN, ARGMAX, ARGMIN])
@pytest.mark.parametrize("n", [1, 63, 4097, 1 << 20])
def test_reduce_all(dev, oracle, dtype, op, n):
    rng = np.random.default_rng(n + op)
    x = rand(rng, 2 * n, dtype)[::2]  # strided
    tx, vx = from_np(dev, x)
    got = getattr(tx, ALL[op])()
    ref = oracle.reduce_all(op, vx)
    if op == SUM and np.dtype(dtype).kind == "f":
        check_sum(got, ref, np.abs(x).sum(), dtype)
    else:
        assert got == ref


def test_complex_sum(dev, oracle):
    rng = np.random.default_rng(2)
    x = (rng.normal(size=(64, 96)) + 1j * rng.normal(size=(64, 96))).T
    tx, vx = from_np(dev, x)
    ref = oracle_reduce(oracle, SUM, vx, [1], np.complex128)
    check_sum(tx.sum_axes([1]).to_numpy(), ref, np.abs(x).sum(), np.complex128)
    check_sum(tx.sum_all(), oracle.reduce_all(SUM, vx), np.abs(x).sum(), np.complex128)


def test_nan_rules(dev):
    t = rt.asarray(np.array([np.nan, 1.0, 2.0]), device=dev)
    assert t.max_all() == 2.0 and t.argmax_all() == 0
    t = rt.asarray(np.array([1.0, np.nan, 5.0, 5.0]), device=dev)
    assert t.argmax_all() == 2 and t.max_all() == 5.0
    t = rt.asarray(np.array([-np.inf, -np.inf]), device=dev)
    assert t.max_all() == -np.finfo(np.float64).max
    # NaN at logical index 0 of each reduced row wins; later NaNs ignored (large row -> split path)
    x = np.random.default_rng(0).normal(size=(3, 50000))
    x[0, 0] = np.nan
    x[1, 17] = np.nan
    x[2, 49999] = 100.0
    got = rt.asarray(x, device=dev).argmax_axes(1).to_vec()
    assert got[0] == 0 and got[1] == int(np.nanargmax(x[1])) and got[2] == 49999


def test_known_answers(dev):
    a = rt.arange(3240, device=dev).reshape([12, 15, 18]).swapaxes(-1, -2)
    assert a[2:-3, 1:-4:2, -1:3:-2].sum_all() == 446586
    s = rt.arange(3240, device=dev).reshape([4, 6, 15, 9]).transpose([2, 0, 3, 1]).sum_axes([0, -2]).to_numpy()
    assert s[0, 1] == 27270 and s[1, 2] == 154845 and s[3, 5] == 428220
    v = rt.asarray(np.array([8, 4, 2, 9, 7, 1, 2, 1, 8, 6, 10, 5], dtype=np.int32).reshape(4, 3), device=dev)
    assert v.argmin_all() == 5
    assert v.argmin_axes(-1).to_vec() == [2, 2, 1, 2]
    assert v.argmin_axes(0).to_vec() == [2, 2, 1]
    b = rt.asarray(np.array([[3, 6, 9], [4, 10, 5], [8, 3, 2]], dtype=np.int32), device=dev)
    assert b.argmax_axes(0).to_vec() == [2, 1, 0] and b.argmax_axes(1).to_vec() == [2, 1, 0]
    m = rt.arange(24.0, device=dev).reshape([2, 3, 4])
    assert (m.sum_axes((0, 2)).to_numpy() / 8).tolist() == [7.5, 11.5, 15.5]
    assert (m[::-1, :, ::-2].sum_axes((-1, 1)).to_numpy() / 6).tolist() == [18.0, 6.0]
    assert m.sum_all() / 24 == 11.5


def test_reduce_errors(dev):
    e = rt.zeros((0,), device=dev)
    assert e.sum_all() == 0.0
    with pytest.raises(rt.RSTSRError) as x:
        e.max_all()
    assert x.value.args[0] == "InvalidValue"
    with pytest.raises(rt.RSTSRError) as x:
        e.argmax_all()
    assert x.value.args[0] == "InvalidLayout"
    z = rt.zeros((3, 0), device=dev).sum_axes([1]).to_numpy()
    assert z.tolist() == [0.0, 0.0, 0.0]
    with pytest.raises(rt.RSTSRError):
        rt.zeros((3, 4), device=dev).sum_axes([0, 0])


# ---------------- prod / mean / var / std / l2_norm (auto_impl/reduction.rs:131-354) ----------------
PROD, MEAN, VAR, STD, L2 = 5, 6, 7, 8, 9
MOM = {PROD: "prod", MEAN: "mean", VAR: "var", STD: "std", L2: "l2_norm"}


def check_moment(op, got, ref, x, axes, dtype):
    """Stated tolerances. mean/l2: both sides are the compensated sum rounded once, then divided or
    square-rooted: 4 ulp of the result. var/std: E|x|^2 - |Ex|^2 cancels; the bound is 8 ulp of
    E|x|^2 (absolute), and for std that bound propagated through the square root. prod: the
    device multiplies in f64 in another order than the oracle's sequential fold; relative error
    <= 2 n u of the element type."""
    kind = np.dtype(dtype).kind
    rdt = np.dtype(dtype).type(0).real.dtype if kind == "c" else np.dtype(dtype)
    if kind == "i":
        assert bits_equal(got, ref)
        return
    eps = float(np.finfo(rdt).eps)
    g = np.asarray(got, dtype=np.complex128 if (kind == "c" and op in (PROD, MEAN)) else np.float64)
    r = np.asarray(ref, dtype=g.dtype)
    nd = x.ndim
    ax = tuple(a % nd for a in axes)
    n = int(np.prod([x.shape[a] for a in ax])) if ax else 1
    if op in (MEAN, L2):
        tol = 4 * eps * np.abs(r) + 1e-300
    elif op == VAR:
        msq = np.mean(np.abs(x.astype(np.complex128 if kind == "c" else np.float64)) ** 2, axis=ax)
        tol = 8 * eps * msq + 1e-300
    elif op == STD:
        msq = np.mean(np.abs(x.astype(np.complex128 if kind == "c" else np.float64)) ** 2, axis=ax)
        tol = 8 * eps * msq / np.maximum(np.abs(r), 1e-300) + 4 * eps * np.abs(r) + 1e-300
    else:
        tol = 2 * n * eps * np.abs(r) + 1e-300
    err = np.abs(g - r)
    assert np.all(err <= tol), float(np.max(err / tol))


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.complex64, np.int32, np.int64])
@pytest.mark.parametrize("op", [PROD, MEAN, VAR, STD, L2])
@pytest.mark.parametrize("ci", [0, 1, 5, 6, 8, 9])
def test_moment_axes(dev, oracle, dtype, op, ci):
    if np.dtype(dtype).kind == "i" and op != PROD:
        pytest.skip("mean/var/std/l2_norm are defined for floating and complex types (ComplexFloat)")
    X, perm, step, axes = CASES[ci]
    rng = np.random.default_rng(zlib.crc32(repr(("mom", ci, op, str(dtype))).encode()))
    n = int(np.prod(X))
    if op == PROD:
        # factors near one keep long products finite; integers in [-3, 3] wrap exactly
        if np.dtype(dtype).kind == "i":
            base = rng.integers(-3, 4, size=n).astype(dtype)
        elif np.dtype(dtype).kind == "c":
            base = (np.exp(rng.normal(0, 1e-3, size=n)) * np.exp(1j * rng.normal(0, 1e-3, size=n))).astype(dtype)
        else:
            base = np.exp(rng.normal(0, 1e-3, size=n)).astype(dtype)
    else:
        base = rand(rng, n, dtype)
        if op != L2:
            base = base + np.asarray(0.25, dtype=base.dtype)  # a nonzero mean exercises the cancellation
    x = base.reshape(X).transpose(perm)[..., ::step]
    tx, vx = from_np(dev, x)
    out = getattr(tx, MOM[op] + "_axes")(axes)
    ref = oracle_reduce(oracle, op, vx, axes, dtype)
    got = out.to_numpy()
    assert got.shape == ref.shape and got.dtype == ref.dtype
    check_moment(op, got, ref, x, axes, dtype)
    ls, lm = rt._rstsr.dim_split_axes(tx.layout, axes)
    assert out.layout == rt._rstsr.layout_for_array_copy(lm, "K")


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32])
@pytest.mark.parametrize("op", [PROD, MEAN, VAR, STD, L2])
@pytest.mark.parametrize("n", [1, 4097, 1 << 20])
def test_moment_all(dev, oracle, dtype, op, n):
    rng = np.random.default_rng(n * 10 + op)
    if op == PROD:
        x = np.exp(rng.normal(0, 1e-4, size=2 * n)).astype(dtype)[::2]
    else:
        x = rand(rng, 2 * n, dtype)[::2]
    tx, vx = from_np(dev, x)
    got = getattr(tx, MOM[op] + "_all")()
    ref = oracle.reduce_all(op, vx)
    check_moment(op, np.asarray(got), np.asarray(ref), x, [0], dtype)


def test_moment_known_answers(dev):
    """tensor/reduction.rs:557-743: mean (arange, flipped slices), var/std (4x3), complex std and
    l2_norm, with the reference's own tolerances."""
    import json
    import os

    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_cases.json")))
    r = ref["mean_arange24"]
    a = rt.asarray(np.arange(24.0).reshape(2, 3, 4), device=dev)
    assert a.mean_all() == r["mean_all"]
    assert list(a.mean_axes([0, 2]).to_numpy()) == r["mean_axes_0_2"]
    assert list(a[::-1, :, ::-2].mean_axes([-1, 1]).to_numpy()) == r["flipped_mean"]
    r = ref["moments_4x3"]
    m = rt.asarray(np.array(r["data"], dtype=np.float64).reshape(r["shape"]), device=dev)

    def allclose_f64(g, want):
        g, want = np.asarray(g), np.asarray(want)
        return g.shape == want.shape and bool(np.all(np.abs(g - want) <= 1e-8 + 1e-5 * np.abs(want)))

    assert abs(m.var_all() - r["var_all"]) < 1e-10
    assert allclose_f64(m.var_axes([0]).to_numpy(), r["var_axes_0"])
    assert allclose_f64(m.var_axes([1]).to_numpy(), r["var_axes_1"])
    assert abs(m.std_all() - r["std_all"]) < 1e-10
    assert allclose_f64(m.std_axes([0]).to_numpy(), r["std_axes_0"])
    assert allclose_f64(m.std_axes([1]).to_numpy(), r["std_axes_1"])
    z = rt.asarray((np.array(r["data"], dtype=np.float64) + 1j * np.array(r["imag"], dtype=np.float64)).reshape(4, 3),
                   device=dev)
    assert abs(z.std_all() - r["std_all_complex"]) < 1e-10
    assert abs(z.l2_norm_all() - r["l2_norm_all_complex"]) < 1e-10
    assert z.std_axes([0]).to_numpy().dtype == np.float64


def test_moment_empty_values(dev):
    """Zero reduced elements: prod 1, l2_norm 0, mean/var/std NaN (f_out(f_init()) in the reference)."""
    e = rt.asarray(np.zeros((3, 0)), device=dev)
    assert np.all(e.prod_axes([1]).to_numpy() == 1.0)
    assert np.all(e.l2_norm_axes([1]).to_numpy() == 0.0)
    assert np.all(np.isnan(e.mean_axes([1]).to_numpy()))
    assert np.isnan(e.std_all())
    assert e.prod_all() == 1.0


# A broadcast (stride-0) axis among the reduced ones is summed as repeated additions of the same
# element (cpu_rayon/reduction.rs:189-193); max/argmax see ties along it (first index wins).
BCAST = [
    ((5, 6, 7), (7, 0, 1), [1]),
    ((5, 6, 7), (7, 0, 1), [0, 1]),
    ((5, 6, 7), (7, 0, 1), [1, 2]),
    ((5, 6, 7), (7, 0, 1), [0, 1, 2]),
    ((6, 7, 5), (0, 1, 7), [0]),           # stride-0 reduced axis first, kept unit axis
    ((6, 7, 5), (0, 1, 7), [2, 0]),
    ((4096, 35), (0, 1), [0]),             # long stride-0 reduced axis (column kernel, split)
    ((35, 4096), (1, 0), [1]),             # long stride-0 reduced axis, the kept axis strided
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.int32])
@pytest.mark.parametrize("op", [SUM, MAX, MIN, ARGMAX, ARGMIN])
@pytest.mark.parametrize("ci", range(len(BCAST)))
def test_reduce_axes_over_broadcast_axis(dev, oracle, dtype, op, ci):
    from gpu_util import pair

    shape, stride, axes = BCAST[ci]
    rng = np.random.default_rng(1000 + ci * 17 + op)
    n = 35
    base = rand(rng, n, dtype)
    tx, vx = pair(dev, base, shape, stride)
    got = getattr(tx, METH[op])(axes).to_numpy()
    ref = oracle_reduce(oracle, op, vx, axes, dtype)
    assert got.shape == ref.shape
    if op == SUM and np.dtype(dtype).kind == "f":
        check_sum(got, ref, float(np.abs(vx.array()).sum()), dtype)
        # repeated additions of one value: the count times the value, rounded once
        if axes == [1] and shape == (5, 6, 7):
            exact = 6 * base.reshape(5, 7).astype(np.float64)
            assert np.all(np.abs(got - exact) <= 2 * np.finfo(dtype).eps * np.abs(exact))
    else:
        assert bits_equal(got, ref)
    if op in (ARGMAX, ARGMIN) and axes == [1] and shape == (5, 6, 7):
        assert np.all(got == 0)  # every position along the broadcast axis ties: the first wins


@pytest.mark.parametrize("op", [SUM, MAX, ARGMAX])
def test_reduce_all_over_broadcast_view(dev, oracle, op):
    from gpu_util import pair

    rng = np.random.default_rng(31 + op)
    base = rng.normal(size=9)
    tx, vx = pair(dev, base, (4, 9, 1000), (0, 1, 0))
    got = getattr(tx, ALL[op])()
    ref = oracle.reduce_all(op, vx)
    if op == SUM:
        assert abs(got - ref) <= 2 * 2.0 ** -52 * abs(ref) + 1e-30 * np.abs(base).sum() * 4000
    else:
        assert got == ref
