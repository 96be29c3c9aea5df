"""The reference's shared core_func test bodies run on the HIP device through the front end.

The reference runs rstsr-core/tests/core_func on each device crate through a symlink
(crates-device/rstsr-openblas/tests/tests_core_row.rs:1-14), comparing with
test_utils/equality.rs:13-27 (equal shapes, then rt::allclose). Here the known answers of
manipulation/test_to_contig.rs, test_transpose.rs, test_broadcast.rs,
operators/test_arithmetic.rs and reduction/test_{min,argmin,mean,prod,std,var}.rs are kept as
data in tests/golden/core_func_cases.json; each test below restates one reference test (cited
by file:line) with the device's default order set as that test sets it (RowMajor or ColMajor).
"""
import json
import os

import numpy as np
import pytest

import rstsr_amd as rt

pytestmark = pytest.mark.gpu

CASES = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "core_func_cases.json")))
DT = {"i32": np.int32, "f64": np.float64}


def device(dev, order):
    d = dev.clone()
    d.set_default_order(order)
    return d


def nested(x, d, dt=None):
    return rt.tensor_from_nested(np.array(x, dtype=dt), device=d)


def assert_equal(a, b):
    """test_utils/equality.rs:13-27: equal shapes, then rt::allclose (rtol 1e-5, atol 1e-8) on the
    device."""
    assert a.shape == b.shape, (a.shape, b.shape)
    if np.dtype(a.dtype) != np.dtype(b.dtype):
        b = b.astype(a.dtype)
    assert rt.allclose(a, b)


# ------------------------------ operators/test_arithmetic.rs ------------------------------

@pytest.mark.parametrize("case", CASES["arithmetic"], ids=lambda c: c["ref"] + ":" + c["op"])
def test_arithmetic(dev, case):
    d = device(dev, "RowMajor")
    dt = DT[case["dtype"]]
    a, b = nested(case["a"], d, dt), nested(case["b"], d, dt)
    fn = {"add": rt.add, "sub": rt.sub, "mul": rt.mul, "div": rt.div, "rem": rt.rem, "matmul": rt.matmul}[case["op"]]
    exp = nested(case["expected"], d, dt)
    assert_equal(fn(a, b), exp)
    op = {"add": "__add__", "sub": "__sub__", "mul": "__mul__", "div": "__truediv__", "matmul": "__mod__"}.get(case["op"])
    if op:  # the operator form (&a + &b, ..., &a % &b as matmul)
        assert_equal(getattr(a, op)(b), exp)


def test_arithmetic_colmajor_same_shape(dev):
    """The same-shape cases under ColMajor (f-contiguous storage, F-order broadcasting rules): the
    values are order independent."""
    d = device(dev, "ColMajor")
    for case in CASES["arithmetic"]:
        if np.array(case["a"]).shape != np.array(case["b"]).shape or case["op"] == "matmul":
            continue
        dt = DT[case["dtype"]]
        a, b = nested(case["a"], d, dt), nested(case["b"], d, dt)
        assert a.f_contig() and not a.c_contig()
        fn = {"add": rt.add, "sub": rt.sub, "mul": rt.mul, "div": rt.div, "rem": rt.rem}[case["op"]]
        assert_equal(fn(a, b), nested(case["expected"], d, dt))


def test_broadcast_colmajor_aligns_leading_axes(dev):
    """ColMajor broadcasting aligns the leading axes (broadcast.rs:21-60 with the shapes reversed):
    [2,3] + [2] adds b along axis 0, which RowMajor rejects."""
    a_h = np.arange(6, dtype=np.float64).reshape(2, 3)
    b_h = np.array([10.0, 20.0])
    dc = device(dev, "ColMajor")
    got = (nested(a_h, dc) + nested(b_h, dc)).to_numpy()
    assert np.array_equal(got, a_h + b_h[:, None])
    dr = device(dev, "RowMajor")
    with pytest.raises(rt.RSTSRError):
        nested(a_h, dr) + nested(b_h, dr)


@pytest.mark.parametrize("shape_a,shape_b", [([4, 6], [6, 5]), ([4, 6, 3], [6, 5]), ([4, 6], [6, 5, 3]),
                                             ([4, 6, 3], [6, 5, 3]), ([6], [6, 5]), ([4, 6], [6])])
def test_matmul_colmajor_reverse_and_swap(dev, oracle, shape_a, shape_b):
    """ColMajor matmul (device_faer/matmul.rs:227-237): the device call reverses every layout and
    swaps the operands; the result equals the ColMajor rules of layout/matmul.rs:423-455, i.e.
    C[..] = A[..] B[..] batched over the TRAILING axes, f-contiguous."""
    rng = np.random.default_rng(20261230)
    A = rng.standard_normal(shape_a)
    B = rng.standard_normal(shape_b)
    d = device(dev, "ColMajor")
    # ColMajor batches over trailing axes: the reversed problem is the row-major one
    got = rt.matmul(nested(A, d), nested(B, d))
    ref = np.matmul(B.T, A.T).T
    assert got.shape == ref.shape
    gn = got.to_numpy()
    k = 6  # the contracted extent in every case
    assert np.all(np.abs(gn - ref) <= 4 * k * 2.0 ** -53 * np.matmul(np.abs(B).T, np.abs(A).T).T + 1e-300)
    if got.ndim >= 2:
        assert got.f_contig()


# ------------------------------ reductions ------------------------------

def _red(t, op, axes):
    if op == "argmin":
        return t.argmin_axes(axes)
    return getattr(t, op + "_axes")(axes)


@pytest.mark.parametrize("case", CASES["reduce_all"], ids=lambda c: c["ref"] + ":" + c["op"])
def test_reduce_all(dev, case):
    d = device(dev, "RowMajor")
    a = nested(case["a"], d, DT[case["dtype"]])
    got = getattr(a, case["op"] + "_all")()
    assert got == case["expected"]


@pytest.mark.parametrize("case", CASES["reduce_axes"], ids=lambda c: c["ref"] + ":" + c["op"] + ":" + str(c["axes"]))
def test_reduce_axes(dev, case):
    d = device(dev, "RowMajor")
    dt = DT[case["dtype"]]
    a = nested(case["a"], d, dt)
    out = _red(a, case["op"], case["axes"])
    if case["op"] == "argmin":
        assert out.to_numpy().tolist() == case["expected"]
        return
    exp = np.array(case["expected"], dtype=dt)
    if exp.ndim == 0:
        assert out.ndim == 0  # axes None reduces to a 0-d tensor
        assert out.to_scalar() == pytest.approx(case["expected"], rel=1e-15, abs=0)
    assert_equal(out, nested(exp, d, dt))


def test_argmin_axes_high_rank(dev):
    """reduction/test_argmin.rs:62-77: argmin over each axis of a rank-5 tensor succeeds (and, for
    arange, is the first index everywhere)."""
    d = device(dev, "RowMajor")
    a = rt.arange(4 * 5 * 6 * 7 * 8, device=d, dtype=np.int32).into_shape([4, 5, 6, 7, 8])
    for i in range(a.ndim):
        r = a.argmin_axes(i)
        assert np.all(r.to_numpy() == 0)


def test_reductions_colmajor(dev):
    """The same reductions on ColMajor tensors (f-contiguous storage): values do not depend on the
    order; the kept axes keep the input's stride order (layout_for_array_copy(K))."""
    d = device(dev, "ColMajor")
    for case in CASES["reduce_axes"]:
        dt = DT[case["dtype"]]
        a = nested(case["a"], d, dt)
        out = _red(a, case["op"], case["axes"])
        if case["op"] == "argmin":
            assert out.to_numpy().tolist() == case["expected"]
        else:
            # var/std use sum|x|^2/n - |mean|^2, whose rounding reaches a few ulp of the exact value
            assert np.allclose(out.to_numpy(), np.array(case["expected"]), rtol=1e-14, atol=0)


# ------------------------------ manipulation/test_transpose.rs ------------------------------

@pytest.mark.parametrize("case", CASES["transpose"], ids=lambda c: c["ref"])
def test_transpose(dev, case):
    d = device(dev, "RowMajor")
    a = nested(case["a"], d, np.int32)
    exp = nested(case["expected"], d, np.int32)
    assert_equal(rt.transpose(a, case["axes"]), exp)
    if case["axes"] is None:
        assert_equal(a.T, exp)


@pytest.mark.parametrize("case", CASES["transpose_errors"], ids=lambda c: str(c["axes"]))
def test_transpose_errors(dev, case):
    d = device(dev, "RowMajor")
    a = rt.zeros(case["shape"], np.int32, device=d)
    with pytest.raises(rt.RSTSRError):
        rt.transpose(a, case["axes"])


def test_transpose_sixteen_axes(dev):
    """manipulation/test_transpose.rs:58-73: a 16-dimensional identity permutation succeeds."""
    d = device(dev, "RowMajor")
    x = rt.arange(65536, device=d).into_shape([2] * 16)
    y = rt.transpose(x, list(range(16)))
    assert y.shape == tuple([2] * 16) and y.as_ptr() == x.as_ptr()


def test_swapaxes_numeric(dev):
    c = CASES["swapaxes"][0]
    d = device(dev, "RowMajor")
    a = nested(c["a"], d, np.int32)
    assert_equal(rt.swapaxes(a, c["i"], c["j"]), nested(c["expected"], d, np.int32))


def test_swapaxes_multiarray(dev):
    """manipulation/test_transpose.rs:99-163: out-of-range axes raise; for every (i, j) in
    [-4, 4)^2 the shape swaps, every element moves accordingly, and the result is a view."""
    d = device(dev, "RowMajor")
    a = rt.arange(24, device=d).into_shape([1, 2, 3, 4])
    for i, j in CASES["swapaxes_errors"][0]["pairs"]:
        with pytest.raises(rt.RSTSRError):
            a.swapaxes(i, j)
    ah = a.to_numpy()
    for i in range(-4, 4):
        for j in range(-4, 4):
            c = a.swapaxes(i, j)
            exp_shape = list(a.shape)
            iu, ju = i % 4, j % 4
            exp_shape[iu], exp_shape[ju] = exp_shape[ju], exp_shape[iu]
            assert list(c.shape) == exp_shape
            assert np.array_equal(c.to_numpy(), np.swapaxes(ah, iu, ju))
            assert c.as_ptr() == a.as_ptr() and not c.is_owned()


# ------------------------------ manipulation/test_broadcast.rs ------------------------------

def _bsource(name, d):
    return {"scalar0": lambda: rt.asarray(np.array(0, dtype=np.int32), device=d),
            "ones1": lambda: rt.ones([1], device=d),
            "ones0": lambda: rt.ones([0], device=d),
            "ones12": lambda: rt.ones([1, 2], device=d),
            "ones21": lambda: rt.ones([2, 1], device=d),
            "arange3": lambda: rt.arange(3, device=d, dtype=np.int32)}[name]()


@pytest.mark.parametrize("case", CASES["broadcast_to"], ids=lambda c: c["input"] + str(c["shape"]))
def test_broadcast_to_succeeds(dev, case):
    d = device(dev, "RowMajor")
    src = _bsource(case["input"], d)
    r = rt.broadcast_to(src, case["shape"])
    assert list(r.shape) == case["shape"]
    if "expected" in case:
        assert_equal(r, nested(case["expected"], d, src.dtype))


@pytest.mark.parametrize("case", CASES["broadcast_to_errors"], ids=lambda c: f"{c['from']}->{c['to']}")
def test_broadcast_to_raises(dev, case):
    d = device(dev, "RowMajor")
    with pytest.raises(rt.RSTSRError):
        rt.broadcast_to(rt.zeros(case["from"], device=d), case["to"])


def test_broadcast_arrays(dev):
    """manipulation/test_broadcast.rs:159-228: shapes (1,1)&(3,4) -> (3,4); 32 copies of (1,2);
    identical inputs unchanged; the one-off row/column case."""
    d = device(dev, "RowMajor")
    r = rt.broadcast_arrays([rt.ones([1, 1], device=d), rt.ones([3, 4], device=d)])
    assert r[0].shape == (3, 4) and r[1].shape == (3, 4)
    a = rt.ones([1, 2], device=d)
    assert all(t.shape == (1, 2) for t in rt.broadcast_arrays([a.view() for _ in range(32)]))
    x = rt.arange(10, device=d)
    bx, by = rt.broadcast_arrays([x.view(), x.view()])
    assert_equal(bx, x)
    assert_equal(by, x)
    c = CASES["broadcast_one_off"]
    bx, by = rt.broadcast_arrays([nested(c["x"], d, np.int32), nested(c["y"], d, np.int32)])
    assert bx.shape == (3, 3) and by.shape == (3, 3)
    assert_equal(bx, nested(c["bx"], d, np.int32))
    assert_equal(by, nested(c["by"], d, np.int32))


@pytest.mark.parametrize("case", CASES["broadcast_arrays_errors"], ids=lambda c: str(c["shapes"]))
def test_broadcast_arrays_raise(dev, case):
    d = device(dev, "RowMajor")
    with pytest.raises(rt.RSTSRError):
        rt.broadcast_arrays([rt.zeros(s, device=d) for s in case["shapes"]])


# ------------------------------ manipulation/test_to_contig.rs ------------------------------

def test_already_c_contig(dev):  # test_to_contig.rs:12-32
    d = device(dev, "RowMajor")
    a = rt.arange(24, device=d).into_shape([2, 3, 4])
    assert a.c_contig()
    r = a.to_contig("RowMajor")
    assert not r.is_owned() and r.as_ptr() == a.as_ptr()
    r = a.to_contig("ColMajor")
    assert r.is_owned() and r.f_contig()


def test_already_f_contig(dev):  # test_to_contig.rs:34-54
    d = device(dev, "ColMajor")
    a = rt.arange(24, device=d).into_shape([2, 3, 4])
    assert a.f_contig()
    r = a.to_contig("ColMajor")
    assert not r.is_owned() and r.as_ptr() == a.as_ptr()
    r = a.to_contig("RowMajor")
    assert r.is_owned() and r.c_contig()
    # F-order reshape: element (i, j, k) is arange's i + 2j + 6k
    i, j, k = np.meshgrid(range(2), range(3), range(4), indexing="ij")
    assert np.array_equal(a.to_numpy(), i + 2 * j + 6 * k)


def test_transposed_tensor(dev):  # test_to_contig.rs:56-85
    d = device(dev, "RowMajor")
    a = rt.arange(12, device=d).into_shape([3, 4])
    t = a.T
    assert not t.c_contig() and t.f_contig()
    rc = t.to_contig("RowMajor")
    assert rc.is_owned() and rc.c_contig() and rc.shape == (4, 3)
    rf = t.to_contig("ColMajor")
    assert not rf.is_owned() and rf.f_contig() and rf.shape == (4, 3)
    exp = CASES["to_contig_values"][0]["expected_t_of_arange12_3x4"]
    assert_equal(rc, nested(exp, d, np.int64))


def test_sliced_tensor(dev):  # test_to_contig.rs:87-114
    d = device(dev, "RowMajor")
    c = CASES["to_contig_values"][1]
    a = rt.arange(24, device=d).into_shape([4, 6])
    s = a[::2, ::2]
    assert list(s.shape) == c["shape"] and list(s.stride) == c["stride"]
    assert not s.c_contig()
    r = s.to_contig("RowMajor")
    assert r.is_owned() and r.c_contig() and list(r.stride) == c["contig_stride"]
    assert_equal(r, nested(c["sliced_arange24_4x6_step2"], d, np.int64))


def test_f_order_sliced(dev):  # test_to_contig.rs:116-137
    d = device(dev, "ColMajor")
    a = rt.arange(24, device=d).into_shape([4, 6])
    s = a.i((slice(None, None, 2), slice(None, None, 2)))
    cc, cf = s.to_contig("RowMajor"), s.to_contig("ColMajor")
    assert cc.c_contig() and cf.f_contig()
    assert list(cc.iter()) == list(cf.iter())


def test_scalar_tensor(dev):  # test_to_contig.rs:139-159
    d = device(dev, "RowMajor")
    a = rt.asarray(np.array(42, dtype=np.int32), device=d)
    cc, cf = a.to_contig("RowMajor"), a.to_contig("ColMajor")
    assert cc.c_contig() and cf.f_contig()
    assert not cc.is_owned() and not cf.is_owned()


def test_1d_tensor(dev):  # test_to_contig.rs:161-184
    d = device(dev, "RowMajor")
    a = rt.arange(10, device=d)
    assert a.c_contig() and a.f_contig()
    assert not a.to_contig("RowMajor").is_owned()
    s = a[::2]
    assert not s.c_contig() and not s.f_contig()
    r = s.to_contig("RowMajor")
    assert r.is_owned() and r.stride == (1,)


def test_variants_equivalence(dev):  # test_to_contig.rs:186-219
    d = device(dev, "RowMajor")
    a = rt.arange(24, device=d).into_shape([4, 6])
    s = a[::2, ::2]
    r1 = rt.to_contig(s, "RowMajor")
    for r in (s.to_contig("RowMajor"), s.to_contig_f("RowMajor"), a.to_owned()[::2, ::2].change_contig("RowMajor"),
              a[::2, ::2].into_contig("RowMajor")):
        assert rt.allclose(r1, r)


def test_preserves_values(dev):  # test_to_contig.rs:221-247
    d = device(dev, "RowMajor")
    data = [1.5, 2.5, 3.5, 4.5, 5.5, 6.5, 7.5, 8.5, 9.5, 10.5, 11.5, 12.5]
    a = rt.asarray(data, shape=[3, 4], device=d)
    s = a[1:, ::2]
    orig = list(s.iter())
    assert orig == list(s.to_contig("RowMajor").iter()) == list(s.to_contig("ColMajor").iter())
    assert orig == [5.5, 7.5, 9.5, 11.5]


@pytest.mark.parametrize("order,prefer,owned", [("RowMajor", "RowMajor", False), ("ColMajor", "ColMajor", False),
                                                ("ColMajor", "RowMajor", True), ("RowMajor", "ColMajor", True)])
def test_to_prefer_contig(dev, order, prefer, owned):  # test_to_contig.rs:255-316
    d = device(dev, order)
    a = rt.arange(24, device=d).into_shape([2, 3, 4])
    r = a.to_prefer(prefer)
    assert r.is_owned() == owned
    if not owned:
        assert r.as_ptr() == a.as_ptr()
    else:
        assert r.c_contig() if prefer == "RowMajor" else r.f_contig()


def test_prefer_on_non_contig_and_vs_contig(dev):  # test_to_contig.rs:318-359
    d = device(dev, "RowMajor")
    a = rt.arange(24, device=d).into_shape([4, 6])
    s = a[::2, ::2]
    assert not s.c_contig() and not s.f_contig()
    assert s.to_prefer("RowMajor").is_owned() and s.to_prefer("ColMajor").is_owned()
    b = rt.arange(24, device=d).into_shape([2, 3, 4])
    p, c = b.to_prefer("RowMajor"), b.to_contig("RowMajor")
    assert not p.is_owned() and not c.is_owned()
    assert rt.allclose(p, c)


def test_edge_cases(dev):  # test_to_contig.rs:366-414
    d = device(dev, "RowMajor")
    z = rt.zeros([0, 5], device=d).to_contig("RowMajor")
    assert z.shape == (0, 5) and z.c_contig()
    one = rt.asarray([42.0], shape=[1, 1], device=d).to_contig("RowMajor")
    assert one.shape == (1, 1) and one.c_contig()
    a = rt.arange(120, device=d).into_shape([2, 3, 4, 5])
    s = a[:, ::2, :, ::2]
    cc, cf = s.to_contig("RowMajor"), s.to_contig("ColMajor")
    assert cc.c_contig() and cf.f_contig() and cc.shape == s.shape and cf.shape == s.shape
    assert np.array_equal(cc.to_numpy(), np.arange(120).reshape(2, 3, 4, 5)[:, ::2, :, ::2])
    assert np.array_equal(cf.to_numpy(), cc.to_numpy())


def test_reverse_stride(dev):  # test_to_contig.rs:416-436
    d = device(dev, "RowMajor")
    a = rt.arange(12, device=d).into_shape([3, 4])
    f = a.flip(0)
    r = f.to_contig("RowMajor")
    assert r.c_contig() and r.stride[0] > 0
    assert_equal(r, nested(CASES["to_contig_values"][2]["flip0_arange12_3x4"], d, np.int64))


def test_colmajor_reshape_copy_pairs_in_f_order(dev):
    """A reshape that must copy under ColMajor pairs elements in F order (reshape.rs:126-160 with
    assign_arbitary in the device order)."""
    d = device(dev, "ColMajor")
    a = rt.arange(24, device=d).into_shape([4, 6])
    s = a[::2, :]          # not f-contiguous: the reshape copies
    r = s.into_shape([3, 4])
    assert r.is_owned() and r.f_contig()
    assert np.array_equal(r.to_numpy(), np.reshape(np.reshape(s.to_numpy(), -1, order="F"), (3, 4), order="F"))


# ------------------------------ reduction/test_{all,any,count_nonzero}.rs ------------------------------

@pytest.mark.parametrize("case", CASES["bool_reduce"], ids=lambda c: c["ref"])
def test_all_any(dev, case):
    d = device(dev, "RowMajor")
    a = nested(case["a"], d, np.bool_)
    if case["op"] == "all_not":  # np.all(~y2): the unary not of a bool tensor, then all
        assert rt.not_(a).all_all() == case["expected"]
        return
    assert getattr(a, case["op"] + "_all")() == case["expected"]
    for ax, exp in case.get("axes", {}).items():
        r = getattr(a, case["op"] + "_axes")(int(ax))
        assert np.dtype(r.dtype) == np.bool_ and r.to_vec() == exp


@pytest.mark.parametrize("case", CASES["count_nonzero"], ids=lambda c: str(c["axes"]))
def test_count_nonzero(dev, case):
    d = device(dev, "RowMajor")
    a = nested(case["a"], d, np.int32)
    r = a.count_nonzero_axes(case["axes"])
    assert np.dtype(r.dtype) == np.uint64  # usize
    if case["axes"] is None:
        assert r.to_scalar() == case["expected"]
    else:
        assert r.to_vec() == case["expected"]
