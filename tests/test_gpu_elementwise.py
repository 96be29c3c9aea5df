"""HIP elementwise / copy / fill kernels vs the oracle: bit-exact (rstsr-native-impl cpu_serial semantics)."""
import zlib

import numpy as np
import pytest

import rstsr_amd as rt
from gpu_util import bits_equal, from_np, oracle_binary, pair, rand
from oracle_ffi import View, c_view

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32, np.int32, np.int64, np.complex128, np.complex64]
OPS = {"add": (0, lambda a, b: a + b), "sub": (1, lambda a, b: a - b), "mul": (2, lambda a, b: a * b),
       "div": (3, lambda a, b: a / b)}


def _strided_case(rng, dtype, X, perm, step):
    """X-shaped base, view = X.transpose(perm)[..., ::step]; b = [1, X1, 1] broadcast the same way."""
    n = int(np.prod(X))
    base = rand(rng, n, dtype)
    x = base.reshape(X).transpose(perm)[..., ::step]
    bshape = [1] * len(X)
    bshape[1] = X[1]
    bb = rand(rng, X[1], dtype)
    if np.dtype(dtype).kind == "f":
        bb = rng.uniform(0.5, 2.0, size=X[1]).astype(dtype)
    y = bb.reshape(bshape).transpose(perm)[..., ::step]
    return x, y


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("op", ["add", "mul", "sub", "div"])
@pytest.mark.parametrize("X", [(16, 32, 64), (20, 37, 150), (4, 8, 2048)])
def test_strided_broadcast_binary(dev, oracle, dtype, op, X):
    """permuted (2,0,1) + step-2 view times a broadcast [1,X1,1] view (the strided config, scaled down)."""
    rng = np.random.default_rng(zlib.crc32(repr((str(dtype), op, X)).encode()))
    x, y = _strided_case(rng, dtype, X, (2, 0, 1), 2)
    ta, va = from_np(dev, x)
    tb, vb = from_np(dev, y)
    code, fn = OPS[op]
    tc = ta._bin(code, tb)
    shape = np.broadcast_shapes(x.shape, y.shape)
    # oracle on the broadcast views
    xb, yb = np.broadcast_arrays(x, y)
    ref = oracle_binary(oracle, code, View.of(xb), View.of(yb), shape, dtype)
    got = tc.to_numpy()
    assert bits_equal(got, ref)
    # the output layout is rstsr's (get_layout_for_binary_op): c-contiguous here
    assert tc.layout == rt._rstsr.c_layout(list(shape))


@pytest.mark.parametrize("dtype", [np.float64, np.int32, np.complex128])
def test_contiguous_broadcast_rows(dev, oracle, dtype):
    """[64,128,256] + [256] (the small config's add), rows kernel."""
    rng = np.random.default_rng(1)
    a = rand(rng, 64 * 128 * 256, dtype).reshape(64, 128, 256)
    b = rand(rng, 256, dtype)
    ta, va = from_np(dev, a)
    tb, vb = from_np(dev, b)
    got = (ta + tb).to_numpy()
    xb, yb = np.broadcast_arrays(a, b)
    ref = oracle_binary(oracle, 0, View.of(xb), View.of(yb), a.shape, dtype)
    assert bits_equal(got, ref)


@pytest.mark.parametrize("case", ["flip", "transposed", "scalar_right", "scalar_left", "inplace", "inplace_scalar",
                                  "size1", "zero_size", "3d_mixed", "long_inner"])
def test_layout_cases(dev, oracle, case):
    rng = np.random.default_rng(5)
    if case == "flip":
        a = rng.normal(size=(33, 70))[::-1, ::-3]
        b = rng.normal(size=(33, 70))[:, ::3][::-1]
        ta, _ = from_np(dev, a)
        tb, _ = from_np(dev, b)
        assert bits_equal((ta * tb).to_numpy(), a * b)
    elif case == "transposed":
        a = rng.normal(size=(130, 90))
        b = rng.normal(size=(90, 130)).T
        ta, _ = from_np(dev, a)
        tb, _ = from_np(dev, b)
        assert bits_equal((ta - tb).to_numpy(), a - b)
    elif case == "scalar_right":
        a = rng.normal(size=(40, 50)).T[::2]
        ta, _ = from_np(dev, a)
        assert bits_equal((ta * 1.5).to_numpy(), a * 1.5)
        assert bits_equal((ta / 3.0).to_numpy(), a / 3.0)
    elif case == "scalar_left":
        a = rng.normal(size=(40, 50)).T
        ta, _ = from_np(dev, a)
        assert bits_equal((2.5 - ta).to_numpy(), 2.5 - a)
    elif case == "inplace":
        a = rng.normal(size=(64, 300))
        b = rng.normal(size=(300,))
        ta = rt.asarray(a, device=dev)
        tb = rt.asarray(b, device=dev)
        ta += tb
        assert bits_equal(ta.to_numpy(), a + b)
    elif case == "inplace_scalar":
        a = rng.integers(-100, 100, size=(7, 9)).astype(np.int32)
        ta = rt.asarray(a, device=dev)
        ta *= 3
        assert bits_equal(ta.to_numpy(), a * np.int32(3))
    elif case == "size1":
        a = rng.normal(size=(1, 5, 1))
        b = rng.normal(size=(4, 1, 1))
        ta, _ = from_np(dev, a)
        tb, _ = from_np(dev, b)
        assert bits_equal((ta + tb).to_numpy(), a + b)
    elif case == "zero_size":
        ta = rt.zeros((0, 5), device=dev)
        tb = rt.zeros((5,), device=dev)
        assert (ta + tb).shape == (0, 5)
    elif case == "3d_mixed":
        a = rng.normal(size=(30, 40, 50)).transpose(1, 2, 0)[:, ::2, 1:]
        b = rng.normal(size=(40, 1, 29))
        ta, _ = from_np(dev, a)
        tb, _ = from_np(dev, b)
        assert bits_equal((ta * tb).to_numpy(), a * b)
    elif case == "long_inner":
        a = rng.normal(size=(3, 100003))
        b = rng.normal(size=(3, 100003))
        ta, _ = from_np(dev, a)
        tb, _ = from_np(dev, b)
        assert bits_equal((ta + tb).to_numpy(), a + b)


def test_complex_mul_is_num_complex(dev):
    rng = np.random.default_rng(3)
    a = (rng.normal(size=4096) + 1j * rng.normal(size=4096))
    b = (rng.normal(size=4096) + 1j * rng.normal(size=4096))
    got = (rt.asarray(a, device=dev) * rt.asarray(b, device=dev)).to_numpy()
    re = a.real * b.real - a.imag * b.imag
    im = a.real * b.imag + a.imag * b.real
    assert np.array_equal(got.real, re) and np.array_equal(got.imag, im)


@pytest.mark.parametrize("X,perm", [((64, 48, 80), (2, 0, 1)), ((33, 65, 17), (1, 2, 0)), ((256, 8, 300), (2, 0, 1)),
                                    ((5, 6, 7, 8), (3, 1, 0, 2))])
@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.int32])
def test_to_contig_permuted(dev, oracle, X, perm, dtype):
    rng = np.random.default_rng(sum(X))
    base = rand(rng, int(np.prod(X)), dtype)
    x = base.reshape(X).transpose(perm)
    tx, vx = from_np(dev, x)
    got = tx.to_contig()
    assert got.layout == rt._rstsr.c_layout(list(x.shape))
    vc = c_view(x.shape, dtype)
    oracle.assign(vc, vx)
    assert bits_equal(got.to_numpy(), vc.array())


def test_to_contig_same_layout_no_copy(dev):
    a = rt.asarray(np.arange(12.0).reshape(3, 4), device=dev)
    b = a.to_contig()
    assert b._t._data_ptr() == a._t._data_ptr()


def test_assign_arbitary_reshape(dev, oracle):
    rng = np.random.default_rng(4)
    x = rng.normal(size=(6, 10, 4)).transpose(2, 0, 1)  # [4,6,10]
    tx, vx = from_np(dev, x)
    got = tx.to_layout((([12, 20]), rt._rstsr.c_layout([12, 20])[1], 0))
    vc = c_view((12, 20), np.float64)
    oracle.assign(vc, vx)
    assert bits_equal(got.to_numpy(), vc.array())


def test_astype_and_fill(dev):
    a = np.arange(-5, 7, dtype=np.int32).reshape(3, 4)
    ta = rt.asarray(a, device=dev)
    assert bits_equal(ta.astype(np.float64).to_numpy(), a.astype(np.float64))
    assert bits_equal(ta.T.astype(np.int64).to_numpy(), a.T.astype(np.int64))
    f = rt.full((5, 7), 2.25, device=dev)
    assert np.all(f.to_numpy() == 2.25)
    z = rt.zeros((3, 3), dtype=np.complex128, device=dev)
    assert np.all(z.to_numpy() == 0)


def test_allclose(dev):
    a = np.linspace(0, 1, 1000)
    ta = rt.asarray(a, device=dev)
    tb = rt.asarray(a + 1e-9, device=dev)
    assert rt.allclose(ta, tb)
    tc = rt.asarray(np.where(np.arange(1000) == 777, a + 1.0, a), device=dev)
    assert not rt.allclose(ta, tc)
    tn = rt.asarray(np.full(4, np.nan), device=dev)
    assert not rt.allclose(tn, tn)
    assert rt.allclose(tn, tn, equal_nan=True)


def test_errors_are_raised_on_host(dev):
    a = rt.asarray(np.zeros((3, 4)), device=dev)
    b = rt.asarray(np.zeros((5,)), device=dev)
    with pytest.raises(rt.RSTSRError) as e:
        a + b
    assert e.value.args[0] == "InvalidLayout"
    # a layout that leaves its buffer is refused before any launch
    with pytest.raises(rt.RSTSRError) as e:
        a._view(([3, 5], [4, 1], 0)).to_contig()
    assert e.value.args[0] == "ValueOutOfRange"
    c = rt.asarray(np.zeros((3, 4), dtype=np.float32), device=dev)
    with pytest.raises(rt.RSTSRError):
        a + c


_ORDER_CHILD = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
import rstsr_amd as rt
dev = rt.DeviceHIP(0)
rng = np.random.default_rng(int(sys.argv[2]))
shape, perm = (4096, 2304), (1, 0)
X = rng.standard_normal(shape)
Xb = rng.standard_normal((1, shape[1]))
t = rt.asarray(X, device=dev).transpose(list(perm))
tb = rt.asarray(Xb, device=dev).transpose(list(perm))
ok = True
for _ in range(2):  # first call tunes (or reads the cache), second uses the choice
    ok &= np.array_equal(t.to_contig().to_numpy(), np.ascontiguousarray(X.T))
    ok &= np.array_equal((t * tb).to_numpy(), X.T * Xb.T)
dev.synchronize()
print("ORDER_OK" if ok else "ORDER_BAD", flush=True)
"""


def test_tiled_tile_orders_all_exact(tmp_path):
    """Every tile order the tuner can pick gives the exact transposed copy and broadcast product:
    a first process tunes and records its choices in a cache file; fresh processes then run with
    each recorded choice forced to order 0, 1 and 2 (the cache file is the tuner's own format)."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cache = tmp_path / "tune.txt"
    env = dict(os.environ, RSTSR_HIP_TUNE_CACHE=str(cache))

    def child(seed):
        r = subprocess.run([sys.executable, "-c", _ORDER_CHILD, root, str(seed)], env=env, capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        return r.stdout

    assert "ORDER_OK" in child(1)
    lines = [ln.split() for ln in cache.read_text().splitlines() if ln.startswith("ew|")]
    assert len(lines) >= 1, "the tiled launch was not tuned"
    for order in (0, 1, 2):
        cache.write_text("".join(f"{k} {order}\n" for k, _ in lines))
        assert "ORDER_OK" in child(10 + order), order
