"""Mixed complex / real operands (num-complex Complex<T> op T and T op Complex<T>) on the device
against the oracle, bit for bit: arrays on strided and broadcast layouts in both operand orders,
and host scalars on either side (a real scalar with a complex tensor, a complex scalar with a real
tensor)."""
import zlib

import numpy as np
import pytest

import rstsr_amd as rt
from gpu_util import bits_equal, from_np
from oracle_ffi import View, c_view

pytestmark = pytest.mark.gpu

PAIRS = [(np.complex128, np.float64), (np.complex64, np.float32)]
OPS = [0, 1, 2, 3]


def oracle_mixed(oracle, op, a, b, tc):
    if isinstance(b, np.ndarray):
        a, b = np.broadcast_arrays(a, b)
        vb = View.of(b)
    else:
        vb = b
    vc = c_view(a.shape, tc)
    oracle.binary_mixed(op, vc, View.of(a), vb)
    return vc.array()


@pytest.mark.parametrize("tc,tr", PAIRS)
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("complex_left", [True, False])
def test_mixed_arrays(dev, oracle, tc, tr, op, complex_left):
    rng = np.random.default_rng(zlib.crc32(repr((str(tc), op, complex_left)).encode()))
    z = (rng.normal(size=(12, 20, 40)) + 1j * rng.normal(size=(12, 20, 40))).astype(tc).transpose(2, 0, 1)[:, :, ::2]
    r = rng.uniform(0.5, 2.0, size=(1, 20, 1)).astype(tr).transpose(2, 0, 1)[:, :, ::2]
    a, b = (z, r) if complex_left else (r, z)
    ta, _ = from_np(dev, a)
    tb, _ = from_np(dev, b)
    got = ta._bin(op, tb)
    assert np.dtype(got.dtype) == np.dtype(tc)
    assert bits_equal(got.to_numpy(), oracle_mixed(oracle, op, a, b, tc))


@pytest.mark.parametrize("tc,tr", PAIRS)
@pytest.mark.parametrize("op", OPS)
def test_mixed_scalars(dev, oracle, tc, tr, op):
    rng = np.random.default_rng(zlib.crc32(repr(("s", str(tc), op)).encode()))
    z = (rng.normal(size=(30, 17)) + 1j * rng.normal(size=(30, 17))).astype(tc)[:, ::-2]
    x = rng.normal(size=(30, 17)).astype(tr).T
    tz, _ = from_np(dev, z)
    tx, _ = from_np(dev, x)
    t = 1.75
    # complex tensor op real scalar, real scalar op complex tensor
    assert bits_equal(tz._bin(op, t).to_numpy(), oracle_mixed(oracle, op, np.ascontiguousarray(z), tr(t), tc))
    zs = np.ascontiguousarray(z)
    ref = oracle_mixed(oracle, op, np.full(zs.shape, t, dtype=tr), zs, tc)
    assert bits_equal(tz._bin(op, t, left=True).to_numpy(), ref)
    # real tensor op complex scalar
    w = complex(0.5, -1.25)
    xs = np.ascontiguousarray(x)
    ref2 = oracle_mixed(oracle, op, np.full(xs.shape, w, dtype=tc), xs, tc)  # w op x  (complex left)
    assert bits_equal(tx._bin(op, w, left=True).to_numpy(), ref2)


def test_mixed_unsupported_pairs(dev):
    a = rt.asarray(np.ones(4, dtype=np.complex128), device=dev)
    b = rt.asarray(np.ones(4, dtype=np.float32), device=dev)
    with pytest.raises(rt.RSTSRError):
        a + b  # complex128 with float32: not a Complex<T> / T pair
