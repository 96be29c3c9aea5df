"""The extended binary operators on the device against the oracle, bit for bit: integer and complex
division, remainder (truncated for integers, fmod for floats, num-complex's formula for complex),
bit operators (integers, bool) and shifts (integers), as arrays, with host scalars on either side
and in place; strided, permuted and broadcast layouts (generic and row kernels). Pairs the
reference does not implement raise UnImplemented."""
import zlib

import numpy as np
import pytest

import rstsr_amd as rt
from gpu_util import bits_equal, from_np, oracle_binary
from oracle_ffi import View

pytestmark = pytest.mark.gpu

DIV, REM, AND, OR, XOR, SHL, SHR = 3, 4, 5, 6, 7, 8, 9
NUM = [np.float64, np.float32, np.int32, np.int64, np.complex128, np.complex64]
CASES = [(op, dt) for op in (DIV, REM) for dt in NUM] + \
    [(op, dt) for op in (AND, OR, XOR) for dt in (np.int32, np.int64, np.bool_)] + \
    [(op, dt) for op in (SHL, SHR) for dt in (np.int32, np.int64)]


def draw(rng, n, dt, op, divisor=False):
    dt = np.dtype(dt)
    if dt == np.bool_:
        return rng.integers(0, 2, n).astype(bool)
    if dt.kind == "i":
        info = np.iinfo(dt)
        if op in (SHL, SHR) and divisor:
            return rng.integers(-3 * info.bits, 3 * info.bits, n).astype(dt)
        if op in (DIV, REM) and divisor:
            v = rng.integers(-40, 40, n).astype(dt)  # zero and -1 included: the documented substitutes
            v[:3] = [0, -1, 1]
            return v
        v = rng.integers(info.min, info.max, n, endpoint=True).astype(dt)
        v[:3] = [info.min, 7, -7]
        return v
    if dt.kind == "f":
        return (rng.normal(size=n) * (3 if divisor else 100)).astype(dt)
    if op == REM:  # Gaussian-integer-valued complex
        lim = 30 if divisor else 300
        v = (rng.integers(-lim, lim + 1, n) + 1j * rng.integers(-lim, lim + 1, n)).astype(dt)
        if divisor:
            v[v == 0] = 2 - 1j
        return v
    return (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(dt)


@pytest.mark.parametrize("op,dt", CASES, ids=lambda x: str(x))
@pytest.mark.parametrize("X", [(12, 20, 40), (3, 5, 2048)])
def test_binop_strided_broadcast(dev, oracle, op, dt, X):
    rng = np.random.default_rng(zlib.crc32(repr((op, str(dt), X)).encode()))
    base = draw(rng, int(np.prod(X)), dt, op)
    x = base.reshape(X).transpose(2, 0, 1)[..., ::2]
    yb = draw(rng, X[1], dt, op, divisor=True)
    y = yb.reshape(1, X[1], 1).transpose(2, 0, 1)[..., ::2]
    ta, _ = from_np(dev, x)
    tb, _ = from_np(dev, y)
    got = ta._bin(op, tb).to_numpy()
    xb, ybb = np.broadcast_arrays(x, y)
    ref = oracle_binary(oracle, op, View.of(xb), View.of(ybb), xb.shape, dt)
    assert bits_equal(got, ref)


@pytest.mark.parametrize("op,dt", CASES, ids=lambda x: str(x))
def test_binop_scalars_and_inplace(dev, oracle, op, dt):
    rng = np.random.default_rng(zlib.crc32(repr(("s", op, str(dt))).encode()))
    x = draw(rng, 40 * 30, dt, op).reshape(40, 30)[:, ::-2]
    y = draw(rng, 40 * 15, dt, op, divisor=True).reshape(40, 15)
    s = draw(rng, 4, dt, op, divisor=True)[3]
    ta, _ = from_np(dev, x)
    # array op scalar / scalar op array
    for left in (False, True):
        got = ta._bin(op, s, left=left).to_numpy()
        sa = np.full(x.shape, s, dtype=dt)
        xa, xb_ = (sa, x) if left else (x, sa)
        ref = oracle_binary(oracle, op, View.of(np.ascontiguousarray(xa)), View.of(np.ascontiguousarray(xb_)),
                            x.shape, dt)
        assert bits_equal(got, ref), left
    # in place on a strided view: other elements of the storage untouched
    full = draw(rng, 40 * 30, dt, op).reshape(40, 30)
    tf = rt.asarray(full, device=dev)
    view = tf[:, 1::2]
    tb = rt.asarray(y, device=dev)
    view._ibin(op, tb)
    ref = full.copy()
    ref[:, 1::2] = oracle_binary(oracle, op, View.of(np.ascontiguousarray(full[:, 1::2])), View.of(y), y.shape, dt)
    assert bits_equal(tf.to_numpy(), ref)


@pytest.mark.parametrize("op,dt", [(AND, np.float64), (SHL, np.float32), (SHR, np.complex128), (XOR, np.complex64),
                                   (REM, np.bool_), (DIV, np.bool_), (SHL, np.bool_)])
def test_undefined_pairs_raise(dev, op, dt):
    t = rt.asarray(np.ones(8, dtype=dt), device=dev)
    with pytest.raises(rt.RSTSRError):
        t._bin(op, t)


def test_python_operators(dev):
    a_h = np.array([[10, 21], [33, 44]], dtype=np.int32)
    b_h = np.array([[3, 4], [5, 7]], dtype=np.int32)
    a, b = rt.asarray(a_h, device=dev), rt.asarray(b_h, device=dev)
    assert np.array_equal(rt.rem(a, b).to_numpy(), a_h % b_h)
    assert np.array_equal((a & b).to_numpy(), a_h & b_h)
    assert np.array_equal((a | b).to_numpy(), a_h | b_h)
    assert np.array_equal((a ^ b).to_numpy(), a_h ^ b_h)
    assert np.array_equal((a << b).to_numpy(), a_h << b_h)
    assert np.array_equal((a >> 1).to_numpy(), a_h >> 1)
    assert np.array_equal((a / b).to_numpy(), np.array([[3, 5], [6, 6]], dtype=np.int32))
    assert np.array_equal((a % b).to_numpy(), a_h @ b_h)  # % is matmul, as in rstsr
