"""all / any / count_nonzero and the sum of a bool array on the device against the oracle, bit for
bit, over the row, column and split kernel paths (auto_impl/reduction.rs:466-551, 678-715)."""
import zlib

import numpy as np
import pytest

import rstsr_amd as rt
from gpu_util import from_np, oracle_reduce

pytestmark = pytest.mark.gpu

ALL, ANY, COUNT, SUM = 10, 11, 12, 0
LAYOUTS = [((64, 48, 40), (2, 0, 1), 2, [0]), ((64, 48, 40), (2, 0, 1), 2, [1, 2]),
           ((8, 3000), (0, 1), 1, [1]), ((3000, 8), (0, 1), 1, [0]), ((6, 7, 9), (1, 2, 0), 1, [2, 0]),
           ((5, 4), (1, 0), 1, [0, 1])]


def draw(rng, n, dt, dense):
    if np.dtype(dt) == np.bool_:
        return rng.random(n) < dense
    v = rng.normal(size=n) if np.dtype(dt).kind in "fc" else rng.integers(-5, 5, n)
    v = np.where(rng.random(n) < dense, v, 0)
    if np.dtype(dt).kind == "c":
        v = v + 1j * np.where(rng.random(n) < 0.1, rng.normal(size=n), 0)
    return v.astype(dt)


@pytest.mark.parametrize("op,dt", [(ALL, np.bool_), (ANY, np.bool_), (SUM, np.bool_), (COUNT, np.bool_),
                                   (COUNT, np.float64), (COUNT, np.float32), (COUNT, np.int32), (COUNT, np.int64),
                                   (COUNT, np.complex128), (COUNT, np.complex64)])
@pytest.mark.parametrize("X,perm,step,axes", LAYOUTS)
@pytest.mark.parametrize("dense", [0.002, 0.5, 0.998])
def test_bool_count_axes(dev, oracle, op, dt, X, perm, step, axes, dense):
    rng = np.random.default_rng(zlib.crc32(repr((op, str(dt), X, axes, dense)).encode()))
    x = draw(rng, int(np.prod(X)), dt, dense).reshape(X).transpose(perm)[..., ::step]
    t, v = from_np(dev, x)
    got = t._red_axes(op, axes).to_numpy()
    ref = oracle_reduce(oracle, op, v, axes, dt)
    assert got.dtype == ref.dtype and np.array_equal(got, ref)


@pytest.mark.parametrize("op,dt", [(ALL, np.bool_), (ANY, np.bool_), (SUM, np.bool_), (COUNT, np.float64),
                                   (COUNT, np.complex64)])
@pytest.mark.parametrize("n", [1, 1000, 1 << 20])
def test_bool_count_all(dev, oracle, op, dt, n):
    rng = np.random.default_rng(zlib.crc32(repr((op, str(dt), n)).encode()))
    for dense in (0.0, 1e-6, 0.5, 1.0):
        x = draw(rng, n, dt, dense)
        t, v = from_np(dev, x)
        assert t._red_all(op) == oracle.reduce_all(op, v).item()


def test_empty_and_special(dev):
    e = rt.asarray(np.zeros((0, 3), dtype=bool), device=dev)
    assert e.all_all() is True and e.any_all() is False and e.sum_all() == 0 and e.count_nonzero_all() == 0
    assert e.all_axes(0).to_vec() == [True, True, True] and e.any_axes(0).to_vec() == [False, False, False]
    f = rt.asarray(np.array([0.0, -0.0, np.nan, np.inf, 1e-300]), device=dev)
    assert f.count_nonzero_all() == 3  # -0.0 is zero, NaN is nonzero
    c = rt.asarray(np.array([0j, 1j, -0.0 + 0j, 2 + 0j]), device=dev)
    assert c.count_nonzero_all() == 2
    with pytest.raises(rt.RSTSRError):
        rt.asarray(np.ones(4), device=dev).all_all()  # all/any take bool arrays
