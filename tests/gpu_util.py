"""Helpers for GPU parity tests: identical strided views on the device and for the oracle."""
import numpy as np

import rstsr_amd as rt
from oracle_ffi import View, c_view, red_out_dtype


def pair(dev, base, shape, stride, offset=0):
    """(device tensor view, oracle View) of one flat host array under one layout."""
    base = np.ascontiguousarray(base).reshape(-1)
    t = rt.asarray(base, device=dev)
    t = t._view((list(shape), list(stride), int(offset)))
    return t, View(base, shape, stride, offset)


def from_np(dev, arr):
    """Device tensor + oracle View reproducing a numpy view's layout over its base buffer."""
    v = View.of(arr)
    t = rt.asarray(v.base, device=dev)._view((list(v.shape), list(v.stride), v.offset))
    return t, v


def bits_equal(x, y):
    x = np.ascontiguousarray(x)
    y = np.ascontiguousarray(y)
    return x.shape == y.shape and x.dtype == y.dtype and np.array_equal(x.view(np.uint8), y.view(np.uint8))


def oracle_binary(oracle, op, va, vb, shape, dtype, scalar_b=None):
    vc = c_view(shape, dtype)
    oracle.binary(op, vc, va, vb, scalar_b=scalar_b)
    return vc.array()


def oracle_reduce(oracle, op, va, axes, dtype):
    nd = len(va.shape)
    ax = [a % nd for a in axes]
    shape = [va.shape[i] for i in range(nd) if i not in ax]
    vo = c_view(shape, red_out_dtype(op, dtype))
    r = oracle.reduce_axes(op, va, list(axes), vo)
    assert r == 0
    return vo.array()


def rand(rng, n, dtype, lo=-1.0, hi=1.0):
    dt = np.dtype(dtype)
    if dt.kind == "f":
        return rng.normal(size=n).astype(dt)
    if dt.kind == "c":
        return (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(dt)
    return rng.integers(-(2 ** 20), 2 ** 20, size=n).astype(dt)
