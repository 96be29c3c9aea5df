"""Tensors past 2^31 elements (a 16 GiB f64 array): a contiguous elementwise op whose merged axis
holds 2^31 + 2^21 elements, reductions over a reduced extent of that size (factored into two
reduced dims) and a reduction producing that many outputs (processed in pieces of the kept axis).
DeviceFaer has no such limit, so neither may the device."""
import math

import numpy as np
import pytest

import rstsr_amd as rt

pytestmark = pytest.mark.gpu

ROWS, COLS = 1 << 21, 1025  # 2,149,580,800 elements > 2^31


@pytest.fixture(scope="module")
def big(dev):
    rng = np.random.default_rng(20261260)
    v = rng.normal(size=COLS)
    tv = rt.asarray(v, device=dev)
    a = tv.broadcast_to([ROWS, COLS]).to_contig()  # c-contiguous 16 GiB
    return v, tv, a


def test_elementwise_past_2_31(dev, big):
    v, tv, a = big
    flat = a.reshape([ROWS * COLS])
    assert flat.shape == (ROWS * COLS,) and ROWS * COLS > 2 ** 31
    c = flat + flat  # one merged axis of 2^31 + 2^21 elements
    c2 = c.reshape([ROWS, COLS])
    for r in (0, ROWS // 2, ROWS - 1):
        assert np.array_equal(c2[r].to_numpy(), 2 * v)
    e = c2 - (tv + tv)  # broadcast row: every element must be exactly zero
    assert np.all(e.max_axes([1]).to_numpy() == 0) and np.all(e.min_axes([1]).to_numpy() == 0)


def test_reduction_over_2_31_elements(dev, big):
    v, tv, a = big
    flat = a.reshape([ROWS * COLS])
    s = flat.sum_axes([0]).to_scalar()
    exact = ROWS * math.fsum(v.tolist())  # a power-of-two multiple: exact
    assert abs(s - exact) <= 4 * 2.0 ** -53 * abs(exact) + 1e-300
    assert flat.argmax_all() == int(np.argmax(v))
    assert flat.max_all() == float(np.max(v))
    # a reversed view starts with v reversed: its first minimum is argmin(v[::-1])
    assert flat[::-1].argmin_all() == int(np.argmin(v[::-1]))


def test_reduction_with_2_31_outputs(dev, big):
    v, tv, a = big
    col = a.reshape([ROWS * COLS, 1])
    out = col.sum_axes([1])  # 2^31 + 2^21 outputs, one element each
    assert out.shape == (ROWS * COLS,)
    e = out.reshape([ROWS, COLS]) - a
    assert np.all(e.max_axes([1]).to_numpy() == 0) and np.all(e.min_axes([1]).to_numpy() == 0)
