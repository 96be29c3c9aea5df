"""rstsr_hip_reduce_axes_noalloc (partials in a fixed per-device array, no hipMalloc) against
rstsr_hip_reduce_axes, the entry point it varies, through the C ABI.

Both run the same plan and kernels, so every result is compared bit for bit. Each case makes
two calls in a row on the shared array, the second on a smaller input, so that leftovers of
the first call cannot pass for results of the second. One case is captured into a hipGraph
and replayed with eager calls of other sizes between the replays.
"""
import ctypes as C

import numpy as np
import pytest

import rstsr_amd as rt
from rstsr_amd import abi

pytestmark = pytest.mark.gpu

SUM, MAX, MIN, PROD, MEAN, VAR, STD, L2 = 0, 1, 2, 5, 6, 7, 8, 9
ES = {np.float64: 8, np.float32: 4, np.int32: 4, np.int64: 8, np.complex128: 16, np.complex64: 8}


def out_esize(op, dt):
    if op in (VAR, STD, L2) and np.dtype(dt).kind == "c":
        return ES[dt] // 2
    return ES[dt]


def layout_of(t):
    return abi.make_layout(t.shape, t.stride, t.offset)


def out_layout(la, axes):
    nd = C.c_size_t()
    sh = (C.c_size_t * max(la.ndim, 1))()
    st = (C.c_ssize_t * max(la.ndim, 1))()
    ax = (C.c_ssize_t * max(len(axes), 1))(*axes)
    abi.check(abi.lib().rstsr_hip_reduce_out_layout(C.byref(la), len(axes), ax, C.byref(nd), sh, st))
    n = nd.value
    return abi.make_layout(list(sh[:n]), list(st[:n]), 0), [int(x) for x in sh[:n]], [int(x) for x in st[:n]]


def alloc(ctx, nbytes):
    p = C.c_void_p()
    abi.check(abi.lib().rstsr_hip_malloc(ctx, max(nbytes, 1), C.byref(p)))
    return p


def span_bytes(shape, stride, es):
    if any(s == 0 for s in shape):
        return 0
    return (1 + sum((n - 1) * abs(s) for n, s in zip(shape, stride))) * es


def fetch(ctx, p, nbytes):
    h = np.empty(max(nbytes, 1), dtype=np.uint8)
    abi.check(abi.lib().rstsr_hip_memcpy_d2h(ctx, h.ctypes.data_as(C.c_void_p), p, nbytes))
    return h[:nbytes]


def both(dev, x, op, axes):
    """(bytes from rstsr_hip_reduce_axes, bytes from the no-alloc variant) for one call each."""
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    la = layout_of(x)
    lo, sh, st = out_layout(la, axes)
    es = out_esize(op, {np.dtype(k): k for k in ES}[np.dtype(x.dtype)])
    nb = span_bytes(sh, st, es)
    o1, o2 = alloc(ctx, nb), alloc(ctx, nb)
    ax = (C.c_ssize_t * max(len(axes), 1))(*axes)
    code = rt._code(x.dtype)
    a = C.c_void_p(x._t._data_ptr())
    abi.check(L.rstsr_hip_reduce_axes(ctx, op, code, a, C.byref(la), len(axes), ax, o1, C.byref(lo)))
    abi.check(L.rstsr_hip_reduce_axes_noalloc(ctx, op, code, a, C.byref(la), len(axes), ax, o2, C.byref(lo)))
    r1, r2 = fetch(ctx, o1, nb), fetch(ctx, o2, nb)
    abi.check(L.rstsr_hip_free(ctx, o1))
    abi.check(L.rstsr_hip_free(ctx, o2))
    return r1, r2


def strided(dev, rng, dt, n0, n1, n2):
    """X[n0, n1, n2] viewed permute(2,0,1)[:, :, ::2] (unit-stride axis 0)."""
    if np.dtype(dt).kind == "c":
        h = (rng.normal(size=(n0, n1, n2)) + 1j * rng.normal(size=(n0, n1, n2))).astype(dt)
    elif np.dtype(dt).kind == "f":
        h = rng.normal(size=(n0, n1, n2)).astype(dt)
    else:
        h = rng.integers(-(2 ** 20), 2 ** 20, size=(n0, n1, n2)).astype(dt)
    return rt.asarray(h, device=dev).transpose([2, 0, 1])[:, :, ::2]


CASES = [
    (SUM, np.float64, [0]), (SUM, np.float64, [1, 2]), (SUM, np.float32, [1, 2]),
    (MAX, np.float64, [1, 2]), (MIN, np.int32, [1, 2]), (SUM, np.int64, [2, 1]),
    (PROD, np.float64, [1]), (MEAN, np.complex128, [1, 2]), (VAR, np.float64, [1, 2]),
    (STD, np.complex64, [0, 2]), (L2, np.float32, [1, 2]), (SUM, np.complex128, [0]),
]


@pytest.mark.parametrize("op,dt,axes", CASES)
def test_noalloc_matches_reduce_axes(dev, op, dt, axes):
    rng = np.random.default_rng(20261210 + 7 * op + len(axes))
    for shape in [(48, 96, 512), (16, 40, 256)]:  # second call on a smaller input
        x = strided(dev, rng, dt, *shape)
        r1, r2 = both(dev, x, op, axes)
        assert r1.size > 0
        assert np.array_equal(r1, r2), (op, dt, axes, shape)


def test_noalloc_ragged_and_row_paths(dev):
    rng = np.random.default_rng(20261211)
    # long contiguous rows (row kernel with splits), then a short ragged one
    for n in [(8, 3 * 65536 + 17), (5, 1000)]:
        h = rng.normal(size=n)
        x = rt.asarray(h, device=dev)
        r1, r2 = both(dev, x, SUM, [1])
        assert np.array_equal(r1, r2)
    # a broadcast (stride-0) reduced axis, then a smaller one
    for n in [(4096, 64), (512, 8)]:
        b = rt.asarray(rng.normal(size=(1, n[1])), device=dev).broadcast_to(list(n))
        r1, r2 = both(dev, b, SUM, [0])
        assert np.array_equal(r1, r2)


def test_noalloc_captured_replays(dev):
    """A no-alloc column reduction with split partials recorded into a graph; eager no-alloc
    calls of other sizes run between the replays and overwrite the shared array."""
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    rng = np.random.default_rng(20261212)
    x = strided(dev, rng, np.float64, 48, 96, 512)
    r_ref, _ = both(dev, x, SUM, [1, 2])
    la = layout_of(x)
    lo, sh, st = out_layout(la, [1, 2])
    nb = span_bytes(sh, st, 8)
    out = alloc(ctx, nb)
    ax = (C.c_ssize_t * 2)(1, 2)
    a = C.c_void_p(x._t._data_ptr())
    with rt.capture(dev) as g:
        abi.check(L.rstsr_hip_reduce_axes_noalloc(ctx, SUM, rt._code(x.dtype), a, C.byref(la), 2, ax, out,
                                                  C.byref(lo)))
    for i in range(3):
        y = strided(dev, rng, np.float64, 16 + 8 * i, 40, 256)
        both(dev, y, SUM, [1, 2])
        g.launch()
        assert np.array_equal(fetch(ctx, out, nb), r_ref), i
    abi.check(L.rstsr_hip_free(ctx, out))


def test_noalloc_errors(dev):
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    rng = np.random.default_rng(20261213)
    x = strided(dev, rng, np.float64, 8, 16, 64)
    la = layout_of(x)
    lo, sh, st = out_layout(la, [0])
    out = alloc(ctx, span_bytes(sh, st, 8))
    ax = (C.c_ssize_t * 1)(0)
    a = C.c_void_p(x._t._data_ptr())
    # argmax goes through the arg entry point, as for rstsr_hip_reduce_axes
    assert L.rstsr_hip_reduce_axes_noalloc(ctx, 3, rt._code(x.dtype), a, C.byref(la), 1, ax, out, C.byref(lo)) != 0
    # an output too small for the layout is refused before launch
    small = alloc(ctx, 8)
    assert L.rstsr_hip_reduce_axes_noalloc(ctx, SUM, rt._code(x.dtype), a, C.byref(la), 1, ax, small,
                                           C.byref(lo)) != 0
    abi.check(L.rstsr_hip_free(ctx, small))
    abi.check(L.rstsr_hip_free(ctx, out))
