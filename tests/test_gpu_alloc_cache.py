"""The context's block cache behind rstsr_hip_malloc / rstsr_hip_free (storage/creation.rs:3-34 uninit,
Drop of the storage's Raw): a freed buffer is reused by the next allocation of its size class on the
same context, the cache is accounted and emptied, a freed pointer is no longer a valid operand, and
results stay exact when temporaries are recycled between operations and across contexts."""
import ctypes as C

import numpy as np
import pytest

import rstsr_amd as rt
from rstsr_amd import abi

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def _cached(L, ctx):
    n = C.c_size_t()
    abi.check(L.rstsr_hip_cached_bytes(ctx, C.byref(n)))
    return n.value


def _malloc(L, ctx, n):
    p = C.c_void_p()
    abi.check(L.rstsr_hip_malloc(ctx, n, C.byref(p)))
    return p


def test_free_then_malloc_reuses_block(dev):
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    abi.check(L.rstsr_hip_empty_cache(ctx))
    assert _cached(L, ctx) == 0
    p = _malloc(L, ctx, 5 * MiB + 3)
    abi.check(L.rstsr_hip_free(ctx, p))
    held = _cached(L, ctx)
    assert held >= 5 * MiB + 3 and held % (2 * MiB) == 0  # 2 MiB classes above 1 MiB
    q = _malloc(L, ctx, 5 * MiB)  # same class
    assert q.value == p.value and _cached(L, ctx) == 0
    r = _malloc(L, ctx, 64 * MiB)  # another class: fresh block
    assert r.value != q.value
    for x in (q, r):
        abi.check(L.rstsr_hip_free(ctx, x))
    assert _cached(L, ctx) >= 69 * MiB
    abi.check(L.rstsr_hip_empty_cache(ctx))
    assert _cached(L, ctx) == 0


def test_small_classes_and_double_free(dev):
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    p = _malloc(L, ctx, 1)
    abi.check(L.rstsr_hip_free(ctx, p))
    q = _malloc(L, ctx, 200)  # both in the 256-byte class
    assert q.value == p.value
    abi.check(L.rstsr_hip_free(ctx, q))
    assert L.rstsr_hip_free(ctx, q) != 0  # a second free of the same pointer is an error
    abi.check(L.rstsr_hip_empty_cache(ctx))


def test_freed_pointer_is_not_an_operand(dev):
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    p = _malloc(L, ctx, 4096)
    abi.check(L.rstsr_hip_free(ctx, p))  # cached, but no longer registered
    lay = abi.make_layout([512], [1], 0)
    one = C.c_double(1.0)
    rc = L.rstsr_hip_fill(ctx, 1, p, C.byref(lay), C.byref(one))
    assert rc != 0
    abi.check(L.rstsr_hip_empty_cache(ctx))


def test_reuse_keeps_results_exact(dev):
    # temporaries of every iteration are freed and recycled by the next; each result is checked
    rng = np.random.default_rng(11)
    for it in range(40):
        n = [4096, 1 << 16, 3 * (1 << 18) + 5][it % 3]
        a, b, c = (rng.standard_normal(n) for _ in range(3))
        ta, tb, tc = (rt.asarray(x, device=dev) for x in (a, b, c))
        got = ((ta * tb) + tc).to_numpy()
        assert np.array_equal(got, (a * b) + c), it
        s = (ta - tc).sum_all()
        assert abs(s - float(np.sum(a - c))) <= 1e-9 * (np.abs(a).sum() + np.abs(c).sum())


def test_matmul_outputs_recycled(dev):
    rng = np.random.default_rng(12)
    A = rng.standard_normal((256, 192))
    B = rng.standard_normal((192, 128))
    tA, tB = rt.asarray(A, device=dev), rt.asarray(B, device=dev)
    first = rt.matmul(tA, tB).to_numpy()
    for _ in range(10):
        out = rt.matmul(tA, tB)  # the previous output's block comes back
        assert np.array_equal(out.to_numpy(), first)
    assert np.allclose(first, A @ B, rtol=1e-12, atol=1e-12)


def test_copy_between_contexts_then_free(dev):
    # the source buffer goes back to its context's cache and is overwritten at once; the copy
    # on the other context's stream must already have read it
    other = rt.DeviceHIP(0)
    rng = np.random.default_rng(13)
    for it in range(6):
        x = rng.standard_normal(3 * MiB // 8 + it)
        src = rt.asarray(x, device=dev)
        moved = src.to_device(other)
        del src
        junk = rt.full([3 * MiB // 8 + it], -7.0, device=dev)  # likely the same block
        assert np.array_equal(moved.to_numpy(), x), it
        assert np.all(junk.to_numpy() == -7.0)
        back = moved.to_device(dev)
        del moved
        assert np.array_equal(back.to_numpy(), x), it


def test_capture_temporaries_not_recycled(dev):
    # (a + b) is a temporary freed during capture: the graph writes and reads it at every replay,
    # so eager allocations made afterwards must not receive its block
    rng = np.random.default_rng(14)
    n = 1 << 18
    a, b, c = (rng.standard_normal(n) for _ in range(3))
    ta, tb, tc = (rt.asarray(x, device=dev) for x in (a, b, c))
    with rt.capture(dev) as g:
        out = (ta + tb) * tc
    held = [rt.full([n], float(i), device=dev) for i in range(4)]  # same size class as the temporary
    for _ in range(3):
        g.launch()
        dev.synchronize()
        assert np.array_equal(out.to_numpy(), (a + b) * c)
        for i, h in enumerate(held):
            assert np.all(h.to_numpy() == float(i)), i
    del g


def test_python_cache_helpers(dev):
    t = rt.zeros([1 << 20], device=dev)
    del t
    assert rt.cached_bytes(dev) >= 8 * MiB
    rt.empty_cache(dev)
    assert rt.cached_bytes(dev) == 0
