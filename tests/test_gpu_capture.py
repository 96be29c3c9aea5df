"""hipGraph capture of the context stream, context-local workspaces, and several contexts on one GPU.

The c128 chain of the captured-graph workload (Z1 <- Z1 + v; Z3 = Z1 * Z2; Z4 = Z3 % Z2;
s = Z4.sum_axes(1)) is checked step by step, eager and replayed, against an extended-precision
host reference: normwise max|s - s*| / max|s*| <= 1e-13 (plain FP64 MFMA and the split GEMM).
"""
import numpy as np
import pytest

import rstsr_amd as rt

pytestmark = pytest.mark.gpu

CHAIN_TOL = 1e-13


def chain_reference(z1h, z2h, vh, steps):
    """s* of each step: Z1 accumulates v in f64 (IEEE adds, as the device does), the product,
    matmul and row sum run in extended precision on the host."""
    z2l = z2h.astype(np.clongdouble)
    z1 = z1h.copy()
    out = []
    for _ in range(steps):
        z1 = z1 + vh[None, :]
        z3 = z1.astype(np.clongdouble) * z2l
        z4 = z3 @ z2l
        out.append(z4.sum(axis=1))
    return out


def normwise(got, ref):
    return float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))


@pytest.mark.parametrize("mode", ["plain", "ozaki"])
def test_graph_chain_against_oracle(dev, mode):
    """Four steps eager and four replays of one captured step, each step's s against the host
    reference; eager work of other sizes runs between the replays."""
    rng = np.random.default_rng(20261206)
    n, steps = 160, 4
    z1h = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    z2h = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    vh = rng.normal(size=n) + 1j * rng.normal(size=n)
    refs = chain_reference(z1h, z2h, vh, steps)
    z2 = rt.asarray(z2h, device=dev)
    v = rt.asarray(vh, device=dev)

    z1 = rt.asarray(z1h, device=dev)
    for t in range(steps):
        z1 += v
        s = rt.matmul(z1 * z2, z2, mode=mode).sum_axes([1]).to_numpy()
        err = normwise(s, refs[t])
        assert err <= CHAIN_TOL, ("eager", t, err)

    z1 = rt.asarray(z1h, device=dev)
    with rt.capture(dev) as g:
        z1 += v
        z3 = z1 * z2
        z4 = rt.matmul(z3, z2, mode=mode)
        s = z4.sum_axes([1])
    for t in range(steps):
        # unrelated eager work of another size between replays
        w = rt.asarray(rng.normal(size=(96, 200)), device=dev)
        _ = rt.matmul(w, w.T, mode=mode).sum_axes([0]).to_numpy()
        g.launch()
        dev.synchronize()
        err = normwise(s.to_numpy(), refs[t])
        assert err <= CHAIN_TOL, ("replay", t, err)


def test_graph_workspace_survives_eager_growth(dev, oracle):
    """A captured split matmul and split reductions keep their workspace addresses in the graph.
    An eager call that needs a much larger workspace afterwards must not free or reuse that memory:
    every replay still matches the oracle."""
    rng = np.random.default_rng(77)
    A = rng.normal(size=(96, 130))
    B = rng.normal(size=(130, 80))
    X = rng.normal(size=(8192, 24))       # column reduction split over the reduced axis
    Y = rng.normal(size=(3, 1 << 16))     # row reduction split along each row
    ta, tb = rt.asarray(A, device=dev), rt.asarray(B, device=dev)
    tx, ty = rt.asarray(X, device=dev), rt.asarray(Y, device=dev)
    with rt.capture(dev) as g:
        c = rt.matmul(ta, tb, mode="ozaki")
        sx = tx.sum_axes([0])
        sy = ty.sum_axes([1])
    cref = oracle.gemm_dot2(A, B)
    for rep in range(3):
        # a much larger eager split product and split sums grow every context workspace
        big = rt.asarray(rng.normal(size=(700 + 300 * rep, 900)), device=dev)
        _ = rt.matmul(big, big.T, mode="ozaki")
        _ = rt.asarray(rng.normal(size=(1 << 18, 40)), device=dev).sum_axes([0])
        g.launch()
        dev.synchronize()
        assert np.max(np.abs(c.to_numpy() - cref)) <= 1e-14 * np.max(np.abs(cref)), rep
        assert np.allclose(sx.to_numpy(), np.sum(X, axis=0), rtol=1e-13, atol=1e-12), rep
        assert np.allclose(sy.to_numpy(), np.sum(Y, axis=1), rtol=1e-13, atol=1e-12), rep
    del g


def test_host_copies_refuse_capture(dev):
    """Uploads and downloads synchronise with the host: under capture they fail before anything is
    enqueued, the block's graph is dropped, and the context stays usable."""
    x = rt.asarray(np.arange(32.0), device=dev)
    with pytest.raises(Exception, match="cannot be captured"):
        with rt.capture(dev):
            rt.asarray(np.ones(8), device=dev)
    with pytest.raises(Exception, match="cannot be captured"):
        with rt.capture(dev):
            x.to_numpy()
    with rt.capture(dev) as g:
        y = x * x
    g.launch()
    dev.synchronize()
    assert np.array_equal(y.to_numpy(), np.arange(32.0) ** 2)


def test_operands_from_another_context_are_ordered(dev):
    """b is produced on a second context of the same GPU and used at once by the first (binary,
    in-place, matmul, assign): the first context's kernels wait for the second's queue."""
    dev2 = rt.DeviceHIP(0)
    rng = np.random.default_rng(5)
    n = 1 << 24
    xh = rng.normal(size=n)
    yh = rng.normal(size=n)
    x2, y2 = rt.asarray(xh, device=dev2), rt.asarray(yh, device=dev2)
    a = rt.asarray(np.ones(n), device=dev)
    for _ in range(3):
        b = x2 * y2                       # queued on dev2
        c = a + b                         # queued on dev
        assert np.array_equal(c.to_numpy(), 1.0 + xh * yh)
        d = rt.asarray(np.zeros(n), device=dev)
        b2 = x2 + y2
        d += b2
        assert np.array_equal(d.to_numpy(), xh + yh)
        e = rt.empty((n,), device=dev)
        e.assign(x2 - y2)
        assert np.array_equal(e.to_numpy(), xh - yh)
    M = rng.normal(size=(256, 256))
    m2 = rt.asarray(M, device=dev2) * 2.0
    got = rt.matmul(rt.asarray(np.eye(256), device=dev), m2).to_numpy()
    assert np.array_equal(got, 2.0 * M)
