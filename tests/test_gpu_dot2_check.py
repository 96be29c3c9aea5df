"""The GPU Dot2 checker used for full-matrix GEMM errors equals the CPU oracle's Dot2 GEMM bit for
bit (same operation order per element), on ragged shapes and strided / transposed operands."""
import numpy as np
import pytest

import rstsr_amd as rt
from dot2_gpu import dot2_matmul, normwise

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (65, 70, 33), (128, 64, 100), (37, 200, 17)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True)])
def test_dot2_checker_equals_oracle(dev, oracle, m, n, k, ta, tb):
    rng = np.random.default_rng(20261220 + m + n + k)
    A = rng.standard_normal((k, m) if ta else (m, k)) * np.exp2(rng.integers(-8, 9, size=(k, m) if ta else (m, k)))
    B = rng.standard_normal((n, k) if tb else (k, n))
    Ah = A.T if ta else A
    Bh = B.T if tb else B
    tA = rt.asarray(A, device=dev)
    tB = rt.asarray(B, device=dev)
    got = dot2_matmul(rt, dev, tA.T if ta else tA, tB.T if tb else tB).to_numpy()
    ref = oracle.gemm_dot2(np.ascontiguousarray(Ah), np.ascontiguousarray(Bh))
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_normwise_reduction(dev):
    rng = np.random.default_rng(20261221)
    r = rng.standard_normal((300, 170))
    g = r.copy()
    g[123, 45] += 1e-3
    tr, tg = rt.asarray(r, device=dev), rt.asarray(g, device=dev)
    e = normwise(dev, tg, tr)
    assert e == pytest.approx(abs(g[123, 45] - r[123, 45]) / np.max(np.abs(r)), rel=1e-12)
    g[7, 7] = np.nan
    assert normwise(dev, rt.asarray(g, device=dev), tr) == np.inf
    # strided views (every other column of the transposed matrix)
    e2 = normwise(dev, tg.T[:, ::2], tr.T[:, ::2])
    ref = np.max(np.abs((g - r)[:, ::2])) / np.max(np.abs(r[:, ::2]))
    assert e2 == pytest.approx(ref, rel=1e-12)
