"""The CPU baseline's OpenMP kernels (b_binary, b_assign, b_reduce_axes in oracle/oracle.c, the
functions bench.py times as cpu_baseline) checked against the oracle's serial restatement
(o_binary, o_assign, o_reduce_axes) on the same strided views the GPU step uses, at 1 and 4
threads: elementwise results, maxima and argmax indices bit for bit, f64 sums within a
summation-order bound (the baseline sums in unrolled blocks and per-thread chunks)."""
import ctypes as C

import numpy as np
import pytest

from oracle_ffi import View, c_view

ADD, MUL = 0, 2
R_SUM, R_MAX, R_ARGMAX = 0, 1, 3
CODE = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.int32): 2}


def _step_views(seed=5):
    """The bench step's view pattern at a small size: X [a, b, c] read as permute (2, 0, 1)
    with step 2 on the last axis, b broadcast as [1, c, 1] before the same transform."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((24, 37, 50))
    b = rng.standard_normal((1, 37, 1))
    Xi = rng.integers(-1000, 1000, size=X.shape).astype(np.int32)
    bi = rng.integers(-1000, 1000, size=b.shape).astype(np.int32)
    X32 = X.astype(np.float32)
    V = X.transpose(2, 0, 1)[:, :, ::2]
    bV = np.broadcast_to(np.broadcast_to(b, X.shape).transpose(2, 0, 1)[:, :, ::2], V.shape)
    V32 = X32.transpose(2, 0, 1)[:, :, ::2]
    bV32 = np.broadcast_to(np.broadcast_to(b.astype(np.float32), X.shape).transpose(2, 0, 1)[:, :, ::2], V32.shape)
    Vi = Xi.transpose(2, 0, 1)[:, :, ::2]
    bVi = np.broadcast_to(np.broadcast_to(bi, X.shape).transpose(2, 0, 1)[:, :, ::2], Vi.shape)
    return V, bV, V32, bV32, Vi, bVi


@pytest.mark.parametrize("threads", [1, 4])
@pytest.mark.parametrize("case", ["mul_f64", "mul_f32", "add_i32"])
def test_baseline_binary_matches_oracle(oracle, threads, case):
    V, bV, V32, bV32, Vi, bVi = _step_views()
    op, a, b = {"mul_f64": (MUL, V, bV), "mul_f32": (MUL, V32, bV32), "add_i32": (ADD, Vi, bVi)}[case]
    va, vb = View.of(a), View.of(b)
    got, ref = c_view(a.shape, a.dtype), c_view(a.shape, a.dtype)
    assert oracle.lib.b_binary(op, CODE[a.dtype], got.ptr, C.byref(got.c), va.ptr, C.byref(va.c), vb.ptr,
                               C.byref(vb.c), threads) == 0
    oracle.binary(op, ref, va, vb)
    assert np.array_equal(got.array().view(np.uint8), ref.array().view(np.uint8))
    # and the oracle itself on this view is NumPy's IEEE result
    assert np.array_equal(ref.array(), (a * b if op == MUL else a + b))


@pytest.mark.parametrize("threads", [1, 4])
def test_baseline_assign_matches_oracle(oracle, threads):
    X = np.random.default_rng(6).standard_normal((24, 37, 50))
    P = X.transpose(2, 0, 1)  # the step's to_contig of the permuted array
    vp = View.of(P)
    got, ref = c_view(P.shape, np.float64), c_view(P.shape, np.float64)
    assert oracle.lib.b_assign(1, got.ptr, C.byref(got.c), vp.ptr, C.byref(vp.c), threads) == 0
    oracle.assign(ref, vp)
    assert np.array_equal(got.array(), ref.array())
    assert np.array_equal(ref.array(), np.ascontiguousarray(P))


def _b_reduce(oracle, op, a, axes, out_dtype, threads):
    va = View.of(a)
    kept = [a.shape[i] for i in range(a.ndim) if i not in axes]
    vo = c_view(kept, out_dtype)
    ax = (C.c_ssize_t * len(axes))(*axes)
    assert oracle.lib.b_reduce_axes(op, CODE[a.dtype], va.ptr, C.byref(va.c), len(axes), ax, vo.ptr, C.byref(vo.c),
                                    threads) == 0
    return vo.array()


def _o_reduce(oracle, op, a, axes, out_dtype):
    kept = [a.shape[i] for i in range(a.ndim) if i not in axes]
    vo = c_view(kept, out_dtype)
    assert oracle.reduce_axes(op, View.of(a), axes, vo) == 0
    return vo.array()


@pytest.mark.parametrize("threads", [1, 4])
@pytest.mark.parametrize("axes", [[0], [1, 2]])
def test_baseline_reductions_match_oracle(oracle, threads, axes):
    V, _, _, _, Vi, _ = _step_views()
    # sums: the baseline's order differs; bound |b - o| by 2 n eps sum|x| per output
    got = _b_reduce(oracle, R_SUM, V, axes, np.float64, threads)
    ref = _o_reduce(oracle, R_SUM, V, axes, np.float64)
    n = int(np.prod([V.shape[i] for i in axes]))
    absum = np.abs(V).sum(axis=tuple(axes))
    assert np.all(np.abs(got - ref) <= 2 * n * np.finfo(np.float64).eps * absum)
    assert np.allclose(ref, V.sum(axis=tuple(axes)), rtol=0, atol=1e-12 * absum.max())
    # max and argmax: exact
    for op, a, dt in ((R_MAX, V, np.float64), (R_ARGMAX, V, np.uint64), (R_ARGMAX, Vi, np.uint64)):
        g = _b_reduce(oracle, op, a, axes, dt, threads)
        r = _o_reduce(oracle, op, a, axes, dt)
        assert np.array_equal(g, r), (op, a.dtype)


def test_planted_wrong_baseline_fails(oracle):
    """A baseline that summed one element twice, or returned the last of tied maxima, would
    fail the checks above: both mistakes are visible at these sizes."""
    V = _step_views()[0]
    ref = _o_reduce(oracle, R_SUM, V, [0], np.float64)
    absum = np.abs(V).sum(axis=0)
    wrong = ref + V[0]  # element 0 of each output counted twice
    assert not np.all(np.abs(wrong - ref) <= 2 * V.shape[0] * np.finfo(np.float64).eps * absum)
    T = np.zeros((6, 4))
    T[[1, 4], :] = 1.0  # ties: the first index (1) is the reference's answer
    r = _o_reduce(oracle, R_ARGMAX, T, [0], np.uint64)
    assert np.all(r == 1)
