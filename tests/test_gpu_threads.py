"""One context driven from several host threads at once through the C ABI (the reference's test
runner calls a device from many threads): each thread runs its own strided multiply, axis sum,
argmax and DGEMM on its own buffers, repeatedly; every result must equal the single-threaded run
bit for bit (same plans, same kernels)."""
import ctypes as C
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import rstsr_amd as rt
from rstsr_amd import abi

pytestmark = pytest.mark.gpu

F64, MUL, SUM, ARGMAX = 1, 2, 0, 3


def run_job(ctx, seed, reps):
    L = abi.lib()
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(24, 40, 64))
    bvec = rng.uniform(0.5, 2.0, size=40)
    A = rng.normal(size=(96, 80))
    B = rng.normal(size=(80, 72))

    def dmalloc(nbytes):
        p = C.c_void_p()
        abi.check(L.rstsr_hip_malloc(ctx, nbytes, C.byref(p)))
        return p

    def up(arr):
        arr = np.ascontiguousarray(arr)
        p = dmalloc(arr.nbytes)
        abi.check(L.rstsr_hip_memcpy_h2d(ctx, p, arr.ctypes.data_as(C.c_void_p), arr.nbytes))
        return p

    def down(p, shape, dt):
        h = np.empty(shape, dtype=dt)
        abi.check(L.rstsr_hip_memcpy_d2h(ctx, h.ctypes.data_as(C.c_void_p), p, h.nbytes))
        return h

    dX, db, dA, dB = up(X), up(bvec), up(A), up(B)
    # V = X.permute(2,0,1)[:, :, ::2] ([64,24,20], strides [1, 2560, 128]); bV broadcast [1,1,20]
    lV = abi.make_layout([64, 24, 20], [1, 2560, 128], 0)
    lb = abi.make_layout([64, 24, 20], [0, 0, 2], 0)
    lY = abi.make_layout([64, 24, 20], [480, 20, 1], 0)
    lS = abi.make_layout([24, 20], [20, 1], 0)
    lI = abi.make_layout([64], [1], 0)
    ax0 = (C.c_ssize_t * 1)(0)
    ax12 = (C.c_ssize_t * 2)(1, 2)
    one, zero = np.array([1.0]), np.array([0.0])
    dY, dS, dI, dCm = dmalloc(64 * 24 * 20 * 8), dmalloc(24 * 20 * 8), dmalloc(64 * 8), dmalloc(96 * 72 * 8)
    outs = []
    for _ in range(reps):
        abi.check(L.rstsr_hip_op_mutc_refa_refb(ctx, MUL, F64, dY, C.byref(lY), dX, C.byref(lV), db, C.byref(lb)))
        abi.check(L.rstsr_hip_reduce_axes(ctx, SUM, F64, dX, C.byref(lV), 1, ax0, dS, C.byref(lS)))
        abi.check(L.rstsr_hip_argreduce_axes(ctx, ARGMAX, F64, dX, C.byref(lV), 2, ax12, dI, C.byref(lI)))
        abi.check(L.rstsr_hip_gemm(ctx, F64, 0, 96, 72, 80, one.ctypes.data_as(C.c_void_p), dA, 80, 1, dB, 72, 1,
                                   zero.ctypes.data_as(C.c_void_p), dCm, 72, 1))
        outs.append((down(dY, (64, 24, 20), np.float64), down(dS, (24, 20), np.float64),
                     down(dI, (64,), np.uint64), down(dCm, (96, 72), np.float64)))
    for p in (dX, db, dA, dB, dY, dS, dI, dCm):
        abi.check(L.rstsr_hip_free(ctx, p))
    # the bit-exact parts against NumPy on the same views
    V = X.transpose(2, 0, 1)[:, :, ::2]
    bV = bvec[::2][None, None, :]
    for y, s, i, c in outs:
        assert np.array_equal(y, V * bV)
        assert np.array_equal(i, np.argmax(V.reshape(64, -1), axis=1))
    return outs


def test_one_context_many_threads(dev):
    ctx = C.c_void_p(dev._ctx_ptr())
    seeds = list(range(20261240, 20261248))
    single = {s: run_job(ctx, s, 1)[0] for s in seeds}
    with ThreadPoolExecutor(max_workers=8) as ex:
        results = list(ex.map(lambda s: run_job(ctx, s, 6), seeds))
    for s, outs in zip(seeds, results):
        for o in outs:
            for got, ref in zip(o, single[s]):
                assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), s


def test_contexts_per_thread(dev):
    """Each thread with its own context on the same GPU, concurrently."""
    seeds = list(range(20261250, 20261254))
    ctx0 = C.c_void_p(dev._ctx_ptr())
    single = {s: run_job(ctx0, s, 1)[0] for s in seeds}
    devs = [rt.DeviceHIP(0) for _ in seeds]
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(lambda sd: run_job(C.c_void_p(sd[1]._ctx_ptr()), sd[0], 4), zip(seeds, devs)))
    for s, outs in zip(seeds, results):
        for o in outs:
            for got, ref in zip(o, single[s]):
                assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), s
