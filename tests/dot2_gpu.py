"""ctypes wrapper of the GPU Dot2 GEMM checker (tests/gpu_check/dot2_check.hip).

Test infrastructure: a full-matrix error reference for the GEMM legs, too large for the CPU
oracle. It follows the oracle's o_gemm_dot2_f64 operation for operation, which
tests/test_gpu_dot2_check.py checks bit for bit."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "gpu_check", "libdot2check.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("tests/gpu_check/libdot2check.so is not built: run `python rstsr_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        i64, vp = C.c_int64, C.c_void_p
        L.dot2_gemm.argtypes = [i64, i64, i64, vp, i64, i64, vp, i64, i64, vp, i64, i64]
        L.dot2_normwise.argtypes = [i64, i64, vp, i64, i64, vp, i64, i64, C.POINTER(C.c_double)]
        L.dot2_gemm.restype = L.dot2_normwise.restype = C.c_int
        _lib = L
    return _lib


def _p(t):
    """Element pointer of a rstsr_amd tensor view (base + offset) and its 2-d strides."""
    return C.c_void_p(t._t._data_ptr() + 8 * t.offset)


def dot2_matmul(rt, dev, a, b):
    """Device tensor C* = a @ b (2-d f64 views) in Dot2; returns a C-contiguous device tensor."""
    m, k = a.shape
    k2, n = b.shape
    assert k == k2 and a.dtype == np.float64 and b.dtype == np.float64
    c = rt.empty([m, n], np.float64, device=dev)
    dev.synchronize()
    rc = lib().dot2_gemm(m, n, k, _p(a), a.stride[0], a.stride[1], _p(b), b.stride[0], b.stride[1], _p(c), n, 1)
    if rc != 0:
        raise RuntimeError(f"dot2_gemm failed with HIP error {rc}")
    return c


def normwise(dev, got, ref):
    """max|got - ref| / max|ref| over two 2-d f64 device views of one shape."""
    assert got.shape == ref.shape
    m, n = got.shape
    out = (C.c_double * 2)()
    dev.synchronize()
    rc = lib().dot2_normwise(m, n, _p(got), got.stride[0], got.stride[1], _p(ref), ref.stride[0], ref.stride[1], out)
    if rc != 0:
        raise RuntimeError(f"dot2_normwise failed with HIP error {rc}")
    return float(out[0] / out[1]) if out[1] > 0 else float(out[0])
