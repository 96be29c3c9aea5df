"""ctypes binding of oracle/liboracle.so (the CPU restatement used as parity checker).

NumPy arrays are passed with their own (shape, strides, offset) expressed in
elements over a base 1-D buffer, exactly as the device sees rstsr layouts.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

CODES = {
    np.dtype(np.float32): 0,
    np.dtype(np.float64): 1,
    np.dtype(np.int32): 2,
    np.dtype(np.int64): 3,
    np.dtype(np.complex64): 4,
    np.dtype(np.complex128): 5,
    np.dtype(np.uint64): 6,
    np.dtype(np.bool_): 7,
}
DTYPES = {v: k for k, v in CODES.items()}

# unary op codes (rstsr_unop_t order)
UNARY_OPS = (
    "acos", "acosh", "asin", "asinh", "atan", "atanh", "ceil", "conj", "cos", "cosh", "exp", "expm1",
    "floor", "inv", "is_finite", "is_inf", "is_nan", "log", "log1p", "log2", "log10", "reciprocal",
    "round", "signbit", "sin", "sinh", "square", "sqrt", "tan", "tanh", "trunc", "abs", "imag", "real",
    "sign", "neg", "not_",
)
UN = {name: i for i, name in enumerate(UNARY_OPS)}


def red_out_dtype(op, dtype):
    """Element type of a reduction's result: u64 (usize) for argmax/argmin, count_nonzero and the
    sum of a bool array; bool for all/any; the real part type for var/std/l2_norm of complex."""
    dt = np.dtype(dtype)
    if op in (3, 4, 12) or (op == 0 and dt == np.bool_):
        return np.dtype(np.uint64)
    if op in (10, 11):
        return np.dtype(np.bool_)
    if op in (7, 8, 9) and dt.kind == "c":
        return np.dtype(np.float32) if dt == np.complex64 else np.dtype(np.float64)
    return dt


class OLayout(C.Structure):
    _fields_ = [
        ("ndim", C.c_size_t),
        ("shape", C.POINTER(C.c_size_t)),
        ("stride", C.POINTER(C.c_ssize_t)),
        ("offset", C.c_size_t),
    ]


class View:
    """A strided view (shape, stride, offset in elements) over a flat 1-D base array."""

    def __init__(self, base, shape, stride, offset):
        self.base = np.ascontiguousarray(base).reshape(-1)
        self.shape = tuple(int(s) for s in shape)
        self.stride = tuple(int(s) for s in stride)
        self.offset = int(offset)
        self._sh = (C.c_size_t * max(len(self.shape), 1))(*self.shape)
        self._st = (C.c_ssize_t * max(len(self.stride), 1))(*self.stride)
        self.c = OLayout(len(self.shape), self._sh, self._st, self.offset)

    @classmethod
    def of(cls, arr):
        """View of an existing numpy array (any strides) without copying its base."""
        it = arr.itemsize
        base = arr
        # walk up to the owning contiguous buffer
        while base.base is not None and isinstance(base.base, np.ndarray):
            base = base.base
        base = base.reshape(-1) if base.flags.c_contiguous else np.ascontiguousarray(base).reshape(-1)
        off_bytes = arr.__array_interface__["data"][0] - base.__array_interface__["data"][0]
        assert off_bytes % it == 0
        v = cls.__new__(cls)
        v.base = base
        v.shape = arr.shape
        v.stride = tuple(s // it for s in arr.strides)
        v.offset = off_bytes // it
        v._sh = (C.c_size_t * max(len(v.shape), 1))(*v.shape)
        v._st = (C.c_ssize_t * max(len(v.stride), 1))(*v.stride)
        v.c = OLayout(len(v.shape), v._sh, v._st, v.offset)
        return v

    def array(self):
        it = self.base.itemsize
        return np.lib.stride_tricks.as_strided(
            self.base[self.offset:], shape=self.shape, strides=[s * it for s in self.stride]
        )

    @property
    def ptr(self):
        return self.base.ctypes.data_as(C.c_void_p)


def c_view(shape, dtype, fill=None):
    n = int(np.prod(shape)) if len(shape) else 1
    base = np.zeros(n, dtype=dtype) if fill is None else np.full(n, fill, dtype=dtype)
    st, acc = [], 1
    for s in reversed(shape):
        st.append(acc)
        acc *= max(s, 1)
    return View(base, shape, list(reversed(st)), 0)


class Oracle:
    def __init__(self):
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so not built: run `make -C oracle`")
        self.lib = C.CDLL(LIB)
        L = C.POINTER(OLayout)
        vp = C.c_void_p
        self.lib.o_binary.argtypes = [C.c_int, C.c_int, vp, L, vp, L, vp, L]
        self.lib.o_assign.argtypes = [C.c_int, vp, L, C.c_int, vp, L]
        self.lib.o_binary_mixed.argtypes = [C.c_int, C.c_int, vp, L, C.c_int, vp, L, C.c_int, vp, L]
        self.lib.o_reduce_axes.argtypes = [C.c_int, C.c_int, vp, L, C.c_size_t, C.POINTER(C.c_ssize_t), vp, L]
        self.lib.o_reduce_all.argtypes = [C.c_int, C.c_int, vp, L, vp]
        self.lib.o_gemm_naive.argtypes = [C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, vp, vp, C.c_ssize_t,
                                          C.c_ssize_t, vp, C.c_ssize_t, C.c_ssize_t, vp, vp, C.c_ssize_t,
                                          C.c_ssize_t]
        self.lib.o_gemm_dot2_f64.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, vp, C.c_ssize_t, C.c_ssize_t, vp,
                                             C.c_ssize_t, C.c_ssize_t, vp, C.c_ssize_t, C.c_ssize_t, C.c_int]
        self.lib.b_binary.argtypes = [C.c_int, C.c_int, vp, L, vp, L, vp, L, C.c_int]
        self.lib.b_assign.argtypes = [C.c_int, vp, L, vp, L, C.c_int]
        self.lib.b_reduce_axes.argtypes = [C.c_int, C.c_int, vp, L, C.c_size_t, C.POINTER(C.c_ssize_t), vp, L,
                                           C.c_int]
        self.lib.o_max_threads.restype = C.c_int
        self.lib.o_unary.argtypes = [C.c_int, C.c_int, vp, L, C.c_int, vp, L]
        self.lib.o_unary_out_dtype.argtypes = [C.c_int, C.c_int]

    # ---- wrappers taking View objects; outputs written into View bases ----
    def binary(self, op, c, a, b, scalar_b=None):
        t = CODES[a.base.dtype]
        if scalar_b is not None:
            sb = np.array([scalar_b], dtype=a.base.dtype)
            r = self.lib.o_binary(op, t, c.ptr, C.byref(c.c), a.ptr, C.byref(a.c), sb.ctypes.data_as(C.c_void_p), None)
        else:
            r = self.lib.o_binary(op, t, c.ptr, C.byref(c.c), a.ptr, C.byref(a.c), b.ptr, C.byref(b.c))
        assert r == 0

    def assign(self, c, a):
        r = self.lib.o_assign(CODES[c.base.dtype], c.ptr, C.byref(c.c), CODES[a.base.dtype], a.ptr, C.byref(a.c))
        assert r == 0

    def reduce_axes(self, op, a, axes, out):
        ax = (C.c_ssize_t * max(len(axes), 1))(*axes)
        return self.lib.o_reduce_axes(op, CODES[a.base.dtype], a.ptr, C.byref(a.c), len(axes), ax, out.ptr,
                                      C.byref(out.c))

    def binary_mixed(self, op, c, a, b):
        """o_binary_mixed: c (complex View) = a op b, one of a / b complex and the other its real type;
        b may be a numpy scalar (one host value)."""
        if isinstance(b, View):
            return self.lib.o_binary_mixed(op, CODES[c.base.dtype], c.ptr, C.byref(c.c), CODES[a.base.dtype], a.ptr,
                                           C.byref(a.c), CODES[b.base.dtype], b.ptr, C.byref(b.c))
        sb = np.array(b)
        return self.lib.o_binary_mixed(op, CODES[c.base.dtype], c.ptr, C.byref(c.c), CODES[a.base.dtype], a.ptr,
                                       C.byref(a.c), CODES[sb.dtype], sb.ctypes.data_as(C.c_void_p), None)

    def allclose(self, a, b, rtol=1e-5, atol=1e-8, equal_nan=False):
        """o_allclose on two Views already broadcast to one shape: True/False, None for zero size."""
        L = self.lib
        L.o_allclose.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                 C.c_int]
        r = L.o_allclose(CODES[a.base.dtype], a.ptr, C.byref(a.c), b.ptr, C.byref(b.c), rtol, atol, int(equal_nan))
        return None if r < 0 else bool(r)

    def reduce_all(self, op, a):
        out = np.zeros(1, dtype=red_out_dtype(op, a.base.dtype))
        r = self.lib.o_reduce_all(op, CODES[a.base.dtype], a.ptr, C.byref(a.c), out.ctypes.data_as(C.c_void_p))
        if r != 0:
            raise ValueError("oracle reduce_all error %d" % r)
        return out[0]

    def unary_out_dtype(self, op, dtype):
        """Output numpy dtype of unary `op` on `dtype`, or None when the reference has none."""
        r = self.lib.o_unary_out_dtype(UN[op] if isinstance(op, str) else op, CODES[np.dtype(dtype)])
        return None if r < 0 else DTYPES[r]

    def unary(self, op, c, a):
        """c = op(a) over two Views of one shape (c's dtype is the op's output dtype)."""
        code = UN[op] if isinstance(op, str) else op
        r = self.lib.o_unary(code, CODES[c.base.dtype], c.ptr, C.byref(c.c), CODES[a.base.dtype], a.ptr,
                             C.byref(a.c))
        if r != 0:
            raise ValueError("oracle unary: no implementation for this op/dtype")

    def unary_array(self, op, x):
        """op(x) for a numpy array x (any strides); returns a new c-contiguous array."""
        dt = self.unary_out_dtype(op, x.dtype)
        if dt is None:
            raise ValueError("oracle unary: no implementation for this op/dtype")
        out = c_view(list(x.shape), dt)
        self.unary(op, out, View.of(x))
        return out.array().copy()

    def gemm_naive(self, A, B, alpha=1, beta=0, C0=None):
        """C = alpha*A@B + beta*C0 with the naive sequential-k order; A, B 2-D numpy (any strides)."""
        dt = A.dtype
        m, k = A.shape
        n = B.shape[1]
        Cm = np.zeros((m, n), dtype=dt) if C0 is None else np.array(C0, dtype=dt, copy=True)
        va, vb = View.of(A), View.of(B)
        al = np.array([alpha], dtype=dt)
        be = np.array([beta], dtype=dt)
        it = dt.itemsize
        pa = C.c_void_p(va.base.ctypes.data + va.offset * it)
        pb = C.c_void_p(vb.base.ctypes.data + vb.offset * it)
        self.lib.o_gemm_naive(CODES[dt], m, n, k, al.ctypes.data_as(C.c_void_p), pa, va.stride[0], va.stride[1],
                              pb, vb.stride[0], vb.stride[1], be.ctypes.data_as(C.c_void_p),
                              Cm.ctypes.data_as(C.c_void_p), n, 1)
        return Cm

    def gemm_dot2(self, A, B, nthreads=0):
        m, k = A.shape
        n = B.shape[1]
        A = np.ascontiguousarray(A, dtype=np.float64)
        B = np.ascontiguousarray(B, dtype=np.float64)
        Cm = np.zeros((m, n))
        vp = lambda x: x.ctypes.data_as(C.c_void_p)
        self.lib.o_gemm_dot2_f64(m, n, k, vp(A), k, 1, vp(B), n, 1, vp(Cm), n, 1, nthreads)
        return Cm
