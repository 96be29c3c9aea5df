"""rstsr_amd: MI355X (gfx950) device backend for RSTSR, driven from Python.

The compute path is the C ABI library ``librstsr_hip.so`` (hand-written HIP
kernels); ``_rstsr`` is the C++ front end over it (tensor layer mirrored from
rstsr-core/src/tensor). This module only adds Python operator syntax and NumPy
conversion. Importing it fails loudly if the native modules are not built.

Names follow rstsr's API (``rt::asarray``, ``sum_axes``, ``argmax_axes``,
``to_contig``, ``%`` for matmul) so that tests read like the reference's:

    >>> import rstsr_amd as rt
    >>> dev = rt.DeviceHIP(0)
    >>> a = rt.arange(3240, device=dev).reshape([12, 15, 18]).swapaxes(-1, -2)
    >>> a[2:-3, 1:-4:2, -1:3:-2].sum_all()
    446586
"""
import numpy as np

try:
    from . import _rstsr
except ImportError as e:  # pragma: no cover - exercised only on a broken build
    raise ImportError(
        "rstsr_amd native modules are not built; run `python rstsr_amd/build.py` "
        "(needs hipcc for gfx950): " + str(e)
    ) from e

RSTSRError = _rstsr.RSTSRError
DeviceHIP = _rstsr.DeviceHIP
device_count = _rstsr.device_count

# dtype codes of include/rstsr_hip.h
F32, F64, I32, I64, C64, C128, U64, BOOL = 0, 1, 2, 3, 4, 5, 6, 7
_NP2CODE = {
    np.dtype(np.float32): F32,
    np.dtype(np.float64): F64,
    np.dtype(np.int32): I32,
    np.dtype(np.int64): I64,
    np.dtype(np.complex64): C64,
    np.dtype(np.complex128): C128,
    np.dtype(np.uint64): U64,
    np.dtype(np.bool_): BOOL,
}
_CODE2NP = {v: k for k, v in _NP2CODE.items()}

OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_REM = 0, 1, 2, 3, 4
OP_BITAND, OP_BITOR, OP_BITXOR, OP_SHL, OP_SHR = 5, 6, 7, 8, 9
RED_SUM, RED_MAX, RED_MIN, RED_ARGMAX, RED_ARGMIN = 0, 1, 2, 3, 4
RED_PROD, RED_MEAN, RED_VAR, RED_STD, RED_L2NORM = 5, 6, 7, 8, 9
RED_ALL, RED_ANY, RED_COUNT_NONZERO = 10, 11, 12
_REAL_OF = {np.dtype(np.complex64): np.dtype(np.float32), np.dtype(np.complex128): np.dtype(np.float64)}
GEMM_PLAIN, GEMM_OZAKI = 0, 1

# unary ops (rstsr_unop_t): operators/ops/op_binary_common.rs:3-39 in order, then neg and not
UNARY_OPS = (
    "acos", "acosh", "asin", "asinh", "atan", "atanh", "ceil", "conj", "cos", "cosh", "exp", "expm1",
    "floor", "inv", "is_finite", "is_inf", "is_nan", "log", "log1p", "log2", "log10", "reciprocal",
    "round", "signbit", "sin", "sinh", "square", "sqrt", "tan", "tanh", "trunc", "abs", "imag", "real",
    "sign", "neg", "not_",
)
UN = {name: i for i, name in enumerate(UNARY_OPS)}

_default_device = None


def default_device():
    global _default_device
    if _default_device is None:
        _default_device = DeviceHIP(0)
    return _default_device


def empty_cache(device=None):
    """Return the device buffers the context keeps for reuse (freed tensors) to the HIP runtime."""
    import ctypes as C
    from . import abi

    dev = device if device is not None else default_device()
    abi.check(abi.lib().rstsr_hip_empty_cache(C.c_void_p(dev._ctx_ptr())))


def cached_bytes(device=None):
    """Bytes of freed device buffers the context keeps for reuse."""
    import ctypes as C
    from . import abi

    dev = device if device is not None else default_device()
    n = C.c_size_t()
    abi.check(abi.lib().rstsr_hip_cached_bytes(C.c_void_p(dev._ctx_ptr()), C.byref(n)))
    return n.value


def _code(dtype):
    dt = np.dtype(dtype)
    if dt not in _NP2CODE:
        raise RSTSRError("UnImplemented", f"dtype {dt} is not supported by the HIP device")
    return _NP2CODE[dt]


def _axes(axes, ndim):
    if axes is None:
        return list(range(ndim))
    if isinstance(axes, (int, np.integer)):
        return [int(axes)]
    return [int(x) for x in axes]


def _mixed_pair(a, b):
    """A complex dtype and its real type, in either order."""
    a, b = np.dtype(a), np.dtype(b)
    return (a.kind == "c" and _REAL_OF.get(a) == b) or (b.kind == "c" and _REAL_OF.get(b) == a)


class Tensor:
    """A strided view of device storage (rstsr-core/src/tensorbase.rs:5-24).

    ``is_owned()`` follows the reference's ownership kinds: results of computations and copies
    own their storage; views (slices, transposes, reshapes and conversions that needed no copy)
    borrow it (TensorView / TensorCow::Borrowed)."""

    __slots__ = ("_t", "_own")

    def __init__(self, t, own=True):
        self._t = t
        self._own = own

    # ---------------- properties ----------------
    @property
    def shape(self):
        return tuple(self._t.shape)

    @property
    def stride(self):
        return tuple(self._t.stride)

    @property
    def offset(self):
        return self._t.offset

    @property
    def layout(self):
        return self._t.layout

    @property
    def ndim(self):
        return self._t.ndim

    @property
    def size(self):
        return self._t.size

    @property
    def dtype(self):
        return _CODE2NP[self._t.dtype_code]

    @property
    def device(self):
        return self._t.device

    def __repr__(self):
        return f"Tensor(shape={self.shape}, stride={self.stride}, offset={self.offset}, dtype={self.dtype})"

    def __str__(self):
        """Display (format/format_tensor.rs:204-226): the storage range the view covers is copied to the
        host and printed with the reference's layout walk and elision (MIN_PRINT 3, MAX_PRINT 8)."""
        if self.size == 0:
            buf, off = np.zeros(0, dtype=self.dtype), 0
        else:
            raw, off = self._t._raw_range()
            buf = np.frombuffer(raw, dtype=self.dtype)
        return format_tensor(buf, self.shape, self.stride, off)

    # ---------------- host transfer ----------------
    def to_numpy(self):
        """Logical contents as a NumPy array (to_vec in logical order)."""
        if self.size == 0:
            return np.zeros(self.shape, dtype=self.dtype)
        raw, off = self._t._raw_range()
        buf = np.frombuffer(raw, dtype=self.dtype)
        it = buf.itemsize
        v = np.lib.stride_tricks.as_strided(
            buf[off:], shape=self.shape, strides=[s * it for s in self.stride]
        )
        return np.array(v, copy=True)

    def to_device(self, device):
        """to_device(&device): a copy of the storage range this view covers on `device`, same layout."""
        return Tensor(self._t.to_device(device))

    def to_vec(self):
        return self.to_numpy().reshape(-1).tolist()

    def to_scalar(self):
        if self.size != 1:
            raise RSTSRError("InvalidLayout", "to_scalar needs exactly one element")
        return self.to_numpy().reshape(-1)[0].item()

    # ---------------- views (layout only) ----------------
    def _view(self, layout):
        return Tensor(self._t._with_layout(layout), own=False)

    def _order(self):
        return self._t.device.default_order

    def transpose(self, axes=None):
        """transpose / permute_dims (tensor/manipulation/transpose.rs): a view with permuted axes;
        the axes must be a permutation of 0..ndim (negative values count from the end)."""
        nd = self.ndim
        if axes is None:
            axes = list(range(nd))[::-1]
        else:
            axes = [int(a) for a in axes]
            if any(a < -nd or a >= nd for a in axes):
                raise RSTSRError("InvalidLayout", "transpose: axis out of range")
            axes = [a + nd if a < 0 else a for a in axes]
        if sorted(axes) != list(range(nd)):
            raise RSTSRError("InvalidLayout", "transpose: invalid permutation")
        sh, st, off = self.layout
        return self._view(([sh[a] for a in axes], [st[a] for a in axes], off))

    permute = transpose

    def reverse_axes(self):
        return self.transpose()

    @property
    def T(self):
        return self.transpose()

    def swapaxes(self, a, b):
        nd = self.ndim
        if not (-nd <= a < nd and -nd <= b < nd):
            raise RSTSRError("InvalidLayout", f"swapaxes: axes ({a}, {b}) out of range for {nd} dimensions")
        a, b = a % nd, b % nd
        p = list(range(nd))
        p[a], p[b] = p[b], p[a]
        return self.transpose(p)

    def flip(self, axis):
        return self._view(_rstsr.dim_narrow(self.layout, axis, None, None, -1))

    def slice(self, axis, start=None, stop=None, step=None):
        return self._view(_rstsr.dim_narrow(self.layout, axis, start, stop, 1 if step is None else step))

    def i(self, idx):
        """rstsr's .i(...) indexing: same as [...]."""
        return self[idx]

    def __getitem__(self, idx):
        if not isinstance(idx, tuple):
            idx = (idx,)
        if any(i is Ellipsis for i in idx):
            k = idx.index(Ellipsis)
            fill = self.ndim - (len(idx) - 1)
            idx = idx[:k] + (slice(None),) * fill + idx[k + 1:]
        lay = self.layout
        ax = 0
        drop = []
        for i in idx:
            if isinstance(i, slice):
                if not (i.start is None and i.stop is None and i.step in (None, 1)):
                    lay = _rstsr.dim_narrow(lay, ax, i.start, i.stop, 1 if i.step is None else i.step)
                ax += 1
            elif isinstance(i, (int, np.integer)):
                n = lay[0][ax]
                j = int(i) + n if i < 0 else int(i)
                if not 0 <= j < n:
                    raise RSTSRError("IndexError", f"index {i} out of range for axis {ax} of size {n}")
                lay = _rstsr.dim_narrow(lay, ax, j, j + 1, 1)
                drop.append(ax)
                ax += 1
            else:
                raise RSTSRError("UnImplemented", "only int / slice / ... indices are supported")
        sh, st, off = lay
        sh = [s for k, s in enumerate(sh) if k not in drop]
        st = [s for k, s in enumerate(st) if k not in drop]
        return self._view((sh, st, off))

    def _reshape(self, shape, order=None):
        """(storage view or copy, new layout, copied?) of a reshape in `order` (default: the device's
        default order, tensor/manipulation/reshape.rs:126): elements pair up in C order for RowMajor
        and F order for ColMajor; a view when the layout is contiguous in that order, else a copy
        into a contiguous layout first."""
        shape = [int(x) for x in shape]
        if shape.count(-1) > 1:
            raise RSTSRError("InvalidLayout", "reshape: at most one -1")
        if -1 in shape:
            known = int(np.prod([x for x in shape if x != -1]))
            shape[shape.index(-1)] = self.size // max(known, 1)
        if int(np.prod(shape)) != self.size or any(x < 0 for x in shape):
            raise RSTSRError("InvalidLayout", "reshape: element counts differ")
        order = order or self._order()
        fo = order in ("F", "ColMajor")
        contig = _rstsr.f_contig(self.layout) if fo else _rstsr.c_contig(self.layout)
        src, copied = (self, False) if contig else (self.to_contig("F" if fo else "C"), True)
        st = (_rstsr.f_layout(shape) if fo else _rstsr.c_layout(shape))[1]
        return src, (shape, st, src.offset), copied

    def reshape(self, shape, order=None):
        """reshape (borrowing): a view when no copy is needed, else an owned copy."""
        src, lay, copied = self._reshape(shape, order)
        return Tensor(src._t._with_layout(lay), own=copied)

    def into_shape(self, shape, order=None):
        """into_shape (consuming): the result owns its storage when this tensor did, or when copied."""
        src, lay, copied = self._reshape(shape, order)
        return Tensor(src._t._with_layout(lay), own=copied or self._own)

    def broadcast_to(self, shape):
        """broadcast_to (tensor/manipulation/broadcast.rs:281): a view broadcast to `shape` under
        the device's default order (ColMajor aligns the leading axes)."""
        lay = _rstsr.broadcast_layout_to_shape(self.layout, [int(x) for x in shape], self._order())
        return self._view(lay)

    # ---------------- contiguity (layout/layoutbase.rs) ----------------
    def c_contig(self):
        return _rstsr.c_contig(self.layout)

    def f_contig(self):
        return _rstsr.f_contig(self.layout)

    def c_prefer(self):
        return _rstsr.c_prefer(self.layout)

    def f_prefer(self):
        return _rstsr.f_prefer(self.layout)

    def is_owned(self):
        return self._own

    def as_ptr(self):
        """Address of the storage (raw().as_ptr()): equal for views of one storage."""
        return self._t._data_ptr()

    def view(self):
        return Tensor(self._t, own=False)

    def view_owned(self):
        return Tensor(self._t, own=True)

    def to_owned(self):
        """An owned copy laid out as layout_for_array_copy(self, K)."""
        return Tensor(self._t.to_layout(_rstsr.layout_for_array_copy(self.layout, "K")))

    def iter(self):
        """Elements in the device's default order (iterator_elem.rs:74-80: C for RowMajor, F for
        ColMajor)."""
        a = self.to_numpy()
        return iter((a.T if self._order() == "ColMajor" else a).reshape(-1).tolist())

    # ---------------- copies ----------------
    def _cow(self, t):
        """TensorCow: borrowed when the conversion returned this very storage and layout."""
        same = t._data_ptr() == self._t._data_ptr() and t.layout == self.layout
        return Tensor(t, own=not same)

    def to_contig(self, order="C"):
        """to_contig (tensor/manipulation/to_contig.rs): a view when the layout already equals the
        contiguous layout of `order` (offset 0), else an owned contiguous copy."""
        return self._cow(self._t.to_contig(order))

    change_contig = to_contig
    to_contig_f = to_contig

    def into_contig(self, order="C"):
        r = self.to_contig(order)
        return Tensor(r._t, own=True)

    def to_prefer(self, order="C"):
        """to_prefer (to_contig.rs:229-243): a view when the layout is c_prefer (RowMajor) /
        f_prefer (ColMajor), else a contiguous copy in that order."""
        fo = order in ("F", "ColMajor")
        if (fo and self.f_prefer()) or (not fo and self.c_prefer()):
            return Tensor(self._t, own=False)
        return Tensor(self._t.to_contig("F" if fo else "C"), own=True)

    def to_layout(self, layout):
        return self._cow(self._t.to_layout(layout))

    def astype(self, dtype):
        return Tensor(self._t.astype(_code(dtype)))

    def assign(self, other):
        self._t.assign(other._t)

    def fill(self, value):
        self._t.fill(np.array(value, dtype=self.dtype).tobytes())

    # ---------------- arithmetic ----------------
    def _bin(self, op, other, left=False):
        if isinstance(other, Tensor):
            x, y = (other, self) if left else (self, other)
            if np.dtype(x.dtype) != np.dtype(y.dtype) and _mixed_pair(x.dtype, y.dtype):
                return Tensor(x._t.binary_mixed(op, y._t))  # Complex<T> op T / T op Complex<T>
            return Tensor(x._t.binary(op, y._t))
        dt = np.dtype(self.dtype)
        if dt.kind == "c" and isinstance(other, (float, int, np.floating, np.integer)) and op <= OP_DIV:
            # a real scalar with a complex tensor: num-complex's Complex<T> op T
            rt_ = _REAL_OF[dt]
            return Tensor(self._t.binary_scalar_mixed(op, np.array(other, dtype=rt_).tobytes(), _code(rt_), left))
        if dt.kind == "f" and isinstance(other, (complex, np.complexfloating)) and op <= OP_DIV:
            ct = np.dtype(np.complex128 if dt == np.float64 else np.complex64)
            return Tensor(self._t.binary_scalar_mixed(op, np.array(other, dtype=ct).tobytes(), _code(ct), left))
        s = np.array(other, dtype=self.dtype).tobytes()
        return Tensor(self._t.binary_scalar(op, s, left))

    def __add__(self, o):
        return self._bin(OP_ADD, o)

    def __radd__(self, o):
        return self._bin(OP_ADD, o, left=True)

    def __sub__(self, o):
        return self._bin(OP_SUB, o)

    def __rsub__(self, o):
        return self._bin(OP_SUB, o, left=True)

    def __mul__(self, o):
        return self._bin(OP_MUL, o)

    def __rmul__(self, o):
        return self._bin(OP_MUL, o, left=True)

    def __truediv__(self, o):
        return self._bin(OP_DIV, o)

    def __rtruediv__(self, o):
        return self._bin(OP_DIV, o, left=True)

    def _ibin(self, op, o):
        if isinstance(o, Tensor):
            self._t.binary_inplace(op, o._t)
        else:
            self._t.binary_inplace_scalar(op, np.array(o, dtype=self.dtype).tobytes())
        return self

    def __iadd__(self, o):
        return self._ibin(OP_ADD, o)

    def __isub__(self, o):
        return self._ibin(OP_SUB, o)

    def __imul__(self, o):
        return self._ibin(OP_MUL, o)

    def __itruediv__(self, o):
        return self._ibin(OP_DIV, o)

    # bit operators (integers and bool) and shifts (integers): operators/ops/op_ternary_arithmetic.rs
    def __and__(self, o):
        return self._bin(OP_BITAND, o)

    def __rand__(self, o):
        return self._bin(OP_BITAND, o, left=True)

    def __or__(self, o):
        return self._bin(OP_BITOR, o)

    def __ror__(self, o):
        return self._bin(OP_BITOR, o, left=True)

    def __xor__(self, o):
        return self._bin(OP_BITXOR, o)

    def __rxor__(self, o):
        return self._bin(OP_BITXOR, o, left=True)

    def __lshift__(self, o):
        return self._bin(OP_SHL, o)

    def __rlshift__(self, o):
        return self._bin(OP_SHL, o, left=True)

    def __rshift__(self, o):
        return self._bin(OP_SHR, o)

    def __rrshift__(self, o):
        return self._bin(OP_SHR, o, left=True)

    def __iand__(self, o):
        return self._ibin(OP_BITAND, o)

    def __ior__(self, o):
        return self._ibin(OP_BITOR, o)

    def __ixor__(self, o):
        return self._ibin(OP_BITXOR, o)

    def __ilshift__(self, o):
        return self._ibin(OP_SHL, o)

    def __irshift__(self, o):
        return self._ibin(OP_SHR, o)

    def rem(self, o):
        """Elementwise remainder (rt::rem; `%` is matmul in rstsr)."""
        return self._bin(OP_REM, o)

    def rem_(self, o):
        return self._ibin(OP_REM, o)

    def __mod__(self, o):
        return matmul(self, o)

    def __matmul__(self, o):
        return matmul(self, o)

    # ---------------- reductions ----------------
    def _red_all(self, op):
        raw = self._t.reduce_all(op)
        if op in (RED_ARGMAX, RED_ARGMIN, RED_COUNT_NONZERO) or (op == RED_SUM and np.dtype(self.dtype) == np.bool_):
            return int(np.frombuffer(raw, dtype=np.uint64)[0])
        if op in (RED_ALL, RED_ANY):
            return bool(np.frombuffer(raw, dtype=np.uint8)[0])
        dt = self.dtype
        if op in (RED_VAR, RED_STD, RED_L2NORM):
            dt = _REAL_OF.get(np.dtype(dt), dt)
        return np.frombuffer(raw, dtype=dt)[0].item()

    def _red_axes(self, op, axes):
        return Tensor(self._t.reduce_axes(op, _axes(axes, self.ndim)))

    def sum_all(self):
        return self._red_all(RED_SUM)

    def max_all(self):
        return self._red_all(RED_MAX)

    def min_all(self):
        return self._red_all(RED_MIN)

    def argmax_all(self):
        return self._red_all(RED_ARGMAX)

    def argmin_all(self):
        return self._red_all(RED_ARGMIN)

    def sum_axes(self, axes=None):
        return self._red_axes(RED_SUM, axes)

    def max_axes(self, axes=None):
        return self._red_axes(RED_MAX, axes)

    def min_axes(self, axes=None):
        return self._red_axes(RED_MIN, axes)

    def argmax_axes(self, axes=None):
        return self._red_axes(RED_ARGMAX, axes)

    def argmin_axes(self, axes=None):
        return self._red_axes(RED_ARGMIN, axes)

    # tensor/reduction.rs:122-126 (prod, mean, var, std, l2_norm); var/std/l2_norm of
    # complex tensors are real
    def prod_all(self):
        return self._red_all(RED_PROD)

    def mean_all(self):
        return self._red_all(RED_MEAN)

    def var_all(self):
        return self._red_all(RED_VAR)

    def std_all(self):
        return self._red_all(RED_STD)

    def l2_norm_all(self):
        return self._red_all(RED_L2NORM)

    def prod_axes(self, axes=None):
        return self._red_axes(RED_PROD, axes)

    def mean_axes(self, axes=None):
        return self._red_axes(RED_MEAN, axes)

    def var_axes(self, axes=None):
        return self._red_axes(RED_VAR, axes)

    def std_axes(self, axes=None):
        return self._red_axes(RED_STD, axes)

    def l2_norm_axes(self, axes=None):
        return self._red_axes(RED_L2NORM, axes)

    # bool folds and counts (operators/reduction.rs: OpAllAPI, OpAnyAPI, OpCountNonZeroAPI); the
    # sum of a bool array counts its true elements (OpSumBoolAPI)
    def all_all(self):
        return self._red_all(RED_ALL)

    def any_all(self):
        return self._red_all(RED_ANY)

    def count_nonzero_all(self):
        return self._red_all(RED_COUNT_NONZERO)

    def all_axes(self, axes=None):
        return self._red_axes(RED_ALL, axes)

    def any_axes(self, axes=None):
        return self._red_axes(RED_ANY, axes)

    def count_nonzero_axes(self, axes=None):
        return self._red_axes(RED_COUNT_NONZERO, axes)

    sum = sum_axes

    # ---------------- unary (tensor/operators/op_unary_common.rs:96-190) ----------------
    def unary(self, op):
        """op(&tensor): a new tensor laid out as layout_for_array_copy(self, K)."""
        return Tensor(self._t.unary(UN[op] if isinstance(op, str) else int(op)))

    def unary_to(self, op, out):
        """op_muta_refb into an existing output view `out` of the same shape."""
        self._t.unary_to(UN[op] if isinstance(op, str) else int(op), out._t)
        return out

    def unary_(self, op):
        """op(tensor) on an owned tensor: in place (op_muta), same-dtype ops only."""
        self._t.unary_inplace(UN[op] if isinstance(op, str) else int(op))
        return self

    def __neg__(self):
        return self.unary("neg")

    def __invert__(self):
        return self.unary("not_")

    def __abs__(self):
        return self.unary("abs")


def _unary_method(name):
    def f(self):
        return self.unary(name)

    f.__name__ = name
    f.__doc__ = f"rt::{name.rstrip('_')}(&tensor)"
    return f


def _unary_func(name):
    def f(x):
        return x.unary(name)

    f.__name__ = name
    f.__doc__ = f"rt::{name.rstrip('_')}(&tensor)"
    return f


for _name in UNARY_OPS:
    setattr(Tensor, _name, _unary_method(_name))
    globals()[_name] = _unary_func(_name)
del _name


MIN_PRINT, MAX_PRINT = 3, 8


def _fmt_float(x):
    """Rust's Display of f32/f64: shortest round-trip digits in positional notation."""
    if np.isnan(x):
        return "NaN"
    if np.isinf(x):
        return "inf" if x > 0 else "-inf"
    return np.format_float_positional(x, unique=True, trim="-")


def _fmt_elem(x):
    if isinstance(x, (np.bool_, bool)):
        return "true" if x else "false"
    if isinstance(x, np.complexfloating):
        re, im = x.real, x.imag
        # num-complex Display: re, then '-' |im| when im < 0, else '+' im, then 'i'
        return _fmt_float(re) + ("-" + _fmt_float(-im) if im < 0 else "+" + _fmt_float(im)) + "i"
    if isinstance(x, np.floating):
        return _fmt_float(x)
    return str(int(x))


def format_tensor(vec, shape, stride, offset, max_print=None, min_print=None, fmt_elem=_fmt_elem):
    """The layout walk of format/format_tensor.rs:25-190 (print_vec_with_layout_dfs_inner) over a
    host buffer `vec` with the view's (shape, stride, offset) in elements; the recursion there is a
    tail call in every branch, so it runs here as a loop."""
    max_print = MAX_PRINT if max_print is None else max_print
    min_print = MIN_PRINT if min_print is None else min_print
    ndim = len(shape)
    out = []
    if ndim == 0:
        return fmt_elem(vec[offset])
    idx, off = [], offset
    while True:
        lp = len(idx)
        if idx and idx[-1] == shape[lp - 1]:
            # hit the end of an axis
            if idx[-1] == 0:
                out.append("[]")
                break
            if lp == 1:
                break
            p1 = idx.pop()
            p2 = idx.pop()
            off = off - p1 * stride[lp - 1] - p2 * stride[lp - 2]
            if shape[lp - 2] < max_print or p2 + min_print >= shape[lp - 2] or p2 + 1 < min_print:
                q = p2 + 1
            else:
                out.append(" " * (lp - 1) + "...\n\n")
                q = shape[lp - 2] - min_print
            idx.append(q)
            off += q * stride[lp - 2]
            continue
        if lp + 1 != ndim:
            idx.append(0)
            continue
        # the innermost axis: one printed line
        nlast, slast = shape[-1], stride[-1]
        if nlast == 0:
            out.append("[]")
            break
        if not idx:
            out.append("[")
        else:
            nbra = 1
            for v in reversed(idx):
                if v != 0:
                    break
                nbra += 1
            out.append(" " * (ndim - nbra) + "[" * nbra)
        if nlast <= max(max_print, 2 * min_print + 1):
            pos = range(nlast)
            out.extend(" " + fmt_elem(vec[off + i * slast]) for i in pos)
        else:
            out.extend(" " + fmt_elem(vec[off + i * slast]) for i in range(min_print))
            out.append(" ...")
            out.extend(" " + fmt_elem(vec[off + i * slast]) for i in range(nlast - min_print, nlast))
        if not idx:
            out.append("]")
        else:
            nket = 1
            for i in range(len(idx) - 1, -1, -1):
                if idx[i] != shape[i] - 1:
                    break
                nket += 1
            out.append("]" * nket)
            if nket != ndim:
                out.append("\n\n" if nket > 1 else "\n")
        if lp == 0:
            break
        p1 = idx.pop()
        off -= p1 * stride[lp - 1]
        if shape[lp - 1] < max_print or p1 + min_print >= shape[lp - 1] or p1 + 1 < min_print:
            q = p1 + 1
        else:
            out.append(" " * lp + "...\n")
            q = shape[lp - 1] - min_print
        idx.append(q)
        off += q * stride[lp - 1]
    return "".join(out)


def _is_f(dev):
    return dev.default_order == "ColMajor"


def asarray(x, device=None, dtype=None, shape=None):
    """rt::asarray: host data uploaded with the device's default order (tensor/asarray.rs:459-470):
    a c-contiguous layout for RowMajor, an f-contiguous one for ColMajor, the same logical
    values either way. With `shape`, x is a flat vector laid out in that order, as
    rt::asarray((vec, shape, &device))."""
    dev = device or default_device()
    a = np.asarray(x, dtype=dtype)
    if shape is not None:
        shape = [int(n) for n in shape]
        a = a.reshape(-1).reshape(shape, order="F" if _is_f(dev) else "C")
    if _is_f(dev) and a.ndim > 1:
        # F-contiguous data of shape S is c-contiguous data of reversed(S) with its axes reversed
        at = np.ascontiguousarray(a.T)
        return Tensor(_rstsr.from_host(dev, at.tobytes(), list(at.shape), _code(at.dtype))).transpose().view_owned()
    a = np.asarray(a, order="C")  # (ascontiguousarray would turn a 0-d array into 1-d)
    return Tensor(_rstsr.from_host(dev, a.tobytes(), list(a.shape), _code(a.dtype)))


def tensor_from_nested(x, device=None, dtype=None):
    """rt::tensor_from_nested!(...): a nested list as a tensor of the device."""
    return asarray(x, device=device, dtype=dtype)


def empty(shape, dtype=np.float64, device=None):
    dev = device or default_device()
    shape = [int(n) for n in shape]
    lay = _rstsr.f_layout(shape) if _is_f(dev) else _rstsr.c_layout(shape)
    return Tensor(_rstsr.empty(dev, lay, _code(dtype)))


def full(shape, value, dtype=np.float64, device=None):
    t = empty(shape, dtype, device)
    t.fill(value)
    return t


def zeros(shape, dtype=np.float64, device=None):
    return full(shape, 0, dtype, device)


def ones(shape, dtype=np.float64, device=None):
    return full(shape, 1, dtype, device)


def arange(n, device=None, dtype=None):
    """rt::arange(n): integers 0..n (i64 by default, like usize/isize), or floats."""
    if dtype is None:
        dtype = np.float64 if isinstance(n, float) else np.int64
    return asarray(np.arange(n, dtype=dtype), device=device)


def linspace(start, stop, num, device=None):
    """rt::linspace((start, stop, n)) as NumPy computes it."""
    return asarray(np.linspace(start, stop, num), device=device)


def transpose(a, axes=None):
    return a.transpose(axes)


permute_dims = transpose


def swapaxes(a, i, j):
    return a.swapaxes(i, j)


def to_contig(a, order="C"):
    return a.to_contig(order)


def to_prefer(a, order="C"):
    return a.to_prefer(order)


def broadcast_to(a, shape):
    return a.broadcast_to(shape)


def broadcast_arrays(tensors):
    """rt::broadcast_arrays (tensor/manipulation/broadcast.rs:240-265): views of every tensor at
    the common broadcast shape, under the first tensor's device order."""
    if not tensors:
        return []
    order = tensors[0]._order()
    lay = tensors[0].layout
    for t in tensors[1:]:
        lay = _rstsr.broadcast_layout(lay, t.layout, order)[0]
    shape = list(lay[0])
    return [t._view(_rstsr.broadcast_layout_to_shape(t.layout, shape, order)) for t in tensors]


def add(a, b):
    return a + b


def sub(a, b):
    return a - b


def mul(a, b):
    return a * b


def div(a, b):
    return a / b


def rem(a, b):
    """rt::rem: elementwise remainder (Rust's %: truncated for integers, fmod for floats)."""
    if isinstance(a, Tensor):
        return a._bin(OP_REM, b)
    return b._bin(OP_REM, a, left=True)


def bitand(a, b):
    return a & b


def bitor(a, b):
    return a | b


def bitxor(a, b):
    return a ^ b


def shl(a, b):
    return a << b


def shr(a, b):
    return a >> b


def matmul(a, b, mode="plain"):
    """`a % b` with rstsr's row-major matmul rules; mode 'plain' (FP64 MFMA) or 'ozaki' (INT8 split)."""
    m = {"plain": GEMM_PLAIN, "ozaki": GEMM_OZAKI}[mode]
    return Tensor(a._t.matmul(b._t, m))


def allclose(a, b, rtol=1e-5, atol=1e-8, equal_nan=False):
    return a._t.allclose(b._t, rtol, atol, equal_nan)


def gemm_grouped(problems, mode="plain", alpha=1.0, beta=0.0):
    """C_g = alpha * A_g @ B_g + beta * C_g for a list of 2-D (A, B, C) tensors of uneven shapes.

    Plain f64 / c128 run as one persistent launch with autotuned tiles."""
    import ctypes as C
    from . import abi

    if not problems:
        return
    dt = problems[0][0].dtype
    isz = np.dtype(dt).itemsize
    descs = (abi.GemmDesc * len(problems))()
    for i, (a, b, c) in enumerate(problems):
        # one layout read per operand (shape, stride, offset): a call builds hundreds of these
        (sa, ta, oa), (sb, tb, ob), (sc, tc, oc) = a._t.layout, b._t.layout, c._t.layout
        if len(sa) != 2 or len(sb) != 2 or len(sc) != 2:
            raise RSTSRError("InvalidLayout", "gemm_grouped takes 2-D operands")
        m, k = sa
        k2, n = sb
        if k != k2 or sc[0] != m or sc[1] != n:
            raise RSTSRError("InvalidLayout", "gemm_grouped shape mismatch")
        descs[i] = abi.GemmDesc(m, n, k, a._t._data_ptr() + oa * isz, ta[0], ta[1], b._t._data_ptr() + ob * isz,
                                tb[0], tb[1], c._t._data_ptr() + oc * isz, tc[0], tc[1])
    al = np.array(alpha, dtype=dt)
    be = np.array(beta, dtype=dt)
    ctx = C.c_void_p(problems[0][0].device._ctx_ptr())
    m = {"plain": GEMM_PLAIN, "ozaki": GEMM_OZAKI}[mode]
    abi.check(abi.lib().rstsr_hip_gemm_grouped(ctx, _code(dt), m, len(problems), descs,
                                               al.ctypes.data_as(C.c_void_p), be.ctypes.data_as(C.c_void_p)))


class capture:
    """Record device work on a DeviceHIP's stream into a hipGraph.

        with rt.capture(dev) as g:
            z1 += v; z3 = z1 * z2; s = (z3 % z2).sum_axes([1])
        g.launch(); dev.synchronize()

    Tensors created inside the block must stay alive while the graph is used."""

    def __init__(self, device):
        import ctypes as C

        self.device = device
        self.ctx = C.c_void_p(device._ctx_ptr())
        self.exec = None

    def __enter__(self):
        from . import abi

        abi.check(abi.lib().rstsr_hip_graph_begin(self.ctx))
        return self

    def __exit__(self, et, ev, tb):
        import ctypes as C
        from . import abi

        g = C.c_void_p()
        rc = abi.lib().rstsr_hip_graph_end(self.ctx, C.byref(g))
        if et is None:
            abi.check(rc)
            self.exec = g
        elif rc == 0 and g:
            # the block raised: nobody will launch this graph; free it and what it holds
            abi.lib().rstsr_hip_graph_destroy(g)
        return False

    def launch(self):
        from . import abi

        abi.check(abi.lib().rstsr_hip_graph_launch(self.ctx, self.exec))

    def __del__(self):
        try:
            if self.exec:
                from . import abi

                abi.lib().rstsr_hip_graph_destroy(self.exec)
        except Exception:
            pass


def add(a, b):
    return a + b


def mul(a, b):
    return a * b
