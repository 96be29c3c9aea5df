"""ctypes binding of the C ABI (include/rstsr_hip.h) for callers that drive the
device entry points directly with preallocated buffers (benchmarks, integration
tests, other language bindings' reference)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librstsr_hip.so")


class Layout(C.Structure):
    _fields_ = [
        ("ndim", C.c_size_t),
        ("shape", C.POINTER(C.c_size_t)),
        ("stride", C.POINTER(C.c_ssize_t)),
        ("offset", C.c_size_t),
    ]


class GemmDesc(C.Structure):
    _fields_ = [
        ("m", C.c_size_t), ("n", C.c_size_t), ("k", C.c_size_t),
        ("a", C.c_void_p), ("rsa", C.c_ssize_t), ("csa", C.c_ssize_t),
        ("b", C.c_void_p), ("rsb", C.c_ssize_t), ("csb", C.c_ssize_t),
        ("c", C.c_void_p), ("rsc", C.c_ssize_t), ("csc", C.c_ssize_t),
    ]


def make_layout(shape, stride, offset=0):
    """rstsr_layout_t owning its arrays (keep the returned object alive across the call)."""
    n = len(shape)
    sh = (C.c_size_t * max(n, 1))(*[int(x) for x in shape])
    st = (C.c_ssize_t * max(n, 1))(*[int(x) for x in stride])
    l = Layout(n, sh, st, int(offset))
    l._keep = (sh, st)
    return l


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("librstsr_hip.so is not built: run `python rstsr_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    vp, P = C.c_void_p, C.POINTER(Layout)
    sig = {
        "rstsr_hip_op_mutc_refa_refb": [vp, C.c_int, C.c_int, vp, P, vp, P, vp, P],
        "rstsr_hip_assign": [vp, C.c_int, vp, P, C.c_int, vp, P],
        "rstsr_hip_op_mutc_mixed": [vp, C.c_int, C.c_int, vp, P, C.c_int, vp, P, C.c_int, vp, P],
        "rstsr_hip_assign_arbitary": [vp, C.c_int, vp, P, C.c_int, vp, P],
        "rstsr_hip_reduce_axes": [vp, C.c_int, C.c_int, vp, P, C.c_size_t, C.POINTER(C.c_ssize_t), vp, P],
        "rstsr_hip_reduce_axes_noalloc": [vp, C.c_int, C.c_int, vp, P, C.c_size_t, C.POINTER(C.c_ssize_t), vp, P],
        "rstsr_hip_argreduce_axes": [vp, C.c_int, C.c_int, vp, P, C.c_size_t, C.POINTER(C.c_ssize_t), vp, P],
        "rstsr_hip_reduce_all": [vp, C.c_int, C.c_int, vp, P, vp],
        "rstsr_hip_unary": [vp, C.c_int, C.c_int, vp, P, C.c_int, vp, P],
        "rstsr_hip_unary_inplace": [vp, C.c_int, C.c_int, vp, P],
        "rstsr_hip_unary_out_dtype": [C.c_int, C.c_int, C.POINTER(C.c_int)],
        "rstsr_hip_matmul": [vp, C.c_int, vp, P, vp, P, vp, P, vp, vp, C.c_int],
        "rstsr_hip_empty_cache": [vp],
        "rstsr_hip_cached_bytes": [vp, C.POINTER(C.c_size_t)],
        "rstsr_hip_gemm": [vp, C.c_int, C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, vp, vp, C.c_ssize_t,
                           C.c_ssize_t, vp, C.c_ssize_t, C.c_ssize_t, vp, vp, C.c_ssize_t, C.c_ssize_t],
        "rstsr_hip_driver_gemm": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_size_t, C.c_size_t,
                                  C.c_size_t, vp, vp, C.c_size_t, vp, C.c_size_t, vp, vp, C.c_size_t],
        "rstsr_hip_driver_syhemm": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_size_t,
                                    C.c_size_t, vp, vp, C.c_size_t, vp, C.c_size_t, vp, vp, C.c_size_t],
        "rstsr_hip_driver_trsm": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_size_t,
                                  C.c_size_t, vp, vp, C.c_size_t, vp, C.c_size_t],
        "rstsr_hip_memcpy_h2d_mode": [vp, vp, vp, C.c_size_t, C.c_int],
        "rstsr_hip_memcpy_d2h_mode": [vp, vp, vp, C.c_size_t, C.c_int],
        "rstsr_hip_memcpy_h2d": [vp, vp, vp, C.c_size_t],
        "rstsr_hip_memcpy_d2h": [vp, vp, vp, C.c_size_t],
        "rstsr_hip_malloc": [vp, C.c_size_t, C.POINTER(C.c_void_p)],
        "rstsr_hip_free": [vp, vp],
        "rstsr_hip_event_record": [vp, C.c_int],
        "rstsr_hip_event_elapsed": [vp, C.c_int, C.c_int, C.POINTER(C.c_float)],
        "rstsr_hip_synchronize": [vp],
        "rstsr_hip_gemm_grouped": [vp, C.c_int, C.c_int, C.c_size_t, C.POINTER(GemmDesc), vp, vp],
        "rstsr_hip_graph_begin": [vp],
        "rstsr_hip_graph_end": [vp, C.POINTER(C.c_void_p)],
        "rstsr_hip_graph_launch": [vp, vp],
        "rstsr_hip_graph_destroy": [vp],
        "rstsr_hip_stream_wait": [vp, vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.rstsr_hip_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().rstsr_hip_last_error().decode()
        raise RuntimeError(f"rstsr_hip error {rc}: {msg}")


class Events:
    """HIP events on a context stream (slots 0..15 of the context)."""

    def __init__(self, ctx):
        self.ctx = C.c_void_p(ctx)

    def record(self, slot):
        check(lib().rstsr_hip_event_record(self.ctx, slot))

    def elapsed_ms(self, s0, s1):
        ms = C.c_float()
        check(lib().rstsr_hip_event_elapsed(self.ctx, s0, s1, C.byref(ms)))
        return ms.value
