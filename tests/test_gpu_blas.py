"""BLAS-style drivers (rstsr_hip_driver_gemm / rstsr_hip_driver_syhemm / rstsr_hip_driver_trsm) against NumPy.

Bars:
  * vs NumPy in extended precision: componentwise |C - C*| <= (K + 2) u (|alpha| |op(A)| |op(B)| + |beta| |C0|),
    u = 2^-53 (f64, c128) or 2^-24 (f32, c64) -- the classical bound of any summation order;
  * vs rstsr_hip_gemm on the same strides (with the conjugated or expanded operand prepared by
    NumPy): bit-identical, since the driver only maps flags to strides and prepares that operand.
"""
import ctypes as C

import numpy as np
import pytest

import rstsr_amd as rt
from rstsr_amd import abi

pytestmark = pytest.mark.gpu

ORD_C, ORD_F = 101, 102
TR = {"N": 111, "T": 112, "C": 113, "CN": 114}
UPLO = {"U": 121, "L": 122}
SIDE = {"L": 141, "R": 142}
CODE = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.complex64): 4, np.dtype(np.complex128): 5}
U = {np.dtype(np.float32): 2.0 ** -24, np.dtype(np.float64): 2.0 ** -53,
     np.dtype(np.complex64): 2.0 ** -24, np.dtype(np.complex128): 2.0 ** -53}


def rnd(rng, shape, dt):
    x = rng.normal(size=shape)
    if np.dtype(dt).kind == "c":
        x = x + 1j * rng.normal(size=shape)
    return x.astype(dt)


def store(S, order, ld):
    """Flat BLAS storage of matrix S (rows x cols) with leading dimension ld, padded with NaN."""
    r, c = S.shape
    if order == ORD_C:
        buf = np.full((r, ld), np.nan, dtype=S.dtype)
        buf[:, :c] = S
    else:
        buf = np.full((c, ld), np.nan, dtype=S.dtype)
        buf[:, :r] = S.T
    return buf.reshape(-1)


def load(flat, shape, order, ld):
    r, c = shape
    if order == ORD_C:
        return flat.reshape(r, ld)[:, :c]
    return flat.reshape(c, ld)[:, :r].T


def op(S, t):
    return {"N": S, "T": S.T, "C": S.conj().T, "CN": S.conj()}[t]


def scal(dt, v):
    return np.array([v], dtype=dt)


def ptr(t):
    return C.c_void_p(t._t._data_ptr())


def within(got, ref, bound):
    return np.all(np.abs(got - ref) <= bound + 1e-300)


@pytest.mark.parametrize("dt", [np.float64, np.complex128, np.float32, np.complex64])
@pytest.mark.parametrize("order", [ORD_C, ORD_F])
def test_driver_gemm_all_flags(dev, dt, order):
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    rng = np.random.default_rng(7)
    m, n, k = 37, 45, 70
    trs = ["N", "T", "C", "CN"]
    alpha, beta = (1.5 - 0.5j, -0.75 + 0.25j) if np.dtype(dt).kind == "c" else (1.5, -0.75)
    al, be = scal(dt, alpha), scal(dt, beta)
    for ta in trs:
        for tb in trs:
            SA = rnd(rng, (k, m) if ta in ("T", "C") else (m, k), dt)
            SB = rnd(rng, (n, k) if tb in ("T", "C") else (k, n), dt)
            C0 = rnd(rng, (m, n), dt)
            lda = (SA.shape[1] if order == ORD_C else SA.shape[0]) + 3
            ldb = (SB.shape[1] if order == ORD_C else SB.shape[0]) + 5
            ldc = (n if order == ORD_C else m) + 2
            da = rt.asarray(store(SA, order, lda), device=dev)
            db = rt.asarray(store(SB, order, ldb), device=dev)
            dc = rt.asarray(store(C0, order, ldc), device=dev)
            r = L.rstsr_hip_driver_gemm(ctx, CODE[np.dtype(dt)], 0, order, TR[ta], TR[tb], m, n, k,
                                        al.ctypes.data_as(C.c_void_p), ptr(da), lda, ptr(db), ldb,
                                        be.ctypes.data_as(C.c_void_p), ptr(dc), ldc)
            assert r == 0, L.rstsr_hip_last_error()
            got = load(dc.to_numpy(), (m, n), order, ldc)
            oa, ob = op(SA, ta).astype(np.complex128 if np.dtype(dt).kind == "c" else np.float64), op(SB, tb)
            ref = alpha * (oa @ ob.astype(oa.dtype)) + beta * C0.astype(oa.dtype)
            bound = (k + 2) * U[np.dtype(dt)] * (abs(alpha) * (np.abs(oa) @ np.abs(ob)) + abs(beta) * np.abs(C0))
            assert within(got, ref, bound), (ta, tb, float(np.max(np.abs(got - ref) / np.maximum(bound, 1e-300))))
            # padding of C (beyond ldc's logical columns / rows) untouched
            flat = dc.to_numpy()
            pad = np.isnan(store(np.zeros((m, n), dt), order, ldc))
            assert np.all(np.isnan(flat[pad])), (ta, tb)


@pytest.mark.parametrize("dt", [np.float64, np.complex128])
def test_driver_gemm_bit_identical_to_strided_gemm(dev, dt):
    """The driver equals rstsr_hip_gemm on the stored strides (conjugated operands prepared by NumPy)."""
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    rng = np.random.default_rng(11)
    m, n, k = 96, 80, 112
    al, be = scal(dt, 1.0), scal(dt, 0.0)
    for ta, tb in [("N", "N"), ("T", "N"), ("C", "T"), ("CN", "C")]:
        SA = rnd(rng, (k, m) if ta in ("T", "C") else (m, k), dt)
        SB = rnd(rng, (n, k) if tb in ("T", "C") else (k, n), dt)
        da = rt.asarray(SA, device=dev)
        db = rt.asarray(SB, device=dev)
        dc = rt.empty((m, n), dtype=dt, device=dev)
        r = L.rstsr_hip_driver_gemm(ctx, CODE[np.dtype(dt)], 0, ORD_C, TR[ta], TR[tb], m, n, k,
                                    al.ctypes.data_as(C.c_void_p), ptr(da), SA.shape[1], ptr(db), SB.shape[1],
                                    be.ctypes.data_as(C.c_void_p), ptr(dc), n)
        assert r == 0, L.rstsr_hip_last_error()
        # same product through the strided GEMM: op() as strides, conjugation done on the host
        cplx = np.dtype(dt).kind == "c"
        oa = np.ascontiguousarray(op(SA, ta)) if (cplx and ta in ("C", "CN")) else None
        ob = np.ascontiguousarray(op(SB, tb)) if (cplx and tb in ("C", "CN")) else None
        ta2 = rt.asarray(oa, device=dev) if oa is not None else da
        tb2 = rt.asarray(ob, device=dev) if ob is not None else db
        rsa, csa = (k, 1) if oa is not None else ((SA.shape[1], 1) if ta in ("N", "CN") else (1, SA.shape[1]))
        rsb, csb = (n, 1) if ob is not None else ((SB.shape[1], 1) if tb in ("N", "CN") else (1, SB.shape[1]))
        d2 = rt.empty((m, n), dtype=dt, device=dev)
        r = L.rstsr_hip_gemm(ctx, CODE[np.dtype(dt)], 0, m, n, k, al.ctypes.data_as(C.c_void_p), ptr(ta2), rsa, csa,
                             ptr(tb2), rsb, csb, be.ctypes.data_as(C.c_void_p), ptr(d2), n, 1)
        assert r == 0, L.rstsr_hip_last_error()
        g1, g2 = dc.to_numpy(), d2.to_numpy()
        assert np.array_equal(g1.view(np.uint8), g2.view(np.uint8)), (ta, tb)


def expand(S, uplo, hermi):
    n = S.shape[0]
    low = np.tril(S) if uplo == "L" else np.triu(S)
    other = (np.tril(S, -1) if uplo == "L" else np.triu(S, 1))
    full = low + (other.conj().T if hermi else other.T)
    if hermi:
        full[np.arange(n), np.arange(n)] = full[np.arange(n), np.arange(n)].real
    return full


@pytest.mark.parametrize("dt,hermi", [(np.float64, 0), (np.complex128, 0), (np.complex128, 1), (np.complex64, 1),
                                      (np.float32, 1)])
@pytest.mark.parametrize("order", [ORD_C, ORD_F])
def test_driver_syhemm(dev, dt, hermi, order):
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    rng = np.random.default_rng(3 + hermi)
    m, n = 41, 29
    alpha, beta = (0.5 + 1j, 2.0 - 0.5j) if np.dtype(dt).kind == "c" else (0.5, 2.0)
    al, be = scal(dt, alpha), scal(dt, beta)
    for side in ("L", "R"):
        for uplo in ("U", "L"):
            na = m if side == "L" else n
            SA = rnd(rng, (na, na), dt)  # the triangle not named by uplo is garbage that must not be read
            B = rnd(rng, (m, n), dt)
            C0 = rnd(rng, (m, n), dt)
            lda, ldb, ldc = na + 1, (n if order == ORD_C else m) + 4, (n if order == ORD_C else m) + 2
            da = rt.asarray(store(SA, order, lda), device=dev)
            db = rt.asarray(store(B, order, ldb), device=dev)
            dc = rt.asarray(store(C0, order, ldc), device=dev)
            r = L.rstsr_hip_driver_syhemm(ctx, CODE[np.dtype(dt)], 0, hermi, order, SIDE[side], UPLO[uplo], m, n,
                                          al.ctypes.data_as(C.c_void_p), ptr(da), lda, ptr(db), ldb,
                                          be.ctypes.data_as(C.c_void_p), ptr(dc), ldc)
            assert r == 0, L.rstsr_hip_last_error()
            got = load(dc.to_numpy(), (m, n), order, ldc)
            hi = np.complex128 if np.dtype(dt).kind == "c" else np.float64
            Af = expand(SA.astype(hi), uplo, bool(hermi) and np.dtype(dt).kind == "c")
            prod = Af @ B.astype(hi) if side == "L" else B.astype(hi) @ Af
            absp = np.abs(Af) @ np.abs(B) if side == "L" else np.abs(B) @ np.abs(Af)
            ref = alpha * prod + beta * C0.astype(hi)
            bound = (na + 2) * U[np.dtype(dt)] * (abs(alpha) * absp + abs(beta) * np.abs(C0))
            assert within(got, ref, bound), (side, uplo)


def test_driver_errors(dev):
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    x = rt.asarray(np.ones(64), device=dev)
    one = scal(np.float64, 1.0)
    p1 = one.ctypes.data_as(C.c_void_p)
    # lda smaller than the stored row length
    assert L.rstsr_hip_driver_gemm(ctx, 1, 0, ORD_C, 111, 111, 4, 4, 4, p1, ptr(x), 3, ptr(x), 4, p1, ptr(x),
                                   4) != 0
    # bad flags
    assert L.rstsr_hip_driver_gemm(ctx, 1, 0, 100, 111, 111, 4, 4, 4, p1, ptr(x), 4, ptr(x), 4, p1, ptr(x), 4) != 0
    assert L.rstsr_hip_driver_gemm(ctx, 1, 0, ORD_C, 115, 111, 4, 4, 4, p1, ptr(x), 4, ptr(x), 4, p1, ptr(x), 4) != 0
    # an operand reaching past its buffer (64 elements; 9 x 9 needed)
    assert L.rstsr_hip_driver_gemm(ctx, 1, 0, ORD_C, 111, 111, 9, 9, 9, p1, ptr(x), 9, ptr(x), 9, p1, ptr(x),
                                   9) != 0
    assert L.rstsr_hip_driver_syhemm(ctx, 1, 0, 0, ORD_C, 143, 121, 4, 4, p1, ptr(x), 4, ptr(x), 4, p1, ptr(x),
                                     4) != 0
    # integer dtypes are not BLAS types
    assert L.rstsr_hip_driver_gemm(ctx, 2, 0, ORD_C, 111, 111, 2, 2, 2, p1, ptr(x), 2, ptr(x), 2, p1, ptr(x), 2) != 0


# ---------------- TRSMDriverAPI::driver_trsm (rstsr-blas-traits/src/blas3/trsm.rs:4-19) ----------------

DIAG = {"N": 131, "U": 132}


def _tri(rng, s, dt, uplo, diag):
    """A well-conditioned triangular matrix (off-diagonal entries ~ 0.5/s, so each row's
    off-diagonal sum stays well under its diagonal of modulus 1-2) and its BLAS storage image
    with the unread parts (other triangle, and a unit diagonal) set to NaN, so that reading
    them would show in the result."""
    T = rnd(rng, (s, s), dt) * (0.5 / s)
    d = 1.0 + rng.uniform(0, 1, s)
    if np.dtype(dt).kind == "c":
        d = d * np.exp(1j * rng.uniform(0, 2 * np.pi, s))
    T[np.diag_indices(s)] = d.astype(dt)
    T = np.tril(T) if uplo == "L" else np.triu(T)
    S = T.copy()
    S[np.triu_indices(s, 1) if uplo == "L" else np.tril_indices(s, -1)] = np.nan
    if diag == "U":
        S[np.diag_indices(s)] = np.nan
        T[np.diag_indices(s)] = 1.0
    return T, S


def _trsm_ref(T, B, side, ta, alpha):
    hi = np.complex128 if np.iscomplexobj(B) else np.float64
    oa = op(T.astype(hi), ta)
    rhs = alpha * B.astype(hi)
    if side == "L":
        return np.linalg.solve(oa, rhs)
    return np.linalg.solve(oa.T, rhs.T).T


def _run_trsm(dev, dt, order, side, uplo, ta, diag, m, n, alpha, S, B, lda, ldb, mode=0):
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    da = rt.asarray(store(S, order, lda), device=dev)
    db = rt.asarray(store(B, order, ldb), device=dev)
    al = scal(dt, alpha)
    r = L.rstsr_hip_driver_trsm(ctx, CODE[np.dtype(dt)], mode, order, SIDE[side], UPLO[uplo], TR[ta], DIAG[diag], m,
                                n, al.ctypes.data_as(C.c_void_p), ptr(da), lda, ptr(db), ldb)
    return r, db


@pytest.mark.parametrize("dt", [np.float64, np.complex128, np.float32, np.complex64])
@pytest.mark.parametrize("order", [ORD_C, ORD_F])
def test_driver_trsm_all_flags(dev, dt, order):
    """Every side x uplo x trans x diag, with padded leading dimensions: X against the extended-
    precision solve of the same triangular system (normwise, 64 s u), the unread parts of A
    (NaN) never used, B's padding untouched. s = 70 crosses one 64-row diagonal block."""
    rng = np.random.default_rng(11)
    m, n = 70, 45
    alpha = (0.75 - 0.5j) if np.dtype(dt).kind == "c" else 0.75
    for side in ("L", "R"):
        s = m if side == "L" else n
        for uplo in ("L", "U"):
            for ta in ("N", "T", "C", "CN"):
                for diag in ("N", "U"):
                    T, S = _tri(rng, s, dt, uplo, diag)
                    B = rnd(rng, (m, n), dt)
                    lda = s + 3
                    ldb = (n if order == ORD_C else m) + 2
                    r, db = _run_trsm(dev, dt, order, side, uplo, ta, diag, m, n, alpha, S, B, lda, ldb)
                    assert r == 0, abi.lib().rstsr_hip_last_error()
                    got = load(db.to_numpy(), (m, n), order, ldb)
                    ref = _trsm_ref(T, B, side, ta, alpha)
                    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
                    assert err <= 64 * s * U[np.dtype(dt)], (side, uplo, ta, diag, err)
                    flat = db.to_numpy()
                    pad = np.isnan(store(np.zeros((m, n), dt), order, ldb))
                    assert np.all(np.isnan(flat[pad])), (side, uplo, ta, diag)


@pytest.mark.parametrize("dt", [np.float64, np.complex128])
@pytest.mark.parametrize("side", ["L", "R"])
def test_driver_trsm_large_blocked(dev, dt, side):
    """A solve of order 1000 (recursive halves down to 64-row blocks, trailing updates on the
    MFMA GEMM), plain and split GEMM modes: normwise against the extended-precision solve."""
    rng = np.random.default_rng(12)
    m, n = (1000, 300) if side == "L" else (300, 1000)
    s = m if side == "L" else n
    for uplo, ta in (("L", "N"), ("U", "T")):
        T, S = _tri(rng, s, dt, uplo, "N")
        B = rnd(rng, (m, n), dt)
        ref = _trsm_ref(T, B, side, ta, 1.0)
        for mode in (0, 1):
            r, db = _run_trsm(dev, dt, ORD_C, side, uplo, ta, "N", m, n, 1.0, S, B, s, n, mode)
            assert r == 0, abi.lib().rstsr_hip_last_error()
            got = load(db.to_numpy(), (m, n), ORD_C, n)
            err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
            assert err <= 1e-12, (uplo, ta, mode, err)


def test_driver_trsm_alpha_zero_and_errors(dev):
    """alpha = 0 sets B to zero without reading it (NaN in B); bad flags and leading dimensions
    are errors, with B untouched."""
    rng = np.random.default_rng(13)
    T, S = _tri(rng, 8, np.float64, "L", "N")
    B = np.full((8, 5), np.nan)
    r, db = _run_trsm(dev, np.float64, ORD_C, "L", "L", "N", "N", 8, 5, 0.0, S, B, 8, 5)
    assert r == 0 and np.array_equal(db.to_numpy(), np.zeros(40))
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    B = rnd(rng, (8, 5), np.float64)
    da = rt.asarray(store(S, ORD_C, 8), device=dev)
    db = rt.asarray(store(B, ORD_C, 5), device=dev)
    one = scal(np.float64, 1.0)
    for args in ((ORD_C, 141, 999, 111, 131, 8), (ORD_C, 141, 122, 111, 999, 8), (ORD_C, 999, 122, 111, 131, 8),
                 (ORD_C, 141, 122, 111, 131, 4)):  # bad uplo, diag, side; lda < 8
        o, sd, ul, tr, dg, lda = args
        r = L.rstsr_hip_driver_trsm(ctx, 1, 0, o, sd, ul, tr, dg, 8, 5, one.ctypes.data_as(C.c_void_p), ptr(da), lda,
                                    ptr(db), 5)
        assert r != 0, args
    assert np.array_equal(db.to_numpy(), B.reshape(-1))
