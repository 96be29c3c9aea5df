"""HIP GEMM / matmul vs the oracle.

f64 / c128 (FP64 MFMA, fused multiply-add accumulation): componentwise bound
|C - C*| <= K * u * (|A| |B|) with u = 2^-53 against the Dot2 / extended-precision oracle
(the classical gamma_K bound of any summation order; stated tolerance).
f32: the same bound with u = 2^-24. Integers: bit-exact (wrapping), vs the naive oracle.
"""
import ctypes as C
import os

import numpy as np
import pytest

import rstsr_amd as rt
from gpu_util import from_np

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def comp_bound_ok(got, A, B, ref, u):
    K = A.shape[-1]
    scale = np.abs(A) @ np.abs(B)
    err = np.abs(got - ref)
    return np.all(err <= K * u * scale + 1e-300), float(np.max(err / np.maximum(K * u * scale, 1e-300)))


def wide_exponent(rng, shape):
    return rng.normal(size=shape) * np.exp2(rng.integers(-8, 9, size=shape))


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 65), (128, 128, 128), (300, 257, 129), (64, 512, 1000),
                                   (512, 512, 512)])
@pytest.mark.parametrize("trans", ["nn", "tn", "nt", "tt"])
def test_dgemm_plain(dev, oracle, m, n, k, trans):
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = wide_exponent(rng, (m, k) if trans[0] == "n" else (k, m))
    B = wide_exponent(rng, (k, n) if trans[1] == "n" else (n, k))
    A = A if trans[0] == "n" else A.T
    B = B if trans[1] == "n" else B.T
    ta, _ = from_np(dev, A)
    tb, _ = from_np(dev, B)
    got = (ta % tb).to_numpy()
    ref = oracle.gemm_dot2(A, B)
    ok, worst = comp_bound_ok(got, A, B, ref, 2.0 ** -53)
    assert ok, worst


def test_dgemm_strided_views(dev, oracle):
    rng = np.random.default_rng(0)
    big = rng.normal(size=(90, 140))
    A = big[::3, 5:105:2]       # 30 x 50, both strides non-unit
    B = rng.normal(size=(100, 70))[::2, ::-1]  # 50 x 70, negative column stride
    ta, _ = from_np(dev, A)
    tb, _ = from_np(dev, B)
    got = (ta % tb).to_numpy()
    ref = oracle.gemm_dot2(np.ascontiguousarray(A), np.ascontiguousarray(B))
    ok, worst = comp_bound_ok(got, A, B, ref, 2.0 ** -53)
    assert ok, worst


def test_zgemm(dev):
    rng = np.random.default_rng(1)
    A = rng.normal(size=(70, 90)) + 1j * rng.normal(size=(70, 90))
    B = rng.normal(size=(90, 130)) + 1j * rng.normal(size=(90, 130))
    got = (rt.asarray(A, device=dev) % rt.asarray(B, device=dev)).to_numpy()
    ref = A.astype(np.clongdouble) @ B.astype(np.clongdouble)
    scale = np.abs(A) @ np.abs(B)
    assert np.all(np.abs(got - ref) <= 2 * 90 * 2.0 ** -53 * scale)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_integer_matmul_exact(dev, oracle, dtype):
    rng = np.random.default_rng(2)
    A = rng.integers(-(2 ** 20), 2 ** 20, size=(37, 53)).astype(dtype)
    B = rng.integers(-(2 ** 20), 2 ** 20, size=(53, 29)).astype(dtype).T.copy().T
    got = (rt.asarray(A, device=dev) % from_np(dev, B)[0]).to_numpy()
    assert np.array_equal(got, oracle.gemm_naive(A, np.ascontiguousarray(B)))


def test_f32_matmul(dev):
    rng = np.random.default_rng(3)
    A = rng.normal(size=(45, 77)).astype(np.float32)
    B = rng.normal(size=(77, 33)).astype(np.float32)
    got = (rt.asarray(A, device=dev) % rt.asarray(B, device=dev)).to_numpy().astype(np.float64)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    scale = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
    assert np.all(np.abs(got - ref) <= 77 * 2.0 ** -24 * scale)


@pytest.mark.parametrize("sa,sb", [((5,), (5,)), ((4,), (3, 4, 6)), ((3, 5, 4), (4,)), ((6, 4), (3, 4, 5)),
                                   ((3, 6, 4), (4, 5)), ((2, 1, 6, 4), (3, 4, 5)), ((64, 32, 48), (48, 40))])
def test_matmul_rules(dev, sa, sb):
    rng = np.random.default_rng(len(sa) * 10 + len(sb))
    A = rng.integers(-50, 50, size=sa).astype(np.int64)
    B = rng.integers(-50, 50, size=sb).astype(np.int64)
    got = rt.asarray(A, device=dev) % rt.asarray(B, device=dev)
    ref = np.matmul(A, B)
    assert got.shape == ref.shape
    assert np.array_equal(got.to_numpy(), ref)
    # the same through the FP64 MFMA path (exact for these small integers)
    gotf = rt.asarray(A.astype(np.float64), device=dev) % rt.asarray(B.astype(np.float64), device=dev)
    assert np.array_equal(gotf.to_numpy(), ref.astype(np.float64))


def test_reference_matmul_cases(dev):
    m1 = rt.asarray(np.array([[1, 2], [3, 4]], dtype=np.int32), device=dev)
    m2 = rt.asarray(np.array([[1, 0], [1, 1]], dtype=np.int32), device=dev)
    assert (m1 % m2).to_numpy().tolist() == [[3, 2], [7, 4]]
    assert (m2 % m1).to_numpy().tolist() == [[1, 2], [4, 6]]
    v = rt.asarray(np.array([1, 2], dtype=np.int32), device=dev)
    assert (v % m1).to_vec() == [7, 10]
    assert (m1 % v).to_vec() == [5, 11]
    m12 = rt.asarray(np.array([[[1, 2], [3, 4]], [[1, 0], [1, 1]]], dtype=np.int32), device=dev)
    assert (m12 % m1).to_numpy().tolist() == [[[7, 10], [15, 22]], [[1, 2], [4, 6]]]
    a = rt.asarray(np.array([0 + 1j, 1 + 2j, 2 + 3j, 3 + 4j, 4 + 5j, 5 + 6j], dtype=np.complex64).reshape(3, 2), device=dev)
    b = rt.asarray(np.array([0 + 1j, 2 + 3j, 4 + 5j, 6 + 7j], dtype=np.complex64).reshape(2, 2), device=dev)
    s = (a % b).to_numpy().sum()
    assert abs(s.real + 78.0) < 1e-5 and abs(s.imag - 270.0) < 1e-5


def test_reference_large_matmul_stats(dev):
    a = rt.linspace(0.0, 1.0, 1048576, device=dev).reshape([16, 256, 256])
    b = rt.linspace(1.0, 2.0, 1048576, device=dev).reshape([16, 256, 256])
    c = (a % b).to_numpy()
    assert abs(c.mean() - 213.2503660477036) < 1e-6
    assert abs(c.std() - 148.88523481701804) < 1e-6
    assert abs(c.std(axis=(0, 1))[0] - 148.8763226818815) < 1e-6
    assert abs(c.std(axis=(0, 1))[255] - 148.8941462322758) < 1e-6
    assert abs(c.std(axis=(1, 2))[0] - 4.763105902995575) < 1e-6
    assert abs(c.std(axis=(1, 2))[15] - 9.093224903569157) < 1e-6


def test_gemm_alpha_beta_c_abi(dev, oracle):
    """rstsr_hip_gemm with alpha/beta on strided C; beta == 0 must not read C (NaN-filled)."""
    lib = C.CDLL(os.path.join(ROOT, "rstsr_amd", "librstsr_hip.so"))
    rng = np.random.default_rng(4)
    m, n, k = 40, 50, 60
    A = rng.normal(size=(m, k))
    B = rng.normal(size=(k, n))
    C0 = rng.normal(size=(n, m)).T  # column-major C
    ta = rt.asarray(A, device=dev)
    tb = rt.asarray(B, device=dev)
    tc, _ = from_np(dev, C0)
    ctx = C.c_void_p(dev._ctx_ptr())
    al = C.c_double(0.5)
    be = C.c_double(-2.0)
    ptr = lambda t: C.c_void_p(t._t._data_ptr() + t.offset * 8)
    lib.rstsr_hip_gemm.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p,
                                   C.c_void_p, C.c_ssize_t, C.c_ssize_t, C.c_void_p, C.c_ssize_t, C.c_ssize_t,
                                   C.c_void_p, C.c_void_p, C.c_ssize_t, C.c_ssize_t]
    r = lib.rstsr_hip_gemm(ctx, 1, 0, m, n, k, C.byref(al), ptr(ta), k, 1, ptr(tb), n, 1, C.byref(be), ptr(tc),
                           tc.stride[0], tc.stride[1])
    assert r == 0
    ref = 0.5 * oracle.gemm_dot2(A, B) - 2.0 * C0
    assert np.max(np.abs(tc.to_numpy() - ref)) < 1e-12 * np.max(np.abs(ref))
    # beta = 0 over a NaN-filled C
    tn = rt.full((m, n), np.nan, device=dev)
    z = C.c_double(0.0)
    r = lib.rstsr_hip_gemm(ctx, 1, 0, m, n, k, C.byref(al), ptr(ta), k, 1, ptr(tb), n, 1, C.byref(z), ptr(tn), n, 1)
    assert r == 0
    assert np.all(np.isfinite(tn.to_numpy()))


# ---------------- Ozaki-II split GEMM (INT8 MFMA + CRT) ----------------
# Stated tolerance: normwise max|C - C*| / max|C*| <= 1e-14 against the Dot2 oracle (the
# project's target), on entries N(0,1) * 2^u, u uniform in [-8, 8].

@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 65), (128, 128, 128), (300, 257, 129), (200, 130, 1500),
                                   (512, 512, 512)])
@pytest.mark.parametrize("trans", ["nn", "tt"])
def test_ozaki_normwise(dev, oracle, m, n, k, trans):
    rng = np.random.default_rng(m + 3 * n + 7 * k)
    A = wide_exponent(rng, (m, k) if trans[0] == "n" else (k, m))
    B = wide_exponent(rng, (k, n) if trans[1] == "n" else (n, k))
    A = A if trans[0] == "n" else A.T
    B = B if trans[1] == "n" else B.T
    ta, _ = from_np(dev, A)
    tb, _ = from_np(dev, B)
    got = rt.matmul(ta, tb, mode="ozaki").to_numpy()
    ref = oracle.gemm_dot2(np.ascontiguousarray(A), np.ascontiguousarray(B))
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err <= 1e-14, err


def test_ozaki_exact_on_integers(dev):
    """Small integer matrices are exact after scaling, so the split product is exact."""
    rng = np.random.default_rng(6)
    A = rng.integers(-1000, 1000, size=(70, 90)).astype(np.float64)
    B = rng.integers(-1000, 1000, size=(90, 50)).astype(np.float64)
    got = rt.matmul(rt.asarray(A, device=dev), rt.asarray(B, device=dev), mode="ozaki").to_numpy()
    assert np.array_equal(got, (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64))


def test_ozaki_batched_and_zero_rows(dev, oracle):
    rng = np.random.default_rng(7)
    A = wide_exponent(rng, (3, 40, 60))
    A[1, 5, :] = 0.0
    B = wide_exponent(rng, (60, 30))
    B[:, 7] = 0.0
    got = rt.matmul(rt.asarray(A, device=dev), rt.asarray(B, device=dev), mode="ozaki").to_numpy()
    for b in range(3):
        ref = oracle.gemm_dot2(A[b], B)
        assert np.max(np.abs(got[b] - ref)) <= 1e-14 * np.max(np.abs(ref))
    assert np.all(got[1, 5, :] == 0.0) and np.all(got[:, :, 7] == 0.0)


@pytest.fixture
def int8_tile():
    """Sets RSTSR_HIP_OZAKI_TILE (read at every call) and restores it afterwards."""
    old = os.environ.get("RSTSR_HIP_OZAKI_TILE")

    def set_tile(t):
        if t is None:
            os.environ.pop("RSTSR_HIP_OZAKI_TILE", None)
        else:
            os.environ["RSTSR_HIP_OZAKI_TILE"] = t

    yield set_tile
    set_tile(old)


def test_ozaki_big_tile_moduli_per_block(dev, oracle, int8_tile):
    """The 256 x 256 direct-load tile runs several moduli per block when the launch stays at four
    blocks per CU (here [32, 1024, 512] % [512, 1024]: 512 tiles, 7 moduli per block): the same
    bits as the 128 x 128 tile (one modulus per block, register-staged loads; residue sums are
    exact), and the Dot2 oracle on sampled rows."""
    rng = np.random.default_rng(21)
    A = wide_exponent(rng, (32, 1024, 512))
    B = wide_exponent(rng, (512, 1000))  # ragged: the residue rows are padded to 1024 bytes
    ta, tb = rt.asarray(A, device=dev), rt.asarray(B, device=dev)
    int8_tile("big")
    big = rt.matmul(ta, tb, mode="ozaki").to_numpy()
    int8_tile("square")
    sq = rt.matmul(ta, tb, mode="ozaki").to_numpy()
    assert np.array_equal(big.view(np.uint64), sq.view(np.uint64))
    for b in (0, 13, 31):
        ref = oracle.gemm_dot2(A[b, ::97], B)
        assert np.max(np.abs(big[b, ::97] - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_ozaki_big_tile_long_k(dev, oracle, int8_tile):
    """K beyond one INT8 launch's 65536 (the second launch adds its residues to the first's) on
    the 256 x 256 tile: the same bits as the 128 x 128 tile, and normwise against Dot2 within
    1e-14 sqrt(K / 4096) = 4e-14 (the split's per-element truncation, g = 54 bits relative to
    each row's and column's largest entry, summed over 16 times the products of K = 4096)."""
    rng = np.random.default_rng(22)
    A = rng.standard_normal((260, 65700))
    B = rng.standard_normal((65700, 260))
    ta, tb = rt.asarray(A, device=dev), rt.asarray(B, device=dev)
    int8_tile("big")
    got = rt.matmul(ta, tb, mode="ozaki").to_numpy()
    int8_tile("square")
    sq = rt.matmul(ta, tb, mode="ozaki").to_numpy()
    assert np.array_equal(got.view(np.uint64), sq.view(np.uint64))
    ref = oracle.gemm_dot2(A, B)
    assert np.max(np.abs(got - ref)) <= 4e-14 * np.max(np.abs(ref))


@pytest.mark.parametrize("nmod", [12, 7])
def test_ozaki_fewer_moduli(dev, oracle, int8_tile, nmod):
    """RSTSR_HIP_OZAKI_MODULI below 14 takes every split kernel's runtime-constant form (residues,
    INT8 epilogue, row and tiled CRT, several moduli per block over a divisor of nmod): the 256 x 256
    and 128 x 128 tiles give the same bits, and each element meets the split's bound. With the scale
    g = floor((log2 M - 1.5) / 2) of the first nmod moduli, every split element is within
    2^-g ||row||_2 of its operand's element (the scale is a power of two at most twice the norm), so
    by Cauchy-Schwarz |C - C*|_ij <= 2 sqrt(K) 2^-g ||a_i|| ||b_j|| (+ the final rounding)."""
    moduli = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199][:nmod]
    g = min(55, int(np.floor((np.sum(np.log2(moduli)) - 1.5) / 2)))
    rng = np.random.default_rng(30 + nmod)
    A = rng.standard_normal((300, 1000))
    B = rng.standard_normal((1000, 520))
    A3 = rng.standard_normal((4, 300, 129))
    B2 = rng.standard_normal((129, 517))
    X = rng.standard_normal((90, 257))
    tA, tB, tA3, tB2, tX = (rt.asarray(v, device=dev) for v in (A, B, A3, B2, X))
    old = os.environ.get("RSTSR_HIP_OZAKI_MODULI")
    os.environ["RSTSR_HIP_OZAKI_MODULI"] = str(nmod)
    try:
        got = {}
        for tile in ("big", "square"):
            int8_tile(tile)
            got[tile] = (rt.matmul(tA, tB, mode="ozaki").to_numpy(), rt.matmul(tA3, tB2, mode="ozaki").to_numpy(),
                         rt.matmul(tX.T, tX, mode="ozaki").to_numpy())
    finally:
        if old is None:
            os.environ.pop("RSTSR_HIP_OZAKI_MODULI", None)
        else:
            os.environ["RSTSR_HIP_OZAKI_MODULI"] = old
    for a, b in zip(got["big"], got["square"]):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))

    def within(c, a, b):
        ref = oracle.gemm_dot2(np.ascontiguousarray(a), np.ascontiguousarray(b))
        bound = 2.0 * np.sqrt(a.shape[1]) * 2.0 ** -g * np.outer(np.linalg.norm(a, axis=1),
                                                                np.linalg.norm(b, axis=0))
        assert np.all(np.abs(c - ref) <= bound + 2.0 ** -52 * np.abs(ref))

    c, c3, cs = got["big"]
    within(c, A, B)
    for bi in (0, 3):
        within(c3[bi], A3[bi], B2)
    assert np.array_equal(cs, cs.T)
    within(cs, X.T, X)


def test_ozaki_nonfinite_falls_back(dev):
    A = np.ones((20, 30))
    A[3, 4] = np.inf
    B = np.ones((30, 10))
    got = rt.matmul(rt.asarray(A, device=dev), rt.asarray(B, device=dev), mode="ozaki").to_numpy()
    assert np.all(np.isinf(got[3])) and np.all(got[:3] == 30.0)


# ---------------- grouped (uneven) GEMM, complex split, graph capture ----------------

@pytest.mark.parametrize("dtype,mode", [(np.float64, "plain"), (np.complex128, "plain"), (np.float64, "ozaki")])
def test_grouped_uneven(dev, oracle, dtype, mode):
    rng = np.random.default_rng(11)
    probs, hosts = [], []
    for g in range(40):
        m, n, k = (int(x) for x in rng.integers(1, 40, size=3) * 8)
        A = rng.normal(size=(m, k)) if g % 3 else rng.normal(size=(k, m)).T
        B = rng.normal(size=(k, n)) if g % 2 else rng.normal(size=(n, k)).T
        if dtype == np.complex128:
            A = A + 1j * rng.normal(size=A.shape)
            B = B + 1j * rng.normal(size=B.shape)
        ta, _ = from_np(dev, A)
        tb, _ = from_np(dev, B)
        tc = rt.empty((m, n), dtype, device=dev)
        probs.append((ta, tb, tc))
        hosts.append((A, B))
    rt.gemm_grouped(probs, mode=mode)
    for (A, B), (_, _, tc) in zip(hosts, probs):
        got = tc.to_numpy()
        if dtype == np.complex128:
            ref = A.astype(np.clongdouble) @ B.astype(np.clongdouble)
            scale = np.abs(A) @ np.abs(B)
            assert np.all(np.abs(got - ref) <= 2 * A.shape[1] * 2.0 ** -53 * scale)
        else:
            ref = oracle.gemm_dot2(np.ascontiguousarray(A), np.ascontiguousarray(B))
            if mode == "plain":
                ok, worst = comp_bound_ok(got, A, B, ref, 2.0 ** -53)
                assert ok, worst
            else:
                assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))


@pytest.mark.parametrize("mode", ["plain", "ozaki"])
def test_grouped_back_to_back_calls(dev, oracle, mode):
    """Five grouped calls with different problem tables queued without a host wait between them:
    the tables travel through two alternating pinned regions, each refilled only after its
    previous copy ran; every call's outputs are checked afterwards."""
    rng = np.random.default_rng(21)
    calls = []
    for c in range(5):
        probs, hosts = [], []
        for g in range(3 + 4 * c):
            m, n, k = (int(x) for x in rng.integers(1, 24, size=3) * 8)
            A, B = rng.normal(size=(m, k)), rng.normal(size=(k, n))
            probs.append((rt.asarray(A, device=dev), rt.asarray(B, device=dev), rt.empty((m, n), np.float64, device=dev)))
            hosts.append((A, B))
        calls.append((probs, hosts))
    for probs, _ in calls:
        rt.gemm_grouped(probs, mode=mode)
    for probs, hosts in calls:
        for (A, B), (_, _, tc) in zip(hosts, probs):
            ref = oracle.gemm_dot2(A, B)
            assert np.max(np.abs(tc.to_numpy() - ref)) <= 1e-13 * max(np.max(np.abs(ref)), 1.0)


def test_grouped_not_capturable(dev):
    A = rt.asarray(np.ones((8, 8)), device=dev)
    C = rt.empty((8, 8), np.float64, device=dev)
    with pytest.raises(Exception, match="cannot be captured"):
        with rt.capture(dev):
            rt.gemm_grouped([(A, A, C)])
    rt.gemm_grouped([(A, A, C)])  # the context works normally afterwards
    assert np.array_equal(C.to_numpy(), np.full((8, 8), 8.0))


def test_grouped_alpha_beta(dev, oracle):
    rng = np.random.default_rng(12)
    A = rng.normal(size=(24, 40))
    B = rng.normal(size=(40, 16))
    C0 = rng.normal(size=(24, 16))
    tc = rt.asarray(C0, device=dev)
    rt.gemm_grouped([(rt.asarray(A, device=dev), rt.asarray(B, device=dev), tc)], alpha=2.0, beta=-1.0)
    ref = 2.0 * oracle.gemm_dot2(A, B) - C0
    assert np.max(np.abs(tc.to_numpy() - ref)) < 1e-13 * np.max(np.abs(ref))


def _grouped_views(rng, g, m, n, k):
    """Operands of problem g in one of several layouts (row/column major, reversed, sliced)."""
    kind = g % 5
    if kind == 0:
        A, B = wide_exponent(rng, (m, k)), wide_exponent(rng, (k, n))
    elif kind == 1:
        A, B = wide_exponent(rng, (k, m)).T, wide_exponent(rng, (n, k)).T
    elif kind == 2:
        A, B = wide_exponent(rng, (m, k))[:, ::-1], wide_exponent(rng, (k, n))[::-1]
    elif kind == 3:
        A, B = wide_exponent(rng, (2 * m, 3 * k))[::2, ::3], wide_exponent(rng, (k, 2 * n))[:, 1::2]
    else:
        A, B = wide_exponent(rng, (k, m))[::-1].T, wide_exponent(rng, (k, n))
    return A, B


@pytest.fixture(params=["split", "auto"])
def grouped_route(request, monkeypatch):
    """Both routes of the grouped split GEMM: every problem through the grouped split kernels,
    and the default per-problem choice between the split path and the plain grouped kernel."""
    monkeypatch.setenv("RSTSR_HIP_OZAKI_GROUPED", request.param)
    return request.param


def test_grouped_ozaki_layouts(dev, oracle, grouped_route):
    """Grouped split GEMM (one launch per stage over the whole group): odd sizes including
    1 and K = 0, wide exponents, every operand layout, alpha/beta; normwise 1e-14 per problem."""
    rng = np.random.default_rng(21)
    sizes = [(1, 1, 1), (1, 7, 3), (5, 1, 9), (33, 17, 0), (129, 65, 257), (8, 300, 40), (200, 3, 131)]
    sizes += [tuple(int(x) for x in rng.integers(1, 300, size=3)) for _ in range(25)]
    probs, hosts = [], []
    for g, (m, n, k) in enumerate(sizes):
        A, B = _grouped_views(rng, g, m, n, k)
        C0 = rng.normal(size=(m, n))
        ta, _ = from_np(dev, A)
        tb, _ = from_np(dev, B)
        probs.append((ta, tb, rt.asarray(C0, device=dev)))
        hosts.append((A, B, C0))
    rt.gemm_grouped(probs, mode="ozaki", alpha=1.5, beta=-0.5)
    for (A, B, C0), (_, _, tc) in zip(hosts, probs):
        got = tc.to_numpy()
        if A.shape[1] == 0:
            assert np.array_equal(got, -0.5 * C0)
            continue
        ab = oracle.gemm_dot2(np.ascontiguousarray(A), np.ascontiguousarray(B))
        ref = 1.5 * ab - 0.5 * C0
        # split-GEMM error of A B (normwise), plus the rounding of the alpha/beta update
        bound = 1e-14 * 1.5 * np.max(np.abs(ab)) + 2.0 ** -52 * (np.abs(1.5 * ab) + np.abs(0.5 * C0))
        assert np.all(np.abs(got - ref) <= bound), (A.shape, B.shape)


def test_grouped_ozaki_nonfinite_falls_back(dev, oracle, grouped_route):
    """A NaN or Inf anywhere in the group sends the whole group to the plain FP64 kernel: IEEE
    special values propagate as in the plain product and finite problems stay accurate."""
    rng = np.random.default_rng(22)
    shapes = [(40, 24, 56), (17, 33, 9), (64, 64, 64)]
    hosts = [(rng.normal(size=(m, k)), rng.normal(size=(k, n))) for m, n, k in shapes]
    hosts[1][0][3, 4] = np.nan
    hosts[2][1][5, 6] = np.inf
    probs = [(rt.asarray(A, device=dev), rt.asarray(B, device=dev), rt.empty((A.shape[0], B.shape[1]), np.float64,
                                                                             device=dev)) for A, B in hosts]
    rt.gemm_grouped(probs, mode="ozaki")
    for (A, B), (_, _, tc) in zip(hosts, probs):
        got = tc.to_numpy()
        with np.errstate(invalid="ignore"):
            ref = A @ B
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        assert np.array_equal(np.isinf(got), np.isinf(ref))
        fin = np.isfinite(ref)
        if fin.any():
            assert np.max(np.abs(got[fin] - ref[fin])) <= 1e-13 * np.max(np.abs(ref[fin]))


def test_grouped_ozaki_routes_large_problem(dev, oracle, monkeypatch):
    """A group with one product large enough for the split path (2048^3) among small ones: the
    default route sends it through the single-problem split path and the rest through the plain
    grouped kernel; every product normwise 1e-14 against Dot2, the large one the same bits as
    its own split call."""
    monkeypatch.delenv("RSTSR_HIP_OZAKI_GROUPED", raising=False)
    rng = np.random.default_rng(23)
    shapes = [(2048, 2048, 2048), (40, 24, 56), (300, 17, 129)]
    hosts = [(rng.normal(size=(m, k)), rng.normal(size=(k, n))) for m, n, k in shapes]
    probs = [(rt.asarray(A, device=dev), rt.asarray(B, device=dev),
              rt.empty((A.shape[0], B.shape[1]), np.float64, device=dev)) for A, B in hosts]
    rt.gemm_grouped(probs, mode="ozaki")
    alone = rt.matmul(probs[0][0], probs[0][1], mode="ozaki").to_numpy()
    assert np.array_equal(probs[0][2].to_numpy(), alone)
    from dot2_gpu import dot2_matmul, normwise
    assert normwise(dev, probs[0][2], dot2_matmul(rt, dev, probs[0][0], probs[0][1])) <= 1e-14
    for (A, B), (_, _, tc) in list(zip(hosts, probs))[1:]:
        ab = oracle.gemm_dot2(A, B)
        assert np.max(np.abs(tc.to_numpy() - ab)) <= 1e-14 * np.max(np.abs(ab))


@pytest.mark.parametrize("shape", [(64, 96, 80), (300, 130, 257)])
def test_zgemm_ozaki(dev, shape):
    m, k, n = shape
    rng = np.random.default_rng(m)
    A = wide_exponent(rng, (m, k)) + 1j * wide_exponent(rng, (m, k))
    B = wide_exponent(rng, (k, n)) + 1j * wide_exponent(rng, (k, n))
    got = rt.matmul(rt.asarray(A, device=dev), rt.asarray(B, device=dev), mode="ozaki").to_numpy()
    ref = A.astype(np.clongdouble) @ B.astype(np.clongdouble)
    err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    assert err <= 1e-14, err


@pytest.mark.parametrize("mode", ["plain", "ozaki"])
def test_graph_capture_chain(dev, mode):
    """add -> mul -> matmul -> sum_axes captured once, replayed 3 times with eager work between
    replays; every replay must equal the eager sequence bit for bit."""
    rng = np.random.default_rng(13)
    n = 192
    z1h = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    z2h = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    vh = rng.normal(size=n) + 1j * rng.normal(size=n)
    z2 = rt.asarray(z2h, device=dev)
    v = rt.asarray(vh, device=dev)

    def eager():
        z1 = rt.asarray(z1h, device=dev)
        z1 += v
        z3 = z1 * z2
        z4 = rt.matmul(z3, z2, mode=mode)
        return z4.sum_axes([1]).to_numpy()

    ref = eager()
    z1 = rt.asarray(z1h, device=dev)
    with rt.capture(dev) as g:
        z1 += v
        z3 = z1 * z2
        z4 = rt.matmul(z3, z2, mode=mode)
        s = z4.sum_axes([1])
    for rep in range(3):
        z1.assign(rt.asarray(z1h, device=dev))  # eager reset between replays
        _ = eager()                              # unrelated eager work between replays
        g.launch()
        dev.synchronize()
        assert np.array_equal(s.to_numpy(), ref), rep


@pytest.mark.parametrize("sa,sb", [((3, 40, 60), (60, 30)), ((40, 60), (3, 60, 30)), ((3, 40, 60), (3, 60, 30)),
                                   ((2, 1, 24, 36), (4, 36, 20))])
def test_ozaki_batch_rules(dev, oracle, sa, sb):
    """Rules 5/6/7 in one batched split launch (a broadcast operand is split once)."""
    rng = np.random.default_rng(len(sa) * 7 + len(sb))
    A = wide_exponent(rng, sa)
    B = wide_exponent(rng, sb)
    got = rt.matmul(rt.asarray(A, device=dev), rt.asarray(B, device=dev), mode="ozaki").to_numpy()
    ref = np.matmul(A.astype(np.longdouble), B.astype(np.longdouble))
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_ri_contraction_small(dev, oracle):
    """(ia|jb) = sum_P B[P,i,a] B[P,j,b] as one rule-6 matmul of permuted views, plain and split."""
    rng = np.random.default_rng(21)
    nP, no, nv = 96, 6, 40
    Bh = rng.normal(0.0, 0.1, size=(nP, no, nv))
    tB = rt.asarray(Bh, device=dev)
    lhs = tB.permute([1, 2, 0])              # [i, a, P], a contiguous
    rhs = tB.reshape([nP, no * nv])          # [P, (j b)]
    ref = np.einsum("pia,pjb->iajb", Bh.astype(np.longdouble), Bh.astype(np.longdouble)).reshape(no, nv, no * nv)
    for mode in ("plain", "ozaki"):
        got = rt.matmul(lhs, rhs, mode=mode).to_numpy()
        assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref)), mode


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
@pytest.mark.parametrize("m,k", [(300, 77), (64, 64), (129, 200)])
def test_syrk_shortcut_bit_identical(dev, oracle, dtype, m, k):
    """A = X^T and B = X in the same buffer with beta = 0 take the symmetric path (lower tiles plus
    a mirror, device_faer/matmul.rs:38-48). Its lower triangle equals, bit for bit, the same
    product with B in a separate buffer (full tile grid); the result is exactly symmetric; for
    real types the whole matrix is bit-identical (the mirrored sums are the same sums)."""
    rng = np.random.default_rng(m + k)
    X = rng.normal(size=(k, m))
    if dtype == np.complex128:
        X = X + 1j * rng.normal(size=(k, m))
    tX = rt.asarray(X, device=dev)
    tY = rt.asarray(X.copy(), device=dev)
    sym = rt.matmul(tX.T, tX).to_numpy()
    full = rt.matmul(tX.T, tY).to_numpy()
    low = np.tril_indices(m)
    assert np.array_equal(sym[low].view(np.uint8), full[low].view(np.uint8))
    assert np.array_equal(sym, sym.T)
    if dtype == np.float64:
        assert np.array_equal(sym.view(np.uint8), full.view(np.uint8))
    else:
        assert np.max(np.abs(sym - full)) <= 4 * k * 2.0 ** -53 * np.max(np.abs(full))
    if dtype == np.float64:
        ref = oracle.gemm_dot2(np.ascontiguousarray(X.T), np.ascontiguousarray(X))
        ok, worst = comp_bound_ok(sym, X.T, X, ref, 2.0 ** -53)
        assert ok, worst


def test_syrk_batched(dev):
    """Batched symmetric products (equal batch strides) on a partial last tile."""
    rng = np.random.default_rng(5)
    X = rng.normal(size=(3, 50, 140))
    tX = rt.asarray(X, device=dev)
    tY = rt.asarray(X.copy(), device=dev)
    sym = rt.matmul(tX.permute([0, 2, 1]), tX).to_numpy()
    full = rt.matmul(tX.permute([0, 2, 1]), tY).to_numpy()
    assert np.array_equal(sym, full)
    ref = np.einsum("bki,bkj->bij", X.astype(np.longdouble), X.astype(np.longdouble))
    assert np.max(np.abs(sym - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_batch_fold_matches_unfolded(dev, oracle):
    """A broadcast B with A and C batches that chain into their rows runs as one taller GEMM;
    the result must equal per-batch products (each batch computed alone)."""
    rng = np.random.default_rng(8)
    A = rng.normal(size=(5, 37, 64))
    B = rng.normal(size=(64, 90))
    tA = rt.asarray(A, device=dev)
    tB = rt.asarray(B, device=dev)
    got = rt.matmul(tA, tB).to_numpy()
    for b in range(5):
        one = rt.matmul(rt.asarray(A[b], device=dev), tB).to_numpy()
        ref = oracle.gemm_dot2(A[b], B)
        ok, worst = comp_bound_ok(got[b], A[b], B, ref, 2.0 ** -53)
        assert ok, worst
        assert np.max(np.abs(one - got[b])) <= 64 * 2.0 ** -52 * np.max(np.abs(ref))


def test_free_during_capture_is_deferred(dev):
    """Dropping a pre-capture tensor inside a capture must not break the capture."""
    a = rt.asarray(np.arange(64.0).reshape(8, 8), device=dev)
    tmp = rt.asarray(np.ones((8, 8)), device=dev)
    with rt.capture(dev) as g:
        tmp = a * a  # rebinding drops the old tmp while the stream is capturing
    g.launch()
    dev.synchronize()
    assert np.array_equal(tmp.to_numpy(), np.arange(64.0).reshape(8, 8) ** 2)


@pytest.mark.parametrize("m,k", [(1, 5), (33, 40), (100, 64), (129, 70), (257, 33), (300, 257), (1024, 512),
                                 (2000, 100)])
def test_ozaki_syrk_bit_identical(dev, oracle, m, k):
    """The split GEMM's symmetric path (X^T X in one buffer, beta = 0: shared scales and residue
    planes, lower-triangle INT8 tiles, a CRT over the lower 32 x 128 tiles that also writes each
    tile's mirror) equals the split GEMM with B in a separate buffer bit for bit (residue sums are
    exact integers), and stays within 1e-14 normwise. Sizes on both sides of the INT8 and CRT tile
    edges."""
    rng = np.random.default_rng(m * 3 + k)
    X = wide_exponent(rng, (k, m))
    tX = rt.asarray(X, device=dev)
    tY = rt.asarray(X.copy(), device=dev)
    sym = rt.matmul(tX.T, tX, mode="ozaki").to_numpy()
    full = rt.matmul(tX.T, tY, mode="ozaki").to_numpy()
    assert np.array_equal(sym.view(np.uint8), full.view(np.uint8))
    assert np.array_equal(sym, sym.T)
    ref = oracle.gemm_dot2(np.ascontiguousarray(X.T), np.ascontiguousarray(X))
    assert np.max(np.abs(sym - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_ozaki_syrk_batched_and_nonfinite(dev):
    rng = np.random.default_rng(9)
    X = rng.normal(size=(3, 70, 130))
    tX = rt.asarray(X, device=dev)
    got = rt.matmul(tX.swapaxes(-1, -2), tX, mode="ozaki").to_numpy()
    ref = np.einsum("bki,bkj->bij", X, X)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    assert np.array_equal(got, got.swapaxes(-1, -2))
    # a NaN input takes the plain fallback inside the same launch sequence
    X[1, 5, 7] = np.nan
    tX = rt.asarray(X, device=dev)
    got = rt.matmul(tX.swapaxes(-1, -2), tX, mode="ozaki").to_numpy()
    ref = np.einsum("bki,bkj->bij", X, X)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    fin = np.isfinite(ref)
    assert np.max(np.abs(got[fin] - ref[fin])) <= 1e-12 * np.max(np.abs(ref[fin]))


def test_zgemm_ozaki_three_products_strided(dev):
    """beta = 0 complex split GEMM (three real products) on transposed and negatively strided
    operands; normwise error against an extended-precision product."""
    rng = np.random.default_rng(21)
    m, k, n = 150, 333, 97
    A0 = wide_exponent(rng, (k, m)) + 1j * wide_exponent(rng, (k, m))
    B0 = wide_exponent(rng, (n, k)) + 1j * wide_exponent(rng, (n, k))
    A = A0.T                 # column-major view
    B = B0[::-1].T           # negative column stride
    ta, _ = from_np(dev, A)
    tb, _ = from_np(dev, B)
    got = rt.matmul(ta, tb, mode="ozaki").to_numpy()
    ref = A.astype(np.clongdouble) @ B.astype(np.clongdouble)
    err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    assert err <= 1e-14, err


def test_zgemm_ozaki_four_products_with_beta(dev):
    """beta != 0 keeps the four-product form: C = a A B + b C through the C ABI."""
    from rstsr_amd import abi

    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    rng = np.random.default_rng(22)
    m, k, n = 70, 90, 60
    A = rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))
    B = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    C0 = rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n))
    ta, tb, tc = rt.asarray(A, device=dev), rt.asarray(B, device=dev), rt.asarray(C0, device=dev)
    al = np.array([1.5 + 0j])
    be = np.array([-0.5 + 0j])
    p = lambda t: C.c_void_p(t._t._data_ptr())  # noqa: E731
    r = L.rstsr_hip_gemm(ctx, 5, 1, m, n, k, al.ctypes.data_as(C.c_void_p), p(ta), k, 1, p(tb), n, 1,
                         be.ctypes.data_as(C.c_void_p), p(tc), n, 1)
    assert r == 0, L.rstsr_hip_last_error()
    ref = 1.5 * (A.astype(np.clongdouble) @ B.astype(np.clongdouble)) - 0.5 * C0
    err = float(np.max(np.abs(tc.to_numpy() - ref)) / np.max(np.abs(ref)))
    assert err <= 1e-14, err


def test_host_value_entry_points_refuse_capture(dev):
    """sum_all and allclose hand a value to the host: under capture they fail before enqueuing
    anything, and the captured graph and the context stay usable."""
    x = rt.asarray(np.arange(16.0), device=dev)
    y = rt.asarray(np.arange(16.0), device=dev)
    with pytest.raises(Exception, match="cannot be captured"):
        with rt.capture(dev):
            x.sum_all()
    with pytest.raises(Exception, match="cannot be captured"):
        with rt.capture(dev):
            rt.allclose(x, y)
    with rt.capture(dev) as g:
        z = x + y
    g.launch()
    dev.synchronize()
    assert np.array_equal(z.to_numpy(), 2 * np.arange(16.0))
    assert x.sum_all() == 120.0 and rt.allclose(x, y)
