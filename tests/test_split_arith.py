"""The split GEMM's integer shortcuts (rstsr_amd/csrc/ozaki.hip), restated in NumPy with the
device's operand widths and f32 rounding, checked against Python's exact `%` over every input
the kernels can produce:

- res_reduce: s mod m for the byte-dot-product sums s <= 8 * 255^2 + 384 of the residue split:
  q = bits [20, 32) of s * ceil(2^20 / m) (the low 32 bits of a 24-bit product), t = s - q m,
  min(t, t + m) as unsigned words. Exhaustive over that range for the 14 moduli.
- mod_res: x mod m for the INT8 products' sums |x| <= 2^30: s = (x >> 16) (2^16 mod m) +
  (x & 0xFFFF) + m 2^14, q = trunc(f32(s) * f32(1/m)(1 - 2^-22)), t = s - q m,
  min(t, t - m) as unsigned words. The f32 quotient step exhaustive over s < 2^24, the whole
  path on random x and the extremes.
"""
import numpy as np

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199]


def res_reduce(s, m):
    mg20 = ((1 << 20) + m - 1) // m
    q = ((s * mg20) & 0xFFFFFFFF) >> 20
    t = s - q * m
    return np.minimum(t.astype(np.uint32), (t + m).astype(np.uint32)).astype(np.int64)


def f32_quotient(s, m):
    inv = np.float32(np.float32(1.0 / m) * np.float32(1.0 - 2.0 ** -22))
    return np.trunc((s.astype(np.float32) * inv).astype(np.float32)).astype(np.int64)


def mod_res(x, m):
    x = x.astype(np.int32)
    c16 = (1 << 16) % m
    s = (x >> 16).astype(np.int64) * c16 + (x & 0xFFFF).astype(np.int64) + (m << 14)
    assert s.min() >= 0 and s.max() < (1 << 24)
    t = s - f32_quotient(s, m) * m
    return np.minimum(t.astype(np.uint32), (t - m).astype(np.int64).astype(np.uint32)).astype(np.int64)


def test_res_reduce_exhaustive():
    # s = two byte dot products (8 bytes times residues of 256^k, each < 256) plus the offset
    # m - (2^56 mod m) + 128 <= 384
    smax = 8 * 255 * 255 + 384
    s = np.arange(0, smax + 1, dtype=np.int64)
    for m in MODULI:
        assert (s * (((1 << 20) + m - 1) // m) < (1 << 32)).all()  # the 24-bit product fits 32 bits
        assert np.array_equal(res_reduce(s, m), s % m), m


def test_mod_res_quotient_exhaustive():
    for m in MODULI:
        for lo in range(0, 1 << 24, 1 << 22):
            s = np.arange(lo, lo + (1 << 22), dtype=np.int64)
            t = s - f32_quotient(s, m) * m
            assert t.min() >= 0 and t.max() < 2 * m, m


def test_mod_res_values():
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.integers(-(1 << 30), (1 << 30) + 1, size=400_000),
                        np.arange(-70_000, 70_000),
                        [-(1 << 30), (1 << 30), (1 << 30) - 1, -(1 << 30) + 1, 0, 1, -1, 65535, -65536]]).astype(np.int64)
    for m in MODULI:
        assert np.array_equal(mod_res(x, m), x % m), m


def test_planted_mistakes_fail():
    """A quotient without the downward bias, or a correction in the wrong direction, fails."""
    s = np.arange(0, 1 << 24, dtype=np.int64)
    over = 0
    for m in MODULI:
        q = np.trunc((s.astype(np.float32) * np.float32(1.0 / m)).astype(np.float32)).astype(np.int64)
        over += int(((s - q * m) < 0).sum())
    assert over > 0  # unbiased f32(1/m): a quotient one too large for some moduli
    m = 199
    x = np.arange(-1000, 1000, dtype=np.int64)
    c16 = (1 << 16) % m
    sx = (x >> 16) * c16 + (x & 0xFFFF) + (m << 14)
    tx = sx - f32_quotient(sx, m) * m
    wrong = np.minimum(tx.astype(np.uint32), (tx + m).astype(np.uint32)).astype(np.int64)
    assert not np.array_equal(wrong, x % m)
