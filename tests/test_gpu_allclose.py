"""rt.allclose on the device (OpAllCloseAPI, isclose.rs:92-106) against the oracle's restatement
over every dtype, strided and broadcast layouts (RowMajor and ColMajor broadcasting), special
values (inf - inf never close, an infinite b makes the bound infinite, NaN only with equal_nan)
and wrapping integer differences; zero size raises."""
import zlib

import numpy as np
import pytest

import rstsr_amd as rt
from gpu_util import from_np
from oracle_ffi import View

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32, np.int32, np.int64, np.complex128, np.complex64]


def base_pair(rng, dt, n):
    dt = np.dtype(dt)
    if dt.kind == "i":
        info = np.iinfo(dt)
        a = rng.integers(-1000, 1000, n).astype(dt)
        a[:2] = [info.max, info.min]
        return a, a.copy()
    a = rng.normal(size=n)
    if dt.kind == "c":
        a = a + 1j * rng.normal(size=n)
    return a.astype(dt), a.astype(dt)


def perturbations(dt):
    dt = np.dtype(dt)
    if dt.kind == "i":
        info = np.iinfo(dt)
        return [("none", None), ("one", 1), ("big", 1000), ("wrap", (0, info.min))]
    return [("none", None), ("tiny", 1e-9), ("small", 3e-6), ("large", 1e-3), ("nan_a", "nan_a"), ("nan_both", "nan_both"),
            ("inf_both", "inf_both"), ("inf_b", "inf_b"), ("inf_a", "inf_a")]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("equal_nan", [False, True])
def test_allclose_against_oracle(dev, oracle, dt, equal_nan):
    rng = np.random.default_rng(zlib.crc32(repr((str(dt), equal_nan)).encode()))
    for name, p in perturbations(dt):
        for layout in ("strided", "broadcast"):
            a, b = base_pair(rng, dt, 6 * 20 * 16)
            a = a.reshape(6, 20, 16)
            b = b.reshape(6, 20, 16)
            k = (3, 4, 5)  # row 4: seen by the stepped view (::2) and used as the broadcast row
            if p is not None:
                if isinstance(p, tuple):  # integer wrap: a = MAX-ish, b = MIN
                    a[k], b[k] = np.iinfo(dt).max, p[1]
                elif isinstance(p, str):
                    if p in ("nan_a", "nan_both"):
                        a[k] = np.nan
                    if p == "nan_both":
                        b[k] = np.nan
                    if p in ("inf_both", "inf_a"):
                        a[k] = np.inf
                    if p in ("inf_both", "inf_b"):
                        b[k] = np.inf
                else:
                    b[k] = b[k] + np.asarray(p, dtype=dt) * (1 if np.dtype(dt).kind == "i" else abs(b[k]) + 1)
            if layout == "strided":
                x, y = a.transpose(2, 0, 1)[:, :, ::2], b.transpose(2, 0, 1)[:, :, ::2]
            else:  # a [6,5,16] (row 4 repeated) against b's row 4 as a [6,1,16] broadcast
                x = np.broadcast_to(a[:, 4:5, :], (6, 5, 16))
                y = b[:, 4:5, :]
            ta, va = from_np(dev, np.ascontiguousarray(x) if layout == "broadcast" else x)
            tb, vb = from_np(dev, y)
            got = rt.allclose(ta, tb, equal_nan=equal_nan)
            xb, yb = np.broadcast_arrays(np.asarray(x), np.asarray(y))
            ref = oracle.allclose(View.of(xb), View.of(yb), equal_nan=equal_nan)
            assert got == ref, (name, layout)


def test_allclose_colmajor_broadcast(dev):
    """Under ColMajor the tensor layer broadcasts leading axes: [2,3] against [2] compares row i
    with b[i]."""
    d = dev.clone()
    d.set_default_order("ColMajor")
    a = rt.asarray(np.array([[1.0, 1.0, 1.0], [2.0, 2.0, 2.0]]), device=d)
    assert rt.allclose(a, rt.asarray(np.array([1.0, 2.0]), device=d))
    assert not rt.allclose(a, rt.asarray(np.array([2.0, 1.0]), device=d))


def test_allclose_zero_size_raises(dev):
    z = rt.asarray(np.zeros((0, 3)), device=dev)
    with pytest.raises(rt.RSTSRError):
        rt.allclose(z, z)
