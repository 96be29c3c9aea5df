"""Every configuration the GEMM autotuner can pick, and every INT8 tile of the split GEMM, is run
through parity.

A first process tunes each problem class and records its choices in a cache file (the tuner's own
format, RSTSR_HIP_TUNE_CACHE); fresh processes then run the same problems with every recorded
choice forced to each candidate index. Each run is checked against the Dot2 / extended-precision
oracle (componentwise K*u*|A||B| for the plain kernels, normwise 1e-14 for the split GEMM), and
all candidates must give the same bits: every configuration accumulates each output over K in
the same order, so the tile shape changes no result (the split GEMM's residue sums are exact
integers). That property is what makes the tuner's choice invisible to callers.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[1] + "/tests")
import rstsr_amd as rt
from oracle_ffi import Oracle
from gpu_util import from_np
out_path, modes = sys.argv[2], sys.argv[3].split(",")
orc = Oracle()
dev = rt.DeviceHIP(0)
rng = np.random.default_rng(4242)
res = {}
bad = []

def wide(shape):
    return rng.normal(size=shape) * np.exp2(rng.integers(-8, 9, size=shape))

def comp_ok(got, A, B, ref, u=2.0 ** -53):
    K = A.shape[-1]
    return bool(np.all(np.abs(got - ref) <= K * u * (np.abs(A) @ np.abs(B)) + 1e-300))

f64_shapes = [(64, 64, 64), (200, 136, 96), (512, 384, 256), (1024, 1024, 512), (33, 700, 129)]
for i, (m, n, k) in enumerate(f64_shapes):
    A, B = wide((m, k)), wide((k, n))
    for tr in ("nn", "tn"):
        Ah = np.ascontiguousarray(A.T).T if tr[0] == "t" else A
        ta, _ = from_np(dev, Ah)
        tb = rt.asarray(B, device=dev)
        ref = orc.gemm_dot2(A, B)
        for mode in modes:
            got = rt.matmul(ta, tb, mode=mode).to_numpy()
            res[f"d{i}{tr}{mode}"] = got
            ok = comp_ok(got, A, B, ref) if mode == "plain" else \
                np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))
            if not ok:
                bad.append(f"d{i}{tr}{mode}")
# symmetric product (SYRK shortcut): X^T X in one buffer
X = rng.normal(size=(300, 260))
tX = rt.asarray(X, device=dev)
ref = orc.gemm_dot2(np.ascontiguousarray(X.T), X)
for mode in modes:
    got = rt.matmul(tX.T, tX, mode=mode).to_numpy()
    res[f"syrk{mode}"] = got
    ok = comp_ok(got, X.T, X, ref) if mode == "plain" else np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))
    if not ok:
        bad.append(f"syrk{mode}")
# complex
for i, (m, n, k) in enumerate([(64, 128, 80), (300, 200, 150), (512, 512, 256)]):
    A = rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))
    B = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    ref = A.astype(np.clongdouble) @ B.astype(np.clongdouble)
    bound = k * 2.0 ** -53 * 2 * (np.abs(A) @ np.abs(B))
    for mode in modes:
        got = rt.matmul(rt.asarray(A, device=dev), rt.asarray(B, device=dev), mode=mode).to_numpy()
        res[f"z{i}{mode}"] = got
        ok = bool(np.all(np.abs(got - ref) <= bound)) if mode == "plain" else \
            np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))
        if not ok:
            bad.append(f"z{i}{mode}")
# grouped (plain persistent kernel: its own tuner key)
for dt in (np.float64, np.complex128):
    sizes = [(int(rng.integers(4, 40)) * 8, int(rng.integers(4, 40)) * 8, int(rng.integers(4, 40)) * 8)
             for _ in range(12)]
    probs, hosts = [], []
    for (m, n, k) in sizes:
        A, B = rng.normal(size=(m, k)), rng.normal(size=(k, n))
        if dt == np.complex128:
            A = A + 1j * rng.normal(size=(m, k))
            B = B + 1j * rng.normal(size=(k, n))
        probs.append((rt.asarray(A, device=dev), rt.asarray(B, device=dev), rt.empty((m, n), dtype=dt, device=dev)))
        hosts.append((A, B))
    rt.gemm_grouped(probs, mode="plain")
    for j, ((A, B), (_, _, tc)) in enumerate(zip(hosts, probs)):
        got = tc.to_numpy()
        ref = A.astype(np.clongdouble) @ B.astype(np.clongdouble)
        bound = A.shape[1] * 2.0 ** -53 * 2 * (np.abs(A) @ np.abs(B))
        res[f"g{np.dtype(dt).char}{j}"] = got
        if not np.all(np.abs(got - ref) <= bound):
            bad.append(f"g{np.dtype(dt).char}{j}")
np.savez(out_path, **res)
dev.synchronize()
print("GEMM_OK" if not bad else "GEMM_BAD " + " ".join(bad), flush=True)
"""


def _child(env, out, modes):
    r = subprocess.run([sys.executable, "-c", _CHILD, ROOT, str(out), modes], env=env, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-3000:]
    return r.stdout


def _same_bits(p, q):
    a, b = np.load(p), np.load(q)
    assert sorted(a.files) == sorted(b.files)
    diff = [k for k in a.files if not np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8))]
    return diff


def test_gemm_tuner_candidates_all_in_parity(tmp_path):
    cache = tmp_path / "tune.txt"
    env = dict(os.environ, RSTSR_HIP_TUNE_CACHE=str(cache))
    env.pop("RSTSR_HIP_NO_TUNE", None)
    first = tmp_path / "tuned.npz"
    out = _child(env, first, "plain")
    assert "GEMM_OK" in out, out
    keys = [ln.split()[0] for ln in cache.read_text().splitlines() if ln.startswith(("gemm:", "grouped:"))]
    assert any(k.startswith("gemm:d") for k in keys) and any(k.startswith("gemm:z") for k in keys)
    assert any(k.startswith("grouped:") for k in keys)

    def ncand(key):
        if key.startswith("gemm:d"):
            return 4
        if key.startswith("gemm:z"):
            return 1 if int(key.rsplit(":f", 1)[1]) & 8 else 3
        return 4 if key.startswith("grouped:d") else 2

    for c in range(4):
        cache.write_text("".join(f"{k} {min(c, ncand(k) - 1)}\n" for k in keys))
        forced = tmp_path / f"forced{c}.npz"
        out = _child(env, forced, "plain")
        assert "GEMM_OK" in out, (c, out)
        assert _same_bits(first, forced) == [], c


def test_ozaki_int8_tiles_in_parity(tmp_path):
    """Each INT8 tile shape of the split GEMM: normwise 1e-14 against Dot2, and the same bits as
    the default tile."""
    base = dict(os.environ, RSTSR_HIP_NO_TUNE="1")
    base.pop("RSTSR_HIP_OZAKI_TILE", None)
    ref = tmp_path / "default.npz"
    out = _child(base, ref, "ozaki")
    assert "GEMM_OK" in out, out
    for tile in ("square", "wide8", "big"):
        forced = tmp_path / f"{tile}.npz"
        out = _child(dict(base, RSTSR_HIP_OZAKI_TILE=tile), forced, "ozaki")
        assert "GEMM_OK" in out, (tile, out)
        assert _same_bits(ref, forced) == [], tile
