"""Benchmark of the strided elementwise / reduction / copy path on one MI355X.

Workload (the bandwidth-bound strided configuration): X f64/f32 [256,512,1024] ~ N(0,1),
b [1,512,1] ~ U(0.5,2), Xi i32 [256,512,1024] ~ U{-2^20..2^20}, bi [1,512,1] likewise.
V = X.permute(2,0,1)[:, :, ::2] ([1024,256,256], strides (1, 524288, 2048)), bV the same
view of b. One step runs the whole suite once on device-resident inputs:

    V*bV (f64), V*bV (f32), Vi+bVi (i32),
    sum / max / argmax of V over axis 0 and over axes (1,2),
    argmax of Vi over axis 0 and over axes (1,2),
    X.permute(2,0,1).to_contig()                            (f64, 2.15 GB moved)

value = algorithmic bytes of one step / measured step time (GB/s). The FP64 GEMM leg
(8192^3, entries N(0,1)*2^u, u in [-8,8]) is reported beside it with its normwise error
against the Dot2 oracle on sampled rows.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-gemm] [--no-cpu] [--no-check]
                    [--only OPNAME]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "HBM GB/s (elemwise/reduce); f64 GEMM TFLOP/s plain vs split at ≤1e-14"
HBM_PEAK = 8000.0  # GB/s, MI355X
FP64_PEAK = 78.6  # TFLOP/s dense FP64 MFMA
INT8_PEAK = 5000.0  # TOP/s dense INT8 MFMA
MODULI = 14  # INT8 products per f64 product in the split GEMM (its default modulus count)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-gemm", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-check", action="store_true")
    p.add_argument("--only", default=None, help="run only this op (for profiling)")
    p.add_argument("--gemm-n", type=int, default=8192)
    p.add_argument("--gemm-modes", default="plain,ozaki")
    p.add_argument("--gemm-only", action="store_true", help="skip the strided suite (profiling the GEMM legs)")
    p.add_argument("--gemm-legs", default="dense,batch,ri,grouped,chain")
    p.add_argument("--no-transfer", action="store_true")
    p.add_argument("--no-small", action="store_true", help="skip the small-configuration leg (profiling)")
    return p.parse_args()


class Op:
    def __init__(self, name, fn, nbytes, in_step=True):
        self.name, self.fn, self.nbytes, self.in_step = name, fn, nbytes, in_step


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(collectives and sharding are design-only). Run with --gpus 1.", flush=True)
        sys.exit(2)

    import rstsr_amd as rt
    from rstsr_amd import abi
    from rstsr_amd import _rstsr as RL

    L = abi.lib()
    dev = rt.DeviceHIP(0)
    if args.gemm_only:
        print(json.dumps({"gemm": gemm_leg(args, dev, rt)}), flush=True)
        return
    ctx = C.c_void_p(dev._ctx_ptr())
    ev = abi.Events(dev._ctx_ptr())

    # ---------------- inputs (seed 20261202 + config index 1) ----------------
    t0 = time.time()
    rng = np.random.default_rng(20261203)
    X = rng.standard_normal((256, 512, 1024))
    b = rng.uniform(0.5, 2.0, size=(1, 512, 1))
    Xi = rng.integers(-(2 ** 20), 2 ** 20, size=(256, 512, 1024), dtype=np.int32, endpoint=True)
    bi = rng.integers(-(2 ** 20), 2 ** 20, size=(1, 512, 1), dtype=np.int32, endpoint=True)
    X32 = X.astype(np.float32)
    b32 = b.astype(np.float32)
    log(f"[bench] inputs generated in {time.time() - t0:.1f}s")
    tX, tX32, tXi = rt.asarray(X, device=dev), rt.asarray(X32, device=dev), rt.asarray(Xi, device=dev)
    tb, tb32, tbi = rt.asarray(b, device=dev), rt.asarray(b32, device=dev), rt.asarray(bi, device=dev)
    log(f"[bench] inputs resident in HBM after {time.time() - t0:.1f}s")

    def strided(t):
        return t.transpose([2, 0, 1])[:, :, ::2]

    V, V32, Vi = strided(tX), strided(tX32), strided(tXi)
    bV, bV32, bVi = strided(tb), strided(tb32), strided(tbi)
    shapeV = list(V.shape)  # [1024, 256, 256]

    def lay(t):
        sh, st, off = t.layout
        return abi.make_layout(sh, st, off)

    def ptr(t):
        return C.c_void_p(t._t._data_ptr())

    # broadcast views of b to V's shape
    def bcast(t):
        sh, st, off = RL.broadcast_layout(V.layout, t.layout, "C")[1]
        return abi.make_layout(sh, st, off)

    outc = rt._rstsr.c_layout(shapeV)
    Y64 = rt.empty(shapeV, np.float64, device=dev)
    Y32 = rt.empty(shapeV, np.float32, device=dev)
    Yi = rt.empty(shapeV, np.int32, device=dev)
    lY = abi.make_layout(*outc)

    def red_out(t, axes, dtype):
        _, lm = RL.dim_split_axes(t.layout, axes)
        lo = RL.layout_for_array_copy(lm, "K")
        o = rt.Tensor(RL.empty(dev, lo, rt._code(dtype)))
        return o, abi.make_layout(*lo)

    red = {}
    for nm, t, axes, dt in [("f64_ax0", V, [0], np.float64), ("f64_ax12", V, [1, 2], np.float64),
                            ("u64_ax0", V, [0], np.uint64), ("u64_ax12", V, [1, 2], np.uint64),
                            ("max_ax0", V, [0], np.float64), ("max_ax12", V, [1, 2], np.float64),
                            ("i_ax0", Vi, [0], np.uint64), ("i_ax12", Vi, [1, 2], np.uint64)]:
        red[nm] = red_out(t, axes, dt)
    # unary outputs laid out as the tensor API allocates them: layout_for_array_copy(V, K)
    lUo = RL.layout_for_array_copy(V.layout, "K")
    U64 = rt.Tensor(RL.empty(dev, lUo, rt._code(np.float64)))
    lU = abi.make_layout(*lUo)
    P = tX.transpose([2, 0, 1])
    Pout = rt.empty(list(P.shape), np.float64, device=dev)
    Cflat = rt.empty([256, 512, 1024], np.float64, device=dev)

    lV, lV32, lVi = lay(V), lay(V32), lay(Vi)
    lbV, lbV32, lbVi = bcast(bV), bcast(bV32), bcast(bVi)
    lP, lPout = lay(P), abi.make_layout(*rt._rstsr.c_layout(list(P.shape)))
    lX, lC = lay(tX), abi.make_layout(*rt._rstsr.c_layout([256, 512, 1024]))
    ax0 = (C.c_ssize_t * 1)(0)
    ax12 = (C.c_ssize_t * 2)(1, 2)

    F32, F64, I32 = 0, 1, 2
    ADD, MUL = 0, 2
    SUM, MAX, ARGMAX = 0, 1, 3
    nV = 1024 * 256 * 256

    def mk_bin(op, t, c, lc, a, la, bb, lb):
        return lambda: abi.check(L.rstsr_hip_op_mutc_refa_refb(ctx, op, t, ptr(c), C.byref(lc), ptr(a), C.byref(la),
                                                               ptr(bb), C.byref(lb)))

    def mk_red(op, t, a, la, axes, na, out):
        o, lo = out
        f = L.rstsr_hip_argreduce_axes if op == ARGMAX else L.rstsr_hip_reduce_axes
        return lambda: abi.check(f(ctx, op, t, ptr(a), C.byref(la), na, axes, ptr(o), C.byref(lo)))

    UN_EXP, UN_SIN, UN_SQRT = 10, 24, 27

    def mk_un(op, c, lc, a, la):
        return lambda: abi.check(L.rstsr_hip_unary(ctx, op, F64, ptr(c), C.byref(lc), F64, ptr(a), C.byref(la)))

    ops = [
        Op("mul_f64", mk_bin(MUL, F64, Y64, lY, tX, lV, tb, lbV), nV * 16 + 512 * 8),
        Op("mul_f32", mk_bin(MUL, F32, Y32, lY, tX32, lV32, tb32, lbV32), nV * 8 + 512 * 4),
        Op("add_i32", mk_bin(ADD, I32, Yi, lY, tXi, lVi, tbi, lbVi), nV * 8 + 512 * 4),
        Op("sum_ax0_f64", mk_red(SUM, F64, tX, lV, ax0, 1, red["f64_ax0"]), nV * 8 + 65536 * 8),
        Op("sum_ax12_f64", mk_red(SUM, F64, tX, lV, ax12, 2, red["f64_ax12"]), nV * 8 + 1024 * 8),
        Op("max_ax0_f64", mk_red(MAX, F64, tX, lV, ax0, 1, red["max_ax0"]), nV * 8 + 65536 * 8),
        Op("max_ax12_f64", mk_red(MAX, F64, tX, lV, ax12, 2, red["max_ax12"]), nV * 8 + 1024 * 8),
        Op("argmax_ax0_f64", mk_red(ARGMAX, F64, tX, lV, ax0, 1, red["u64_ax0"]), nV * 8 + 65536 * 8),
        Op("argmax_ax12_f64", mk_red(ARGMAX, F64, tX, lV, ax12, 2, red["u64_ax12"]), nV * 8 + 1024 * 8),
        Op("argmax_ax0_i32", mk_red(ARGMAX, I32, tXi, lVi, ax0, 1, red["i_ax0"]), nV * 4 + 65536 * 8),
        Op("argmax_ax12_i32", mk_red(ARGMAX, I32, tXi, lVi, ax12, 2, red["i_ax12"]), nV * 4 + 1024 * 8),
        Op("to_contig_f64", lambda: abi.check(L.rstsr_hip_assign(ctx, F64, ptr(Pout), C.byref(lPout), F64, ptr(tX),
                                                                 C.byref(lP))), 2 * 8 * 256 * 512 * 1024),
        # unary ops on the strided view (not part of the step): sqrt is a single IEEE op, exp and
        # sin are the device's f64 math library
        Op("sqrt_f64", mk_un(UN_SQRT, U64, lU, tX, lV), nV * 16, in_step=False),
        Op("exp_f64", mk_un(UN_EXP, U64, lU, tX, lV), nV * 16, in_step=False),
        Op("sin_f64", mk_un(UN_SIN, U64, lU, tX, lV), nV * 16, in_step=False),
        # calibration: contiguous copy of the same 1.07 GB (not part of the step)
        Op("contig_copy_f64", lambda: abi.check(L.rstsr_hip_assign(ctx, F64, ptr(Cflat), C.byref(lC), F64, ptr(tX),
                                                                   C.byref(lX))), 2 * 8 * 256 * 512 * 1024,
           in_step=False),
    ]
    if args.only:
        ops = [o for o in ops if o.name == args.only]
        if not ops:
            sys.exit(f"unknown op {args.only}")

    step_ops = [o for o in ops if o.in_step]
    step_bytes = sum(o.nbytes for o in step_ops)

    # ---------------- timed region ----------------
    for _ in range(args.warmup):
        for o in ops:
            o.fn()
    dev.synchronize()
    log("[bench] warmup done")
    K = args.steps
    ev.record(0)
    for _ in range(K):
        for o in step_ops:
            o.fn()
    ev.record(1)
    dev.synchronize()
    step_ms = ev.elapsed_ms(0, 1) / max(K, 1)
    log(f"[bench] {K} steps, {step_ms:.3f} ms/step")

    per_op = {}
    for o in ops:
        ev.record(2)
        for _ in range(K):
            o.fn()
        ev.record(3)
        ms = ev.elapsed_ms(2, 3) / max(K, 1)
        per_op[o.name] = {"ms": round(ms, 4), "GB_per_launch": round(o.nbytes / 1e9, 4),
                          "GB/s": round(o.nbytes / 1e9 / (ms / 1e3), 1),
                          "frac_hbm": round(o.nbytes / 1e9 / (ms / 1e3) / HBM_PEAK, 3)}
    dev.synchronize()
    log("[bench] per-op: " + json.dumps(per_op))

    # ---------------- parity at full size ----------------
    parity = {}
    if not args.no_check and not args.only:
        t1 = time.time()
        Vn = X.transpose(2, 0, 1)[:, :, ::2]
        bVn = b.transpose(2, 0, 1)[:, :, ::2]
        parity["mul_f64_bitexact"] = bool(np.array_equal(Y64.to_numpy(), Vn * bVn))
        parity["mul_f32_bitexact"] = bool(np.array_equal(Y32.to_numpy(), X32.transpose(2, 0, 1)[:, :, ::2] * b32.transpose(2, 0, 1)[:, :, ::2]))
        parity["add_i32_bitexact"] = bool(np.array_equal(Yi.to_numpy(), Xi.transpose(2, 0, 1)[:, :, ::2] + bi.transpose(2, 0, 1)[:, :, ::2]))
        parity["to_contig_bitexact"] = bool(np.array_equal(Pout.to_numpy(), np.ascontiguousarray(X.transpose(2, 0, 1))))
        parity["argmax_ax0_bitexact"] = bool(np.array_equal(red["u64_ax0"][0].to_numpy(), np.argmax(Vn, axis=0)))
        parity["argmax_ax12_bitexact"] = bool(np.array_equal(red["u64_ax12"][0].to_numpy(), np.argmax(Vn.reshape(1024, -1), axis=1)))
        Vin = Xi.transpose(2, 0, 1)[:, :, ::2]
        parity["argmax_i32_ax0_bitexact"] = bool(np.array_equal(red["i_ax0"][0].to_numpy(), np.argmax(Vin, axis=0)))
        parity["argmax_i32_ax12_bitexact"] = bool(np.array_equal(red["i_ax12"][0].to_numpy(), np.argmax(Vin.reshape(1024, -1), axis=1)))
        parity["max_ax0_bitexact"] = bool(np.array_equal(red["max_ax0"][0].to_numpy(), Vn.max(axis=0)))
        parity["max_ax12_bitexact"] = bool(np.array_equal(red["max_ax12"][0].to_numpy(), Vn.reshape(1024, -1).max(axis=1)))
        s0 = np.sum(Vn.astype(np.longdouble), axis=0)
        parity["sum_ax0_normwise_err"] = float(np.max(np.abs(red["f64_ax0"][0].to_numpy() - s0)) / np.max(np.abs(s0)))
        s12 = np.sum(Vn.reshape(1024, -1).astype(np.longdouble), axis=1)
        parity["sum_ax12_normwise_err"] = float(np.max(np.abs(red["f64_ax12"][0].to_numpy() - s12)) / np.max(np.abs(s12)))
        parity["sum_tolerance"] = 1e-13
        parity["all_ok"] = all(v for k, v in parity.items() if k.endswith("bitexact")) and \
            parity["sum_ax0_normwise_err"] <= 1e-13 and parity["sum_ax12_normwise_err"] <= 1e-13
        log(f"[bench] parity checked in {time.time() - t1:.1f}s: {json.dumps(parity)}")

    # ---------------- roofline of the dominant kernel ----------------
    dom = max(step_ops, key=lambda o: per_op[o.name]["ms"]) if step_ops else ops[0]
    dms = per_op[dom.name]["ms"]
    achieved = dom.nbytes / 1e9 / (dms / 1e3)
    # HBM bytes per launch measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs, scaled by
    # the same counters on a contiguous copy of known size (profiles/pmc_traffic.json)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            t = json.load(f).get("ops", {}).get(dom.name)
        if t:
            traffic = t["hbm_bytes_per_launch"]
    roofline = {"bound": "hbm", "kernel": dom.name, "achieved": round(achieved, 1), "peak": HBM_PEAK,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                "algorithmic_bytes_per_launch": dom.nbytes, "avg_launch_ms": dms}
    if not args.only:
        ceil = _stream_ceilings()
        if ceil:
            roofline["stream_ceilings"] = ceil
            roofline["frac_of_copy_ceiling"] = round(achieved / ceil["copy_GBps"], 4)
            log("[bench] HBM stream ceilings: " + json.dumps(ceil))

    # ---------------- host <-> device transfers (PCIe-inclusive; not part of the step) ----------------
    transfer = {}
    if not args.no_transfer and not args.only:
        transfer = transfer_leg(dev)
        log("[bench] transfer: " + json.dumps(transfer))

    # ---------------- the small configuration, whole (GPU and host port) ----------------
    small = {}
    if not args.only and not args.no_small:
        small = c1_leg(dev, rt, args)

    # ---------------- FP64 GEMM leg ----------------
    gemm = {}
    if not args.no_gemm and not args.only:
        gemm = gemm_leg(args, dev, rt)

    # ---------------- CPU baseline (oracle port of the cpu_rayon control flow) ----------------
    cpu = None
    if not args.no_cpu and not args.only:
        cpu = cpu_leg(X, b, Xi, bi)

    value = step_bytes / 1e9 / (step_ms / 1e3)
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "strided broadcast elementwise + axis reductions + permuted copy: "
                               "X[256,512,1024] viewed permute(2,0,1)[:,:,::2], b[1,512,1]; f64/f32/i32",
                   "step_bytes": step_bytes, "ops": [o.name for o in step_ops]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "ops": per_op,
        "parity": parity,
        "small_config": small,
        "gemm": gemm,
        "transfer": transfer,
    }
    print(json.dumps(out), flush=True)


def _stream_ceilings(nbytes=1 << 30):
    """Best GB/s of a contiguous 16-byte-lane copy, read-only and write-only pass over 1 GiB, with
    ordinary and with nontemporal accesses (tests/gpu_check/stream_ceiling.hip): what this GPU's
    HBM sustains for each traffic mix."""
    path = os.path.join(ROOT, "tests", "gpu_check", "libstreamcheck.so")
    if not os.path.exists(path):
        log("[bench] stream-ceiling probe not built; skipped")
        return None
    lib = C.CDLL(path)
    lib.stream_ceiling.argtypes = [C.c_int, C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_float)]
    out = {}
    for kind, name, mult in ((0, "copy", 2), (1, "read", 1), (2, "write", 1)):
        for nt in (0, 1):
            ms = C.c_float(0)
            rc = lib.stream_ceiling(kind, nt, nbytes, 10, C.byref(ms))
            if rc != 0:
                raise RuntimeError(f"stream_ceiling({kind}, {nt}) failed with HIP error {rc}")
            out[f"{name}_{'nt' if nt else 'plain'}_GBps"] = round(mult * nbytes / 1e9 / (ms.value / 1e3), 1)
        # the ceiling: the better of ordinary and nontemporal accesses
        out[f"{name}_GBps"] = max(out[f"{name}_plain_GBps"], out[f"{name}_nt_GBps"])
    out["bytes_per_buffer"] = nbytes
    return out


def _timed(ev, fn, reps, warmups=3):
    for _ in range(warmups):
        fn()  # warmup (also settles the autotuner)
    ev.record(4)
    for _ in range(reps):
        fn()
    ev.record(5)
    return ev.elapsed_ms(4, 5) / reps


def _gemm_rates(flop, ms, mode, int8_ops):
    """Rates of one GEMM call: effective FP64 TFLOP/s, and for the split mode the INT8 operations
    the call issues per second; each with its roofline fraction."""
    tf = flop / (ms / 1e3) / 1e12
    r = {"ms": round(ms, 3), "TFLOP/s": round(tf, 2)}
    if mode == "plain":
        r["roofline"] = {"bound": "mfma", "unit": "TFLOP/s (fp64)", "achieved": round(tf, 2), "peak": FP64_PEAK,
                         "frac": round(tf / FP64_PEAK, 3)}
    else:
        tops = int8_ops / (ms / 1e3) / 1e12
        r["fp64_equivalent_frac"] = round(tf / FP64_PEAK, 3)
        r["roofline"] = {"bound": "mfma", "unit": "TOP/s (int8, whole call)", "achieved": round(tops, 1),
                         "peak": INT8_PEAK, "frac": round(tops / INT8_PEAK, 3), "int8_ops": int8_ops}
    return r


def _np_gemm_rate(fn, flop, sample, min_s=2.0):
    """NumPy matmul (its bundled OpenBLAS, all host threads) on a bounded sample of a GEMM leg: a
    labelled stand-in for DeviceFaer's faer GEMM, which cannot be built here (no Rust toolchain)."""
    fn()  # thread pool and page warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        el = time.perf_counter() - t0
        if el >= min_s or reps >= 20:
            break
    tf = flop * reps / el / 1e12
    return {"value": round(tf, 3), "unit": "TFLOP/s", "cores": _host_threads(), "kind": "numpy-openblas",
            "sample": sample, "reps": reps}


def _host_threads():
    for k in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS"):
        if os.environ.get(k, "").isdigit():
            return int(os.environ[k])
    return os.cpu_count() or 1


def gemm_leg(args, dev, rt):
    """Dense FP64 GEMM configurations: plain FP64 MFMA vs Ozaki-II INT8 split, error vs Dot2."""
    from rstsr_amd import abi

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from dot2_gpu import dot2_matmul, normwise

    ev = abi.Events(dev._ctx_ptr())
    modes = args.gemm_modes.split(",")
    out = {}

    def wide(rng, shape):
        return rng.standard_normal(shape) * np.exp2(rng.integers(-8, 9, size=shape))

    legs = set(args.gemm_legs.split(","))
    # ---- square n^3 with wide exponents ----
    n = args.gemm_n
    rng = np.random.default_rng(20261204)
    t0 = time.time()
    A = wide(rng, (n, n))
    B = wide(rng, (n, n))
    tA, tB = rt.asarray(A, device=dev), rt.asarray(B, device=dev)
    log(f"[bench] gemm inputs {n}^3 ready in {time.time() - t0:.1f}s")
    res = {"shape": [n, n, n], "flop": 2.0 * n ** 3}
    if "dense" in legs:
        t0 = time.time()
        ref = dot2_matmul(rt, dev, tA, tB)
        log(f"[bench] dense Dot2 reference on the GPU in {time.time() - t0:.1f}s")
    for mode in (modes if "dense" in legs else []):
        holder = {}

        def run():
            holder["c"] = rt.matmul(tA, tB, mode=mode)

        ms = _timed(ev, run, 3)
        err = normwise(dev, holder["c"], ref)
        res[mode] = _gemm_rates(res["flop"], ms, mode, MODULI * res["flop"])
        res[mode]["normwise_err_full"] = err
        log(f"[bench] dense {n}^3 {mode}: {ms:.2f} ms {res[mode]['TFLOP/s']:.2f} TFLOP/s err {err:.2e}")
        holder.clear()
    if "dense" in legs and not args.no_cpu:
        res["cpu_baseline"] = _np_gemm_rate(lambda: A[:512] @ B, 2.0 * 512 * n * n, "A[:512] @ B")
    out["dense_square"] = res
    if "dense" in legs:
        del ref
    del tA, tB, A, B

    # ---- broadcast batch [64,1024,1024] % [1024,1024] (rule 6) ----
    rng = np.random.default_rng(20261205)
    A3 = wide(rng, (64, 1024, 1024))
    B2 = wide(rng, (1024, 1024))
    tA3, tB2 = rt.asarray(A3, device=dev), rt.asarray(B2, device=dev)
    res = {"shape": [64, 1024, 1024, 1024], "flop": 64 * 2.0 * 1024 ** 3}
    if "batch" in legs:
        # the batch of row blocks is one 65536 x 1024 operand (A3 is contiguous)
        ref = dot2_matmul(rt, dev, tA3.reshape([65536, 1024]), tB2)
    for mode in (modes if "batch" in legs else []):
        holder = {}

        def run():
            holder["c"] = rt.matmul(tA3, tB2, mode=mode)

        ms = _timed(ev, run, 3)
        err = normwise(dev, holder["c"].reshape([65536, 1024]), ref)
        res[mode] = _gemm_rates(res["flop"], ms, mode, MODULI * res["flop"])
        res[mode]["normwise_err_full"] = err
        log(f"[bench] batch64 {mode}: {ms:.2f} ms {res[mode]['TFLOP/s']:.2f} TFLOP/s err {err:.2e}")
        holder.clear()
    if "batch" in legs and not args.no_cpu:
        res["cpu_baseline"] = _np_gemm_rate(lambda: A3[:4] @ B2, 4 * 2.0 * 1024 ** 3, "A3[:4] @ B2 (4 of 64 batches)")
    out["batched_broadcast"] = res
    if "batch" in legs:
        del ref
    del tA3, tB2, A3

    # ---- RI contraction (ia|jb) = sum_P B[P,i,a] B[P,j,b], B [1600,60,400] ~ N(0,0.1) ----
    rng = np.random.default_rng(20261206)
    Bri = rng.normal(0.0, 0.1, size=(1600, 60, 400))
    tB = rt.asarray(Bri, device=dev)
    lhs = tB.permute([1, 2, 0])            # [60, 400, 1600]  (i, a, P)
    rhs = tB.reshape([1600, 24000])        # [1600, 24000]    (P, (j, b))
    res = {"shape": [60, 400, 24000, 1600], "flop": 60 * 2.0 * 400 * 24000 * 1600}
    if "ri" in legs:
        # (i, a) rows of lhs merge into one axis of unit stride: a 24000 x 1600 operand
        ref = dot2_matmul(rt, dev, tB._view(([24000, 1600], [1, 24000], 0)), rhs)
    for mode in (modes if "ri" in legs else []):
        holder = {}

        def run():
            holder["c"] = rt.matmul(lhs, rhs, mode=mode)

        ms = _timed(ev, run, 2)
        err = normwise(dev, holder["c"].reshape([24000, 24000]), ref)
        # the symmetric (SYRK) shortcut computes the lower triangle only; rates count the full product
        res[mode] = _gemm_rates(res["flop"], ms, mode, MODULI * res["flop"])
        res[mode]["normwise_err_full"] = err
        log(f"[bench] RI {mode}: {ms:.2f} ms {res[mode]['TFLOP/s']:.2f} TFLOP/s err {err:.2e}")
        holder.clear()
    if "ri" in legs and not args.no_cpu:
        Bm = Bri.reshape(1600, 24000)
        res["cpu_baseline"] = _np_gemm_rate(lambda: Bri[:, 5, :].T @ Bm, 2.0 * 400 * 24000 * 1600,
                                            "B[:, 5, :].T @ B.reshape(1600, 24000) (1 of 60 batches)")
    out["ri_contraction"] = res
    if "ri" in legs:
        del ref
    del tB, lhs, rhs, Bri

    # ---- 512 uneven GEMMs, m,n,k ~ U{32..1024 step 8}, one grouped launch ----
    rng = np.random.default_rng(20261207)
    dims = rng.integers(4, 129, size=(512, 3)) * 8
    pool = rng.standard_normal(1 << 24)  # 128 MiB operand pool; operands are views into it
    tpool = rt.asarray(pool, device=dev)
    offs = rng.integers(0, (1 << 24) - 1024 * 1024, size=(512, 2))
    tot_c = int(sum(int(m) * int(n) for m, n, _ in dims))
    tC = rt.empty([tot_c], np.float64, device=dev)
    probs, coff = [], 0
    for g, (m, n, k) in enumerate(dims):
        m, n, k = int(m), int(n), int(k)
        ta = tpool._view(([m, k], [k, 1], int(offs[g, 0])))
        tb = tpool._view(([k, n], [n, 1], int(offs[g, 1])))
        tc = tC._view(([m, n], [n, 1], coff))
        coff += m * n
        probs.append((ta, tb, tc))
    flop = float(sum(2.0 * int(m) * int(n) * int(k) for m, n, k in dims))
    res = {"count": 512, "flop": flop}
    refs = [dot2_matmul(rt, dev, ta, tb) for ta, tb, _ in probs] if "grouped" in legs else []
    for mode in (modes if "grouped" in legs else []):
        # six warm-up calls: the error checks of the previous mode leave the GPU idle long
        # enough for its clock to drop, and it takes about six 3.5 ms calls to come back
        ms = _timed(ev, lambda: rt.gemm_grouped(probs, mode=mode), 3, warmups=6)
        # every problem, every element: the largest normwise error over the 512 products
        err = max(normwise(dev, tc, r) for (_, _, tc), r in zip(probs, refs))
        res[mode] = _gemm_rates(flop, ms, mode, MODULI * flop)
        res[mode]["normwise_err_full_max_over_problems"] = err
        log(f"[bench] grouped512 {mode}: {ms:.2f} ms {res[mode]['TFLOP/s']:.2f} TFLOP/s err {err:.2e}")
    if "grouped" in legs and "ozaki" in modes:
        # the split mode routes each problem (none of these pays for the split); the grouped
        # split kernels forced on every problem, for the record
        os.environ["RSTSR_HIP_OZAKI_GROUPED"] = "split"
        try:
            ms = _timed(ev, lambda: rt.gemm_grouped(probs, mode="ozaki"), 3, warmups=6)
            err = max(normwise(dev, tc, r) for (_, _, tc), r in zip(probs, refs))
        finally:
            del os.environ["RSTSR_HIP_OZAKI_GROUPED"]
        res["ozaki_all_split"] = _gemm_rates(flop, ms, "ozaki", MODULI * flop)
        res["ozaki_all_split"]["normwise_err_full_max_over_problems"] = err
        log(f"[bench] grouped512 ozaki, every problem split: {ms:.2f} ms "
            f"{res['ozaki_all_split']['TFLOP/s']:.2f} TFLOP/s err {err:.2e}")
    if "grouped" in legs and not args.no_cpu:
        sel = list(range(0, 512, 16))

        def np_group():
            for g in sel:
                m, n, k = (int(x) for x in dims[g])
                pool[offs[g, 0]:offs[g, 0] + m * k].reshape(m, k) @ pool[offs[g, 1]:offs[g, 1] + k * n].reshape(k, n)

        res["cpu_baseline"] = _np_gemm_rate(np_group, float(sum(2.0 * int(m) * int(n) * int(k) for m, n, k in dims[sel])),
                                            "32 of the 512 problems (every 16th), one NumPy call each")
    out["uneven_grouped"] = res
    del probs, tpool, tC, refs

    # ---- hipGraph-captured c128 chain, 20 steps ----
    if "chain" in legs:
        out["graph_chain_c128"] = chain_leg(dev, rt, modes)
    return out


def chain_reference(z1h, z2h, vh, steps):
    """s* of every step of Z1 += v; Z3 = Z1 * Z2; s = (Z3 % Z2).sum_axes(1), all rows.

    Z1 accumulates v in f64 exactly as the device does (IEEE adds); the row sums of the product
    use sum_j (Z3 Z2)[i, j] = sum_k Z3[i, k] w[k] with w = Z2.sum_axes(1), both in extended
    precision (64-bit significand), so each s* is far more accurate than the 1e-13 target."""
    z2l = z2h.astype(np.clongdouble)
    w = z2l.sum(axis=1)
    wr, wi = np.ascontiguousarray(w.real), np.ascontiguousarray(w.imag)
    z = z1h.copy()
    out = []
    for _ in range(steps):
        z = z + vh[None, :]
        z3 = z.astype(np.clongdouble) * z2l
        zr, zi = np.ascontiguousarray(z3.real), np.ascontiguousarray(z3.imag)
        s = (zr @ wr - zi @ wi) + 1j * (zr @ wi + zi @ wr)
        out.append(s.astype(np.complex128))
    return out


def chain_leg(dev, rt, modes, n=4096, steps=20):
    """Z1 += v; Z3 = Z1 * Z2; Z4 = Z3 % Z2; s = Z4.sum_axes(1): uncaptured vs one captured graph;
    the normwise error of s against the extended-precision reference at every step, both ways."""
    from rstsr_amd import abi

    ev = abi.Events(dev._ctx_ptr())
    rng = np.random.default_rng(20261208)
    z1h = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    z2h = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    vh = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Z2, V = rt.asarray(z2h, device=dev), rt.asarray(vh, device=dev)
    res = {"shape": [n, n], "steps": steps, "flop_per_step": 8.0 * n ** 3}
    t0 = time.time()
    s_ref = chain_reference(z1h, z2h, vh, steps)
    log(f"[bench] chain reference ({steps} steps, every row) in {time.time() - t0:.1f}s")

    def errs(ss):
        return [float(np.max(np.abs(x.to_numpy() - r)) / np.max(np.abs(r))) for x, r in zip(ss, s_ref)]

    for mode in modes:
        Z1 = rt.asarray(z1h, device=dev)
        kept = []

        def step():
            nonlocal Z1
            Z1 += V
            Z3 = Z1 * Z2
            Z4 = rt.matmul(Z3, Z2, mode=mode)
            kept.append(Z4.sum_axes([1]))

        step()  # settles autotuner / workspaces
        Z1.assign(rt.asarray(z1h, device=dev))
        kept.clear()
        dev.synchronize()
        t0 = time.perf_counter()
        ev.record(6)
        for _ in range(steps):
            step()
        ev.record(7)
        eager_ms = ev.elapsed_ms(6, 7) / steps
        wall_eager = (time.perf_counter() - t0) * 1e3 / steps
        err_eager = errs(kept)
        kept.clear()
        # captured: the whole 20-step chain as one graph
        Z1.assign(rt.asarray(z1h, device=dev))
        keep = []
        with rt.capture(dev) as g:
            for _ in range(steps):
                Z1 += V
                z3 = Z1 * Z2
                z4 = rt.matmul(z3, Z2, mode=mode)
                keep.append((z3, z4, z4.sum_axes([1])))
            del z3, z4
        Z1.assign(rt.asarray(z1h, device=dev))
        dev.synchronize()
        t0 = time.perf_counter()
        ev.record(6)
        g.launch()
        ev.record(7)
        graph_ms = ev.elapsed_ms(6, 7) / steps
        wall_graph = (time.perf_counter() - t0) * 1e3 / steps
        err_graph = errs([k[2] for k in keep])
        zg = 8.0 * n ** 3 / (graph_ms / 1e3) / 1e12
        res[mode] = {"eager_ms_per_step": round(eager_ms, 3), "graph_ms_per_step": round(graph_ms, 3),
                     "eager_wall_ms_per_step": round(wall_eager, 3), "graph_wall_ms_per_step": round(wall_graph, 3),
                     "launch_overhead_saved_ms_per_step": round(eager_ms - graph_ms, 4),
                     "err_per_step_eager": err_eager, "err_per_step_graph": err_graph,
                     "err_max_eager": max(err_eager), "err_max_graph": max(err_graph), "target_err": 1e-13,
                     "all_steps_within_target": max(err_eager + err_graph) <= 1e-13,
                     "zgemm_TFLOP/s_graph": round(zg, 2)}
        log(f"[bench] chain {mode}: eager {eager_ms:.3f} ms/step, graph {graph_ms:.3f} ms/step, "
            f"max err {max(err_eager):.2e}/{max(err_graph):.2e}")
        del keep, g
    return res


def c1_leg(dev, rt, args, reps=50):
    """The small configuration whole, on the GPU and on the host port: a f64 [64,128,256] ~ U(-1,1)
    + b [256] ~ U(-1,1) (row broadcast), s = c.sum_axes(1), i = c.argmax_axes(2), and
    C = A % B with A, B f64 [512,512] ~ N(0,1); seed 20261202. Parity: c and i bit for bit,
    s within 1e-13 normwise of an extended-precision sum, C against the oracle's Dot2 GEMM."""
    from rstsr_amd import abi

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ffi import Oracle, View, c_view

    rng = np.random.default_rng(20261202)
    a = rng.uniform(-1.0, 1.0, size=(64, 128, 256))
    b = rng.uniform(-1.0, 1.0, size=256)
    A = rng.standard_normal((512, 512))
    B = rng.standard_normal((512, 512))
    ta, tb, tA, tB = (rt.asarray(x, device=dev) for x in (a, b, A, B))
    ev = abi.Events(dev._ctx_ptr())
    hold = {}

    def add():
        hold["c"] = ta + tb

    def ssum():
        hold["s"] = hold["c"].sum_axes([1])

    def amax():
        hold["i"] = hold["c"].argmax_axes([2])

    def mm():
        hold["C"] = tA % tB

    res = {"workload": "a[64,128,256] + b[256]; sum over axis 1; argmax over axis 2; 512^3 matmul (f64)"}
    gpu = {}
    for nm, fn in (("add", add), ("sum_ax1", ssum), ("argmax_ax2", amax), ("matmul_512", mm)):
        gpu[nm + "_ms"] = round(_timed(ev, fn, reps), 4)

    def step():
        add()
        ssum()
        amax()
        mm()

    gpu["step_ms"] = round(_timed(ev, step, reps), 4)
    nbytes = (2 * a.size + b.size) * 8 + (a.size * 8 + 64 * 256 * 8) + (a.size * 8 + 64 * 128 * 8)
    gpu["elementwise_reduce_GBps"] = round(nbytes / 1e9 / ((gpu["add_ms"] + gpu["sum_ax1_ms"] + gpu["argmax_ax2_ms"]) / 1e3), 1)
    gpu["matmul_TFLOP/s"] = round(2.0 * 512 ** 3 / (gpu["matmul_512_ms"] / 1e3) / 1e12, 2)
    res["gpu"] = gpu

    c_ref = a + b
    par = {"add_bitexact": bool(np.array_equal(hold["c"].to_numpy(), c_ref)),
           "argmax_bitexact": bool(np.array_equal(hold["i"].to_numpy(), np.argmax(c_ref, axis=2)))}
    s_ref = np.sum(c_ref.astype(np.longdouble), axis=1)
    par["sum_normwise_err"] = float(np.max(np.abs(hold["s"].to_numpy() - s_ref)) / np.max(np.abs(s_ref)))
    orc = Oracle()
    C_ref = orc.gemm_dot2(A, B)
    par["matmul_normwise_err"] = float(np.max(np.abs(hold["C"].to_numpy() - C_ref)) / np.max(np.abs(C_ref)))
    par["all_ok"] = par["add_bitexact"] and par["argmax_bitexact"] and par["sum_normwise_err"] <= 1e-13 and \
        par["matmul_normwise_err"] <= 1e-13
    res["parity"] = par

    if not args.no_cpu:
        # the oracle's port of the cpu_rayon control flow on the full inputs, NumPy (OpenBLAS) for the matmul
        cores = _host_threads()
        lib = orc.lib
        bb = np.broadcast_to(b, a.shape)
        va, vb2 = View.of(a), View.of(bb)
        vc = c_view(a.shape, np.float64)
        vs = c_view([64, 256], np.float64)
        vi = c_view([64, 128], np.uint64)
        ax1 = (C.c_ssize_t * 1)(1)
        ax2 = (C.c_ssize_t * 1)(2)
        cpu_fns = {
            "add": lambda: lib.b_binary(0, 1, vc.ptr, C.byref(vc.c), va.ptr, C.byref(va.c), vb2.ptr, C.byref(vb2.c), cores),
            "sum_ax1": lambda: lib.b_reduce_axes(0, 1, vc.ptr, C.byref(vc.c), 1, ax1, vs.ptr, C.byref(vs.c), cores),
            "argmax_ax2": lambda: lib.b_reduce_axes(3, 1, vc.ptr, C.byref(vc.c), 1, ax2, vi.ptr, C.byref(vi.c), cores),
            "matmul_512": lambda: A @ B,
        }
        cpu = {}
        for nm, fn in cpu_fns.items():
            fn()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            cpu[nm + "_ms"] = round(1e3 * float(np.median(ts)), 4)
        cpu["step_ms"] = round(sum(v for k, v in cpu.items() if k.endswith("_ms")), 4)
        cpu["parity_add_bitexact"] = bool(np.array_equal(vc.array(), c_ref))
        cpu["parity_argmax_bitexact"] = bool(np.array_equal(vi.array(), np.argmax(c_ref, axis=2)))
        res["cpu_baseline"] = {"ms": cpu, "cores": cores, "kind": "port",
                               "sample": "the whole small configuration (full inputs); elementwise and reductions "
                                         "by the oracle's OpenMP port of the cpu_rayon control flow, the matmul by "
                                         "NumPy's OpenBLAS (stand-in for faer)"}
    log("[bench] C1: " + json.dumps(res))
    return res


def transfer_leg(dev, nbytes=1 << 30, reps=3):
    """Host <-> device GB/s for 1 GiB: pageable memory through the HIP runtime (mode 1), through the
    library's pinned double-buffered staging (mode 2), and from pinned memory (hipHostMalloc) by
    direct DMA; wall clock around each synchronous call, best of `reps`."""
    from rstsr_amd import abi

    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    d = C.c_void_p()
    abi.check(L.rstsr_hip_malloc(ctx, nbytes, C.byref(d)))
    host = np.ones(nbytes // 8)
    hp = host.ctypes.data_as(C.c_void_p)
    hip = C.CDLL("libamdhip64.so")
    pinned = C.c_void_p()
    res = {"bytes": nbytes}
    try:
        if hip.hipHostMalloc(C.byref(pinned), C.c_size_t(nbytes), C.c_uint(0)) != 0:
            pinned = None

        def best(fn):
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            return round(nbytes / 1e9 / min(ts), 1)

        for name, mode in (("runtime", 1), ("staged", 2)):
            res[f"h2d_{name}_GBps"] = best(lambda: abi.check(L.rstsr_hip_memcpy_h2d_mode(ctx, d, hp, nbytes, mode)))
            res[f"d2h_{name}_GBps"] = best(lambda: abi.check(L.rstsr_hip_memcpy_d2h_mode(ctx, hp, d, nbytes, mode)))
        if pinned:
            res["h2d_pinned_GBps"] = best(lambda: abi.check(L.rstsr_hip_memcpy_h2d(ctx, d, pinned, nbytes)))
            res["d2h_pinned_GBps"] = best(lambda: abi.check(L.rstsr_hip_memcpy_d2h(ctx, pinned, d, nbytes)))
        res["h2d_default_GBps"] = best(lambda: abi.check(L.rstsr_hip_memcpy_h2d(ctx, d, hp, nbytes)))
        res["d2h_default_GBps"] = best(lambda: abi.check(L.rstsr_hip_memcpy_d2h(ctx, hp, d, nbytes)))
    finally:
        if pinned:
            hip.hipHostFree(pinned)
        abi.check(L.rstsr_hip_free(ctx, d))
    return res


def cpu_leg(X, b, Xi, bi):
    """Time the oracle's OpenMP port of the cpu_rayon control flow on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ffi import Oracle, View, c_view
    import ctypes as C

    orc = Oracle()
    cores = _host_threads()
    # sample: the same ops on X[:, :, :128] (1/8 of the data)
    Xs, Xis = X[:, :, :128], Xi[:, :, :128]
    X32s = np.ascontiguousarray(Xs).astype(np.float32)
    Vs = Xs.transpose(2, 0, 1)[:, :, ::2]
    bVs = np.broadcast_to(b.transpose(2, 0, 1)[:, :, ::2], Vs.shape)
    V32s = X32s.transpose(2, 0, 1)[:, :, ::2]
    bV32s = np.broadcast_to(b.astype(np.float32).transpose(2, 0, 1)[:, :, ::2], V32s.shape)
    Vis = Xis.transpose(2, 0, 1)[:, :, ::2]
    bVis = np.broadcast_to(bi.transpose(2, 0, 1)[:, :, ::2], Vis.shape)
    Ps = Xs.transpose(2, 0, 1)
    n = Vs.size
    vV, vb, vVi, vbi, vP = View.of(Vs), View.of(bVs), View.of(Vis), View.of(bVis), View.of(Ps)
    vV32, vb32 = View.of(V32s), View.of(bV32s)
    vY = c_view(Vs.shape, np.float64)
    vY32 = c_view(V32s.shape, np.float32)
    vYi = c_view(Vis.shape, np.int32)
    vPo = c_view(Ps.shape, np.float64)
    lib = orc.lib
    code = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.int32): 2}

    def red(op, v, axes, dt):
        nd = 3
        kept = [v.shape[i] for i in range(nd) if i not in axes]
        vo = c_view(kept, dt)
        ax = (C.c_ssize_t * len(axes))(*axes)
        return lambda: lib.b_reduce_axes(op, code[v.base.dtype], v.ptr, C.byref(v.c), len(axes),
                                         ax, vo.ptr, C.byref(vo.c), cores)

    def binop(op, t, vc, va, vbb):
        return lambda: lib.b_binary(op, t, vc.ptr, C.byref(vc.c), va.ptr, C.byref(va.c), vbb.ptr, C.byref(vbb.c), cores)

    # the twelve ops of the GPU step, in its order, on the sample
    tasks = [
        ("mul_f64", binop(2, 1, vY, vV, vb), n * 16),
        ("mul_f32", binop(2, 0, vY32, vV32, vb32), n * 8),
        ("add_i32", binop(0, 2, vYi, vVi, vbi), n * 8),
        ("sum_ax0_f64", red(0, vV, [0], np.float64), n * 8),
        ("sum_ax12_f64", red(0, vV, [1, 2], np.float64), n * 8),
        ("max_ax0_f64", red(1, vV, [0], np.float64), n * 8),
        ("max_ax12_f64", red(1, vV, [1, 2], np.float64), n * 8),
        ("argmax_ax0_f64", red(3, vV, [0], np.uint64), n * 8),
        ("argmax_ax12_f64", red(3, vV, [1, 2], np.uint64), n * 8),
        ("argmax_ax0_i32", red(3, vVi, [0], np.uint64), n * 4),
        ("argmax_ax12_i32", red(3, vVi, [1, 2], np.uint64), n * 4),
        ("to_contig_f64", lambda: lib.b_assign(1, vPo.ptr, C.byref(vPo.c), vP.ptr, C.byref(vP.c), cores), Ps.size * 16),
    ]
    tot_bytes, tot_s, reps = 0, 0.0, 0
    per = {}
    t_start = time.time()
    while True:
        for nm, fn, nb in tasks:
            t0 = time.perf_counter()
            fn()
            dt = time.perf_counter() - t0
            per.setdefault(nm, []).append(nb / 1e9 / dt)
            tot_bytes += nb
            tot_s += dt
        reps += 1
        if time.time() - t_start > 10.0 or reps >= 200:
            break
    value = tot_bytes / 1e9 / tot_s
    log(f"[bench] cpu baseline {value:.2f} GB/s over {reps} reps, {cores} threads")
    return {"value": round(value, 3), "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": "oracle OpenMP port of rstsr cpu_rayon control flow (greedy order, contiguous inner run, "
                      "parallel outer loop) on X[:, :, :128] of the same workload (1/8 of the data), "
                      f"{len(tasks)} ops (the whole step) x {reps} reps, f64/f32/i32",
            "per_op_GBps": {k: round(float(np.median(v)), 3) for k, v in per.items()}}


if __name__ == "__main__":
    main()
