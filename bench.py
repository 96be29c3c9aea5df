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

    # ---------------- host <-> device transfers (PCIe-inclusive; not part of the step) ----------------
    transfer = {}
    if not args.no_transfer and not args.only:
        transfer = transfer_leg(dev)
        log("[bench] transfer: " + json.dumps(transfer))

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
        "gemm": gemm,
        "transfer": transfer,
    }
    print(json.dumps(out), flush=True)


def _timed(ev, fn, reps):
    fn()  # warmup (also settles the autotuner)
    ev.record(4)
    for _ in range(reps):
        fn()
    ev.record(5)
    return ev.elapsed_ms(4, 5) / reps


def gemm_leg(args, dev, rt):
    """Dense FP64 GEMM configurations: plain FP64 MFMA vs Ozaki-II INT8 split, error vs Dot2."""
    from rstsr_amd import abi

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ffi import Oracle

    orc = Oracle()
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
    rows = np.array([0, n // 3, (2 * n) // 3, n - 1])
    ref = orc.gemm_dot2(A[rows], B, nthreads=0)
    res = {"shape": [n, n, n], "flop": 2.0 * n ** 3}
    for mode in (modes if "dense" in legs else []):
        holder = {}

        def run():
            holder["c"] = rt.matmul(tA, tB, mode=mode)

        ms = _timed(ev, run, 3)
        err = float(np.max(np.abs(holder["c"].to_numpy()[rows] - ref)) / np.max(np.abs(ref)))
        tf = 2.0 * n ** 3 / (ms / 1e3) / 1e12
        res[mode] = {"ms": round(ms, 3), "TFLOP/s": round(tf, 2), "frac_fp64_peak": round(tf / FP64_PEAK, 3),
                     "normwise_err_sampled_rows": err}
        log(f"[bench] dense {n}^3 {mode}: {ms:.2f} ms {tf:.2f} TFLOP/s err {err:.2e}")
        holder.clear()
    out["dense_square"] = res
    del tA, tB, A, B

    # ---- broadcast batch [64,1024,1024] % [1024,1024] (rule 6) ----
    rng = np.random.default_rng(20261205)
    A3 = wide(rng, (64, 1024, 1024))
    B2 = wide(rng, (1024, 1024))
    tA3, tB2 = rt.asarray(A3, device=dev), rt.asarray(B2, device=dev)
    ref = orc.gemm_dot2(A3[7, :4], B2)
    res = {"shape": [64, 1024, 1024, 1024], "flop": 64 * 2.0 * 1024 ** 3}
    for mode in (modes if "batch" in legs else []):
        holder = {}

        def run():
            holder["c"] = rt.matmul(tA3, tB2, mode=mode)

        ms = _timed(ev, run, 3)
        err = float(np.max(np.abs(holder["c"].to_numpy()[7, :4] - ref)) / np.max(np.abs(ref)))
        tf = res["flop"] / (ms / 1e3) / 1e12
        res[mode] = {"ms": round(ms, 3), "TFLOP/s": round(tf, 2), "normwise_err_sampled_rows": err}
        log(f"[bench] batch64 {mode}: {ms:.2f} ms {tf:.2f} TFLOP/s err {err:.2e}")
        holder.clear()
    out["batched_broadcast"] = res
    del tA3, tB2, A3

    # ---- RI contraction (ia|jb) = sum_P B[P,i,a] B[P,j,b], B [1600,60,400] ~ N(0,0.1) ----
    rng = np.random.default_rng(20261206)
    Bri = rng.normal(0.0, 0.1, size=(1600, 60, 400))
    tB = rt.asarray(Bri, device=dev)
    lhs = tB.permute([1, 2, 0])            # [60, 400, 1600]  (i, a, P)
    rhs = tB.reshape([1600, 24000])        # [1600, 24000]    (P, (j, b))
    Ai = np.ascontiguousarray(Bri[:, 5, :].T[:4])  # rows a=0..3 of C_5
    ref = orc.gemm_dot2(Ai, Bri.reshape(1600, 24000))
    res = {"shape": [60, 400, 24000, 1600], "flop": 60 * 2.0 * 400 * 24000 * 1600}
    for mode in (modes if "ri" in legs else []):
        holder = {}

        def run():
            holder["c"] = rt.matmul(lhs, rhs, mode=mode)

        ms = _timed(ev, run, 2)
        got = holder["c"][5, 0:4].to_numpy()
        err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        tf = res["flop"] / (ms / 1e3) / 1e12
        res[mode] = {"ms": round(ms, 3), "TFLOP/s": round(tf, 2), "normwise_err_sampled_rows": err}
        log(f"[bench] RI {mode}: {ms:.2f} ms {tf:.2f} TFLOP/s err {err:.2e}")
        holder.clear()
    out["ri_contraction"] = res
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
    check = [0, 100, 511]
    refs = {}
    for g in check:
        m, n, k = (int(x) for x in dims[g])
        Ah = pool[offs[g, 0]:offs[g, 0] + m * k].reshape(m, k)
        Bh = pool[offs[g, 1]:offs[g, 1] + k * n].reshape(k, n)
        refs[g] = orc.gemm_dot2(Ah, Bh)
    for mode in (modes if "grouped" in legs else []):
        ms = _timed(ev, lambda: rt.gemm_grouped(probs, mode=mode), 3)
        err = max(float(np.max(np.abs(probs[g][2].to_numpy() - refs[g])) / np.max(np.abs(refs[g]))) for g in check)
        tf = flop / (ms / 1e3) / 1e12
        res[mode] = {"ms": round(ms, 3), "TFLOP/s": round(tf, 2), "normwise_err_sampled": err}
        log(f"[bench] grouped512 {mode}: {ms:.2f} ms {tf:.2f} TFLOP/s err {err:.2e}")
    out["uneven_grouped"] = res
    del probs, tpool, tC

    # ---- hipGraph-captured c128 chain, 20 steps ----
    if "chain" in legs:
        out["graph_chain_c128"] = chain_leg(dev, rt, orc, modes)
    return out


def chain_leg(dev, rt, orc, modes, n=4096, steps=20):
    """Z1 += v; Z3 = Z1 * Z2; Z4 = Z3 % Z2; s = Z4.sum_axes(1): uncaptured vs one captured graph."""
    from rstsr_amd import abi

    ev = abi.Events(dev._ctx_ptr())
    rng = np.random.default_rng(20261208)
    z1h = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    z2h = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    vh = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Z2, V = rt.asarray(z2h, device=dev), rt.asarray(vh, device=dev)
    res = {"shape": [n, n], "steps": steps, "flop_per_step": 8.0 * n ** 3}
    # oracle for the last step, sampled rows of s
    z = z1h.copy()
    for _ in range(steps):
        z = z + vh
    z3 = z * z2h
    rows = [0, n // 2, n - 1]
    r_re = orc.gemm_dot2(np.ascontiguousarray(z3.real[rows]), z2h.real) - orc.gemm_dot2(
        np.ascontiguousarray(z3.imag[rows]), z2h.imag)
    r_im = orc.gemm_dot2(np.ascontiguousarray(z3.real[rows]), z2h.imag) + orc.gemm_dot2(
        np.ascontiguousarray(z3.imag[rows]), z2h.real)
    s_ref = (r_re + 1j * r_im).sum(axis=1)
    for mode in modes:
        Z1 = rt.asarray(z1h, device=dev)
        holder = {}

        def step():
            nonlocal Z1
            Z1 += V
            Z3 = Z1 * Z2
            Z4 = rt.matmul(Z3, Z2, mode=mode)
            holder["s"] = Z4.sum_axes([1])
            holder["keep"] = (Z3, Z4)

        step()  # settles autotuner / workspaces
        Z1.assign(rt.asarray(z1h, device=dev))
        dev.synchronize()
        t0 = time.perf_counter()
        ev.record(6)
        for _ in range(steps):
            step()
        ev.record(7)
        eager_ms = ev.elapsed_ms(6, 7) / steps
        wall_eager = (time.perf_counter() - t0) * 1e3 / steps
        err_eager = float(np.max(np.abs(holder["s"].to_numpy()[rows] - s_ref)) / np.max(np.abs(s_ref)))
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
        err_graph = float(np.max(np.abs(keep[-1][2].to_numpy()[rows] - s_ref)) / np.max(np.abs(s_ref)))
        res[mode] = {"eager_ms_per_step": round(eager_ms, 3), "graph_ms_per_step": round(graph_ms, 3),
                     "eager_wall_ms_per_step": round(wall_eager, 3), "graph_wall_ms_per_step": round(wall_graph, 3),
                     "launch_overhead_saved_ms_per_step": round(eager_ms - graph_ms, 4),
                     "err_last_step_eager": err_eager, "err_last_step_graph": err_graph, "target_err": 1e-13,
                     "zgemm_TFLOP/s_graph": round(8.0 * n ** 3 / (graph_ms / 1e3) / 1e12, 2)}
        log(f"[bench] chain {mode}: eager {eager_ms:.3f} ms/step, graph {graph_ms:.3f} ms/step, "
            f"err {err_eager:.2e}/{err_graph:.2e}")
        del keep, g, holder
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
    cores = os.cpu_count() or 1
    # sample: the same ops on X[:, :, :128] (1/8 of the data)
    Xs, Xis = X[:, :, :128], Xi[:, :, :128]
    Vs = Xs.transpose(2, 0, 1)[:, :, ::2]
    bVs = np.broadcast_to(b.transpose(2, 0, 1)[:, :, ::2], Vs.shape)
    Vis = Xis.transpose(2, 0, 1)[:, :, ::2]
    bVis = np.broadcast_to(bi.transpose(2, 0, 1)[:, :, ::2], Vis.shape)
    Ps = Xs.transpose(2, 0, 1)
    n = Vs.size
    vV, vb, vVi, vbi, vP = View.of(Vs), View.of(bVs), View.of(Vis), View.of(bVis), View.of(Ps)
    vY = c_view(Vs.shape, np.float64)
    vYi = c_view(Vis.shape, np.int32)
    vPo = c_view(Ps.shape, np.float64)
    lib = orc.lib

    def red(op, v, axes, dt):
        nd = 3
        kept = [v.shape[i] for i in range(nd) if i not in axes]
        vo = c_view(kept, dt)
        ax = (C.c_ssize_t * len(axes))(*axes)
        return lambda: lib.b_reduce_axes(op, 1 if v.base.dtype == np.float64 else 2, v.ptr, C.byref(v.c), len(axes),
                                         ax, vo.ptr, C.byref(vo.c), cores)

    tasks = [
        ("mul_f64", lambda: lib.b_binary(2, 1, vY.ptr, C.byref(vY.c), vV.ptr, C.byref(vV.c), vb.ptr, C.byref(vb.c), cores), n * 16),
        ("add_i32", lambda: lib.b_binary(0, 2, vYi.ptr, C.byref(vYi.c), vVi.ptr, C.byref(vVi.c), vbi.ptr, C.byref(vbi.c), cores), n * 8),
        ("sum_ax0_f64", red(0, vV, [0], np.float64), n * 8),
        ("sum_ax12_f64", red(0, vV, [1, 2], np.float64), n * 8),
        ("max_ax0_f64", red(1, vV, [0], np.float64), n * 8),
        ("argmax_ax0_f64", red(3, vV, [0], np.uint64), n * 8),
        ("argmax_ax12_f64", red(3, vV, [1, 2], np.uint64), n * 8),
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
        if time.time() - t_start > 10.0 or reps >= 20:
            break
    value = tot_bytes / 1e9 / tot_s
    log(f"[bench] cpu baseline {value:.2f} GB/s over {reps} reps, {cores} threads")
    return {"value": round(value, 3), "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": "oracle OpenMP port of rstsr cpu_rayon control flow (greedy order, contiguous inner run, "
                      "parallel outer loop) on X[:, :, :128] of the same workload (1/8 of the data), "
                      f"{len(tasks)} ops x {reps} reps, f64/i32",
            "per_op_GBps": {k: round(float(np.median(v)), 3) for k, v in per.items()}}


if __name__ == "__main__":
    main()
