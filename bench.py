"""Benchmark of the cuMat hot path on one MI355X.

Headline (value): Jacobi-PCG iterations per second on the 2048x2048 2D
Poisson system (n = 4,194,304, nnz = 20,963,328, FP32), the largest
single-GPU bandwidth configuration; a step is one CG iteration of the fused
solver (tol = 0, so exactly K iterations run in the timed region).

roofline: the CSR SpMV kernel on the same matrix, timed with HIP events on
the stream it runs on, with four rotating copies of (matrix, x, y) so that
successive launches cannot be served from the 256 MB last-level cache.
Algorithmic bytes per launch = 8*nnz + 12*n.

Also reported (bounded, after the timed region): the linear combination of 8
vectors of 2^26 floats, dot of 2^26, general cwise programs over 2^26
(generated kernels beside the interpreter; also FP64 over 2^25, broadcasts
and reductions of broadcasts over 4096x4096x4),
row/column/batch sums of 1024x1024x64, the 16384^2 dense product (both
layouts), ELLPACK SpMV and cwise evaluation into CSR values on the 2048^2
matrix, the block-scalar (3x3 blocks) SpMV and CG on a 96^3 grid, the 128^2
plain CG solve (iters/s, time to tol 1e-6) and small systems, the batched, mixed-precision and adjoint solvers, the reference's
published benchmark workloads, and the CPU baseline: the reference's
vendored Eigen (oracle/_ref) on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras] [--no-cpu]
                    [--kernels-only | --solvers-only]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


class HipTimer:
    """hipEvent pair on a given stream (the library's context stream)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(self.e0)) == 0
        assert self.hip.hipEventCreate(C.byref(self.e1)) == 0

    def start(self, stream):
        self.hip.hipEventRecord(self.e0, C.c_void_p(stream))

    def stop_ms(self, stream):
        self.hip.hipEventRecord(self.e1, C.c_void_p(stream))
        self.hip.hipEventSynchronize(self.e1)
        ms = C.c_float()
        self.hip.hipEventElapsedTime(C.byref(ms), self.e0, self.e1)
        return float(ms.value)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cg-path", type=int, default=1, choices=[1, 3],
                    help="multi-kernel CG form: 1 = three launches per iteration, 3 = two (fused p update)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the SpMV roofline launches and a copy kernel of known size (counter runs)")
    ap.add_argument("--solvers-only", action="store_true",
                    help="only the batched, mixed-precision and adjoint workloads, one pass each (profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU sharding is "
              "a design, not run.", flush=True)
        sys.exit(2)

    import cumat_amd as cm
    from cumat_amd import capi
    from cumat_amd.problems import poisson2d

    info = cm.device_info(0)
    log("device:", info)
    ctx = cm.Context(0)
    stream = ctx.stream()
    timer = HipTimer()
    if args.solvers_only:
        out = {}
        for name, fn in (("batched_cg_4096", bench_batched), ("mixed_ir_160cubed", bench_mixed),
                         ("adjoint_64x512sq", bench_adjoint)):
            out[name] = fn(cm, capi, ctx)
            log(name, out[name])
        print(json.dumps(out), flush=True)
        return

    # ------------------------------------------------------------ workload
    g = 2048
    JA, IA, V = poisson2d(g)
    n, nnz = g * g, IA.size
    rng = np.random.default_rng(20261202)
    b = rng.uniform(-1, 1, n).astype(np.float32)
    A = cm.SparseMatrix.from_csr(n, n, JA, IA, V, ctx=ctx)
    bd = cm.DeviceArray.from_numpy(b, ctx=ctx)
    xd = cm.DeviceArray(n, dtype=np.float32, ctx=ctx)
    cg = cm.ConjugateGradient(A, path=args.cg_path).setTolerance(0.0)

    # ------------------------------------------------------------ SpMV roofline
    # (measured first: the CG below then runs on a GPU that has been busy,
    # as in continuous use, rather than straight out of idle)
    copies = []
    for k in range(4):
        Ak = cm.SparseMatrix.from_csr(n, n, JA, IA, V, ctx=ctx)
        xk = cm.DeviceArray.from_numpy(rng.uniform(-1, 1, n).astype(np.float32), ctx=ctx)
        yk = cm.DeviceArray(n, dtype=np.float32, ctx=ctx)
        copies.append((Ak, xk, yk))
    for k in range(8):
        Ak, xk, yk = copies[k % 4]
        Ak.matvec(xk, yk, ctx=ctx)
    ctx.sync()
    # median over 20 groups of 12 launches (three rounds of the four copies)
    group_ms = []
    for _ in range(20):
        timer.start(stream)
        for k in range(12):
            Ak, xk, yk = copies[k % 4]
            Ak.matvec(xk, yk, ctx=ctx)
        group_ms.append(timer.stop_ms(stream) / 12)
    spmv_ms = float(np.median(group_ms))
    spmv_bytes = 8 * nnz + 12 * n
    spmv_gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    log(f"SpMV 2048^2: {spmv_ms * 1000:.1f} us/launch, {spmv_gbs:.0f} GB/s ({spmv_gbs / HBM_PEAK_GBS:.1%})")
    del copies

    # ------------------------------------------------------------ timed CG
    if args.kernels_only:
        args.steps = args.warmup = 3  # a few CG iterations, so every kernel is counted mid-solve
    cg.setMaxIterations(args.warmup)
    cg.solve(bd, xd, ctx=ctx)
    ctx.sync()
    cg.setMaxIterations(args.steps)
    ctx.sync()
    t0 = time.perf_counter()
    cg.solve(bd, xd, ctx=ctx)
    ctx.sync()
    t1 = time.perf_counter()
    assert cg.iterations() == args.steps, cg.iterations()
    sec = t1 - t0
    iters_per_s = args.steps / sec
    ms_per_step = 1000.0 * sec / args.steps
    cg_bytes = 8 * nnz + 56 * n
    log(f"CG 2048^2: {iters_per_s:.1f} it/s, {ms_per_step * 1000:.1f} us/it, "
        f"{cg_bytes / (ms_per_step * 1e-3) / 1e9:.0f} GB/s (fused-minimum bytes)")

    if args.kernels_only:
        # calibration for the HBM byte counters: dst = src over 2^26 floats
        # (one cwise launch, 256 MiB read + 256 MiB written)
        src = cm.DeviceArray.from_numpy(rng.uniform(-1, 1, 1 << 26).astype(np.float32), ctx=ctx)
        dst = cm.DeviceArray(1 << 26, dtype=np.float32, ctx=ctx)
        cm.assign(dst, src.expr(), ctx=ctx)
        ctx.sync()
        timer.start(stream)
        for _ in range(20):
            cm.assign(dst, src.expr(), ctx=ctx)
        copy_ms = timer.stop_ms(stream) / 20
        log(f"copy 2^26 f32: {copy_ms * 1000:.1f} us/launch, {2 * 4 * (1 << 26) / (copy_ms * 1e-3) / 1e9:.0f} GB/s")
        print(json.dumps({"spmv_us": round(spmv_ms * 1000, 2), "spmv_algorithmic_bytes": spmv_bytes,
                          "copy_us": round(copy_ms * 1000, 2), "copy_bytes_read": 4 << 26,
                          "copy_bytes_written": 4 << 26}), flush=True)
        return
    cfg1 = bench_config1(cm, capi, ctx)
    log("128^2 CG (tol 1e-6):", cfg1)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "spmv_poisson2048_traffic.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("hbm_bytes_per_launch")

    extras = {}
    if not args.no_extras:
        extras = run_extras(cm, capi, ctx, stream, timer, rng)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(JA, IA, V, b, n, nnz)

    line = {
        "metric": "CG iters/s + SpMV HBM GB/s (% of peak); batched/mixed vs plain at rel-res <=1e-6",
        "value": round(iters_per_s, 2),
        "unit": "CG iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (2D 5-point Poisson, b ~ U[-1,1] seed 20261202)",
        "config": {"workload": "Jacobi-PCG iteration on 2D Poisson 2048x2048 (n=4194304, nnz=20963328), FP32 CSR",
                   "n": n, "nnz": nnz, "cg_bytes_per_iteration": cg_bytes,
                   "cg_effective_GBps": round(cg_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                   "cg_128x128_tol1e-6": cfg1},
        "roofline": {"bound": "hbm", "kernel": "csr_spmv_kernel<float,float> (2048^2 Poisson)",
                     "achieved": round(spmv_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(spmv_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": spmv_bytes, "avg_launch_us": round(spmv_ms * 1000, 2)},
        "cpu_baseline": cpu,
        "extras": extras,
        "device": info["name"],
    }
    print(json.dumps(line), flush=True)


def run_extras(cm, capi, ctx, stream, timer, rng):
    out = {}
    # ---- linear combination 8 x 2^26 (fused: one kernel)
    N = 1 << 26
    vecs = [cm.DeviceArray.from_numpy(rng.random(N, dtype=np.float32) * 2 - 1, ctx=ctx) for _ in range(8)]
    coefs = rng.uniform(-1, 1, 8).astype(np.float32)
    outv = cm.DeviceArray(N, dtype=np.float32, ctx=ctx)
    e = vecs[0].expr() * float(coefs[0])
    for i in range(1, 8):
        e = e + vecs[i].expr() * float(coefs[i])
    ptrs = (C.c_void_p * 8)(*[v.ptr for v in vecs])
    cf = (C.c_float * 8)(*coefs)
    lc_bytes = 9 * 4 * N
    for label, fn in (("lincomb8_expr", lambda: cm.assign(outv, e, ctx=ctx)),
                      ("lincomb8_kernel", lambda: capi.call("cm_lincomb_f32", ctx.handle, 8, ptrs, cf, outv.ptr, N))):
        fn()
        ctx.sync()
        reps = 20
        timer.start(stream)
        for _ in range(reps):
            fn()
        ms = timer.stop_ms(stream) / reps
        out[label] = {"us": round(ms * 1000, 1), "GBps": round(lc_bytes / (ms * 1e-3) / 1e9, 1),
                      "frac": round(lc_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(label, out[label])
    # ---- dot 2^26
    d_out = cm.DeviceArray(1, dtype=np.float32, ctx=ctx)
    cm.dot(vecs[0], vecs[1], out=d_out, ctx=ctx)
    ctx.sync()
    reps = 50
    timer.start(stream)
    for k in range(reps):
        cm.dot(vecs[k % 8], vecs[(k + 1) % 8], out=d_out, ctx=ctx)
    ms = timer.stop_ms(stream) / reps
    out["dot_2^26"] = {"us": round(ms * 1000, 1), "GBps": round(8 * N / (ms * 1e-3) / 1e9, 1),
                       "frac": round(8 * N / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    log("dot", out["dot_2^26"])
    # ---- reductions of unevaluated expressions (one launch, no temporary),
    # 2^26 elements: squared norm of a difference and the sum of a*b+c
    red = {}
    for label, fn, bpe in (("(a-b).squaredNorm()", lambda: cm.squared_norm(vecs[0].expr() - vecs[1].expr(), out=d_out,
                                                                           ctx=ctx), 8),
                           ("dot(a-b, c)", lambda: cm.dot(vecs[0].expr() - vecs[1].expr(), vecs[2], out=d_out, ctx=ctx),
                            12),
                           ("(a*b+c).sum()", lambda: cm.reduce(vecs[0].expr() * vecs[1].expr() + vecs[2].expr(),
                                                              "sum", capi.AXIS_ALL, out=d_out, ctx=ctx), 12)):
        fn()
        ctx.sync()
        timer.start(stream)
        for _ in range(10):
            fn()
        ms = timer.stop_ms(stream) / 10
        red[label] = {"us": round(ms * 1000, 1), "GBps": round(bpe * N / (ms * 1e-3) / 1e9, 1),
                      "frac": round(bpe * N / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    out["reduce_expr_2^26"] = red
    log("reduce_expr_2^26", red)
    # ---- general programs (not linear combinations), 2^26 elements: the
    # kernel generated and compiled for each program at run time (the
    # default from 2^18 elements on; the compilation happens in the untimed
    # first call) and the built-in program interpreter beside it
    ga, gb, gc, gd = vecs[:4]
    ia_ = cm.DeviceArray.from_numpy(rng.integers(-1000, 1000, N).astype(np.int32), ctx=ctx)
    ib_ = cm.DeviceArray.from_numpy(rng.integers(1, 1000, N).astype(np.int32), ctx=ctx)
    io_ = cm.DeviceArray(N, dtype=np.int32, ctx=ctx)
    for key, jm in (("cwise_generated_2^26", 2), ("cwise_interpreter_2^26", 0)):
        old_mode = cm.jit_mode(jm)
        res = {}
        for label, dst, e, bpe in (("a*b+c", outv, ga.expr() * gb.expr() + gc.expr(), 16),
                                   ("(a+b)*(c-d)", outv, (ga.expr() + gb.expr()) * (gc.expr() - gd.expr()), 20),
                                   ("int32 a%b", io_, ia_.expr() % ib_.expr(), 12)):
            cm.assign(dst, e, ctx=ctx)
            ctx.sync()
            timer.start(stream)
            for _ in range(10):
                cm.assign(dst, e, ctx=ctx)
            ms = timer.stop_ms(stream) / 10
            res[label] = {"us": round(ms * 1000, 1), "GBps": round(bpe * N / (ms * 1e-3) / 1e9, 1),
                          "frac": round(bpe * N / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        cm.jit_mode(old_mode)
        out[key] = res
        log(key, res)
    del vecs, outv, ia_, ib_, io_
    # ---- FP64 program a*b+c over 2^25 doubles (64-bit registers), generated
    # kernel against the interpreter
    N64 = 1 << 25
    da, db, dc = (cm.DeviceArray.from_numpy(rng.standard_normal(N64), ctx=ctx) for _ in range(3))
    do_ = cm.DeviceArray(N64, dtype=np.float64, ctx=ctx)
    e64 = da.expr() * db.expr() + dc.expr()
    for key, jm in (("cwise_f64_generated_2^25", 2), ("cwise_f64_interpreter_2^25", 0)):
        old_mode = cm.jit_mode(jm)
        cm.assign(do_, e64, ctx=ctx)
        ctx.sync()
        timer.start(stream)
        for _ in range(10):
            cm.assign(do_, e64, ctx=ctx)
        ms = timer.stop_ms(stream) / 10
        cm.jit_mode(old_mode)
        out[key] = {"a*b+c": {"us": round(ms * 1000, 1), "GBps": round(32 * N64 / (ms * 1e-3) / 1e9, 1),
                              "frac": round(32 * N64 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
        log(key, out[key])
    del da, db, dc, do_
    # ---- broadcasts over a 4096 x 4096 x 4 FP32 tensor (strided leaves read
    # at the element's coordinates), generated kernel against the interpreter
    Rb, Cb, Bb = 4096, 4096, 4
    nb = Rb * Cb * Bb
    TA = cm.DeviceArray.from_numpy(rng.standard_normal(nb, dtype=np.float32), rows=Rb, cols=Cb, batches=Bb, ctx=ctx)
    TV = cm.DeviceArray.from_numpy(rng.standard_normal(Rb, dtype=np.float32), rows=Rb, cols=1, batches=1, ctx=ctx)
    TW = cm.DeviceArray.from_numpy(rng.standard_normal(Cb, dtype=np.float32), rows=1, cols=Cb, batches=1, ctx=ctx)
    TO = cm.DeviceArray(Rb, Cb, Bb, dtype=np.float32, ctx=ctx)
    for key, jm in (("broadcast_generated_4096x4096x4", 2), ("broadcast_interpreter_4096x4096x4", 0)):
        old_mode = cm.jit_mode(jm)
        res = {}
        for label, e in (("A*colvec", TA.expr() * TV.expr()), ("A*rowvec+colvec", TA.expr() * TW.expr() + TV.expr())):
            cm.assign(TO, e, ctx=ctx)
            ctx.sync()
            timer.start(stream)
            for _ in range(10):
                cm.assign(TO, e, ctx=ctx)
            ms = timer.stop_ms(stream) / 10
            res[label] = {"us": round(ms * 1000, 1), "GBps": round(8 * nb / (ms * 1e-3) / 1e9, 1),
                          "frac": round(8 * nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        cm.jit_mode(old_mode)
        out[key] = res
        log(key, res)
    # reductions of broadcast expressions over the same tensor (4 bytes read
    # per element)
    for key, jm in (("reduce_broadcast_generated_4096x4096x4", 2), ("reduce_broadcast_interpreter_4096x4096x4", 0)):
        old_mode = cm.jit_mode(jm)
        res = {}
        for label, fn in (("(A*colvec).sum()", lambda: cm.reduce(TA.expr() * TV.expr(), "sum", 7, ctx=ctx)),
                          ("(A-rowvec).sum<Row>()", lambda: cm.reduce(TA.expr() - TW.expr(), "sum", 1, ctx=ctx))):
            fn()
            ctx.sync()
            timer.start(stream)
            for _ in range(10):
                fn()
            ms = timer.stop_ms(stream) / 10
            res[label] = {"us": round(ms * 1000, 1), "GBps": round(4 * nb / (ms * 1e-3) / 1e9, 1),
                          "frac": round(4 * nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        cm.jit_mode(old_mode)
        out[key] = res
        log(key, res)
    del TA, TV, TW, TO
    # ---- axis sums of 1024 x 1024 x 64 (two rotating tensors of 256 MiB)
    R, Cc, B = 1024, 1024, 64
    ts = [cm.DeviceArray.from_numpy(rng.standard_normal(R * Cc * B, dtype=np.float32), rows=R, cols=Cc, batches=B,
                                    ctx=ctx) for _ in range(2)]
    for name, axis, ob in (("row_sum", 1, Cc * B), ("col_sum", 2, R * B), ("batch_sum", 4, R * Cc)):
        res = [cm.reduce(t, "sum", axis, ctx=ctx) for t in ts]
        ctx.sync()
        reps = 20
        timer.start(stream)
        for k in range(reps):
            cm.reduce(ts[k % 2], "sum", axis, out=res[k % 2], ctx=ctx)
        ms = timer.stop_ms(stream) / reps
        by = 4 * R * Cc * B + 4 * ob
        out[name] = {"us": round(ms * 1000, 1), "GBps": round(by / (ms * 1e-3) / 1e9, 1),
                     "frac": round(by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(name, out[name])
    del ts
    # ---- dense matrix-vector product, 16384 x 16384 FP32 (1 GiB; two
    # rotating matrices so launches stream from HBM)
    Nd = 16384
    for rm, label in ((False, "dense_gemv_16384_colmajor"), (True, "dense_gemv_16384_rowmajor")):
        mats = [cm.DeviceArray(Nd, Nd, dtype=np.float32, row_major=rm, ctx=ctx) for _ in range(2)]
        for m_ in mats:
            cm.assign(m_, cm.as_expr(0.5), ctx=ctx)
        xv = cm.DeviceArray.from_numpy(rng.uniform(-1, 1, Nd).astype(np.float32), ctx=ctx)
        yv = cm.DeviceArray(Nd, dtype=np.float32, ctx=ctx)
        for k in range(4):
            cm.dense.matvec(mats[k % 2], xv, out=yv, ctx=ctx)
        ctx.sync()
        reps = 20
        timer.start(stream)
        for k in range(reps):
            cm.dense.matvec(mats[k % 2], xv, out=yv, ctx=ctx)
        ms = timer.stop_ms(stream) / reps
        by = 4 * Nd * Nd + 8 * Nd
        out[label] = {"us": round(ms * 1000, 1), "GBps": round(by / (ms * 1e-3) / 1e9, 1),
                      "frac": round(by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(label, out[label])
        del mats
    # ---- ELLPACK SpMV and component-wise evaluation into CSR values on the
    # 2048^2 Poisson matrix (two rotating copies, more than the last-level
    # cache together)
    from cumat_amd.problems import poisson2d
    JA, IA, V = poisson2d(2048)
    n = 2048 * 2048
    W = int(np.max(np.diff(JA)))
    lens = np.diff(JA)
    rows = np.repeat(np.arange(n), lens)
    pos = np.arange(IA.size) - np.repeat(JA[:-1], lens)
    ind = np.full((n, W), -1, np.int32)
    val = np.zeros((n, W), np.float32)
    ind[rows, pos] = IA
    val[rows, pos] = V
    ells = [cm.ELLMatrix(n, n, ind, val, ctx=ctx) for _ in range(2)]
    del ind, val
    xs = [cm.DeviceArray.from_numpy(rng.uniform(-1, 1, n).astype(np.float32), ctx=ctx) for _ in range(2)]
    ys = [cm.DeviceArray(n, dtype=np.float32, ctx=ctx) for _ in range(2)]
    for k in range(4):
        ells[k % 2].matvec(xs[k % 2], out=ys[k % 2], ctx=ctx)
    ctx.sync()
    reps = 20
    timer.start(stream)
    for k in range(reps):
        ells[k % 2].matvec(xs[k % 2], out=ys[k % 2], ctx=ctx)
    ms = timer.stop_ms(stream) / reps
    by = 8 * n * W + 8 * n
    out["ell_spmv_2048^2"] = {"us": round(ms * 1000, 1), "GBps": round(by / (ms * 1e-3) / 1e9, 1),
                              "frac": round(by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "width": W}
    log("ell_spmv_2048^2", out["ell_spmv_2048^2"])
    del ells
    sms = [cm.SparseMatrix.from_csr(n, n, JA, IA, V, ctx=ctx) for _ in range(2)]
    rv = cm.DeviceArray.from_numpy(rng.uniform(-1, 1, n).astype(np.float32), rows=1, cols=n, batches=1, ctx=ctx)
    nnz = IA.size
    # ---- the fused sparse-expression product (2*A).sparseView(pattern) * x:
    # the same bytes as the plain SpMV (values, IA, JA, x, y), one launch
    views = [cm.sparse_view(2.0 * S.direct(), S.pattern) for S in sms]
    for k in range(4):
        views[k % 2].matvec(xs[k % 2], out=ys[k % 2], ctx=ctx)
    ctx.sync()
    timer.start(stream)
    for k in range(reps):
        views[k % 2].matvec(xs[k % 2], out=ys[k % 2], ctx=ctx)
    ms = timer.stop_ms(stream) / reps
    by = 8 * nnz + 12 * n
    out["csr_expr_spmv_2048^2"] = {"expr": "(2*A).sparseView(pattern) * x", "us": round(ms * 1000, 1),
                                   "GBps": round(by / (ms * 1e-3) / 1e9, 1),
                                   "frac": round(by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    log("csr_expr_spmv_2048^2", out["csr_expr_spmv_2048^2"])
    csr_cw = {}
    # bytes: values read + written; the column-dependent program also reads
    # IA and gathers the row vector (counted once)
    for label, mk, by in (("A = A*2+1", lambda S: S.direct() * 2.0 + 1.0, 8 * nnz),
                          ("A = A*r(col)", lambda S: S.direct() * rv.expr(), 12 * nnz + 4 * n)):
        exprs = [mk(S) for S in sms]
        for k in range(2):
            sms[k].assign(exprs[k], ctx=ctx)
        ctx.sync()
        timer.start(stream)
        for k in range(reps):
            sms[k % 2].assign(exprs[k % 2], ctx=ctx)
        ms = timer.stop_ms(stream) / reps
        csr_cw[label] = {"us": round(ms * 1000, 1), "GBps": round(by / (ms * 1e-3) / 1e9, 1),
                         "frac": round(by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    out["csr_cwise_2048^2"] = csr_cw
    log("csr_cwise_2048^2", csr_cw)
    del sms, xs, ys
    # ---- plain CG on 128^2 (the reference's CPU-runnable configuration)
    from cumat_amd.problems import poisson2d
    JA, IA, V = poisson2d(128)
    n = 128 * 128
    b = np.random.default_rng(20261203).uniform(-1, 1, n).astype(np.float32)
    A = cm.SparseMatrix.from_csr(n, n, JA, IA, V, ctx=ctx)
    bd = cm.DeviceArray.from_numpy(b, ctx=ctx)
    xd = cm.DeviceArray(n, dtype=np.float32, ctx=ctx)
    for path, label in ((capi.CG_RESIDENT, "cg128_resident"), (capi.CG_MULTIKERNEL, "cg128_multikernel")):
        cg = cm.ConjugateGradient(A, path=path).setTolerance(1e-6).setMaxIterations(1000)
        cg.solve(bd, xd, ctx=ctx)
        ts_ = []
        for _ in range(5):
            t0 = time.perf_counter()
            cg.solve(bd, xd, ctx=ctx)
            ts_.append(time.perf_counter() - t0)
        t = float(np.median(ts_))
        out[label] = {"iterations": cg.iterations(), "error": cg.error(), "ms_to_solution": round(t * 1e3, 3),
                      "iters_per_s": round(cg.iterations() / t, 1)}
        log(label, out[label])
    # ---- small systems: iterations/s of both launch shapes (200 iterations, tol 0)
    sweep = {}
    for g in (16, 32, 48, 64, 96, 128, 192):
        JA, IA, V = poisson2d(g)
        Ag = cm.SparseMatrix.from_csr(g * g, g * g, JA, IA, V, ctx=ctx)
        bg = cm.DeviceArray.from_numpy(np.ones(g * g, np.float32), ctx=ctx)
        xg = cm.DeviceArray(g * g, dtype=np.float32, ctx=ctx)
        row = {}
        for path, label in ((capi.CG_RESIDENT, "resident"), (capi.CG_MULTIKERNEL, "multikernel")):
            if path == capi.CG_RESIDENT and g * g > 16384:
                continue
            cg = cm.ConjugateGradient(Ag, path=path).setTolerance(0.0).setMaxIterations(200)
            cg.solve(bg, xg, ctx=ctx)
            t0 = time.perf_counter()
            cg.solve(bg, xg, ctx=ctx)
            row[label] = round(cg.iterations() / (time.perf_counter() - t0), 1)
        sweep[f"{g}^2"] = row
    out["cg_small_iters_per_s"] = sweep
    log("cg_small_iters_per_s", sweep)
    for name, fn in (("batched_cg_4096", bench_batched), ("mixed_ir_160cubed", bench_mixed),
                     ("adjoint_64x512sq", bench_adjoint),
                     ("reference_published_workloads", bench_reference_workloads)):
        try:
            out[name] = fn(cm, capi, ctx)
            log(name, out[name])
        except Exception as ex:  # reported, the headline line is still printed
            out[name] = {"error": repr(ex)}
            log(name, "failed:", repr(ex))
    out["block_cg_3x3_96^3"] = bench_block_cg(cm, ctx, stream, timer)
    log("block_cg_3x3_96^3", out["block_cg_3x3_96^3"])
    return out


def bench_block_cg(cm, ctx, stream, timer):
    """Block-scalar path (the reference's SparseMatrix<real3x3> CG,
    tests/TestBlockedConjugateGradient.cu): 3D 7-point block Poisson with
    3x3 blocks (three float4 rows, 48 bytes each) on a 96^3 grid. Block SpMV
    bytes = 52*nnz (blocks + column indices) + 4*(rows+1) + 32*rows (x, y);
    one CG iteration moves those (with p for x and Ap for y) plus 160*rows
    for the two vector kernels (float4 entries: r, Ap, M^-1 read and r
    written; x, r, M^-1, p read and x, p written)."""
    from cumat_amd.problems import block_poisson
    g = 96
    JA, IA, V = block_poisson((g, g, g), 3, seed=20261203)
    rows, nnz = g ** 3, int(IA.size)
    A = cm.BlockSparseMatrix.from_bcsr(rows, rows, JA, IA, V, ctx=ctx)
    del V
    rng = np.random.default_rng(5)
    x = cm.DeviceArray.from_numpy(rng.uniform(-1, 1, 4 * rows).astype(np.float32), ctx=ctx)
    y = cm.DeviceArray(4 * rows, dtype=np.float32, ctx=ctx)
    for _ in range(3):
        A.matvec(x, out=y, ctx=ctx)
    ctx.sync()
    group = []
    for _ in range(10):
        timer.start(stream)
        for _ in range(5):
            A.matvec(x, out=y, ctx=ctx)
        group.append(timer.stop_ms(stream) / 5)
    ms = float(np.median(group))
    sp_bytes = 52 * nnz + 4 * (rows + 1) + 32 * rows
    res = {"grid": f"{g}^3", "block_rows": rows, "blocks": nnz,
           "spmv_us": round(ms * 1000, 1), "spmv_GBps": round(sp_bytes / (ms * 1e-3) / 1e9, 1),
           "spmv_frac": round(sp_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    b = cm.DeviceArray.from_numpy(rng.uniform(-1, 1, 4 * rows).astype(np.float32), ctx=ctx)
    cg = cm.ConjugateGradient(A).setTolerance(0.0)
    cg.setMaxIterations(5)
    cg.solve(b, y, ctx=ctx)
    ctx.sync()
    K = 50
    cg.setMaxIterations(K)
    t0 = time.perf_counter()
    cg.solve(b, y, ctx=ctx)
    ctx.sync()
    sec = time.perf_counter() - t0
    it_bytes = sp_bytes + 160 * rows
    res.update({"cg_iterations": cg.iterations(), "cg_us_per_iteration": round(sec / K * 1e6, 1),
                "cg_iters_per_s": round(K / sec, 1),
                "cg_GBps": round(it_bytes / (sec / K) / 1e9, 1),
                "cg_frac": round(it_bytes / (sec / K) / 1e9 / HBM_PEAK_GBS, 4)})
    # convergence at the reference test's tolerance, against the FP64 residual
    cg.setTolerance(1e-5).setMaxIterations(4000)
    t0 = time.perf_counter()
    xs = cg.solve(b, ctx=ctx)
    ctx.sync()
    res.update({"solve_tol1e-5_ms": round((time.perf_counter() - t0) * 1000, 2),
                "solve_iterations": cg.iterations(), "solve_error": cg.error()})
    ax = A.matvec(xs, ctx=ctx).numpy().astype(np.float64).reshape(rows, 4)[:, :3]
    bh = b.numpy().astype(np.float64).reshape(rows, 4)[:, :3]
    res["true_rel_residual_fp64"] = float(np.linalg.norm(bh - ax) / np.linalg.norm(bh))
    return res


def bench_config1(cm, capi, ctx):
    """The plain Jacobi-PCG of the 128^2 Poisson system (n = 16,384, nnz =
    81,408, FP32, b ~ U[-1,1], x0 = 0, tol 1e-6, max 1000 iterations) through
    the default (automatic) path: time to solution and iterations/s, median of
    7 solves after one warm-up."""
    from cumat_amd.problems import poisson2d
    JA, IA, V = poisson2d(128)
    n = 128 * 128
    b = np.random.default_rng(20261203).uniform(-1, 1, n).astype(np.float32)
    A = cm.SparseMatrix.from_csr(n, n, JA, IA, V, ctx=ctx)
    bd = cm.DeviceArray.from_numpy(b, ctx=ctx)
    xd = cm.DeviceArray(n, dtype=np.float32, ctx=ctx)
    cg = cm.ConjugateGradient(A).setTolerance(1e-6).setMaxIterations(1000)
    cg.solve(bd, xd, ctx=ctx)
    ts = []
    for _ in range(7):
        ctx.sync()
        t0 = time.perf_counter()
        cg.solve(bd, xd, ctx=ctx)
        ctx.sync()
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    return {"iterations": cg.iterations(), "error": cg.error(), "ms_to_solution": round(t * 1e3, 3),
            "iters_per_s": round(cg.iterations() / t, 1), "path": cg.last_stats.get("path")}


def _sub_csr(P, s):
    ro = P["row_off"]
    lo, hi = int(ro[s]), int(ro[s + 1])
    a, b = int(P["JA"][lo]), int(P["JA"][hi])
    return (P["JA"][lo:hi + 1] - a).astype(np.int32), (P["IA"][a:b] - lo).astype(np.int32), P["vals"][a:b], \
        P["b"][lo:hi], hi - lo


def bench_batched(cm, capi, ctx):
    """4096 independent shifted 2D Poisson systems, side ~ U{16..96}, one
    persistent launch (autotuned block size, graph replay) vs looping the
    plain GPU solver over a sample of the systems."""
    from cumat_amd.problems import batched_poisson
    from cumat_amd.solvers import pcg_batched, pcg_batched_autotune
    nsys, tol, maxit = 4096, 1e-6, 5000
    P = batched_poisson(nsys, 16, 96, seed=20261205)
    d = {k: cm.DeviceArray.from_numpy(P[k], ctx=ctx) for k in ("row_off", "JA", "IA", "vals", "b")}
    x = cm.DeviceArray(P["b"].size, dtype=np.float32, ctx=ctx)
    its = cm.DeviceArray(nsys, dtype=np.int32, ctx=ctx)
    errs = cm.DeviceArray(nsys, dtype=np.float32, ctx=ctx)
    bs, gpc, tune_ms = pcg_batched_autotune(d["row_off"], d["JA"], d["IA"], d["vals"], d["b"], tol, maxit, ctx=ctx,
                                            with_groups=True)
    pcg_batched(d["row_off"], d["JA"], d["IA"], d["vals"], d["b"], x, tol, maxit, its, errs, use_graph=True, ctx=ctx)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        pcg_batched(d["row_off"], d["JA"], d["IA"], d["vals"], d["b"], x, tol, maxit, its, errs, use_graph=True,
                    ctx=ctx)
        ts.append(time.perf_counter() - t0)
    tb = min(ts)
    ih, eh = its.numpy(), errs.numpy()
    # looping the plain solver (cm_pcg_solve, its automatic path) on every 64th system
    sample = list(range(0, nsys, 64))
    loop_t = 0.0
    for s in sample:
        JA, IA, V, b, n = _sub_csr(P, s)
        A = cm.SparseMatrix.from_csr(n, n, JA, IA, V, ctx=ctx)
        bd = cm.DeviceArray.from_numpy(b, ctx=ctx)
        xd = cm.DeviceArray(n, dtype=np.float32, ctx=ctx)
        cg = cm.ConjugateGradient(A).setTolerance(tol).setMaxIterations(maxit)
        cg.solve(bd, xd, ctx=ctx)
        ctx.sync()
        t0 = time.perf_counter()
        cg.solve(bd, xd, ctx=ctx)
        ctx.sync()
        loop_t += time.perf_counter() - t0
    loop_sps = len(sample) / loop_t
    return {"systems": nsys, "rows": int(P["b"].size), "nnz": int(P["IA"].size), "block_size": bs,
            "workgroups_per_cu": gpc,
            "seconds": round(tb, 4), "solves_per_s": round(nsys / tb, 1),
            "plain_loop_solves_per_s": round(loop_sps, 1), "speedup_vs_plain_loop": round(nsys / tb / loop_sps, 2),
            "iterations_total": int(ih.sum()), "iterations_max": int(ih.max()), "max_error": float(eh.max())}


def bench_mixed(cm, capi, ctx):
    """3D 7-point variable-coefficient Poisson 160^3, log-normal(0,1) edges:
    FP64 iterative refinement around FP32 / BF16-storage inner CG against
    plain FP64 CG, all to relative residual 1e-12."""
    from cumat_amd.problems import grid_laplacian, lognormal_sampler
    from cumat_amd.solvers import pcg_ir
    P = grid_laplacian((160, 160, 160), lognormal_sampler(np.random.default_rng(20261206), 0.0, 1.0), edges=False)
    n = P["n"]
    b = np.random.default_rng(20261207).uniform(-1, 1, n)
    dV = cm.DeviceArray.from_numpy(P["vals"], ctx=ctx)
    db = cm.DeviceArray.from_numpy(b, ctx=ctx)
    A = cm.SparseMatrix(cm.SparsityPattern(n, n, P["JA"], P["IA"], ctx=ctx), dV)
    dJA, dIA = A.pattern.JA, A.pattern.IA
    out = {"n": n, "nnz": P["nnz"]}

    def true_rel(xd):
        r = A.matvec(xd, ctx=ctx).numpy() - b
        return float(np.linalg.norm(r) / np.linalg.norm(b))
    x64 = cm.DeviceArray(n, dtype=np.float64, ctx=ctx)
    cg = cm.ConjugateGradient(A, path=capi.CG_MULTIKERNEL).setTolerance(1e-12).setMaxIterations(40000)
    t0 = time.perf_counter()
    cg.solve(db, x64, ctx=ctx)
    ctx.sync()
    t64 = time.perf_counter() - t0
    out["fp64_cg"] = {"seconds": round(t64, 3), "iterations": cg.iterations(), "rel_residual": true_rel(x64)}
    xref = x64.numpy()
    # inner tolerances: the fastest of a sweep on this system (FP32 inner CG
    # stalls near 1e-6, BF16 storage near 1e-3)
    for storage, itol in (("f32", 1e-5), ("bf16", 1e-3)):
        xd = cm.DeviceArray(n, dtype=np.float64, ctx=ctx)
        st = pcg_ir(dJA, dIA, dV, db, xd, storage=storage, tol=1e-12, inner_tol=itol, max_outer=60, ctx=ctx)
        xh = xd.numpy()
        out["ir_" + storage] = {"seconds": round(st["seconds"], 3), "outer": st["outer"], "inner": st["inner"],
                                "rel_residual": true_rel(xd),
                                "err_vs_fp64_cg": float(np.linalg.norm(xh - xref) / np.linalg.norm(xref)),
                                "speedup_vs_fp64_cg": round(t64 / st["seconds"], 3)}
    return out


def bench_reference_workloads(cm, capi, ctx):
    """The reference's own published benchmark workloads (NVIDIA RTX 2070,
    numbers quoted from its benchmark result files), run here unchanged in
    size and method so the two can be set side by side (different hardware:
    informational, not a target)."""
    import ctypes as C
    from cumat_amd.problems import poisson2d
    out = {}
    # CG: 2D Poisson, rhs = rand()/RAND_MAX on the boundary nodes after
    # srand(42) in the reference's loop order, tol 1e-4, default max iterations
    # (benchmarks/conjugate_gradient/Implementation_cuMat.cu)
    libc = C.CDLL("libc.so.6")
    published_cg = {500: 201.583, 1000: 599.841}  # ms, Conjugate Gradient - 2D-Poisson Matrix.json:19,22
    for g in (500, 1000):
        JA, IA, V = poisson2d(g)
        n = g * g
        rhs = np.zeros(n, np.float32)
        libc.srand(42)
        for x in range(g):
            for y in range(g):
                if x == 0 or y == 0 or x == g - 1 or y == g - 1:
                    rhs[y + x * g] = libc.rand() / float(2147483647)
        A = cm.SparseMatrix.from_csr(n, n, JA, IA, V, ctx=ctx)
        bd = cm.DeviceArray.from_numpy(rhs, ctx=ctx)
        xd = cm.DeviceArray(n, dtype=np.float32, ctx=ctx)
        cm.ConjugateGradient(A).setTolerance(1e-4).solve(bd, xd, ctx=ctx)
        ctx.sync()
        ts = []
        for _ in range(2):  # timed as the reference does: set-up (diagonal) + solve, 2 runs
            t0 = time.perf_counter()
            cg = cm.ConjugateGradient(A).setTolerance(1e-4)
            cg.solve(bd, xd, ctx=ctx)
            ctx.sync()
            ts.append(time.perf_counter() - t0)
        ms = 1000 * sum(ts) / len(ts)
        out[f"cg_poisson{g}_tol1e-4"] = {"ms": round(ms, 3), "iterations": cg.iterations(), "error": cg.error(),
                                          "reference_rtx2070_ms": published_cg[g],
                                          "ratio": round(published_cg[g] / ms, 1)}
    # CSR SpMV, grid 1000 (CSRMV - 2D-Poisson Matrix.json:24): 0.163467 ms
    JA, IA, V = poisson2d(1000)
    n = 1000 * 1000
    A = cm.SparseMatrix.from_csr(n, n, JA, IA, V, ctx=ctx)
    xd = cm.DeviceArray.from_numpy(np.random.default_rng(1).uniform(-1, 1, n).astype(np.float32), ctx=ctx)
    yd = cm.DeviceArray(n, dtype=np.float32, ctx=ctx)
    A.matvec(xd, yd, ctx=ctx)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(100):
        A.matvec(xd, yd, ctx=ctx)
    ctx.sync()
    ms = (time.perf_counter() - t0) * 10
    out["spmv_poisson1000"] = {"ms": round(ms, 4), "reference_rtx2070_ms": 0.163467, "ratio": round(0.163467 / ms, 1)}
    # dot, n = 1e8 FP32, result copied to the host each call (Dotproduct.json:77)
    n = 100_000_000
    a = cm.DeviceArray.from_numpy(np.ones(n, np.float32), ctx=ctx)
    b = cm.DeviceArray.from_numpy(np.full(n, 0.5, np.float32), ctx=ctx)
    d = cm.DeviceArray(1, dtype=np.float32, ctx=ctx)
    cm.dot(a, b, out=d, ctx=ctx)
    d.numpy()
    t0 = time.perf_counter()
    for _ in range(20):
        cm.dot(a, b, out=d, ctx=ctx)
        v = d.numpy()
    ms = (time.perf_counter() - t0) * 1000 / 20
    out["dot_1e8_with_host_copy"] = {"ms": round(ms, 4), "value": float(v[0]), "reference_rtx2070_ms": 2.37392,
                                     "ratio": round(2.37392 / ms, 1)}
    del a, b
    # linear combination k = 2, n = 5e7 (Linear Combination - Constant Count.json:45)
    n = 50_000_000
    v0 = cm.DeviceArray.from_numpy(np.ones(n, np.float32), ctx=ctx)
    v1 = cm.DeviceArray.from_numpy(np.full(n, 2.0, np.float32), ctx=ctx)
    o = cm.DeviceArray(n, dtype=np.float32, ctx=ctx)
    e = v0.expr() * 0.5 + v1.expr() * 0.25
    cm.assign(o, e, ctx=ctx)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(50):
        cm.assign(o, e, ctx=ctx)
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1000 / 50
    out["lincomb_k2_5e7"] = {"ms": round(ms, 4), "reference_rtx2070_ms": 2.00751, "ratio": round(2.00751 / ms, 1)}
    return out


def bench_adjoint(cm, capi, ctx):
    """64 variable-coefficient 2D 512^2 Poisson systems, theta = exp(N(0, 0.5^2))
    per edge: FP32 forward CG (tol 1e-7) + adjoint CG + fused edge gradient of
    L = 0.5 |x - x*|^2, through the C ABI (the kernels the autograd function
    calls)."""
    import ctypes as C
    from cumat_amd.problems import grid_laplacian
    B, g, tol = 64, 512, 1e-7
    P = grid_laplacian((g, g), lambda s: np.ones(s))
    rng = np.random.default_rng(20261208)
    n, nnz, ne = P["n"], P["nnz"], P["nedges"]
    theta = np.exp(0.5 * rng.standard_normal((B, ne))).astype(np.float32)
    b = rng.uniform(-1, 1, (B, n)).astype(np.float32)
    xs = rng.standard_normal((B, n)).astype(np.float32)
    d = {k: cm.DeviceArray.from_numpy(P[k], ctx=ctx) for k in ("JA", "IA", "eptr", "eidx", "ei", "ej")}
    dth = cm.DeviceArray.from_numpy(theta.reshape(-1), ctx=ctx)
    db = cm.DeviceArray.from_numpy(b.reshape(-1), ctx=ctx)
    dxs = cm.DeviceArray.from_numpy(xs.reshape(-1), ctx=ctx)
    vals = cm.DeviceArray(B * nnz, dtype=np.float32, ctx=ctx)
    x = cm.DeviceArray(B * n, dtype=np.float32, ctx=ctx)
    gx = cm.DeviceArray(B * n, dtype=np.float32, ctx=ctx)
    lam = cm.DeviceArray(B * n, dtype=np.float32, ctx=ctx)
    grad = cm.DeviceArray(B * ne, dtype=np.float32, ctx=ctx)
    its = (C.c_int * B)()
    errs = (C.c_double * B)()
    st = capi.CgStats()

    def solve(rhs, out):
        capi.call("cm_pcg_solve_ex", ctx.handle, capi.F32, n, nnz, B, d["JA"].ptr, d["IA"].ptr, vals.ptr, 1, None,
                  rhs.ptr, out.ptr, 0, tol, 20000, 1, its, errs, C.byref(st))
        return list(its), list(errs)
    ctx.sync()
    t0 = time.perf_counter()
    capi.call("cm_edge_assemble", ctx.handle, capi.F32, nnz, B, ne, d["eptr"].ptr, d["eidx"].ptr, dth.ptr, vals.ptr)
    f_its, f_errs = solve(db, x)
    ctx.sync()
    t1 = time.perf_counter()
    cm.assign(gx, x.expr() - dxs.expr(), ctx=ctx)
    b_its, b_errs = solve(gx, lam)
    capi.call("cm_edge_grad", ctx.handle, capi.F32, ne, B, n, d["ei"].ptr, d["ej"].ptr, lam.ptr, x.ptr, grad.ptr)
    ctx.sync()
    t2 = time.perf_counter()
    gh = grad.numpy()
    return {"systems": B, "n_each": n, "nnz_each": nnz, "edges_each": ne, "forward_s": round(t1 - t0, 3),
            "backward_s": round(t2 - t1, 3), "forward_iters_max": max(f_its), "backward_iters_max": max(b_its),
            "forward_err_max": max(f_errs), "backward_err_max": max(b_errs),
            "grad_finite": bool(np.isfinite(gh).all())}


def cpu_baseline(JA, IA, V, b, n, nnz):
    """The reference's vendored Eigen ConjugateGradient (oracle/_ref) on the
    host cores: a bounded sample of 2048^2 CG iterations (tol 0)."""
    try:
        import oracle
        if not oracle.ref_available():
            return None
        L = oracle.ref()
        threads = len(os.sched_getaffinity(0))
        JA = np.ascontiguousarray(JA)
        IA = np.ascontiguousarray(IA)
        h = L.eref_csr_create_f32(n, n, nnz, JA.ctypes.data, IA.ctypes.data, V.ctypes.data)
        x = np.zeros(n, np.float32)
        err = C.c_double()
        L.eref_pcg_handle_f32(h, b.ctypes.data, x.ctypes.data, 0.0, 5, threads, C.byref(err))  # warm
        sample = 100
        t0 = time.perf_counter()
        it = L.eref_pcg_handle_f32(h, b.ctypes.data, x.ctypes.data, 0.0, sample, threads, C.byref(err))
        t = time.perf_counter() - t0
        v = it / t
        log(f"CPU Eigen CG 2048^2: {v:.2f} it/s on {threads} threads")
        rows = {}
        # the other measured rows on the same Eigen build, bounded samples
        y = np.zeros(n, np.float32)
        L.eref_spmv_f32(h, b.ctypes.data, y.ctypes.data, threads)
        t0 = time.perf_counter()
        for _ in range(10):
            L.eref_spmv_f32(h, b.ctypes.data, y.ctypes.data, threads)
        dt = (time.perf_counter() - t0) / 10
        rows["spmv_2048^2"] = {"ms": round(dt * 1e3, 3), "GBps": round((8 * nnz + 12 * n) / dt / 1e9, 1),
                               "threads": threads}
        L.eref_csr_destroy_f32(h)
        N = 1 << 26
        rng = np.random.default_rng(5)
        vecs = [rng.random(N, dtype=np.float32) for _ in range(8)]
        coef = np.linspace(-1, 1, 8).astype(np.float32)
        out = np.empty(N, np.float32)
        ptrs = (C.c_void_p * 8)(*[v.ctypes.data for v in vecs])
        t0 = time.perf_counter()
        L.eref_lincomb8_f32(N, ptrs, coef.ctypes.data, out.ctypes.data)
        dt = time.perf_counter() - t0
        rows["lincomb8_2^26"] = {"ms": round(dt * 1e3, 2), "GBps": round(9 * 4 * N / dt / 1e9, 1), "threads": 1}
        t0 = time.perf_counter()
        L.eref_dot_f32(N, vecs[0].ctypes.data, vecs[1].ctypes.data)
        dt = time.perf_counter() - t0
        rows["dot_2^26"] = {"ms": round(dt * 1e3, 2), "GBps": round(8 * N / dt / 1e9, 1), "threads": 1}
        R_, C_, B_ = 1024, 1024, 64
        for name, axis, osz in (("row_sum", 1, C_ * B_), ("col_sum", 2, R_ * B_), ("batch_sum", 4, R_ * C_)):
            o = np.empty(osz, np.float32)
            t0 = time.perf_counter()
            L.eref_axis_sum_f32(axis, R_, C_, B_, vecs[2].ctypes.data, o.ctypes.data)
            dt = time.perf_counter() - t0
            rows[name + "_1024x1024x64"] = {"ms": round(dt * 1e3, 2), "GBps": round(4 * N / dt / 1e9, 1),
                                            "threads": 1}
        del vecs
        # the same 2048^2 CG on one thread (30 iterations)
        h = L.eref_csr_create_f32(n, n, nnz, JA.ctypes.data, IA.ctypes.data, V.ctypes.data)
        t0 = time.perf_counter()
        it1 = L.eref_pcg_handle_f32(h, b.ctypes.data, x.ctypes.data, 0.0, 30, 1, C.byref(err))
        rows["cg_2048^2_1thread"] = {"iters_per_s": round(it1 / (time.perf_counter() - t0), 2), "threads": 1}
        L.eref_csr_destroy_f32(h)
        log("CPU Eigen rows:", rows)
        configs = cpu_other_configs(oracle, L, threads)
        log("CPU Eigen configs:", configs)
        return {"value": round(v, 3), "unit": "CG iterations/s", "cores": threads, "kind": "reference",
                "sample": f"{it} Jacobi-PCG iterations of Eigen 3.3.90 (reference third-party) on the same "
                          f"2048^2 FP32 system, Eigen::setNbThreads({threads})",
                "cpu_model": cpu_model(), "other_rows": rows, "configs": configs}
    except Exception as ex:  # the baseline is reported, never required
        log("cpu baseline failed:", ex)
        return None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _median_time(fn, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), r


def cpu_other_configs(oracle, L, threads):
    """Eigen (oracle/_ref) on the host cores for the other workloads, each on
    a bounded sample: the 128^2 solve (1 thread and all threads, median of 5),
    the 4096-system batch (all systems over all threads, every 16th system on
    one thread), 20 FP64 iterations of the 160^3 system (time to 1e-12
    extrapolated to the 1000 iterations FP64 CG needs there), and one 512^2
    system's FP64 forward + adjoint solves (extrapolated to 64 systems)."""
    from cumat_amd.problems import batched_poisson, grid_laplacian, lognormal_sampler, poisson2d
    out = {}
    err = C.c_double()
    # config 1: 128^2 FP32, tol 1e-6, max 1000
    JA, IA, V = poisson2d(128)
    n = 128 * 128
    b = np.random.default_rng(20261203).uniform(-1, 1, n).astype(np.float32)
    x = np.zeros(n, np.float32)
    for nt in (1, threads):
        t, it = _median_time(lambda: L.eref_pcg_f32(n, IA.size, JA.ctypes.data, IA.ctypes.data, V.ctypes.data,
                                                    b.ctypes.data, x.ctypes.data, 0, 1e-6, 1000, nt,
                                                    C.byref(err)), 5)
        out[f"cg128_{nt}threads"] = {"iterations": it, "ms_to_solution": round(t * 1e3, 3),
                                     "iters_per_s": round(it / t, 1), "threads": nt}
    # config 3: 4096 systems, side 16..96 (the GPU bench's seed), tol 1e-6
    P = batched_poisson(4096, 16, 96, seed=20261205)
    xb = np.zeros(P["b"].size, np.float32)
    its = np.zeros(4096, np.int32)
    args = (P["row_off"].ctypes.data, P["JA"].ctypes.data, P["IA"].ctypes.data, P["vals"].ctypes.data,
            P["b"].ctypes.data, xb.ctypes.data, 1e-6, 5000)
    t0 = time.perf_counter()
    ns = L.eref_batched_pcg_f32(4096, *args, threads, 1, its.ctypes.data)
    t = time.perf_counter() - t0
    out["batched_4096_all_threads"] = {"systems": ns, "seconds": round(t, 3), "solves_per_s": round(ns / t, 1),
                                       "threads": threads}
    t0 = time.perf_counter()
    ns = L.eref_batched_pcg_f32(4096, *args, 1, 16, its.ctypes.data)
    t = time.perf_counter() - t0
    out["batched_4096_1thread"] = {"systems": ns, "sample": "every 16th system", "seconds": round(t, 3),
                                   "solves_per_s": round(ns / t, 1), "threads": 1}
    del P, xb
    # config 4: 160^3 log-normal(0,1), FP64 CG (the bench's seeds)
    P = grid_laplacian((160, 160, 160), lognormal_sampler(np.random.default_rng(20261206), 0.0, 1.0), edges=False)
    n = P["n"]
    b = np.random.default_rng(20261207).uniform(-1, 1, n)
    x = np.zeros(n)

    def run(k):
        return L.eref_pcg_f64(n, P["nnz"], P["JA"].ctypes.data, P["IA"].ctypes.data, P["vals"].ctypes.data,
                              b.ctypes.data, x.ctypes.data, 0, 1e-300, k, threads, C.byref(err))
    t_1, _ = _median_time(lambda: run(1), 1)    # matrix build + set-up + 1 iteration
    t_21, it = _median_time(lambda: run(21), 1)
    per = (t_21 - t_1) / 20
    out["fp64_cg_160cubed"] = {"sample_iterations": 20, "ms_per_iteration": round(per * 1e3, 2),
                               "est_seconds_to_1e-12_at_1000_iterations": round(per * 1000, 2),
                               "threads": threads}
    del P, x
    # config 5: one 512^2 system, theta = exp(N(0, 0.5^2)), FP64 forward + adjoint at tol 1e-7
    rng = np.random.default_rng(20261208)
    P0 = grid_laplacian((512, 512), lambda s: np.ones(s), edges=False)
    ne = P0["nedges"]
    theta = np.exp(0.5 * rng.standard_normal(ne))
    state = {"o": 0}

    def samp(shape):
        k = int(np.prod(shape))
        o = theta[state["o"]:state["o"] + k].reshape(shape)
        state["o"] += k
        return o
    P = grid_laplacian((512, 512), samp, edges=False)
    n = P["n"]
    b = rng.uniform(-1, 1, n)
    xs = rng.standard_normal(n)
    x = np.zeros(n)
    lam = np.zeros(n)
    t0 = time.perf_counter()
    itf = L.eref_pcg_f64(n, P["nnz"], P["JA"].ctypes.data, P["IA"].ctypes.data, P["vals"].ctypes.data,
                         b.ctypes.data, x.ctypes.data, 0, 1e-7, 20000, threads, C.byref(err))
    t1 = time.perf_counter()
    gx = x - xs
    itb = L.eref_pcg_f64(n, P["nnz"], P["JA"].ctypes.data, P["IA"].ctypes.data, P["vals"].ctypes.data,
                         gx.ctypes.data, lam.ctypes.data, 0, 1e-7, 20000, threads, C.byref(err))
    t2 = time.perf_counter()
    out["adjoint_512sq_fp64"] = {"systems_sampled": 1, "forward_s": round(t1 - t0, 3), "backward_s": round(t2 - t1, 3),
                                 "forward_iterations": itf, "backward_iterations": itb,
                                 "est_seconds_64_systems": round(64 * (t2 - t0), 1), "threads": threads}
    return out


if __name__ == "__main__":
    main()
