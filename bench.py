"""Benchmark: LSQR time-to-1e-6 on one MI355X, mixed precision + sketch-QR
preconditioner vs plain FP32, plus the packed GEMV's fraction of HBM peak.

Workload (one "step"): A = G diag(d), 2,000,000 x 1000 FP32, d log-spaced
1 .. 1e-6 (kappa ~ 1e6), stored as FP32_BF16 (planar wh/wr, 8 GB); b = A
x_true in FP64 (a consistent system, so the FP64 least-squares solution is
x_true).  A step builds the preconditioner (hashed CountSketch with s = 4000
rows, FP64 Householder QR, R^-1) and runs restarted LSQR whose inner
iterations use the hi-only (bf16) GEMV and whose restart residual uses the
exact (wh+wr) GEMV, to relative residual 1e-6.  All inputs are resident in
HBM before timing.  Plain FP32 LSQR (FP32 A, no preconditioner, same
stopping rule) is timed once beside it, capped at --plain-cap iterations.

The roofline object reports the LSQR iteration's dominant kernel, the
single-pass A p / A^T u kernel on the hi plane (tomos_matvec_pair) at the
full 2e6 x 1000 shape, timed with HIP events on the launch stream.  The CPU
baseline is the oracle's FP32 LSQR with OpenBLAS sgemv (the reference's CPU
path) on a row sample of the same matrix, scaled to the full row count.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--m", type=int, default=2_000_000)
    p.add_argument("--n", type=int, default=1000)
    p.add_argument("--s", type=int, default=4000)
    p.add_argument("--kappa-exp", type=float, default=6.0)
    p.add_argument("--tol", type=float, default=1e-6)
    p.add_argument("--inner", type=int, default=8,
                   help="LSQR iterations per restart (mixed): a short restart cycle lets the exact "
                        "residual correct the hi-only iterations early")
    p.add_argument("--tune", type=int, default=1,
                   help="autotune the single-pass kernel's launch shape at the workload shapes first")
    p.add_argument("--restarts", type=int, default=40)
    p.add_argument("--plain-cap", type=int, default=1500)
    p.add_argument("--gemv-reps", type=int, default=20)
    p.add_argument("--cpu-rows", type=int, default=0,
                   help="rows of the workload the CPU baseline uses (0 = all)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--check-every", type=int, default=8)
    p.add_argument("--fused", type=int, default=1,
                   help="single-pass LSQR matvec kernel (A read once per iteration)")
    p.add_argument("--gemm", type=int, default=8192, help="split GEMM size (0 = skip)")
    p.add_argument("--c2", type=int, default=1, help="FP32_16 GEMV / GEMV^T rows at 1e6 x 1000 (0 = skip)")
    p.add_argument("--sparse-row", type=int, default=50,
                   help="entries per row of the CSR LSQR leg (0 = skip)")
    p.add_argument("--batched", type=int, default=64,
                   help="batched differentiable least squares: number of uneven problems (0 = skip)")
    p.add_argument("--batched-sketch-factor", type=int, default=8,
                   help="sketch rows of the batched problems: s = factor * n (8: fewer forward and adjoint "
                        "iterations than 4 for a slightly dearer sketch QR)")
    p.add_argument("--batched-restarts", type=int, default=12,
                   help="restart cap of the batched solve (its graph holds that many restarts)")
    p.add_argument("--adjoint-workers", type=int, default=8,
                   help="host threads (one HIP stream each) of the batched backward's host-thread form "
                        "(0: that form is not run)")
    p.add_argument("--stream-loads", type=int, default=0, choices=(0, 1, 2),
                   help="matrix load policy: 0 auto (non-temporal above 192 MiB per pass), "
                        "1 always non-temporal, 2 always cacheable")
    p.add_argument("--qr-gemm-only", action="store_true",
                   help="only the sketch-QR preconditioner at 4000 x 1000 and the packed GEMMs at --gemm^3 "
                        "(split and hi-only, FP32_BF16), two calls each (for counter runs); prints their timings")
    p.add_argument("--kernels-only", action="store_true",
                   help="only the hi-only GEMV-N at the full shape and a copy of the same bytes "
                        "(for hardware-counter runs); prints their timings")
    return p.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


# Untimed launches before each kernel timing: after an idle gap the device
# clocks go through a transient of ~15 ms (a 4 GB stream measured 0.67 ms,
# then 0.82, then 0.68 per launch); the timings are of the steady state.
WARM_LAUNCHES = 20


def time_call(fn, reps, ev0, ev1):
    """Average ms of fn over reps launches on the NULL stream (after warm-up launches)."""
    from tomos_amd import _capi as C
    for _ in range(WARM_LAUNCHES):
        fn()
    C.call("tomos_event_record", ev0, None)
    for _ in range(reps):
        fn()
    C.call("tomos_event_record", ev1, None)
    f = ctypes.c_float()
    C.call("tomos_event_elapsed_ms", ctypes.byref(f), ev0, ev1)
    return f.value / reps


def time_pair(C, DeviceBuffer, args, m, n, wh, ev0, ev1):
    """Average ms of tomos_matvec_pair (u <- A p - 0.5 u, t <- A^T u) on the
    hi plane, over --gemv-reps launches on the launch stream."""
    rng = np.random.default_rng(4)
    pd = DeviceBuffer.from_numpy(rng.standard_normal(n))
    ud = DeviceBuffer.from_numpy(rng.standard_normal(m).astype(np.float32))
    td, nd = DeviceBuffer(n * 8), DeviceBuffer(8)
    sz = C.lib().tomos_matvec_pair_scratch_size(C.FP32_BF16, m, n)
    sc = DeviceBuffer(sz)
    call = lambda: C.call("tomos_matvec_pair", C.FP32_BF16, m, n, wh.ptr, m, pd.ptr, -0.5, ud.ptr,
                          td.ptr, nd.ptr, sc.ptr, sz, None)
    for _ in range(WARM_LAUNCHES):
        call()
    C.call("tomos_event_record", ev0, None)
    for _ in range(args.gemv_reps):
        call()
    C.call("tomos_event_record", ev1, None)
    f = ctypes.c_float()
    C.call("tomos_event_elapsed_ms", ctypes.byref(f), ev0, ev1)
    return f.value / args.gemv_reps


def sparse_leg(args, C, DeviceBuffer, ev0, ev1):
    """CSR LSQR at the tall shape: m x n with args.sparse_row entries per row
    at uniform random columns (seed 8), N(0,1) FP32 values, int32 indices;
    b = A x_true.  Times A p (the CSR row kernel, HBM-bound), the LSQR
    iteration (A p + A^T u from the panel-blocked transpose + vector
    kernels: a 40-iteration run minus a 0-iteration run), the setup
    (validation, transpose, first residual) and a full unpreconditioned
    FP32/FP64 solve to 1e-6 (the matrix is well conditioned: m >> n, i.i.d.
    entries)."""
    import ctypes
    m, n, k = args.m, args.n, args.sparse_row
    nnz = m * k
    rng = np.random.default_rng(8)
    rp = DeviceBuffer.from_numpy((np.arange(m + 1, dtype=np.int64) * k).astype(np.int32))
    ci = DeviceBuffer.from_numpy(rng.integers(0, n, nnz, dtype=np.int32))
    vals = DeviceBuffer(nnz * 4)
    C.call("tomos_fill_normal_f32", vals.ptr, nnz, 1, nnz, 21, None, 0, 0, None)
    xt = rng.standard_normal(n)
    xd = DeviceBuffer.from_numpy(xt)
    bd = DeviceBuffer(m * 8)
    C.call("tomos_csr_mv", C.F32, C.INDEX_INT32, 0, 1.0, m, n, nnz, vals.ptr, rp.ptr, ci.ptr, C.F64, xd.ptr, 0.0,
           C.F64, bd.ptr, None, 0, None)
    # A p: rows of A
    y = DeviceBuffer(m * 8)
    sn = C.lib().tomos_csr_mv_scratch_size(C.F32, C.INDEX_INT32, 0, m, n, nnz)
    scn = DeviceBuffer(sn)
    t_n = time_call(lambda: C.call("tomos_csr_mv", C.F32, C.INDEX_INT32, 0, 1.0, m, n, nnz, vals.ptr, rp.ptr,
                                   ci.ptr, C.F64, xd.ptr, 0.0, C.F64, y.ptr, scn.ptr, sn, None), 10, ev0, ev1)
    del y, scn
    # the solve
    opt = C.LsqrOptions(C.F32, C.TIER_DEFAULT, C.TIER_DEFAULT, C.F32, C.F64, -1, 1e-6, 1e-6, 200, 1, 8, 0)
    ab = C.lib().tomos_lsqr_arena_size(ctypes.byref(opt), m, n)
    sb = C.lib().tomos_lsqr_csr_scratch_size(ctypes.byref(opt), C.INDEX_INT32, m, n, nnz)
    arena, scr = DeviceBuffer(ab), DeviceBuffer(sb)
    x = DeviceBuffer(n * 8)
    its = ctypes.c_int(0)
    solve = lambda: C.call("tomos_lsqr_solve_csr", ctypes.byref(opt), C.INDEX_INT32, m, n, nnz, vals.ptr, rp.ptr,
                           ci.ptr, -1, None, bd.ptr, x.ptr, None, 1, arena.ptr, ab, scr.ptr, sb, ctypes.byref(its),
                           None, 0, None, None)
    solve()
    C.call("tomos_device_synchronize")
    t0 = time.perf_counter()
    solve()
    C.call("tomos_device_synchronize")
    t_solve = (time.perf_counter() - t0) * 1e3
    iters = its.value
    err = float(np.linalg.norm(x.to_numpy(np.float64, (n,)) - xt) / np.linalg.norm(xt))

    # per-iteration cost: a 40-iteration run (tolerance out of reach) minus
    # a 0-iteration run (validation, transpose, first residual)
    def timed(max_iters):
        o = C.LsqrOptions(C.F32, C.TIER_DEFAULT, C.TIER_DEFAULT, C.F32, C.F64, -1, 1e-30, 1e-30, max_iters, 1,
                          1000, 0)
        C.call("tomos_device_synchronize")
        t = time.perf_counter()
        C.call("tomos_lsqr_solve_csr", ctypes.byref(o), C.INDEX_INT32, m, n, nnz, vals.ptr, rp.ptr, ci.ptr, -1,
               None, bd.ptr, x.ptr, None, 1, arena.ptr, ab, scr.ptr, sb, ctypes.byref(its), None, 0, None, None)
        C.call("tomos_device_synchronize")
        return (time.perf_counter() - t) * 1e3
    timed(0)
    t_setup = timed(0)
    t_iter = (timed(40) - t_setup) / 40.0
    gbs = lambda byts, ms: byts / (ms * 1e-3) / 1e9
    n_bytes = nnz * 8.0 + (m + 1) * 4.0 + n * 8.0 + m * 8.0
    t_bytes = nnz * 8.0 + (n + 1) * 4.0 + m * 8.0 + n * 8.0
    t_t = t_iter - t_n  # the iteration's A^T u (plus its few vector kernels)
    return {"m": m, "n": n, "nnz": nnz, "values": "f32", "index": "int32",
            "Ap_ms": round(t_n, 4), "Ap_GBps": round(gbs(n_bytes, t_n), 1),
            "iteration_ms": round(t_iter, 4), "iteration_GBps": round(gbs(n_bytes + t_bytes, t_iter), 1),
            "ATu_ms_iteration_minus_Ap": round(t_t, 4), "ATu_GBps": round(gbs(t_bytes, t_t), 1),
            "setup_ms": round(t_setup, 2),
            "solve_ms": round(t_solve, 2), "iterations": iters, "err_vs_x_true": err,
            "bytes_per_product": {"Ap": n_bytes, "ATu": t_bytes}}


def batched_leg(args, C, DeviceBuffer, ev0, ev1):
    """64 uneven problems: m ~ U{20000..200000}, n ~ U{64..512}, kappa
    log-uniform in [1e2, 1e5] (seed 7), A = G diag(d) stored FP32_BF16,
    b = A x_true.  Forward: batched CountSketch (s = 8n by default) + batched sketch-QR +
    batched LSQR to 1e-6 (one hipGraph per restart).  Backward: for
    g ~ N(0,1), w = (A^T A)^-1 g per problem (adjoint CG) and grad_b, grad_A
    (rank-2 kernel, m x n FP32 each).  Also the same forward one problem at
    a time with the single-problem solver."""
    import ctypes
    rng = np.random.default_rng(7)
    P = args.batched
    ms = rng.integers(20000, 200001, P)
    ns = rng.integers(64, 513, P)
    kap = 10.0 ** rng.uniform(2, 5, P)
    probs, keep = [], []
    for i in range(P):
        m, n = int(ms[i]), int(ns[i])
        lda = (m + 63) // 64 * 64  # 128-byte columns: the tiles' column runs start on cache lines
        d = DeviceBuffer.from_numpy(np.logspace(0, -np.log10(kap[i]), n).astype(np.float32))
        A = DeviceBuffer(lda * n * 4)
        C.call("tomos_fill_normal_f32", A.ptr, m, n, lda, 100 + i, d.ptr, 0, 0, None)
        wh, wr = DeviceBuffer(lda * n * 2), DeviceBuffer(lda * n * 2)
        C.call("tomos_pack", C.PACK_FP32_BF16, m, n, A.ptr, lda, wh.ptr, wr.ptr, lda, None)
        del A
        xt = DeviceBuffer.from_numpy(rng.standard_normal(n))
        bd = DeviceBuffer(m * 8)
        C.call("tomos_gemv", C.FP32_BF16, C.TIER_EXACT, 0, 1.0, m, n, wh.ptr, wr.ptr, lda, C.F64, xt.ptr, 0.0,
               C.F64, bd.ptr, None, 0, None)
        s = args.batched_sketch_factor * n
        probs.append(dict(m=m, n=n, lda=lda, wh=wh, wr=wr, b=bd, xt=xt, s=s, sk=DeviceBuffer(s * n * 8),
                          x=DeviceBuffer(n * 8)))
    cp = (C.CsProblem * P)(*[C.CsProblem(p["m"], p["n"], p["lda"], p["s"], p["s"], p["wh"].ptr, p["wr"].ptr,
                                         p["sk"].ptr) for p in probs])
    cs_bytes = C.lib().tomos_countsketch_apply_batched_scratch_size(P, cp)
    cs = DeviceBuffer(cs_bytes)
    qp = (C.QrProblem * P)(*[C.QrProblem(p["s"], p["n"], p["s"], p["sk"].ptr) for p in probs])
    qsb = C.lib().tomos_sketch_qr_precond_batched_scratch_size(P, qp)
    qs = DeviceBuffer(qsb)
    # the whole solve is one graph launch with blind restarts: the restart cap
    # bounds the launches enqueued (12 x 8 iterations, over twice what C5 needs)
    o = C.LsqrOptions(C.FP32_BF16, C.TIER_DEFAULT, C.TIER_EXACT, C.F32, C.F64, C.F64, 1e-6, 1e-6, args.inner,
                      args.batched_restarts, 8, 1, 1)
    lp = (C.LsqProblem * P)(*[C.LsqProblem(p["m"], p["n"], p["lda"], p["wh"].ptr, p["wr"].ptr, p["b"].ptr,
                                           p["x"].ptr, p["sk"].ptr, p["s"]) for p in probs])
    sb = C.lib().tomos_lsqr_batched_scratch_size(ctypes.byref(o), P, lp, 0, 0, 0)
    scr = DeviceBuffer(sb)
    its = (ctypes.c_int * P)()
    rr = (ctypes.c_double * P)()
    C.call("tomos_device_synchronize")
    log("batched: %d problems, %.2f G elements; data ready" % (P, float((ms * ns).sum()) / 1e9))

    def forward():
        t0 = time.perf_counter()
        C.call("tomos_countsketch_apply_batched", C.FP32_BF16, C.TIER_EXACT, P, cp, cs.ptr, cs_bytes, None)
        C.call("tomos_sketch_qr_precond_batched_f64", P, qp, qs.ptr, qsb, None)
        t1 = time.perf_counter()
        C.call("tomos_lsqr_batched_solve", ctypes.byref(o), P, lp, 0, 0, 0, 1, scr.ptr, sb, its, rr, None)
        C.call("tomos_device_synchronize")
        t2 = time.perf_counter()
        return (t1 - t0) * 1e3, (t2 - t1) * 1e3

    forward()  # warm-up (graph instantiation, code objects)
    pre_ms, solve_ms = forward()
    # true relative residuals (exact A, FP64)
    worst = 0.0
    for p in probs:
        r = DeviceBuffer(p["m"] * 8)
        C.call("tomos_memcpy_dtod", r.ptr, p["b"].ptr, p["m"] * 8, None)
        C.call("tomos_gemv", C.FP32_BF16, C.TIER_EXACT, 0, -1.0, p["m"], p["n"], p["wh"].ptr, p["wr"].ptr,
               p["lda"], C.F64, p["x"].ptr, 1.0, C.F64, r.ptr, None, 0, None)
        rn = np.linalg.norm(r.to_numpy(np.float64, (p["m"],)))
        bn = np.linalg.norm(p["b"].to_numpy(np.float64, (p["m"],)))
        worst = max(worst, rn / bn)
    # the same forward solves one problem at a time (single-problem solver)
    so = C.LsqrOptions(C.FP32_BF16, C.TIER_DEFAULT, C.TIER_EXACT, C.F32, C.F64, C.F64, 1e-6, 1e-6, args.inner, 60, 8, 1,
                       1)
    mmax = int(ms.max())
    nmax = int(ns.max())
    sarena = DeviceBuffer(C.lib().tomos_lsqr_arena_size(ctypes.byref(so), mmax, nmax))
    sscr = DeviceBuffer(max(C.lib().tomos_lsqr_scratch_size(ctypes.byref(so), p["m"], p["n"]) for p in probs))
    xs = DeviceBuffer(nmax * 8)
    C.call("tomos_device_synchronize")
    t0 = time.perf_counter()
    for p in probs:
        it = ctypes.c_int(0)
        C.call("tomos_lsqr_solve", ctypes.byref(so), p["m"], p["n"], p["wh"].ptr, p["wr"].ptr, p["lda"], p["b"].ptr,
               xs.ptr, p["sk"].ptr, p["s"], sarena.ptr, sarena.nbytes, sscr.ptr, sscr.nbytes, ctypes.byref(it),
               None, 0, None, None)
    C.call("tomos_device_synchronize")
    seq_ms = (time.perf_counter() - t0) * 1e3
    # backward: adjoint solve + gradients for every problem, at the adjoint
    # tolerance the gradient tests hold (|g - A^T A w| / |g| < 1e-10)
    adj_tol = 1e-10
    ao = C.LsqrOptions(C.FP32_BF16, C.TIER_DEFAULT, C.TIER_EXACT, C.F32, C.F64, C.F64, adj_tol, adj_tol, 30, 20, 8, 1)
    asb = max(C.lib().tomos_lsq_adjoint_scratch_size(ctypes.byref(ao), p["m"], p["n"]) for p in probs)
    ascr = DeviceBuffer(asb)
    gsb = max(C.lib().tomos_lsq_grad_scratch_size(C.FP32_BF16, C.TIER_EXACT, p["m"], p["n"]) for p in probs)
    gscr = DeviceBuffer(gsb)
    gA = DeviceBuffer(max(p["m"] * p["n"] for p in probs) * 4)
    gb = DeviceBuffer(mmax * 8)
    wd = DeviceBuffer(nmax * 8)
    gs = [DeviceBuffer.from_numpy(rng.standard_normal(p["n"])) for p in probs]
    adj_its = []
    C.call("tomos_device_synchronize")
    t0 = time.perf_counter()
    for p, g in zip(probs, gs):
        it = ctypes.c_int(0)
        C.call("tomos_lsq_adjoint_solve", ctypes.byref(ao), p["m"], p["n"], p["wh"].ptr, p["wr"].ptr, p["lda"],
               g.ptr, p["sk"].ptr, p["s"], wd.ptr, ascr.ptr, asb, ctypes.byref(it), None, 0, None, None)
        C.call("tomos_lsq_grad", C.FP32_BF16, C.TIER_EXACT, p["m"], p["n"], p["wh"].ptr, p["wr"].ptr, p["lda"],
               p["b"].ptr, p["x"].ptr, wd.ptr, gb.ptr, gA.ptr, p["m"], gscr.ptr, gsb, None)
        adj_its.append(it.value)
    C.call("tomos_device_synchronize")
    bwd_seq_ms = (time.perf_counter() - t0) * 1e3
    # the same backward with tomos_lsq_adjoint_batched: problems spread over
    # host threads, one HIP stream each (every problem its own w, grad_b, grad_A)
    P_ = len(probs)
    wds = [DeviceBuffer(p["n"] * 8) for p in probs]
    gbds = [DeviceBuffer(p["m"] * 8) for p in probs]
    gads = [DeviceBuffer(p["m"] * p["n"] * 4) for p in probs]
    ptrs = lambda bufs: (ctypes.c_void_p * P_)(*[x.ptr for x in bufs])
    ldg = (ctypes.c_int64 * P_)(*[p["m"] for p in probs])
    def batched_backward(workers):
        sbytes = C.lib().tomos_lsq_adjoint_batched_scratch_size(ctypes.byref(ao), P_, lp, workers)
        sc_ = DeviceBuffer(sbytes)
        its_ = (ctypes.c_int * P_)()
        bargs = (ctypes.byref(ao), P_, lp, ptrs(gs), ptrs(wds), ptrs(gbds), ptrs(gads), ldg, workers, sc_.ptr, sbytes,
                 its_, None)
        C.call("tomos_lsq_adjoint_batched", *bargs)  # warm-up (graphs, threads and streams)
        C.call("tomos_device_synchronize")
        t0_ = time.perf_counter()
        C.call("tomos_lsq_adjoint_batched", *bargs)
        C.call("tomos_device_synchronize")
        return (time.perf_counter() - t0_) * 1e3, list(its_)

    aw = args.adjoint_workers
    bwd_thr_ms = None
    if aw > 0:  # 0 skips the host-thread form (kernel-trace profiling runs)
        bwd_thr_ms, thr_its = batched_backward(aw)
        assert thr_its == adj_its, "threaded batched adjoint iterations differ from the one-problem calls"
    # all problems in grouped device launches; its results are the ones checked below
    bwd_ms, grp_its = batched_backward(0)
    # gradients of the two smallest problems against the dense FP64 QR gradient
    # (host numpy QR of the exactly decoded A), for the loss g . x
    grad_err = []
    for i in np.argsort(ms * ns)[:2]:
        p = probs[i]
        m_, n_ = p["m"], p["n"]
        Af = DeviceBuffer(m_ * n_ * 4)
        C.call("tomos_unpack", C.DECODE_FP32_BF16_EXACT, m_, n_, p["wh"].ptr, p["wr"].ptr, p["lda"], Af.ptr, m_, None)
        A64 = Af.to_numpy(np.float32, (m_, n_)).astype(np.float64)
        b64 = p["b"].to_numpy(np.float64, (m_,))
        g64 = gs[i].to_numpy(np.float64, (n_,))
        Q, R = np.linalg.qr(A64)
        x_ref = np.linalg.solve(R, Q.T @ b64)
        w_ref = np.linalg.solve(R, np.linalg.solve(R.T, g64))
        gb_ref = A64 @ w_ref
        x_got = p["x"].to_numpy(np.float64, (n_,))
        gA_ref = np.outer(b64 - A64 @ x_got, w_ref) - np.outer(gb_ref, x_got)
        rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
        grad_err.append({"m": m_, "n": n_, "w": rel(wds[i].to_numpy(np.float64, (n_,)), w_ref),
                         "grad_b": rel(gbds[i].to_numpy(np.float64, (m_,)), gb_ref),
                         "grad_A": rel(gads[i].to_numpy(np.float32, (m_, n_)).astype(np.float64), gA_ref),
                         "x_vs_fp64_lstsq": rel(x_got, x_ref)})
        del Af, A64, Q, R
    del wds, gbds, gads
    its_l = list(its)
    return {"problems": P, "elements": int((ms * ns).sum()), "storage": "FP32_BF16 planar",
            "sketch_rows": "%d n" % args.batched_sketch_factor,
            "precond_ms": round(pre_ms, 2), "solve_ms": round(solve_ms, 2), "forward_ms": round(pre_ms + solve_ms, 2),
            "solve_one_at_a_time_ms": round(seq_ms, 2), "batching_speedup": round(seq_ms / solve_ms, 2),
            "iterations_min": min(its_l), "iterations_max": max(its_l), "worst_relres": worst,
            "backward_ms": round(bwd_ms, 2),
            "backward_threads_ms": round(bwd_thr_ms, 2) if bwd_thr_ms is not None else None,
            "backward_one_at_a_time_ms": round(bwd_seq_ms, 2), "adjoint_tol": adj_tol,
            "adjoint_iterations_max": max(grp_its), "adjoint_iterations_max_one_problem": max(adj_its),
            "gradient_errors_vs_dense_fp64_qr": grad_err,
            "forward": "forward_ms = tomos_countsketch_apply_batched + tomos_sketch_qr_precond_batched_f64 + "
                       "tomos_lsqr_batched_solve to relres 1e-6 (one graph launch); tomos_amd.autograd.lstsq_batched "
                       "runs the same and, with its default normal_correction, one more batched adjoint solve "
                       "(about backward_ms without the gradients)",
            "backward": "tomos_lsq_adjoint_batched, workers = 0: every problem's adjoint CG (tol %g) in grouped "
                        "device launches (replayed graph segments), then grad_b and grad_A (m x n FP32) per problem; "
                        "backward_threads_ms: the same over %d host threads / HIP streams" % (adj_tol, aw)}


def qr_gemm_only(args, C, DeviceBuffer):
    """The sketch-QR preconditioner (FP64 QR + R^-1; CholeskyQR2 panels, then
    the column-by-column Householder panels) of a 4000 x 1000 sketch with
    column scales 1 .. 1e-6, three calls each, and the FP32_BF16
    GEMMs at G^3 (split with packed B, hi-only with FP32 B), two calls each."""
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    for e in (ev0, ev1):
        C.call("tomos_event_create", ctypes.byref(e))
    f = ctypes.c_float()
    out = {}
    k, n = 4000, 1000
    A0 = np.asfortranarray(np.random.default_rng(1).standard_normal((k, n)) * np.logspace(0, -6, n))
    src = DeviceBuffer.from_numpy(A0)
    A = DeviceBuffer(k * n * 8)
    for mode, key in ((0, "sketch_qr_precond_4000x1000_ms"), (1, "sketch_qr_precond_householder_panels_4000x1000_ms")):
        C.call("tomos_qr_panels_set", mode)
        ts = []
        for _ in range(3):
            C.call("tomos_memcpy_dtod", A.ptr, src.ptr, k * n * 8, None)
            C.call("tomos_event_record", ev0, None)
            C.call("tomos_sketch_qr_precond_f64", k, n, A.ptr, k, None)
            C.call("tomos_event_record", ev1, None)
            C.call("tomos_event_elapsed_ms", ctypes.byref(f), ev0, ev1)
            ts.append(round(f.value, 3))
        out[key] = ts
    C.call("tomos_qr_panels_set", 0)
    G = args.gemm
    if G > 0:
        GA, GB = DeviceBuffer(G * G * 4), DeviceBuffer(G * G * 4)
        C.call("tomos_fill_normal_f32", GA.ptr, G, G, G, 11, None, 0, 0, None)
        C.call("tomos_fill_normal_f32", GB.ptr, G, G, G, 12, None, 0, 0, None)
        P = [DeviceBuffer(G * G * 2) for _ in range(4)]
        Cg = DeviceBuffer(G * G * 4)
        C.call("tomos_pack", C.PACK_FP32_BF16, G, G, GA.ptr, G, P[0].ptr, P[1].ptr, G, None)
        C.call("tomos_pack", C.PACK_FP32_BF16, G, G, GB.ptr, G, P[2].ptr, P[3].ptr, G, None)
        for name, fn in (
            ("gemm_split_ms", lambda: C.call("tomos_gemm_split", C.FP32_BF16, G, G, G, 1.0, P[0].ptr, P[1].ptr, G,
                                             P[2].ptr, P[3].ptr, G, 0.0, Cg.ptr, G, None, 0, None)),
            ("gemm_hi_only_ms", lambda: C.call("tomos_gemm", C.FP32_BF16, C.TIER_DEFAULT, 0, 0, 1.0, G, G, G,
                                               P[0].ptr, P[1].ptr, G, GB.ptr, G, 0.0, Cg.ptr, G, None))):
            ts = []
            for _ in range(2):
                C.call("tomos_event_record", ev0, None)
                fn()
                C.call("tomos_event_record", ev1, None)
                C.call("tomos_event_elapsed_ms", ctypes.byref(f), ev0, ev1)
                ts.append(round(f.value, 3))
            out[name] = ts
    print(json.dumps({"qr_gemm_only": out}), flush=True)


def kernels_only(args, C, DeviceBuffer, m, n, wh):
    """GEMV-N (hi-only FP32_BF16, y = A x) and tomos_copy_kernel over the
    same 2*m*n bytes, each launched --gemv-reps times: the copy's counter
    readings calibrate FETCH_SIZE / WRITE_SIZE into bytes."""
    rng = np.random.default_rng(3)
    xv = DeviceBuffer.from_numpy(rng.standard_normal(n).astype(np.float32))
    y = DeviceBuffer(m * 4)
    nbytes = 2 * m * n
    dst = DeviceBuffer(nbytes)
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    for e in (ev0, ev1):
        C.call("tomos_event_create", ctypes.byref(e))
    f = ctypes.c_float()
    out = {}
    for name, fn in (
        ("gemv_n_bf16_hi", lambda: C.call("tomos_gemv", C.FP32_BF16, C.TIER_DEFAULT, 0, 1.0, m, n,
                                          wh.ptr, None, m, C.F32, xv.ptr, 0.0, C.F32, y.ptr, None, 0, None)),
        ("copy", lambda: C.call("tomos_copy_kernel", dst.ptr, wh.ptr, nbytes, None)),
        ("read", lambda: C.call("tomos_read_kernel", wh.ptr, nbytes, dst.ptr, None)),
    ):
        fn()
        C.call("tomos_event_record", ev0, None)
        for _ in range(args.gemv_reps):
            fn()
        C.call("tomos_event_record", ev1, None)
        C.call("tomos_event_elapsed_ms", ctypes.byref(f), ev0, ev1)
        out[name] = {"ms": round(f.value / args.gemv_reps, 4), "bytes": nbytes,
                     "launches": args.gemv_reps + 1}
    if args.tune:
        from tomos_amd import autotune
        best, _ = autotune.tune_fused(C.FP32_BF16, m, n, wh.ptr, m, reps=10)
        out["tuned_shape"] = best
    out["matvec_pair_bf16_hi"] = {"ms": round(time_pair(C, DeviceBuffer, args, m, n, wh, ev0, ev1), 4),
                                  "bytes": 2 * m * n + 8 * m + 16 * n, "launches": args.gemv_reps + 2}
    print(json.dumps({"kernels_only": out, "m": m, "n": n}), flush=True)


def c2_leg(C, DeviceBuffer, ev0, ev1, reps):
    """C2: packed GEMV and GEMV^T at 1,000,000 x 1000 with A_ij = z 2^k
    (z ~ N(0,1), k uniform in [-8, 8], device generator seed 4) stored as
    FP32_16; x, y ~ N(0,1) (numpy seeds 5, 6).  The FP32_16 pure-half tier
    (DEFAULT: half_only decode of wh, the reference's registered 16-bit
    tier) and the exact tier (wh+wr), each against the FP64 product of the
    exact values (EXACT tier with FP64 output)."""
    m, n = 1_000_000, 1000
    A = DeviceBuffer(m * n * 4)
    C.call("tomos_fill_normal_f32", A.ptr, m, n, m, 4, None, -8, 8, None)
    wh, wr = DeviceBuffer(m * n * 2), DeviceBuffer(m * n * 2)
    C.call("tomos_pack", C.PACK_FP32_16, m, n, A.ptr, m, wh.ptr, wr.ptr, m, None)
    del A
    x = np.random.default_rng(5).standard_normal(n).astype(np.float32)
    y = np.random.default_rng(6).standard_normal(m).astype(np.float32)
    out = {"m": m, "n": n, "storage": "FP32_16 planar", "data": "z 2^k, k in [-8, 8]"}
    for trans, vin, rows in ((0, x, m), (1, y, n)):
        vd = DeviceBuffer.from_numpy(vin)
        ref = DeviceBuffer(rows * 8)
        C.call("tomos_gemv", C.FP32_16, C.TIER_EXACT, trans, 1.0, m, n, wh.ptr, wr.ptr, m, C.F32, vd.ptr, 0.0, C.F64,
               ref.ptr, None, 0, None)
        r64 = ref.to_numpy(np.float64, (rows,))
        yd = DeviceBuffer(rows * 4)
        for name, tier, esz in (("pure_half", C.TIER_DEFAULT, 2), ("exact", C.TIER_EXACT, 4)):
            sb = C.lib().tomos_gemv_scratch_size(C.FP32_16, tier, trans, m, n)
            sc = DeviceBuffer(max(sb, 1))
            call = lambda: C.call("tomos_gemv", C.FP32_16, tier, trans, 1.0, m, n, wh.ptr, wr.ptr, m, C.F32, vd.ptr,
                                  0.0, C.F32, yd.ptr, sc.ptr, sb, None)
            call()
            t = time_call(call, reps, ev0, ev1)
            got = yd.to_numpy(np.float32, (rows,)).astype(np.float64)
            byts = esz * m * n + 4.0 * (m + n)
            key = "%s_%s" % (name, "T" if trans else "N")
            out[key] = {"ms": round(t, 4), "GBps": round(byts / (t * 1e-3) / 1e9, 1),
                        "GFlops": round(2.0 * m * n / (t * 1e-3) / 1e9, 1),
                        "l2_rel_err_vs_fp64": float(np.linalg.norm(got - r64) / np.linalg.norm(r64)),
                        "max_rel_err_vs_fp64": float(np.max(np.abs(got - r64)) / np.max(np.abs(r64)))}
    return out


def c1_problem(m=20000, n=200):
    """C1: A = G diag(d), G ~ N(0,1) (numpy seed 1), d log-spaced 1..1e-3,
    x_true ~ N(0,1) (seed 2), b = A x_true + 1e-6 N(0,1) (seed 3) in FP64."""
    G = np.random.default_rng(1).standard_normal((m, n))
    A = (G * np.logspace(0, -3, n)).astype(np.float32, order="F")
    xt = np.random.default_rng(2).standard_normal(n)
    b = A.astype(np.float64) @ xt + 1e-6 * np.random.default_rng(3).standard_normal(m)
    return A, b


def c1_leg(C, DeviceBuffer, O):
    """C1 (plain FP32 LSQR, 20000 x 200, kappa 1e3, tol 1e-6, <= 400
    iterations, one restart) on the host cores (median of 3) and on the GPU."""
    A, b = c1_problem()
    m, n = A.shape
    b32 = b.astype(np.float32)
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        xc, itc, _ = O.lsqr_f32(A, b32, 1e-6, 400, 1e-6)
        runs.append((time.perf_counter() - t0) * 1e3)
    rr_cpu = float(np.linalg.norm(A.astype(np.float64) @ xc - b) / np.linalg.norm(b))
    Ad, bd, xd = DeviceBuffer.from_numpy(A), DeviceBuffer.from_numpy(b32), DeviceBuffer(n * 4)
    po = C.LsqrOptions(C.F32, C.TIER_DEFAULT, C.TIER_DEFAULT, C.F32, C.F32, -1, 1e-6, 1e-6, 400, 1, 64, 1)
    pa = C.lib().tomos_lsqr_arena_size(ctypes.byref(po), m, n)
    pw = C.lib().tomos_lsqr_scratch_size(ctypes.byref(po), m, n)
    parena, pws = DeviceBuffer(pa), DeviceBuffer(pw)
    it, hl = ctypes.c_int(0), ctypes.c_int(0)
    gpu_runs = []
    for _ in range(4):
        C.call("tomos_device_synchronize")
        t0 = time.perf_counter()
        C.call("tomos_lsqr_solve", ctypes.byref(po), m, n, Ad.ptr, None, m, bd.ptr, xd.ptr, None, n, parena.ptr, pa,
               pws.ptr, pw, ctypes.byref(it), None, 0, ctypes.byref(hl), None)
        C.call("tomos_device_synchronize")
        gpu_runs.append((time.perf_counter() - t0) * 1e3)
    xg = xd.to_numpy(np.float32, (n,)).astype(np.float64)
    rr_gpu = float(np.linalg.norm(A.astype(np.float64) @ xg - b) / np.linalg.norm(b))
    return {"shape": [m, n], "cpu_ms_median_of_3": round(float(np.median(runs)), 3),
            "cpu_ms_runs": [round(r, 3) for r in runs], "cpu_iterations": itc, "cpu_relres": rr_cpu,
            "gpu_ms_median_of_3_after_warmup": round(float(np.median(gpu_runs[1:])), 3),
            "gpu_iterations": it.value, "gpu_relres": rr_gpu}


def cpu_leg(args, C, DeviceBuffer, A, b, m, n, s, b_host, bnorm, x_true):
    """The reference's CPU path on the headline workload, timed on the host
    cores: CountSketch of the FP32 matrix (s rows), FP64 Householder QR of
    the sketch and R^-1 (LAPACK dgeqrf / dtrtri in numpy's OpenBLAS 0.3.29),
    then FP32 LSQR right-preconditioned by R^-1 (cblas_sgemv and BLAS-1) with
    the GPU solve's restart cycle, to the same tolerance.  Also: C1 on the
    host (median of 3) and on the GPU, host sgemv rates at the workload
    shape, and a STREAM-style host copy rate."""
    from oracle import oracle as O
    import scipy.linalg.lapack as lapack
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    rc = O.blas_load(threads)
    detail = {"threads": threads, "openblas": "loaded" if rc == 0 else "NOT loaded (plain loops)"}
    detail["stream_copy_GBps"] = round(O.stream_copy_gbs(1 << 31, 3), 1)
    detail["c1"] = c1_leg(C, DeviceBuffer, O)
    log("cpu c1: %s" % detail["c1"])
    rows = m if args.cpu_rows <= 0 else min(args.cpu_rows, m)
    Ahost = np.empty((rows, n), np.float32, order="F")
    for j in range(n):  # the first `rows` rows of every column of the device matrix
        C.call("tomos_memcpy_dtoh", Ahost[:, j].ctypes.data, A.ptr + j * m * 4, rows * 4, None)
    gb = 4.0 * rows * n / 1e9
    tn, tt = O.sgemv_ms(0, Ahost, 2), O.sgemv_ms(1, Ahost, 2)
    detail["sgemv"] = {"rows": rows, "N_ms": round(tn, 2), "N_GBps": round(gb / (tn * 1e-3), 1),
                       "T_ms": round(tt, 2), "T_GBps": round(gb / (tt * 1e-3), 1)}
    bh = b_host[:rows].astype(np.float32)
    t0 = time.perf_counter()
    sk = O.countsketch_apply_f32(Ahost, s)
    t1 = time.perf_counter()
    R = np.linalg.qr(sk, mode="r")
    Rinv, info = lapack.dtrtri(R, lower=0)
    t2 = time.perf_counter()
    xc, itc, hist = O.lsqr_f32(Ahost, bh, args.tol, args.inner, args.tol, restarts=args.restarts,
                               Rinv=Rinv.astype(np.float32))
    t3 = time.perf_counter()
    # true relative residual in FP64 on the GPU (exact FP32 A)
    rr = None
    if rows == m:
        xd = DeviceBuffer.from_numpy(xc.astype(np.float64))
        r = DeviceBuffer(m * 8)
        C.call("tomos_memcpy_dtod", r.ptr, b.ptr, m * 8, None)
        C.call("tomos_gemv", C.F32, C.TIER_EXACT, 0, -1.0, m, n, A.ptr, None, m, C.F64, xd.ptr, 1.0, C.F64,
               r.ptr, None, 0, None)
        rr = float(np.linalg.norm(r.to_numpy(np.float64, (m,)))) / bnorm
    total = (t3 - t0) * 1e3
    detail["precond_lsqr"] = {"sketch_ms": round((t1 - t0) * 1e3, 1), "qr_trtri_ms": round((t2 - t1) * 1e3, 1),
                              "lsqr_ms": round((t3 - t2) * 1e3, 1), "iterations": itc,
                              "ms_per_iteration": round((t3 - t2) * 1e3 / max(itc, 1), 2), "relres": rr,
                              "err_vs_x_true": float(np.linalg.norm(xc - x_true) / np.linalg.norm(x_true)),
                              "lapack_info": int(info)}
    del Ahost
    converged = rr is not None and rr <= args.tol * (1 + 1e-6)
    cpu = {"value": round(total, 1),
           "unit": "ms (time to relres %g on the host cores, preconditioner build included)" % args.tol,
           "cores": threads, "kind": "port",
           "converged": converged,
           "sample": "the whole workload (%d x %d FP32, %s rows used): oracle CountSketch (s=%d, OpenMP), numpy "
                     "LAPACK dgeqrf + scipy dtrtri (OpenBLAS 0.3.29), oracle FP32 LSQR with the FP32 R^-1 right "
                     "preconditioner (cblas_sgemv + BLAS-1, OpenBLAS %s), %d iterations per restart, "
                     "one run" % (m, n, rows, s, detail["openblas"], args.inner)}
    return cpu, detail


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the LSQR path here is single-GPU "
              "(row-sharding across GPUs is a design note, not built); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from tomos_amd import _capi as C
    from tomos_amd._capi import DeviceBuffer

    C.call("tomos_set_device", 0)
    C.call("tomos_stream_loads_set", args.stream_loads)
    cu, amaj, amin, mem = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
    C.call("tomos_get_device_info", ctypes.byref(cu), ctypes.byref(amaj), ctypes.byref(amin),
           ctypes.byref(mem))
    m, n, s = args.m, args.n, args.s
    log("device: %d CUs, %.0f GB; workload %d x %d, s=%d" % (cu.value, mem.value / 1e9, m, n, s))
    if args.qr_gemm_only:
        return qr_gemm_only(args, C, DeviceBuffer)

    # ---- synthetic data, resident in HBM ------------------------------------
    d = np.logspace(0, -args.kappa_exp, n).astype(np.float32)
    dd = DeviceBuffer.from_numpy(d)
    A = DeviceBuffer(m * n * 4)
    C.call("tomos_fill_normal_f32", A.ptr, m, n, m, 1, dd.ptr, 0, 0, None)
    wh, wr = DeviceBuffer(m * n * 2), DeviceBuffer(m * n * 2)
    C.call("tomos_pack", C.PACK_FP32_BF16, m, n, A.ptr, m, wh.ptr, wr.ptr, m, None)
    rng = np.random.default_rng(2)
    x_true = rng.standard_normal(n)
    xt_d = DeviceBuffer.from_numpy(x_true)
    b = DeviceBuffer(m * 8)
    C.call("tomos_gemv", C.F32, C.TIER_EXACT, 0, 1.0, m, n, A.ptr, None, m, C.F64, xt_d.ptr,
           0.0, C.F64, b.ptr, None, 0, None)
    C.call("tomos_device_synchronize")
    b_host = b.to_numpy(np.float64, (m,))
    bnorm = float(np.linalg.norm(b_host))
    log("data ready")
    if args.kernels_only:
        return kernels_only(args, C, DeviceBuffer, m, n, wh)
    tuned = {}
    if args.tune:
        # the project's autotuner picks the single-pass kernel's launch shape
        # for both storage types at this shape (set-up, outside the timed region)
        from tomos_amd import autotune
        for name, at, ptr in (("fp32_bf16", C.FP32_BF16, wh.ptr), ("f32", C.F32, A.ptr)):
            best, _ = autotune.tune_fused(at, m, n, ptr, m, reps=10)
            tuned[name] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in best.items()}
        log("tuned single-pass shapes: %s" % tuned)

    # ---- buffers for the solve -----------------------------------------------
    sketch = DeviceBuffer(s * n * 8)
    cs_bytes = C.lib().tomos_countsketch_scratch_size(m, n, s)
    cs_scratch = DeviceBuffer(cs_bytes)
    x = DeviceBuffer(n * 8)
    # confirm = 1: the solve stops only when a restart's exact residual is below tol
    opt = C.LsqrOptions(C.FP32_BF16, C.TIER_DEFAULT, C.TIER_EXACT, C.F32, C.F64, C.F64,
                        args.tol, args.tol, args.inner, args.restarts, args.check_every, args.fused, 1)
    ar_bytes = C.lib().tomos_lsqr_arena_size(ctypes.byref(opt), m, n)
    sc_bytes = C.lib().tomos_lsqr_scratch_size(ctypes.byref(opt), m, n)
    arena, scr = DeviceBuffer(ar_bytes), DeviceBuffer(sc_bytes)
    hist_cap = args.inner * args.restarts + args.restarts + 1
    hist = np.zeros(hist_cap)
    ev0, ev1, ev2 = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    for e in (ev0, ev1, ev2):
        C.call("tomos_event_create", ctypes.byref(e))

    def solve_mixed():
        it = ctypes.c_int(0)
        hl = ctypes.c_int(0)
        C.call("tomos_event_record", ev0, None)
        C.call("tomos_countsketch_apply", C.FP32_BF16, C.TIER_EXACT, m, n, wh.ptr, wr.ptr, m, s,
               sketch.ptr, s, cs_scratch.ptr, cs_bytes, None)
        C.call("tomos_sketch_qr_precond_f64", s, n, sketch.ptr, s, None)
        # the leading n x n block of the sketch now holds R^-1 (ld = s)
        C.call("tomos_event_record", ev1, None)
        C.call("tomos_lsqr_solve", ctypes.byref(opt), m, n, wh.ptr, wr.ptr, m, b.ptr, x.ptr,
               sketch.ptr, s, arena.ptr, ar_bytes, scr.ptr, sc_bytes, ctypes.byref(it),
               hist.ctypes.data, hist_cap,
               ctypes.byref(hl), None)
        C.call("tomos_event_record", ev2, None)
        f1, f2 = ctypes.c_float(), ctypes.c_float()
        C.call("tomos_event_elapsed_ms", ctypes.byref(f1), ev0, ev1)
        C.call("tomos_event_elapsed_ms", ctypes.byref(f2), ev1, ev2)
        return it.value, hl.value, f1.value, f2.value

    for w in range(args.warmup):
        it, hl, pms, sms = solve_mixed()
        log("warmup %d: iters=%d precond=%.1f ms solve=%.1f ms" % (w, it, pms, sms))
    C.call("tomos_device_synchronize")
    # HIP events around every single-pass kernel launch of the timed solves
    # (the roofline's kernel, measured where it runs)
    C.call("tomos_kernel_timer_begin", 4096)
    t0 = time.perf_counter()
    per = []
    for k in range(args.steps):
        per.append(solve_mixed())
        log("step %d: iters=%d precond=%.1f ms solve=%.1f ms" % (k, per[-1][0], per[-1][2], per[-1][3]))
    C.call("tomos_device_synchronize")
    t1 = time.perf_counter()
    k_n = ctypes.c_int()
    k_dur = np.zeros(4096, np.float32)
    C.call("tomos_kernel_timer_end", k_dur.ctypes.data, k_dur.size, ctypes.byref(k_n))
    k_dur = k_dur[:min(k_n.value, k_dur.size)].astype(np.float64)
    # launches enqueued past convergence exit at the stop flag (a few us);
    # the roofline averages the launches that streamed the matrix
    k_full = k_dur[k_dur > 0.25 * np.median(k_dur)] if k_dur.size else k_dur
    fused_in_solve = {"launches": int(k_n.value), "streamed_launches": int(k_full.size),
                      "avg_ms": round(float(k_full.mean()), 4) if k_full.size else None,
                      "min_ms": round(float(k_full.min()), 4) if k_full.size else None,
                      "max_ms": round(float(k_full.max()), 4) if k_full.size else None}
    log("single-pass kernel in the timed solves: %s" % fused_in_solve)
    ms_per_step = (t1 - t0) * 1e3 / max(args.steps, 1)
    iters, hist_len, precond_ms, solve_ms = per[-1]

    # ---- accuracy of the last solve (FP64 residual with the exact A) --------
    xs = x.to_numpy(np.float64, (n,))
    r = DeviceBuffer(m * 8)
    C.call("tomos_memcpy_dtod", r.ptr, b.ptr, m * 8, None)
    C.call("tomos_gemv", C.FP32_BF16, C.TIER_EXACT, 0, -1.0, m, n, wh.ptr, wr.ptr, m, C.F64,
           x.ptr, 1.0, C.F64, r.ptr, None, 0, None)
    relres = float(np.linalg.norm(r.to_numpy(np.float64, (m,)))) / bnorm
    err = float(np.linalg.norm(xs - x_true) / np.linalg.norm(x_true))
    log("mixed+precond: relres=%.3e err_vs_fp64=%.3e iters=%d" % (relres, err, iters))

    # ---- roofline: hi-only packed GEMV at the full shape ---------------------
    y = DeviceBuffer(m * 4)
    xv = DeviceBuffer.from_numpy(rng.standard_normal(n).astype(np.float32))
    uv = DeviceBuffer.from_numpy(rng.standard_normal(m).astype(np.float32))
    zv = DeviceBuffer(n * 4)
    gsc = max(C.lib().tomos_gemv_scratch_size(t, tier, tr, m, n)
              for t, tier in ((C.FP32_BF16, 0), (C.FP32_BF16, 2), (C.F32, 0)) for tr in (0, 1))
    gs = DeviceBuffer(max(gsc, 1))

    def time_gemv(trans, a_type, tier, a_ptr, wr_ptr):
        xin, yout = (uv, zv) if trans else (xv, y)
        for _ in range(WARM_LAUNCHES):
            C.call("tomos_gemv", a_type, tier, trans, 1.0, m, n, a_ptr, wr_ptr, m, C.F32, xin.ptr,
                   0.0, C.F32, yout.ptr, gs.ptr, gsc, None)
        C.call("tomos_event_record", ev0, None)
        for _ in range(args.gemv_reps):
            C.call("tomos_gemv", a_type, tier, trans, 1.0, m, n, a_ptr, wr_ptr, m, C.F32, xin.ptr,
                   0.0, C.F32, yout.ptr, gs.ptr, gsc, None)
        C.call("tomos_event_record", ev1, None)
        f = ctypes.c_float()
        C.call("tomos_event_elapsed_ms", ctypes.byref(f), ev0, ev1)
        return f.value / args.gemv_reps

    gemv_bytes_hi = 2.0 * m * n + 4.0 * n + 4.0 * m  # beta = 0
    gemv_bytes_f32 = 4.0 * m * n + 4.0 * n + 4.0 * m
    t_n = time_gemv(0, C.FP32_BF16, 0, wh.ptr, None)
    t_t = time_gemv(1, C.FP32_BF16, 0, wh.ptr, None)
    t_nx = time_gemv(0, C.FP32_BF16, 2, wh.ptr, wr.ptr)
    t_f = time_gemv(0, C.F32, 0, A.ptr, None)
    t_ft = time_gemv(1, C.F32, 0, A.ptr, None)
    gbs = lambda byts, ms: byts / (ms * 1e-3) / 1e9
    # achievable read rate: a plain read-only stream over the same hi plane
    rflag = DeviceBuffer(4)
    C.call("tomos_memset", rflag.ptr, 0, 4, None)
    t_rd = time_call(lambda: C.call("tomos_read_kernel", wh.ptr, 2 * m * n, rflag.ptr, None), args.gemv_reps,
                     ev0, ev1)

    # the LSQR iteration's dominant kernel: A p and A^T u from one read of the
    # hi plane (tomos_matvec_pair = lsqr_fused_kernel + its partial sums)
    t_pair = time_pair(C, DeviceBuffer, args, m, n, wh, ev0, ev1)
    pair_bytes = 2.0 * m * n + 8.0 * m + 16.0 * n  # hi plane, u read + written, p and t in FP64
    # roofline: the kernel's average device time over its launches inside
    # the timed solves (tomos_kernel_timer); the standalone entry-point
    # timing (back-to-back calls, with its set-up and finishing kernels) is
    # reported beside it
    t_kernel = float(k_full.mean()) if k_full.size else t_pair
    achieved = gbs(pair_bytes, t_kernel)
    gemv_detail = {
        "pair_bf16_hi_Ap_ATu": {"ms": round(t_pair, 4), "GBps": round(gbs(pair_bytes, t_pair), 1),
                                "note": "tomos_matvec_pair entry point, back-to-back calls"},
        "single_pass_kernel_in_timed_solves": dict(fused_in_solve, GBps=round(achieved, 1)),
        "bf16_hi_N": {"ms": round(t_n, 4), "GBps": round(gbs(gemv_bytes_hi, t_n), 1)},
        "bf16_hi_T": {"ms": round(t_t, 4), "GBps": round(gbs(gemv_bytes_hi, t_t), 1)},
        "bf16_exact_N": {"ms": round(t_nx, 4), "GBps": round(gbs(gemv_bytes_f32, t_nx), 1)},
        "fp32_N": {"ms": round(t_f, 4), "GBps": round(gbs(gemv_bytes_f32, t_f), 1)},
        "fp32_T": {"ms": round(t_ft, 4), "GBps": round(gbs(gemv_bytes_f32, t_ft), 1)},
        "read_stream_same_bytes": {"ms": round(t_rd, 4), "GBps": round(gbs(2.0 * m * n, t_rd), 1)},
    }
    log("gemv: %s" % gemv_detail)
    c2 = None
    if args.c2:
        c2 = c2_leg(C, DeviceBuffer, ev0, ev1, args.gemv_reps)
        log("c2 fp32_16 gemv: %s" % c2)

    # dense Gaussian sketch fill (create_gaussian_sketch): 2n x 131072 FP32, 1 GiB written
    gk, gm = 2 * n, 131072
    gbuf = DeviceBuffer(gk * gm * 4)
    t_gf = time_call(lambda: C.call("tomos_gaussian_fill", C.F32, gk, gm, gbuf.ptr, gk, 5, None), 5, ev0, ev1)
    gaussian = {"rows": gk, "cols": gm, "type": "f32", "ms": round(t_gf, 4),
                "GBps_written": round(gbs(4.0 * gk * gm, t_gf), 1),
                "Gsamples_per_s": round(gk * gm / (t_gf * 1e-3) / 1e9, 1)}
    del gbuf
    log("gaussian sketch fill: %s" % gaussian)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "lsqr_fused_bf16_2e6x1000_traffic.md")
    if os.path.exists(tf) and (m, n) == (2_000_000, 1000):
        for line in open(tf):
            if line.startswith("traffic_bytes_per_launch:"):
                traffic = int(line.split(":")[1])

    # ---- plain FP32 LSQR on the same GPU --------------------------------------
    # (after the kernel timings: its seconds of back-to-back streaming would
    # otherwise precede them)
    plain = None
    if not args.no_plain:
        bf = DeviceBuffer.from_numpy(b_host.astype(np.float32))
        xf = DeviceBuffer(n * 4)
        po = C.LsqrOptions(C.F32, C.TIER_DEFAULT, C.TIER_DEFAULT, C.F32, C.F32, -1, args.tol,
                           args.tol, args.plain_cap, 1, 64, args.fused)
        pa = C.lib().tomos_lsqr_arena_size(ctypes.byref(po), m, n)
        pw = C.lib().tomos_lsqr_scratch_size(ctypes.byref(po), m, n)
        parena, pws = DeviceBuffer(pa), DeviceBuffer(pw)
        it = ctypes.c_int(0)
        hl = ctypes.c_int(0)
        ph = np.zeros(args.plain_cap + 2)
        C.call("tomos_device_synchronize")
        t0 = time.perf_counter()
        C.call("tomos_lsqr_solve", ctypes.byref(po), m, n, A.ptr, None, m, bf.ptr, xf.ptr, None,
               n, parena.ptr, pa, pws.ptr, pw, ctypes.byref(it), ph.ctypes.data, ph.size, ctypes.byref(hl), None)
        C.call("tomos_device_synchronize")
        pt = (time.perf_counter() - t0) * 1e3
        xp = xf.to_numpy(np.float32, (n,)).astype(np.float64)
        xpd = DeviceBuffer.from_numpy(xp)
        C.call("tomos_memcpy_dtod", r.ptr, b.ptr, m * 8, None)
        C.call("tomos_gemv", C.F32, C.TIER_EXACT, 0, -1.0, m, n, A.ptr, None, m, C.F64, xpd.ptr,
               1.0, C.F64, r.ptr, None, 0, None)
        prr = float(np.linalg.norm(r.to_numpy(np.float64, (m,)))) / bnorm
        plain = {"ms": round(pt, 2), "iterations": it.value, "relres": prr,
                 "converged": bool(prr < args.tol), "cap": args.plain_cap,
                 "ms_per_iteration": round(pt / max(it.value, 1), 4)}
        log("plain fp32: %s" % plain)
        del bf, xf, parena, pws, xpd

    # ---- split-precision MFMA GEMM (C3 square shape) -------------------------
    gemm = None
    if args.gemm > 0:
        del y, uv, gs
        G = args.gemm
        gemm = {"M": G, "N": G, "K": G}
        GA, GB = DeviceBuffer(G * G * 4), DeviceBuffer(G * G * 4)
        C.call("tomos_fill_normal_f32", GA.ptr, G, G, G, 11, None, 0, 0, None)
        C.call("tomos_fill_normal_f32", GB.ptr, G, G, G, 12, None, 0, 0, None)
        planes = [DeviceBuffer(G * G * 2) for _ in range(4)]
        Cg = DeviceBuffer(G * G * 4)
        gsz = C.lib().tomos_gemm_split_scratch_size(G, G, G)
        gscr = DeviceBuffer(gsz)
        cols = [0, 1, G // 2, G - 1]
        for name, at, var in (("fp32_bf16", C.FP32_BF16, C.PACK_FP32_BF16),
                              ("fp32_16", C.FP32_16, C.PACK_FP32_16)):
            C.call("tomos_pack", var, G, G, GA.ptr, G, planes[0].ptr, planes[1].ptr, G, None)
            C.call("tomos_pack", var, G, G, GB.ptr, G, planes[2].ptr, planes[3].ptr, G, None)
            args_g = (at, G, G, G, 1.0, planes[0].ptr, planes[1].ptr, G, planes[2].ptr,
                      planes[3].ptr, G, 0.0, Cg.ptr, G, gscr.ptr, gsz, None)
            C.call("tomos_gemm_split", *args_g)
            C.call("tomos_event_record", ev0, None)
            reps = 3
            for _ in range(reps):
                C.call("tomos_gemm_split", *args_g)
            C.call("tomos_event_record", ev1, None)
            f = ctypes.c_float()
            C.call("tomos_event_elapsed_ms", ctypes.byref(f), ev0, ev1)
            ms = f.value / reps
            useful = 2.0 * G * G * G / (ms * 1e-3) / 1e12
            # FP64-accumulated reference columns: C[:, j] = A_exact @ B_exact[:, j]
            dec = C.DECODE_FP32_BF16_EXACT if at == C.FP32_BF16 else C.DECODE_FP32_16_EXACT
            errs = []
            for j in cols:
                bcol = DeviceBuffer(G * 4)
                C.call("tomos_unpack", dec, G, 1, planes[2].ptr + j * G * 2, planes[3].ptr + j * G * 2,
                       G, bcol.ptr, G, None)
                refc = DeviceBuffer(G * 8)
                C.call("tomos_gemv", at, C.TIER_EXACT, 0, 1.0, G, G, planes[0].ptr, planes[1].ptr, G,
                       C.F32, bcol.ptr, 0.0, C.F64, refc.ptr, None, 0, None)
                rc_ = refc.to_numpy(np.float64, (G,))
                gc_ = np.empty(G, np.float32)
                C.call("tomos_memcpy_dtoh", gc_.ctypes.data, Cg.ptr + j * G * 4, G * 4, None)
                errs.append(np.linalg.norm(gc_ - rc_) / np.linalg.norm(rc_))
            gemm[name] = {"ms": round(ms, 3), "useful_TFLOPs": round(useful, 1),
                          "mfma_TFLOPs": round(3 * useful, 1),
                          "mfma_frac_of_2500": round(3 * useful / 2500.0, 4),
                          "l2_rel_err_sampled_cols": float(max(errs))}
            # mm_apply on packed A with FP32 B (tomos_gemm): the hi-only tier
            # (16-bit matrix cores on the hi word) and the materialize tier
            # (3-term split); whole calls, operand staging included
            for tname, tier, terms in (("mm_hi_only", C.TIER_DEFAULT, 1), ("mm_materialize", C.TIER_MATERIALIZE, 3)):
                args_m = (at, tier, 0, 0, 1.0, G, G, G, planes[0].ptr, planes[1].ptr, G, GB.ptr, G, 0.0,
                          Cg.ptr, G, None)
                C.call("tomos_gemm", *args_m)
                C.call("tomos_event_record", ev0, None)
                for _ in range(reps):
                    C.call("tomos_gemm", *args_m)
                C.call("tomos_event_record", ev1, None)
                C.call("tomos_event_elapsed_ms", ctypes.byref(f), ev0, ev1)
                mms = f.value / reps
                useful_m = 2.0 * G * G * G / (mms * 1e-3) / 1e12
                errs_m = []
                for j in cols:
                    refc = DeviceBuffer(G * 8)
                    C.call("tomos_gemv", at, C.TIER_EXACT, 0, 1.0, G, G, planes[0].ptr, planes[1].ptr, G,
                           C.F32, GB.ptr + j * G * 4, 0.0, C.F64, refc.ptr, None, 0, None)
                    rc_ = refc.to_numpy(np.float64, (G,))
                    gc_ = np.empty(G, np.float32)
                    C.call("tomos_memcpy_dtoh", gc_.ctypes.data, Cg.ptr + j * G * 4, G * 4, None)
                    errs_m.append(np.linalg.norm(gc_ - rc_) / np.linalg.norm(rc_))
                gemm[name + "_" + tname] = {
                    "ms": round(mms, 3), "useful_TFLOPs": round(useful_m, 1),
                    "mfma_TFLOPs": round(terms * useful_m, 1),
                    "mfma_frac_of_2500": round(terms * useful_m / 2500.0, 4),
                    "l2_rel_err_vs_exact_A_sampled_cols": float(max(errs_m))}
        log("split gemm: %s" % gemm)
        del GA, GB, planes, Cg, gscr

    # ---- batched differentiable least squares (64 uneven problems) ------------
    batched = None
    if args.batched > 0:
        batched = batched_leg(args, C, DeviceBuffer, ev0, ev1)
        log("batched: %s" % batched)

    # ---- sparse (CSR) LSQR at the tall shape ----------------------------------
    sparse = None
    if args.sparse_row > 0:
        sparse = sparse_leg(args, C, DeviceBuffer, ev0, ev1)
        log("sparse: %s" % sparse)

    # ---- CPU baselines on the host cores (OpenBLAS 0.3.29 from numpy) ---------
    cpu, cpu_detail = None, None
    if not args.no_cpu:
        cpu, cpu_detail = cpu_leg(args, C, DeviceBuffer, A, b, m, n, s, b_host, bnorm, x_true)
        log("cpu: %s %s" % (cpu, cpu_detail))

    speedup = None
    if plain:
        speedup = round(plain["ms"] / ms_per_step, 2)
    out = {
        "metric": "LSQR time-to-1e-6 residual, mixed+precond vs plain FP32 on MI355X; GEMV %HBM",
        "value": round(ms_per_step, 3),
        "unit": "ms (time to relres 1e-6, preconditioner build included)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 (hi-only GEMV) / fp32 vectors / fp64 scalars+QR",
        "data": "synthetic: A = G diag(d), G ~ N(0,1) (device hash Box-Muller), d log-spaced 1..1e-%g; b = A x_true in FP64" % args.kappa_exp,
        "config": {"workload": "C4 sketch-preconditioned mixed-precision LSQR", "m": m, "n": n,
                   "storage": "FP32_BF16 planar", "sketch_rows": s, "tol": args.tol,
                   "inner_iters": args.inner, "restarts_max": args.restarts},
        "iterations": iters,
        "precond_ms": round(precond_ms, 3),
        "solve_ms": round(solve_ms, 3),
        "relres": relres,
        "err_vs_fp64": err,
        "plain_fp32": plain,
        "speedup_vs_plain_fp32": speedup,
        "speedup_is_lower_bound": bool(plain and not plain["converged"]),
        "gemv": gemv_detail,
        "c2_fp32_16_gemv": c2,
        "tuned_shapes": tuned,
        "gemm_split": gemm,
        "gaussian_sketch": gaussian,
        "batched_lstsq": batched,
        "sparse_lsqr": sparse,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "lsqr_fused_kernel<FP32_BF16 hi-only> (u = A p + c u, t = A^T u from one "
                               "read of A), %dx%d, %.3f GB algorithmic per launch; average of its %d "
                               "launches that streamed A inside the timed solves (HIP events on the "
                               "solve stream)" % (m, n, pair_bytes / 1e9, k_full.size)},
        "cpu_baseline": cpu,
        "cpu_detail": cpu_detail,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
