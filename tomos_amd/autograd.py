"""Differentiable least squares on the GPU: x = argmin |A x - b|.

``lstsq(A, b)`` is a ``torch.autograd.Function`` over the C ABI:

forward
    packs A (FP32, (m, n)) as FP32_BF16 planes, builds the sketch-QR
    preconditioner (hashed CountSketch with s = 4n rows, FP64 Householder
    QR, R^-1) and runs the restarted mixed-precision LSQR: hi-only GEMVs
    inside a restart, exact (wh+wr) FP64-accumulated residual at each
    restart.  Restarting on b - A x alone converges to A_hi^T r = 0, not
    A^T r = 0: for an inconsistent system (r != 0) that leaves a bias of
    the order of the hi-plane rounding times |r|.  So the forward ends
    with one normal-equation correction: s = A^T (b - A x) exactly (FP64),
    delta = (A^T A)^-1 s by the adjoint's preconditioned CG (whose restarts
    also use exact products), x += delta.  Returns x in FP64.
backward
    implicit differentiation of the normal equations A^T A x = A^T b:
    w = (A^T A)^-1 g by R-preconditioned CG with the same packed kernels
    (``tomos_lsq_adjoint_solve``), then ``tomos_lsq_grad``: grad_b = A w and
    grad_A = r w^T - (A w) x^T (r = b - A x) from one fused rank-2 kernel.

``lstsq_batched(As, bs)`` does the same for many uneven problems at once:
one ``tomos_lsqr_batched_solve`` (grouped kernels, a hipGraph per restart)
for the forward, and ``tomos_lsq_adjoint_batched`` (problems over worker
threads with one HIP stream each) for the normal-equation correction and
the backward.

The truncating FP32_BF16 split is lossless (hi and lo halves of the FP32
bits), so the solve and the gradients refer to A itself.  Everything runs
on torch's current stream (the batched adjoint waits for it and returns
when its own streams are done); there is no CPU path.
"""
import ctypes
from dataclasses import dataclass

import torch

from . import _capi as C


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _bytes(nbytes, device):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


@dataclass
class SolveOptions:
    tol: float = 1e-6            # relative residual target (forward and adjoint)
    inner: int = 30              # Krylov iterations per restart
    restarts: int = 40
    sketch_factor: int = 8       # sketch rows s = sketch_factor * n (8: fewer adjoint iterations than 4)
    adjoint_tol: float = 1e-10   # |g - A^T A w| / |g| for the adjoint solve
    check_every: int = 8
    normal_correction: bool = True  # forward: x += (A^T A)^-1 A^T (b - A x), exact products
    adjoint_workers: int = 0     # lstsq_batched's adjoint: 0 = all problems in grouped device launches,
                                 # 1..16 = host threads (one HIP stream each) running one-problem solves
    confirm: bool = True         # forward: stop only when a restart's true residual is below tol


class PackedProblem:
    """A packed as FP32_BF16 planes plus its sketch-QR preconditioner."""

    def __init__(self, A, opts, precondition=True):
        if not A.is_cuda or A.dtype != torch.float32 or A.dim() != 2:
            raise ValueError("A must be a 2-D float32 CUDA tensor")
        m, n = A.shape
        if m < n or n < 1:
            raise ValueError("need m >= n >= 1")
        self.m, self.n, self.opts = m, n, opts
        dev = A.device
        # 128-byte columns: the single-pass kernels read 64-row (128-byte) runs
        # of every column, which then start on cache lines (16 bytes is the
        # minimum; a 64-byte offset made the batched passes fetch ~1.4x)
        self.ld = (m + 63) // 64 * 64
        acm = A.t().contiguous()    # (n, m) row-major == column-major (m, n)
        self.wh = torch.empty(self.ld * n, dtype=torch.int16, device=dev)
        self.wr = torch.empty(self.ld * n, dtype=torch.int16, device=dev)
        s = _stream()
        C.call("tomos_pack", C.PACK_FP32_BF16, m, n, _p(acm), m, _p(self.wh), _p(self.wr), self.ld, s)
        self.s_rows = max(opts.sketch_factor * n, n)
        self.sketch = torch.empty((n, self.s_rows), dtype=torch.float64, device=dev)
        if precondition:
            self.precondition()

    def precondition(self):
        """sketch = S A (hashed CountSketch), then R^-1 of its QR over the
        leading n x n block."""
        m, n, s = self.m, self.n, _stream()
        cs = C.lib().tomos_countsketch_scratch_size(m, n, self.s_rows)
        scr = _bytes(cs, self.wh.device)
        C.call("tomos_countsketch_apply", C.FP32_BF16, C.TIER_EXACT, m, n, _p(self.wh), _p(self.wr),
               self.ld, self.s_rows, _p(self.sketch), self.s_rows, _p(scr), cs, s)
        # R^-1 overwrites the leading n x n block (ld = s_rows)
        C.call("tomos_sketch_qr_precond_f64", self.s_rows, n, _p(self.sketch), self.s_rows, s)

    def options(self, tol, iters, restarts, confirm=False):
        return C.LsqrOptions(C.FP32_BF16, C.TIER_DEFAULT, C.TIER_EXACT, C.F32, C.F64, C.F64,
                             tol, tol, iters, restarts, self.opts.check_every, 1, int(confirm))

    def solve(self, b64):
        o = self.options(self.opts.tol, self.opts.inner, self.opts.restarts, self.opts.confirm)
        m, n, dev = self.m, self.n, b64.device
        x = torch.empty(n, dtype=torch.float64, device=dev)
        ab = C.lib().tomos_lsqr_arena_size(ctypes.byref(o), m, n)
        sb = C.lib().tomos_lsqr_scratch_size(ctypes.byref(o), m, n)
        arena, scr = _bytes(ab, dev), _bytes(sb, dev)
        it = ctypes.c_int(0)
        C.call("tomos_lsqr_solve", ctypes.byref(o), m, n, _p(self.wh), _p(self.wr), self.ld, _p(b64), _p(x),
               _p(self.sketch), self.s_rows, _p(arena), ab, _p(scr), sb, ctypes.byref(it), None, 0, None,
               _stream())
        self.last_iterations = it.value
        if self.opts.normal_correction:
            x = x + self.adjoint(self.normal_residual(b64, x))
        return x

    def normal_residual(self, b64, x):
        """s = A^T (b - A x) with the exact (wh+wr) tier, FP64 throughout."""
        m, n, dev = self.m, self.n, x.device
        r = b64.clone()
        s_ = torch.empty(n, dtype=torch.float64, device=dev)
        st = _stream()
        C.call("tomos_gemv", C.FP32_BF16, C.TIER_EXACT, C.NO_TRANSPOSE, -1.0, m, n, _p(self.wh), _p(self.wr),
               self.ld, C.F64, _p(x), 1.0, C.F64, _p(r), None, 0, st)
        C.call("tomos_gemv", C.FP32_BF16, C.TIER_EXACT, C.TRANSPOSE, 1.0, m, n, _p(self.wh), _p(self.wr),
               self.ld, C.F64, _p(r), 0.0, C.F64, _p(s_), None, 0, st)
        return s_

    def adjoint(self, g64):
        """w = (A^T A)^-1 g."""
        o = self.options(self.opts.adjoint_tol, self.opts.inner, self.opts.restarts)
        m, n, dev = self.m, self.n, g64.device
        w = torch.empty(n, dtype=torch.float64, device=dev)
        sb = C.lib().tomos_lsq_adjoint_scratch_size(ctypes.byref(o), m, n)
        scr = _bytes(sb, dev)
        it = ctypes.c_int(0)
        C.call("tomos_lsq_adjoint_solve", ctypes.byref(o), m, n, _p(self.wh), _p(self.wr), self.ld, _p(g64),
               _p(self.sketch), self.s_rows, _p(w), _p(scr), sb, ctypes.byref(it), None, 0, None, _stream())
        self.last_adjoint_iterations = it.value
        return w

    def grads(self, b64, x, w, want_a):
        m, n, dev = self.m, self.n, x.device
        gb = torch.empty(m, dtype=torch.float64, device=dev)
        ga = torch.empty((n, m), dtype=torch.float32, device=dev) if want_a else None
        sb = C.lib().tomos_lsq_grad_scratch_size(C.FP32_BF16, C.TIER_EXACT, m, n)
        scr = _bytes(sb, dev)
        C.call("tomos_lsq_grad", C.FP32_BF16, C.TIER_EXACT, m, n, _p(self.wh), _p(self.wr), self.ld,
               _p(b64), _p(x), _p(w), _p(gb), _p(ga), m, _p(scr), sb, _stream())
        return gb, (ga.t() if want_a else None)


class _Lstsq(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, b, opts):
        prob = PackedProblem(A.detach(), opts)
        b64 = b.detach().to(torch.float64).contiguous()
        x = prob.solve(b64)
        ctx.prob, ctx.b64, ctx.b_dtype = prob, b64, b.dtype
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        prob = ctx.prob
        w = prob.adjoint(g.detach().to(torch.float64).contiguous())
        gb, ga = prob.grads(ctx.b64, x, w, ctx.needs_input_grad[0])
        grad_b = gb.to(ctx.b_dtype) if ctx.needs_input_grad[1] else None
        return ga, grad_b, None


def lstsq(A, b, opts=None):
    """x = argmin |A x - b| (FP64 result), differentiable in A and b."""
    return _Lstsq.apply(A, b, opts or SolveOptions())


def _problem_table(probs, b64s, xs):
    return (C.LsqProblem * len(probs))(*[
        C.LsqProblem(p.m, p.n, p.ld, p.wh.data_ptr(), p.wr.data_ptr(), b.data_ptr(), x.data_ptr(),
                     p.sketch.data_ptr(), p.s_rows) for p, b, x in zip(probs, b64s, xs)])


def _precondition_batched(probs):
    """Every problem's sketch and R^-1 in two batched calls: the CountSketches
    with one host synchronisation (tomos_countsketch_apply_batched), then the
    sketch QRs and inverses in grouped launches
    (tomos_sketch_qr_precond_batched_f64).  Equal to each problem's
    precondition() bit for bit."""
    P, dev, s = len(probs), probs[0].wh.device, _stream()
    cp = (C.CsProblem * P)(*[C.CsProblem(p.m, p.n, p.ld, p.s_rows, p.s_rows, _p(p.wh), _p(p.wr), _p(p.sketch))
                             for p in probs])
    cb = C.lib().tomos_countsketch_apply_batched_scratch_size(P, cp)
    if cb == 0:
        raise RuntimeError(C.lib().tomos_last_error().decode())
    scr = _bytes(cb, dev)
    C.call("tomos_countsketch_apply_batched", C.FP32_BF16, C.TIER_EXACT, P, cp, _p(scr), cb, s)
    del scr
    qp = (C.QrProblem * P)(*[C.QrProblem(p.s_rows, p.n, p.s_rows, _p(p.sketch)) for p in probs])
    qb = C.lib().tomos_sketch_qr_precond_batched_scratch_size(P, qp)
    if qb == 0:
        raise RuntimeError(C.lib().tomos_last_error().decode())
    qs = _bytes(qb, dev)
    C.call("tomos_sketch_qr_precond_batched_f64", P, qp, _p(qs), qb, s)


def _ptrs(ts):
    return (ctypes.c_void_p * len(ts))(*[t.data_ptr() if t is not None else None for t in ts])


def _adjoint_batched(probs, b64s, xs, gs, tol, want_a):
    """w_i = (A_i^T A_i)^-1 g_i, grad_b_i = A_i w_i and (want_a) grad_A_i for
    every problem, through tomos_lsq_adjoint_batched."""
    opts = probs[0].opts
    o = probs[0].options(tol, opts.inner, opts.restarts)
    dev = gs[0].device
    P = len(probs)
    table = _problem_table(probs, b64s, xs)
    ws = [torch.empty(p.n, dtype=torch.float64, device=dev) for p in probs]
    gbs = [torch.empty(p.m, dtype=torch.float64, device=dev) for p in probs]
    gas = [torch.empty((p.n, p.m), dtype=torch.float32, device=dev) if want_a else None for p in probs]
    ldg = (ctypes.c_int64 * P)(*[p.m for p in probs])
    workers = max(0, min(16, opts.adjoint_workers))
    sb = C.lib().tomos_lsq_adjoint_batched_scratch_size(ctypes.byref(o), P, table, workers)
    if sb == 0:
        raise RuntimeError(C.lib().tomos_last_error().decode())
    scr = _bytes(sb, dev)
    its = (ctypes.c_int * P)()
    C.call("tomos_lsq_adjoint_batched", ctypes.byref(o), P, table, _ptrs(gs), _ptrs(ws), _ptrs(gbs),
           _ptrs(gas) if want_a else None, ldg, workers, _p(scr), sb, its, _stream())
    return ws, gbs, [g.t() if g is not None else None for g in gas], list(its)


class _LstsqBatched(torch.autograd.Function):
    @staticmethod
    def forward(ctx, opts, nprob, *tensors):
        As, bs = tensors[:nprob], tensors[nprob:]
        probs = [PackedProblem(A.detach(), opts, precondition=False) for A in As]
        _precondition_batched(probs)
        dev = As[0].device
        b64s = [b.detach().to(torch.float64).contiguous() for b in bs]
        xs = [torch.zeros(p.n, dtype=torch.float64, device=dev) for p in probs]
        o = probs[0].options(opts.tol, opts.inner, opts.restarts, opts.confirm)
        table = _problem_table(probs, b64s, xs)
        sb = C.lib().tomos_lsqr_batched_scratch_size(ctypes.byref(o), nprob, table, 0, 0, 0)
        if sb == 0:
            raise RuntimeError(C.lib().tomos_last_error().decode())
        scr = _bytes(sb, dev)
        its = (ctypes.c_int * nprob)()
        C.call("tomos_lsqr_batched_solve", ctypes.byref(o), nprob, table, 0, 0, 0, 1, _p(scr), sb, its, None,
               _stream())
        if opts.normal_correction:
            ss = [p.normal_residual(b, x) for p, b, x in zip(probs, b64s, xs)]
            deltas, _, _, _ = _adjoint_batched(probs, b64s, xs, ss, opts.adjoint_tol, False)
            xs = [x + d for x, d in zip(xs, deltas)]
        ctx.probs, ctx.b64s, ctx.nprob = probs, b64s, nprob
        ctx.b_dtypes = [b.dtype for b in bs]
        ctx.last_iterations = list(its)
        ctx.save_for_backward(*xs)
        return tuple(xs)

    @staticmethod
    def backward(ctx, *gs):
        xs = ctx.saved_tensors
        probs, P = ctx.probs, ctx.nprob
        dev = xs[0].device
        g64 = [(g if g is not None else torch.zeros(p.n, device=dev)).detach().to(torch.float64).contiguous()
               for g, p in zip(gs, probs)]
        want_a = any(ctx.needs_input_grad[2:2 + P])
        _, gbs, gas, _ = _adjoint_batched(probs, ctx.b64s, list(xs), g64, probs[0].opts.adjoint_tol, want_a)
        grad_As = [ga if ctx.needs_input_grad[2 + i] else None for i, ga in enumerate(gas)]
        grad_bs = [gb.to(dt) if ctx.needs_input_grad[2 + P + i] else None
                   for i, (gb, dt) in enumerate(zip(gbs, ctx.b_dtypes))]
        return (None, None, *grad_As, *grad_bs)


def lstsq_batched(As, bs, opts=None):
    """[argmin |A_i x - b_i|] for lists of problems of any shapes (FP64
    results), differentiable in every A_i and b_i."""
    if len(As) != len(bs) or not As:
        raise ValueError("need as many right-hand sides as matrices (at least one)")
    return list(_LstsqBatched.apply(opts or SolveOptions(), len(As), *As, *bs))
