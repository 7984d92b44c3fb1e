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

The truncating FP32_BF16 split is lossless (hi and lo halves of the FP32
bits), so the solve and the gradients refer to A itself.  Everything runs
on torch's current stream; there is no CPU path.
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
    sketch_factor: int = 4       # sketch rows s = sketch_factor * n
    adjoint_tol: float = 1e-10   # |g - A^T A w| / |g| for the adjoint solve
    check_every: int = 8
    normal_correction: bool = True  # forward: x += (A^T A)^-1 A^T (b - A x), exact products


class PackedProblem:
    """A packed as FP32_BF16 planes plus its sketch-QR preconditioner."""

    def __init__(self, A, opts):
        if not A.is_cuda or A.dtype != torch.float32 or A.dim() != 2:
            raise ValueError("A must be a 2-D float32 CUDA tensor")
        m, n = A.shape
        if m < n or n < 1:
            raise ValueError("need m >= n >= 1")
        self.m, self.n, self.opts = m, n, opts
        dev = A.device
        self.ld = (m + 7) // 8 * 8  # 16-byte row vectors for the single-pass kernel
        acm = A.t().contiguous()    # (n, m) row-major == column-major (m, n)
        self.wh = torch.empty(self.ld * n, dtype=torch.int16, device=dev)
        self.wr = torch.empty(self.ld * n, dtype=torch.int16, device=dev)
        s = _stream()
        C.call("tomos_pack", C.PACK_FP32_BF16, m, n, _p(acm), m, _p(self.wh), _p(self.wr), self.ld, s)
        self.s_rows = max(opts.sketch_factor * n, n)
        self.sketch = torch.empty((n, self.s_rows), dtype=torch.float64, device=dev)
        cs = C.lib().tomos_countsketch_scratch_size(m, n, self.s_rows)
        scr = _bytes(cs, dev)
        C.call("tomos_countsketch_apply", C.FP32_BF16, C.TIER_EXACT, m, n, _p(self.wh), _p(self.wr),
               self.ld, self.s_rows, _p(self.sketch), self.s_rows, _p(scr), cs, s)
        # R^-1 overwrites the leading n x n block (ld = s_rows)
        C.call("tomos_sketch_qr_precond_f64", self.s_rows, n, _p(self.sketch), self.s_rows, s)

    def options(self, tol, iters, restarts):
        return C.LsqrOptions(C.FP32_BF16, C.TIER_DEFAULT, C.TIER_EXACT, C.F32, C.F64, C.F64,
                             tol, tol, iters, restarts, self.opts.check_every, 1)

    def solve(self, b64):
        o = self.options(self.opts.tol, self.opts.inner, self.opts.restarts)
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
