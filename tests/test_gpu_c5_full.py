"""Batched differentiable least squares at its full size: 64 problems,
m ~ U{20000..200000}, n ~ U{64..512}, kappa log-uniform in [1e2, 1e5]
(numpy seed 7), A = G diag(d) (G ~ N(0,1) from the device generator,
seed 100 + i), stored FP32_BF16; b = A x_true in FP64.

Forward: hashed CountSketch (s = 4n), batched FP64 sketch-QR, the batched
LSQR to 1e-6 with the stop confirmed on the exact restart residual; every
problem's true FP64 relative residual is asserted <= 1e-6.

Backward (tomos_amd.autograd.lstsq_batched, upstream g ~ N(0,1)): w,
grad_b = A w and grad_A = r w^T - (A w) x^T of the 4 largest and the 4
smallest problems against the dense FP64 QR gradient (tests/adjoint_ref.py),
and central finite differences (h = 1e-3) of L = g^T x along random b and A
directions for three problems."""
import ctypes

import numpy as np
import pytest

from tests.adjoint_ref import dense_grads
from tomos_amd import _capi as C
from tomos_amd._capi import DeviceBuffer

pytestmark = pytest.mark.gpu

P = 64
TRUE_TOL = 1e-6 * (1 + 1e-6)


def c5_shapes():
    rng = np.random.default_rng(7)
    ms = rng.integers(20000, 200001, P)
    ns = rng.integers(64, 513, P)
    kap = 10.0 ** rng.uniform(2, 5, P)
    return rng, ms, ns, kap


def test_c5_forward_full_size(capi):
    rng, ms, ns, kap = c5_shapes()
    probs = []
    for i in range(P):
        m, n = int(ms[i]), int(ns[i])
        lda = (m + 7) // 8 * 8
        d = DeviceBuffer.from_numpy(np.logspace(0, -np.log10(kap[i]), n).astype(np.float32))
        A = DeviceBuffer(lda * n * 4)
        C.call("tomos_fill_normal_f32", A.ptr, m, n, lda, 100 + i, d.ptr, 0, 0, None)
        wh, wr = DeviceBuffer(lda * n * 2), DeviceBuffer(lda * n * 2)
        C.call("tomos_pack", C.PACK_FP32_BF16, m, n, A.ptr, lda, wh.ptr, wr.ptr, lda, None)
        del A
        xt = DeviceBuffer.from_numpy(rng.standard_normal(n))
        b = DeviceBuffer(m * 8)
        C.call("tomos_gemv", C.FP32_BF16, C.TIER_EXACT, 0, 1.0, m, n, wh.ptr, wr.ptr, lda, C.F64, xt.ptr, 0.0,
               C.F64, b.ptr, None, 0, None)
        s = 4 * n
        sk = DeviceBuffer(s * n * 8)
        C.call("tomos_countsketch_apply", C.FP32_BF16, C.TIER_EXACT, m, n, wh.ptr, wr.ptr, lda, s, sk.ptr, s,
               None, 0, None)
        probs.append(dict(m=m, n=n, lda=lda, wh=wh, wr=wr, b=b, s=s, sk=sk, x=DeviceBuffer(n * 8)))
    qp = (C.QrProblem * P)(*[C.QrProblem(p["s"], p["n"], p["s"], p["sk"].ptr) for p in probs])
    C.call("tomos_sketch_qr_precond_batched_f64", P, qp, None, 0, None)
    o = C.LsqrOptions(C.FP32_BF16, C.TIER_DEFAULT, C.TIER_EXACT, C.F32, C.F64, C.F64, 1e-6, 1e-6, 8, 20, 8, 1, 1)
    lp = (C.LsqProblem * P)(*[C.LsqProblem(p["m"], p["n"], p["lda"], p["wh"].ptr, p["wr"].ptr, p["b"].ptr,
                                           p["x"].ptr, p["sk"].ptr, p["s"]) for p in probs])
    sb = C.lib().tomos_lsqr_batched_scratch_size(ctypes.byref(o), P, lp, 0, 0, 0)
    assert sb > 0, C.lib().tomos_last_error()
    scr = DeviceBuffer(sb)
    its = (ctypes.c_int * P)()
    rr = (ctypes.c_double * P)()
    C.call("tomos_lsqr_batched_solve", ctypes.byref(o), P, lp, 0, 0, 0, 1, scr.ptr, sb, its, rr, None)
    worst = 0.0
    for i, p in enumerate(probs):
        r = DeviceBuffer(p["m"] * 8)
        C.call("tomos_memcpy_dtod", r.ptr, p["b"].ptr, p["m"] * 8, None)
        C.call("tomos_gemv", C.FP32_BF16, C.TIER_EXACT, 0, -1.0, p["m"], p["n"], p["wh"].ptr, p["wr"].ptr,
               p["lda"], C.F64, p["x"].ptr, 1.0, C.F64, r.ptr, None, 0, None)
        rel = np.linalg.norm(r.to_numpy(np.float64, (p["m"],))) / np.linalg.norm(p["b"].to_numpy(np.float64,
                                                                                                  (p["m"],)))
        worst = max(worst, rel)
        assert rel <= TRUE_TOL, (i, p["m"], p["n"], kap[i], rel, its[i])
    print("C5 forward: iterations %d..%d, worst true relres %.4e" % (min(its), max(its), worst), flush=True)
    assert max(its) <= 8 * 20


@pytest.fixture(scope="module")
def c5_torch(capi):
    import torch
    rng, ms, ns, kap = c5_shapes()
    dev = torch.device("cuda:0")
    As, bs = [], []
    for i in range(P):
        m, n = int(ms[i]), int(ns[i])
        At = torch.empty((n, m), dtype=torch.float32, device=dev)  # column-major (m, n)
        d = torch.from_numpy(np.logspace(0, -np.log10(kap[i]), n).astype(np.float32)).to(dev)
        C.call("tomos_fill_normal_f32", At.data_ptr(), m, n, m, 100 + i, d.data_ptr(), 0, 0,
               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        x_true = torch.from_numpy(rng.standard_normal(n)).to(dev)
        A = At.t()
        As.append(A)
        bs.append(A.to(torch.float64) @ x_true)
    torch.cuda.synchronize()
    return ms, ns, kap, As, bs


def test_c5_gradients_vs_dense_qr(c5_torch):
    import torch
    from tomos_amd.autograd import SolveOptions, lstsq_batched
    ms, ns, kap, As, bs = c5_torch
    opts = SolveOptions(tol=1e-6, inner=8, restarts=20, adjoint_tol=1e-10)
    Ar = [A.detach().clone().requires_grad_(True) for A in As]
    br = [b.detach().clone().requires_grad_(True) for b in bs]
    xs = lstsq_batched(Ar, br, opts)
    gen = torch.Generator(device="cpu").manual_seed(11)
    gs = [torch.randn(int(n), generator=gen, dtype=torch.float64).to(As[0].device) for n in ns]
    loss = sum((x * g).sum() for x, g in zip(xs, gs))
    loss.backward()
    size = ms.astype(np.int64) * ns
    order = np.argsort(size)
    pick = list(order[:4]) + list(order[-4:])
    for i in pick:
        A = As[i].detach().cpu().numpy().astype(np.float64)
        b = bs[i].detach().cpu().numpy()
        g = gs[i].cpu().numpy()
        w_ref, gb_ref, _, x_ref, _ = dense_grads(A, b, g)
        x = xs[i].detach().cpu().numpy()
        gb = br[i].grad.cpu().numpy()
        gA = Ar[i].grad.cpu().numpy().astype(np.float64)
        # grad_A at the solver's own x: r w^T - (A w) x^T with the FP64 QR's w
        gA_ref = np.outer(b - A @ x, w_ref) - np.outer(gb_ref, x)
        ex = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
        eb = np.linalg.norm(gb - gb_ref) / np.linalg.norm(gb_ref)
        eA = np.linalg.norm(gA - gA_ref) / np.linalg.norm(gA_ref)
        print("C5 problem %d (%d x %d, kappa %.1e): x %.2e, grad_b %.2e, grad_A %.2e vs dense FP64 QR"
              % (i, ms[i], ns[i], kap[i], ex, eb, eA), flush=True)
        assert ex < 1e-6, ex
        assert eb < 1e-5, eb
        assert eA < 1e-5, eA


def test_c5_gradients_central_differences(c5_torch):
    import torch
    from tomos_amd.autograd import SolveOptions, lstsq_batched
    ms, ns, kap, As, bs = c5_torch
    opts = SolveOptions(tol=1e-6, inner=8, restarts=20, adjoint_tol=1e-10)
    order = np.argsort(kap)
    pick = [int(order[0]), int(order[1]), int(order[len(order) // 2])]
    sub_A = [As[i].detach().clone().requires_grad_(True) for i in pick]
    sub_b = [bs[i].detach().clone().requires_grad_(True) for i in pick]
    gen = torch.Generator(device="cpu").manual_seed(13)
    gs = [torch.randn(int(ns[i]), generator=gen, dtype=torch.float64).to(As[0].device) for i in pick]
    xs = lstsq_batched(sub_A, sub_b, opts)
    sum((x * g).sum() for x, g in zip(xs, gs)).backward()
    h = 1e-3

    def loss(Alist, blist):
        with torch.no_grad():
            ys = lstsq_batched(Alist, blist, opts)
        return [float((y * g).sum()) for y, g in zip(ys, gs)]

    # b direction: unit-norm random db, step h * |b|
    dbs = [torch.randn(b.shape, generator=gen, dtype=torch.float64).to(b.device) for b in sub_b]
    dbs = [d / d.norm() * b.norm() for d, b in zip(dbs, sub_b)]
    Ad = [A.detach() for A in sub_A]
    lp = loss(Ad, [b.detach() + h * d for b, d in zip(sub_b, dbs)])
    lm = loss(Ad, [b.detach() - h * d for b, d in zip(sub_b, dbs)])
    for k, i in enumerate(pick):
        fd = (lp[k] - lm[k]) / (2 * h)
        an = float((sub_b[k].grad * dbs[k]).sum())
        print("C5 problem %d (kappa %.1e): b-direction FD %.10e, adjoint %.10e" % (i, kap[i], fd, an), flush=True)
        assert abs(fd - an) <= 1e-4 * abs(an) + 1e-9 * float(dbs[k].norm()), (fd, an)
    # A direction: E with |E| = 1e-2 |A|, step h; the FP32 rounding of A +- hE
    # is absorbed by using the effective perturbation (A+ - A-) / 2h
    Es = [torch.randn(A.shape, generator=gen, dtype=torch.float32).to(A.device) for A in sub_A]
    Es = [E / E.norm() * (1e-2 * A.detach().norm()) for E, A in zip(Es, sub_A)]
    Ap = [(A.detach() + h * E).t().contiguous().t() for A, E in zip(sub_A, Es)]
    Am = [(A.detach() - h * E).t().contiguous().t() for A, E in zip(sub_A, Es)]
    bd = [b.detach() for b in sub_b]
    lp, lm = loss(Ap, bd), loss(Am, bd)
    for k, i in enumerate(pick):
        eff = (Ap[k].double() - Am[k].double()) / (2 * h)
        fd = (lp[k] - lm[k]) / (2 * h)
        an = float((sub_A[k].grad.double() * eff).sum())
        print("C5 problem %d (kappa %.1e): A-direction FD %.10e, adjoint %.10e" % (i, kap[i], fd, an), flush=True)
        assert abs(fd - an) <= 1e-3 * abs(an) + 1e-9, (fd, an)
