"""Adjoint of the least-squares solve on the GPU (C ABI and the torch
autograd function) vs the dense FP64 QR gradient and central finite
differences."""
import ctypes

import numpy as np
import pytest

from tests import hipcall as H
from tests.adjoint_ref import dense_grads
from tomos_amd import _capi as C
from tomos_amd._capi import DeviceBuffer

pytestmark = pytest.mark.gpu


def problem(m, n, kappa_exp, seed, noise=1e-2):
    rng = np.random.default_rng(seed)
    A = (rng.standard_normal((m, n)) * np.logspace(0, -kappa_exp, n)).astype(np.float32, order="F")
    b = A.astype(np.float64) @ rng.standard_normal(n) + noise * rng.standard_normal(m)
    g = rng.standard_normal(n)
    return A, b, g


def adjoint_w(a_type, planes, m, n, lda, g, rinv_buf=None, ldr=0, tol=1e-10, fused=1,
              inner=30, restarts=20):
    opt = C.LsqrOptions(a_type, C.TIER_DEFAULT, C.TIER_EXACT, C.F32, C.F64,
                        C.F64 if rinv_buf is not None else -1, tol, tol, inner, restarts, 4, fused)
    sb = C.lib().tomos_lsq_adjoint_scratch_size(ctypes.byref(opt), m, n)
    assert sb > 0, C.lib().tomos_last_error()
    scr, gd, wd = DeviceBuffer(sb), H.dev(g), DeviceBuffer(n * 8)
    it, hl = ctypes.c_int(0), ctypes.c_int(0)
    hist = np.zeros(4096)
    C.call("tomos_lsq_adjoint_solve", ctypes.byref(opt), m, n, planes[0].ptr,
           planes[1].ptr if planes[1] is not None else None, lda, gd.ptr,
           rinv_buf.ptr if rinv_buf is not None else None, ldr, wd.ptr, scr.ptr, sb,
           ctypes.byref(it), hist.ctypes.data, hist.size, ctypes.byref(hl), None)
    return wd.to_numpy(np.float64, (n,)), it.value, hist[:hl.value]


@pytest.mark.parametrize("m,n,kexp,fused", [(20000, 64, 2, 1), (5003, 37, 2, 1), (20000, 64, 2, 0)])
def test_adjoint_w_preconditioned_packed(capi, m, n, kexp, fused):
    A, b, g = problem(m, n, kexp, m + n)
    lda = (m + 7) // 8 * 8
    wh, wr = H.pack(C.PACK_FP32_BF16, A, ld=lda)
    _, sk = H.sketch_rinv(C.FP32_BF16, C.TIER_EXACT, m, n, wh, wr, lda, 4 * n)
    w, iters, hist = adjoint_w(C.FP32_BF16, (wh, wr), m, n, lda, g, sk, 4 * n, fused=fused)
    w_ref = dense_grads(A, b, g)[0]
    assert hist[-1] < 1e-10, hist[-5:]
    # |w - w*| <= kappa(A^T A) |res| / |g| |w*|-ish; kappa(A) = 1e2 here
    assert np.linalg.norm(w - w_ref) / np.linalg.norm(w_ref) < 1e-5
    assert iters < 200


def test_adjoint_w_unpreconditioned_f32(capi):
    m, n = 3000, 20
    A, b, g = problem(m, n, 0.5, 5)
    w, iters, hist = adjoint_w(C.F32, (H.dev(A), None), m, n, m, g, None, 0, tol=1e-9, fused=1,
                               inner=60, restarts=10)
    w_ref = dense_grads(A, b, g)[0]
    assert np.linalg.norm(w - w_ref) / np.linalg.norm(w_ref) < 1e-7


def test_grad_kernel_no_columns(capi):
    """m > 0, n == 0: x and w are empty, so grad_b = A w = 0 is written over
    whatever grad_b held, and grad_A (m x 0) is untouched."""
    m = 1000
    gbd = H.dev(np.full(m, 7.0))
    a = DeviceBuffer(m * 2)
    C.call("tomos_lsq_grad", C.FP32_BF16, C.TIER_EXACT, m, 0, a.ptr, a.ptr, m, gbd.ptr, None, None, gbd.ptr,
           None, m, None, 0, None)
    assert np.array_equal(gbd.to_numpy(np.float64, (m,)), np.zeros(m))


@pytest.mark.parametrize("ldg_pad", [0, 1])
def test_grad_kernel(capi, ldg_pad):
    m, n = 10007, 33
    A, b, g = problem(m, n, 1, 9)
    wh, wr = H.pack(C.PACK_FP32_BF16, A)
    rng = np.random.default_rng(3)
    x, w = rng.standard_normal(n), rng.standard_normal(n)
    ldg = m + ldg_pad  # odd ldg takes the element-wise store path
    gbd, gad = DeviceBuffer(m * 8), DeviceBuffer(ldg * n * 4)
    bd, xd, wd = H.dev(b), H.dev(x), H.dev(w)
    C.call("tomos_lsq_grad", C.FP32_BF16, C.TIER_EXACT, m, n, wh.ptr, wr.ptr, m, bd.ptr, xd.ptr, wd.ptr,
           gbd.ptr, gad.ptr, ldg, None, 0, None)
    A64 = A.astype(np.float64)
    gb_ref = A64 @ w
    gA_ref = np.outer(b - A64 @ x, w) - np.outer(gb_ref, x)
    gb = gbd.to_numpy(np.float64, (m,))
    gA = gad.to_numpy(np.float32, (ldg, n))[:m]
    assert np.abs(gb - gb_ref).max() < 1e-12 * np.abs(A64).max() * np.abs(w).sum()
    assert np.abs(gA - gA_ref).max() < 2e-7 * np.abs(gA_ref).max()


def test_torch_autograd_vs_dense_and_fd(capi):
    torch = pytest.importorskip("torch")
    from tomos_amd.autograd import lstsq, SolveOptions
    m, n = 20000, 64
    A_np, b_np, g_np = problem(m, n, 3, 21)
    dev = torch.device("cuda:0")
    A = torch.tensor(A_np, device=dev, requires_grad=True)
    b = torch.tensor(b_np, device=dev, requires_grad=True)
    g = torch.tensor(g_np, device=dev)
    opts = SolveOptions(tol=1e-10, adjoint_tol=1e-10)
    x = lstsq(A, b, opts)
    (x @ g).backward()
    w_ref, gb_ref, gA_ref, x_ref, _ = dense_grads(A_np, b_np, g_np)
    xe = np.linalg.norm(x.detach().cpu().numpy() - x_ref) / np.linalg.norm(x_ref)
    assert xe < 1e-6, xe  # kappa = 1e3, FP64-refined solve
    gb = b.grad.cpu().numpy()
    gA = A.grad.cpu().numpy()
    assert np.linalg.norm(gb - gb_ref) / np.linalg.norm(gb_ref) < 1e-5
    assert np.linalg.norm(gA - gA_ref) / np.linalg.norm(gA_ref) < 1e-5
    # central differences (h = 1e-3) of the GPU forward along random directions
    rng = np.random.default_rng(2)
    h = 1e-3
    with torch.no_grad():
        for _ in range(2):
            db = torch.tensor(rng.standard_normal(m), device=dev)
            fd = ((lstsq(A, b + h * db, opts) - lstsq(A, b - h * db, opts)) @ g).item() / (2 * h)
            an = float(gb_ref @ db.cpu().numpy())
            assert abs(fd - an) < 1e-4 * (abs(an) + np.abs(gb_ref).sum() * 1e-3), (fd, an)
        dA = torch.tensor(rng.standard_normal((m, n)).astype(np.float32), device=dev) * A.detach().abs()
        fd = ((lstsq(A.detach() + h * dA, b, opts) - lstsq(A.detach() - h * dA, b, opts)) @ g).item() / (2 * h)
        an = float((gA_ref * dA.cpu().numpy().astype(np.float64)).sum())
        assert abs(fd - an) < 1e-2 * abs(an) + 1e-6, (fd, an)


@pytest.mark.parametrize("workers", [1, 3, 16])
def test_adjoint_batched_matches_one_problem_calls(capi, workers):
    """tomos_lsq_adjoint_batched (problems spread over host threads, one HIP
    stream each) gives bit-identical w, grad_b and grad_A to the
    one-problem tomos_lsq_adjoint_solve + tomos_lsq_grad calls, for uneven
    shapes (packed FP32_BF16 with the sketch-QR preconditioner)."""
    shapes = [(20000, 64), (5003, 37), (12000, 128), (4096, 16), (9001, 90)]
    opt = C.LsqrOptions(C.FP32_BF16, C.TIER_DEFAULT, C.TIER_EXACT, C.F32, C.F64, C.F64, 1e-8, 1e-8, 30, 20, 4, 1)
    keep, probs, gs, ws, gbs, gas = [], [], [], [], [], []
    for k, (m, n) in enumerate(shapes):
        A, b, g = problem(m, n, 2, 100 + k)
        lda = (m + 7) // 8 * 8
        wh, wr = H.pack(C.PACK_FP32_BF16, A, ld=lda)
        _, sk = H.sketch_rinv(C.FP32_BF16, C.TIER_EXACT, m, n, wh, wr, lda, 4 * n)
        bd, xd = H.dev(b), H.dev(np.random.default_rng(k).standard_normal(n))
        gd, wd, gbd, gad = H.dev(g), DeviceBuffer(n * 8), DeviceBuffer(m * 8), DeviceBuffer(m * n * 4)
        keep += [wh, wr, sk, bd, xd]
        probs.append(C.LsqProblem(m, n, lda, wh.ptr, wr.ptr, bd.ptr, xd.ptr, sk.ptr, 4 * n))
        gs.append(gd)
        ws.append(wd)
        gbs.append(gbd)
        gas.append(gad)
    P = len(shapes)
    lp = (C.LsqProblem * P)(*probs)
    ptrs = lambda bufs: (ctypes.c_void_p * P)(*[x.ptr for x in bufs])
    ldg = (ctypes.c_int64 * P)(*[m for m, _ in shapes])
    sb = C.lib().tomos_lsq_adjoint_batched_scratch_size(ctypes.byref(opt), P, lp, workers)
    assert sb > 0, C.lib().tomos_last_error()
    scr = DeviceBuffer(sb)
    its = (ctypes.c_int * P)()
    C.call("tomos_lsq_adjoint_batched", ctypes.byref(opt), P, lp, ptrs(gs), ptrs(ws), ptrs(gbs), ptrs(gas), ldg,
           workers, scr.ptr, sb, its, None)
    for k, (m, n) in enumerate(shapes):
        q = probs[k]
        asb = C.lib().tomos_lsq_adjoint_scratch_size(ctypes.byref(opt), m, n)
        ascr = DeviceBuffer(asb)
        w1, it1 = DeviceBuffer(n * 8), ctypes.c_int(0)
        C.call("tomos_lsq_adjoint_solve", ctypes.byref(opt), m, n, q.a, q.wr, q.lda, gs[k].ptr, q.rinv, q.ldr,
               w1.ptr, ascr.ptr, asb, ctypes.byref(it1), None, 0, None, None)
        gsb = C.lib().tomos_lsq_grad_scratch_size(C.FP32_BF16, C.TIER_EXACT, m, n)
        gscr, gb1, ga1 = DeviceBuffer(gsb), DeviceBuffer(m * 8), DeviceBuffer(m * n * 4)
        C.call("tomos_lsq_grad", C.FP32_BF16, C.TIER_EXACT, m, n, q.a, q.wr, q.lda, q.b, q.x, w1.ptr, gb1.ptr,
               ga1.ptr, m, gscr.ptr, gsb, None)
        assert its[k] == it1.value
        assert np.array_equal(ws[k].to_numpy(np.float64, (n,)), w1.to_numpy(np.float64, (n,)))
        assert np.array_equal(gbs[k].to_numpy(np.float64, (m,)), gb1.to_numpy(np.float64, (m,)))
        assert np.array_equal(gas[k].to_numpy(np.float32, (m, n)), ga1.to_numpy(np.float32, (m, n)))


def test_adjoint_grouped_matches_one_problem_calls(capi):
    """tomos_lsq_adjoint_batched with workers = 0 (every problem's CG in the
    same grouped launches, replayed graph segments) against the one-problem
    calls: w, grad_b and grad_A agree to the CG tolerance (the partial sums
    of A^T A d are taken in another order, so not bit for bit), for uneven
    shapes, a problem with g = 0 and restarts forced by a short inner cap;
    and w against the dense FP64 QR adjoint."""
    shapes = [(20000, 64), (5003, 37), (12000, 128), (4096, 16), (9001, 90), (3000, 5)]
    opt = C.LsqrOptions(C.FP32_BF16, C.TIER_DEFAULT, C.TIER_EXACT, C.F32, C.F64, C.F64, 1e-10, 1e-10, 12, 20, 4, 1)
    keep, probs, gs, ws, gbs, gas, dense = [], [], [], [], [], [], []
    for k, (m, n) in enumerate(shapes):
        A, b, g = problem(m, n, 2, 400 + k)
        if k == 5:
            g = np.zeros(n)  # w = 0, grad_b = 0
        lda = (m + 7) // 8 * 8
        wh, wr = H.pack(C.PACK_FP32_BF16, A, ld=lda)
        _, sk = H.sketch_rinv(C.FP32_BF16, C.TIER_EXACT, m, n, wh, wr, lda, 4 * n)
        xv = np.random.default_rng(k).standard_normal(n)
        bd, xd = H.dev(b), H.dev(xv)
        gd, wd, gbd, gad = H.dev(g), DeviceBuffer(n * 8), DeviceBuffer(m * 8), DeviceBuffer(m * n * 4)
        keep += [wh, wr, sk, bd, xd]
        probs.append(C.LsqProblem(m, n, lda, wh.ptr, wr.ptr, bd.ptr, xd.ptr, sk.ptr, 4 * n))
        gs.append(gd)
        ws.append(wd)
        gbs.append(gbd)
        gas.append(gad)
        dense.append(dense_grads(A, b, g)[0] if k != 5 else np.zeros(n))
    P = len(shapes)
    lp = (C.LsqProblem * P)(*probs)
    ptrs = lambda bufs: (ctypes.c_void_p * P)(*[x.ptr for x in bufs])
    ldg = (ctypes.c_int64 * P)(*[m for m, _ in shapes])
    sb = C.lib().tomos_lsq_adjoint_batched_scratch_size(ctypes.byref(opt), P, lp, 0)
    assert sb > 0, C.lib().tomos_last_error()
    scr = DeviceBuffer(sb)
    its = (ctypes.c_int * P)()
    for rep in range(2):  # the second call replays the cached segment graphs
        C.call("tomos_lsq_adjoint_batched", ctypes.byref(opt), P, lp, ptrs(gs), ptrs(ws), ptrs(gbs), ptrs(gas),
               ldg, 0, scr.ptr, sb, its, None)
        for k, (m, n) in enumerate(shapes):
            q = probs[k]
            asb = C.lib().tomos_lsq_adjoint_scratch_size(ctypes.byref(opt), m, n)
            ascr = DeviceBuffer(asb)
            w1, it1 = DeviceBuffer(n * 8), ctypes.c_int(0)
            C.call("tomos_lsq_adjoint_solve", ctypes.byref(opt), m, n, q.a, q.wr, q.lda, gs[k].ptr, q.rinv, q.ldr,
                   w1.ptr, ascr.ptr, asb, ctypes.byref(it1), None, 0, None, None)
            gsb = C.lib().tomos_lsq_grad_scratch_size(C.FP32_BF16, C.TIER_EXACT, m, n)
            gscr, gb1, ga1 = DeviceBuffer(gsb), DeviceBuffer(m * 8), DeviceBuffer(m * n * 4)
            C.call("tomos_lsq_grad", C.FP32_BF16, C.TIER_EXACT, m, n, q.a, q.wr, q.lda, q.b, q.x, w1.ptr, gb1.ptr,
                   ga1.ptr, m, gscr.ptr, gsb, None)
            w0, wref = ws[k].to_numpy(np.float64, (n,)), w1.to_numpy(np.float64, (n,))
            gb0, gbr = gbs[k].to_numpy(np.float64, (m,)), gb1.to_numpy(np.float64, (m,))
            ga0, gar = gas[k].to_numpy(np.float32, (m, n)), ga1.to_numpy(np.float32, (m, n))
            if k == 5:
                assert np.array_equal(w0, np.zeros(n)) and np.array_equal(gb0, np.zeros(m)) and its[k] == 0
                continue
            rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
            # a 1e-10 normal-equation residual bounds w to ~kappa(A^T A) 1e-10
            assert rel(w0, wref) < 1e-6, (k, rel(w0, wref))
            assert rel(w0, dense[k]) < 1e-6, (k, rel(w0, dense[k]))
            assert rel(gb0, gbr) < 1e-6 and rel(ga0, gar) < 1e-5, k
            assert abs(its[k] - it1.value) <= max(3, it1.value // 5), (k, its[k], it1.value)


@pytest.mark.parametrize("workers", [0, 3])
def test_torch_lstsq_batched_vs_dense(capi, workers):
    """lstsq_batched (batched LSQR forward, batched adjoint backward) on
    uneven problems: x, grad_b and grad_A of every problem vs the dense FP64
    QR solution and gradients, for the loss sum_i x_i . g_i."""
    torch = pytest.importorskip("torch")
    from tomos_amd.autograd import lstsq_batched, SolveOptions
    dev = torch.device("cuda:0")
    shapes = [(20000, 64, 3), (7001, 40, 2), (12000, 100, 3), (4096, 16, 1)]
    As, bs, gs, refs = [], [], [], []
    for k, (m, n, kexp) in enumerate(shapes):
        A_np, b_np, g_np = problem(m, n, kexp, 300 + k)
        As.append(torch.tensor(A_np, device=dev, requires_grad=True))
        bs.append(torch.tensor(b_np, device=dev, requires_grad=True))
        gs.append(torch.tensor(g_np, device=dev))
        refs.append(dense_grads(A_np, b_np, g_np))
    opts = SolveOptions(tol=1e-10, adjoint_tol=1e-10, adjoint_workers=workers)
    xs = lstsq_batched(As, bs, opts)
    sum((x * g).sum() for x, g in zip(xs, gs)).backward()
    for k, (x, A, b) in enumerate(zip(xs, As, bs)):
        w_ref, gb_ref, gA_ref, x_ref, _ = refs[k]
        xe = np.linalg.norm(x.detach().cpu().numpy() - x_ref) / np.linalg.norm(x_ref)
        assert xe < 1e-6, (k, xe)
        gb, gA = b.grad.cpu().numpy(), A.grad.cpu().numpy()
        assert np.linalg.norm(gb - gb_ref) / np.linalg.norm(gb_ref) < 1e-5, k
        assert np.linalg.norm(gA - gA_ref) / np.linalg.norm(gA_ref) < 1e-5, k


@pytest.mark.parametrize("n2", [100, 200, 400, 700, 1030])
def test_adjoint_grouped_restart_paths_and_new_matrices(capi, n2):
    """The grouped adjoint's exact restart products come from one grouped
    launch over row items: register tiles for n <= 512 (n2 = 100, 200, 400:
    one, two and four 128-column rounds; the other problem, n = 40, takes
    one round in the same launch), one row
    vector a lane for n <= 1024 (n2 = 700), or from a per-problem GEMV pair
    (a wider problem in the batch, n2 = 1030).  Each is checked against the
    dense FP64 QR adjoint, on two calls with the same scratch and shapes but
    different matrices (the second call must not replay launches bound to
    the first call's matrices).  The gradients (one fused launch over the
    same row items when n <= 512) are checked against tomos_lsq_grad with
    the same w."""
    shapes = [(6000, 40), (5000, n2)]
    opt = C.LsqrOptions(C.FP32_BF16, C.TIER_DEFAULT, C.TIER_EXACT, C.F32, C.F64, C.F64, 1e-10, 1e-10, 12, 20, 4, 1)
    P = len(shapes)
    scr = None
    for call in range(2):
        keep, probs, gs, ws, gbs, gas, dense = [], [], [], [], [], [], []
        for k, (m, n) in enumerate(shapes):
            A, b, g = problem(m, n, 2, 700 + 10 * call + k)
            lda = (m + 7) // 8 * 8
            wh, wr = H.pack(C.PACK_FP32_BF16, A, ld=lda)
            _, sk = H.sketch_rinv(C.FP32_BF16, C.TIER_EXACT, m, n, wh, wr, lda, 4 * n)
            bd, xd = H.dev(b), H.dev(np.random.default_rng(90 + k).standard_normal(n))
            gd, wd, gbd, gad = H.dev(g), DeviceBuffer(n * 8), DeviceBuffer(m * 8), DeviceBuffer(m * n * 4)
            keep += [wh, wr, sk, bd, xd]
            probs.append(C.LsqProblem(m, n, lda, wh.ptr, wr.ptr, bd.ptr, xd.ptr, sk.ptr, 4 * n))
            gs.append(gd)
            ws.append(wd)
            gbs.append(gbd)
            gas.append(gad)
            dense.append(dense_grads(A, b, g)[0])
        lp = (C.LsqProblem * P)(*probs)
        ptrs = lambda bufs: (ctypes.c_void_p * P)(*[x.ptr for x in bufs])
        ldg = (ctypes.c_int64 * P)(*[m for m, _ in shapes])
        sb = C.lib().tomos_lsq_adjoint_batched_scratch_size(ctypes.byref(opt), P, lp, 0)
        assert sb > 0, C.lib().tomos_last_error()
        if scr is None:
            scr = DeviceBuffer(sb)
        its = (ctypes.c_int * P)()
        C.call("tomos_lsq_adjoint_batched", ctypes.byref(opt), P, lp, ptrs(gs), ptrs(ws), ptrs(gbs), ptrs(gas), ldg,
               0, scr.ptr, sb, its, None)
        for k, (m, n) in enumerate(shapes):
            w0 = ws[k].to_numpy(np.float64, (n,))
            rel = np.linalg.norm(w0 - dense[k]) / np.linalg.norm(dense[k])
            assert rel < 1e-6, (call, k, rel)
            # the batch's gradients (one fused launch for n <= 512) against
            # the one-problem gradient kernel with the same w: the FP64 row
            # sums differ only in order, grad_A by FP32 rounding of those
            q = probs[k]
            gsb = C.lib().tomos_lsq_grad_scratch_size(C.FP32_BF16, C.TIER_EXACT, m, n)
            gscr, gb1, ga1 = DeviceBuffer(gsb), DeviceBuffer(m * 8), DeviceBuffer(m * n * 4)
            C.call("tomos_lsq_grad", C.FP32_BF16, C.TIER_EXACT, m, n, q.a, q.wr, q.lda, q.b, q.x, ws[k].ptr,
                   gb1.ptr, ga1.ptr, m, gscr.ptr, gsb, None)
            gb0, gbr = gbs[k].to_numpy(np.float64, (m,)), gb1.to_numpy(np.float64, (m,))
            ga0, gar = gas[k].to_numpy(np.float32, (m, n)), ga1.to_numpy(np.float32, (m, n))
            assert np.linalg.norm(gb0 - gbr) <= 1e-12 * np.linalg.norm(gbr), (call, k)
            assert np.max(np.abs(ga0.astype(np.float64) - gar)) <= 2.5e-7 * np.max(np.abs(gar)), (call, k)
