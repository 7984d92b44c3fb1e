"""Batched LSQR over uneven problems vs FP64 least squares and the
single-problem solver; graph replay vs plain launches."""
import ctypes

import numpy as np
import pytest

from tests import hipcall as H
from tomos_amd import _capi as C
from tomos_amd._capi import DeviceBuffer

pytestmark = pytest.mark.gpu


def make_batch(shapes, seed, a_type=C.FP32_BF16):
    rng = np.random.default_rng(seed)
    out = []
    for (m, n, kexp) in shapes:
        A = (rng.standard_normal((m, n)) * np.logspace(0, -kexp, n)).astype(np.float32, order="F")
        x_true = rng.standard_normal(n)
        b = A.astype(np.float64) @ x_true
        lda = (m + 7) // 8 * 8
        if a_type == C.F32:
            Ap = np.zeros((lda, n), np.float32, order="F")
            Ap[:m] = A
            wh, wr = H.dev(Ap), None
        else:
            wh, wr = H.pack(C.PACK_FP32_BF16 if a_type == C.FP32_BF16 else C.PACK_FP32_16, A, ld=lda)
        s = 4 * n
        sk = DeviceBuffer(s * n * 8)
        C.call("tomos_countsketch_apply", a_type, C.TIER_EXACT if a_type != C.F32 else C.TIER_DEFAULT,
               m, n, wh.ptr, wr.ptr if wr else None, lda, s, sk.ptr, s, None, 0, None)
        out.append(dict(A=A, b=b, x_true=x_true, m=m, n=n, lda=lda, wh=wh, wr=wr, sk=sk, s=s,
                        bd=H.dev(b), xd=DeviceBuffer(n * 8)))
    probs = (C.QrProblem * len(out))(*[C.QrProblem(p["s"], p["n"], p["s"], p["sk"].ptr) for p in out])
    C.call("tomos_sketch_qr_precond_batched_f64", len(out), probs, None, 0, None)
    return out


# tol = 1e-6 on the true residual (the stop is confirmed on the exact-tier
# restart residual); the factor covers the FP32 rounding of the stored
# residual vector the solver's norm is taken from
TRUE_TOL = 1e-6 * (1 + 1e-6)


def opts(a_type, tol=1e-6, inner=30, restarts=40, confirm=1):
    return C.LsqrOptions(a_type, C.TIER_DEFAULT, C.TIER_EXACT if a_type != C.F32 else C.TIER_FAST,
                         C.F32, C.F64, C.F64, tol, tol, inner, restarts, 8, 1, confirm)


def solve_batched(batch, o, use_graph=1, groups=0, tiles=0, threads=0, scr=None):
    P = (C.LsqProblem * len(batch))(*[
        C.LsqProblem(p["m"], p["n"], p["lda"], p["wh"].ptr, p["wr"].ptr if p["wr"] else None, p["bd"].ptr,
                     p["xd"].ptr, p["sk"].ptr, p["s"]) for p in batch])
    sb = C.lib().tomos_lsqr_batched_scratch_size(ctypes.byref(o), len(batch), P, groups, threads, tiles)
    assert sb > 0, C.lib().tomos_last_error()
    if scr is None:
        scr = DeviceBuffer(sb)
    assert scr.nbytes >= sb
    its = (ctypes.c_int * len(batch))()
    rr = (ctypes.c_double * len(batch))()
    C.call("tomos_lsqr_batched_solve", ctypes.byref(o), len(batch), P, groups, threads, tiles, use_graph, scr.ptr,
           sb, its, rr, None)
    xs = [p["xd"].to_numpy(np.float64, (p["n"],)) for p in batch]
    return xs, list(its), list(rr)


SHAPES = [(20000, 64, 2), (3001, 17, 1), (40000, 300, 4), (9000, 128, 3), (12345, 1, 0), (25000, 200, 2)]


@pytest.mark.parametrize("a_type", [C.FP32_BF16, C.FP32_16, C.F32])
def test_batched_matches_fp64(capi, a_type):
    batch = make_batch(SHAPES, 3, a_type)
    xs, its, rr = solve_batched(batch, opts(a_type))
    for p, x, it in zip(batch, xs, its):
        A64 = p["A"].astype(np.float64)
        relres = np.linalg.norm(p["b"] - A64 @ x) / np.linalg.norm(p["b"])
        assert relres <= TRUE_TOL, (p["m"], p["n"], relres, it)
        assert it < 200, it


def test_batched_graph_equals_plain_and_single(capi):
    batch = make_batch(SHAPES, 4)
    o = opts(C.FP32_BF16)
    xg, itg, _ = solve_batched(batch, o, use_graph=1)
    xp, itp, _ = solve_batched(batch, o, use_graph=0)
    assert itg == itp
    for a, b in zip(xg, xp):
        assert np.array_equal(a, b)
    # the single-problem solver reaches the same tolerance in a similar count
    for p, x, it in zip(batch, xg, itg):
        rinv = p["sk"].to_numpy(np.float64, (p["s"], p["n"]))[:p["n"], :p["n"]]
        xs, its, _ = H.lsqr((p["wh"], p["wr"]), p["b"], a_type=C.FP32_BF16, restart_tier=C.TIER_EXACT,
                            value_type=C.F32, refine_type=C.F64, rinv=rinv, precond_type=C.F64, tol=1e-6,
                            max_iters=30, restarts=40, check_every=8, lda=p["lda"], m=p["m"], n=p["n"], fused=1,
                            confirm=1)
        assert abs(its - it) <= max(3, 0.2 * its), (p["m"], p["n"], its, it)
        ref = np.linalg.lstsq(p["A"].astype(np.float64), p["b"], rcond=None)[0]
        assert np.linalg.norm(x - xs) <= 1e-3 * np.linalg.norm(ref) + 1e-12


@pytest.mark.parametrize("groups,threads,tiles", [(2, 256, 0), (8, 256, 3), (4, 256, 1), (8, 512, 0), (4, 512, 2),
                                                  (32, 1024, 0), (32, 1024, 5)])
def test_batched_launch_shapes(capi, groups, threads, tiles):
    shapes = [(5000, 40, 1), (7001, 100, 2), (3000, 300, 1)]
    batch = make_batch(shapes, 5)
    xs, its, _ = solve_batched(batch, opts(C.FP32_BF16), groups=groups, tiles=tiles, threads=threads)
    for p, x in zip(batch, xs):
        relres = np.linalg.norm(p["b"] - p["A"].astype(np.float64) @ x) / np.linalg.norm(p["b"])
        assert relres <= TRUE_TOL, relres


def test_batched_graph_cache_reuse(capi):
    """The instantiated whole-solve graph is cached per launch configuration
    and scratch address: a second batch of the same shapes but other data,
    solved with the same scratch, reuses it and still matches plain launches;
    so does a batch of other shapes (another graph) and the first batch again."""
    shapes2 = [(m + 5, n, k) for (m, n, k) in SHAPES]
    b1, b2, b3 = make_batch(SHAPES, 6), make_batch(SHAPES, 7), make_batch(shapes2, 8)
    o = opts(C.FP32_BF16, inner=8, restarts=12)
    P1 = (C.LsqProblem * len(b1))(*[
        C.LsqProblem(p["m"], p["n"], p["lda"], p["wh"].ptr, p["wr"].ptr, p["bd"].ptr, p["xd"].ptr, p["sk"].ptr,
                     p["s"]) for p in b1])
    sb = max(C.lib().tomos_lsqr_batched_scratch_size(ctypes.byref(o), len(b), P1, 0, 0, 0) for b in (b1,)) * 2
    scr = DeviceBuffer(sb)
    for batch in (b1, b2, b3, b1):
        xg, itg, _ = solve_batched(batch, o, use_graph=1, scr=scr)
        xp, itp, _ = solve_batched(batch, o, use_graph=0)
        assert itg == itp
        for a, b in zip(xg, xp):
            assert np.array_equal(a, b)
        for p, x in zip(batch, xg):
            relres = np.linalg.norm(p["b"] - p["A"].astype(np.float64) @ x) / np.linalg.norm(p["b"])
            assert relres <= TRUE_TOL, relres


def test_batched_confirm_off_keeps_estimate_stop(capi):
    """confirm = 0 is the reference's rule: a problem stops on the |phi_bar|
    estimate; its history ends below tol."""
    batch = make_batch(SHAPES[:3], 9)
    xs, its, rr = solve_batched(batch, opts(C.FP32_BF16, confirm=0))
    for p, x, r in zip(batch, xs, rr):
        if p["n"] > 0:
            assert r < 1e-6
        relres = np.linalg.norm(p["b"] - p["A"].astype(np.float64) @ x) / np.linalg.norm(p["b"])
        assert relres < 5e-6, relres
