"""Sketch-preconditioned mixed-precision LSQR at its full size: A = G diag(d),
2,000,000 x 1000 FP32 (G ~ N(0,1) from the device generator, seed 1; d
log-spaced 1 .. 1e-6, kappa ~ 1e6), stored FP32_BF16; b = A x_true in FP64
(x_true ~ N(0,1), numpy seed 2), so the FP64 least-squares solution is
x_true.  Hashed CountSketch with s = 4000 rows, FP64 Householder QR, R^-1;
LSQR with hi-only matvecs (FP32 vectors, FP64 scalars), exact restart
residuals, 8 iterations per restart, stop confirmed on the true residual.

Asserted: the true FP64 relative residual of the exact matrix is <= 1e-6,
the iteration count lies in a stated range, and the error vs x_true is
printed and bounded (relres 1e-6 bounds it only to ~kappa * 1e-6 in the
worst direction).  The same solve through the two-pass kernels (A read
twice per iteration) is held to the same bar."""
import ctypes

import numpy as np
import pytest

from tomos_amd import _capi as C
from tomos_amd._capi import DeviceBuffer

pytestmark = pytest.mark.gpu

M, N, S = 2_000_000, 1000, 4000
TRUE_TOL = 1e-6 * (1 + 1e-6)


@pytest.fixture(scope="module")
def c4(capi):
    d = DeviceBuffer.from_numpy(np.logspace(0, -6, N).astype(np.float32))
    A = DeviceBuffer(M * N * 4)
    C.call("tomos_fill_normal_f32", A.ptr, M, N, M, 1, d.ptr, 0, 0, None)
    wh, wr = DeviceBuffer(M * N * 2), DeviceBuffer(M * N * 2)
    C.call("tomos_pack", C.PACK_FP32_BF16, M, N, A.ptr, M, wh.ptr, wr.ptr, M, None)
    del A
    x_true = np.random.default_rng(2).standard_normal(N)
    xt = DeviceBuffer.from_numpy(x_true)
    b = DeviceBuffer(M * 8)
    # FP32_BF16's truncating split is lossless: the exact tier is A itself
    C.call("tomos_gemv", C.FP32_BF16, C.TIER_EXACT, 0, 1.0, M, N, wh.ptr, wr.ptr, M, C.F64, xt.ptr, 0.0, C.F64,
           b.ptr, None, 0, None)
    sk = DeviceBuffer(S * N * 8)
    C.call("tomos_countsketch_apply", C.FP32_BF16, C.TIER_EXACT, M, N, wh.ptr, wr.ptr, M, S, sk.ptr, S, None, 0,
           None)
    C.call("tomos_sketch_qr_precond_f64", S, N, sk.ptr, S, None)
    C.call("tomos_device_synchronize")
    return dict(wh=wh, wr=wr, b=b, x_true=x_true, sk=sk)


def true_relres(c, x):
    r = DeviceBuffer(M * 8)
    C.call("tomos_memcpy_dtod", r.ptr, c["b"].ptr, M * 8, None)
    C.call("tomos_gemv", C.FP32_BF16, C.TIER_EXACT, 0, -1.0, M, N, c["wh"].ptr, c["wr"].ptr, M, C.F64, x.ptr, 1.0,
           C.F64, r.ptr, None, 0, None)
    rn = np.linalg.norm(r.to_numpy(np.float64, (M,)))
    bn = np.linalg.norm(c["b"].to_numpy(np.float64, (M,)))
    return rn / bn


@pytest.mark.parametrize("fused", [1, 0])
def test_c4_full_size(c4, fused):
    o = C.LsqrOptions(C.FP32_BF16, C.TIER_DEFAULT, C.TIER_EXACT, C.F32, C.F64, C.F64, 1e-6, 1e-6, 8, 40, 8,
                      fused, 1)
    ab = C.lib().tomos_lsqr_arena_size(ctypes.byref(o), M, N)
    sb = C.lib().tomos_lsqr_scratch_size(ctypes.byref(o), M, N)
    arena, scr, x = DeviceBuffer(ab), DeviceBuffer(sb), DeviceBuffer(N * 8)
    it, hl = ctypes.c_int(0), ctypes.c_int(0)
    hist = np.zeros(400)
    C.call("tomos_lsqr_solve", ctypes.byref(o), M, N, c4["wh"].ptr, c4["wr"].ptr, M, c4["b"].ptr, x.ptr,
           c4["sk"].ptr, S, arena.ptr, ab, scr.ptr, sb, ctypes.byref(it), hist.ctypes.data, hist.size,
           ctypes.byref(hl), None)
    rr = true_relres(c4, x)
    xs = x.to_numpy(np.float64, (N,))
    err = np.linalg.norm(xs - c4["x_true"]) / np.linalg.norm(c4["x_true"])
    print("C4 fused=%d: iterations %d, true relres %.3e, |x - x_FP64| / |x_FP64| %.3e"
          % (fused, it.value, rr, err), flush=True)
    assert rr <= TRUE_TOL, rr
    assert 8 <= it.value <= 64, it.value
    assert err < 0.1, err
    # the last history entry is the estimate that ended the last restart
    assert hist[min(hl.value, hist.size) - 1] < 1e-6 or hl.value > hist.size
