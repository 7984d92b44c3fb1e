"""GPU packed GEMM tiers against the CPU oracle and FP64 numpy."""
import numpy as np
import pytest

from tests import hipcall as H
from tomos_amd import _capi as C

pytestmark = pytest.mark.gpu


def run_gemm(a_type, tier, ta, tb, alpha, M, N, K, a_buf, wr_buf, lda, B, beta, C0):
    Bd = H.dev(np.asarray(B, np.float32))
    Cd = H.dev(np.asarray(C0, np.float32))
    C.call("tomos_gemm", a_type, tier, ta, tb, alpha, M, N, K, a_buf.ptr, wr_buf.ptr if wr_buf else None,
           lda, Bd.ptr, B.shape[0], beta, Cd.ptr, M, None)
    return Cd.to_numpy(np.float32, (M, N))


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("tier,dec", [(0, 1), (1, 1), (2, 0)])
def test_fp32_16_gemm_vs_oracle(capi, orc, ta, tb, tier, dec):
    M, N, K = 70, 45, 130
    rng = np.random.default_rng(ta * 2 + tb)
    Ash = (K, M) if ta else (M, K)
    A = rng.uniform(-8, 8, Ash).astype(np.float32, order="F")
    B = rng.uniform(-8, 8, (N, K) if tb else (K, N)).astype(np.float32, order="F")
    C0 = rng.uniform(-8, 8, (M, N)).astype(np.float32, order="F")
    wh, wr = H.pack(C.PACK_FP32_16, A)
    got = run_gemm(C.FP32_16, tier, ta, tb, 1.5, M, N, K, wh, wr, Ash[0], B, 0.7, C0)
    eh, er = orc.pack(0, A)
    exp = C0.copy(order="F")
    # DEFAULT runs on fp16 matrix cores: B is rounded to fp16 first
    Bo = np.asfortranarray(B.astype(np.float16).astype(np.float32)) if tier == 0 else B
    orc.lib().orc_mm_packed(dec, ta, tb, 1.5, M, N, K,
                            eh.ctypes.data_as(orc.u16p), er.ctypes.data_as(orc.u16p), Ash[0],
                            Bo.ctypes.data_as(orc.f32p), B.shape[0], 0.7,
                            exp.ctypes.data_as(orc.f32p), M)
    Ad = orc.unpack(dec, eh, er).astype(np.float64)
    Aop = Ad.T if ta else Ad
    Bop = B.T.astype(np.float64) if tb else B.astype(np.float64)
    scale = 1.5 * (np.abs(Aop) @ np.abs(Bop)) + 0.7 * np.abs(C0)
    rtol = 1e-5 if tier == 0 else 1e-6
    assert np.all(np.abs(got - exp) <= rtol * scale)
    ref = 1.5 * (A.T.astype(np.float64) if ta else A.astype(np.float64)) @ Bop + 0.7 * C0
    tol = {0: (0.15, 0.15), 1: (0.05, 0.05), 2: (1e-3, 1e-4)}[tier]  # test_fp32_16.cpp:253-257 style
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert rel < tol[0], rel
