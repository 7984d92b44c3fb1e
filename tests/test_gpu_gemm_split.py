"""Split-precision MFMA GEMM (hi*hi + hi*lo + lo*hi, FP32 accumulate) vs an
FP64 GEMM of the exactly decoded packed operands, over the exponent sweep
s in {-24,-12,0,12,24} of the C3 configuration (smaller shapes here).

Error model: with hi = rn16(a), lo = rn16(a - hi) each operand keeps
~2x(8) bits (bf16) or ~2x(11) bits (fp16) of significand, and the dropped
lo*lo term is of the same order; L2-relative bounds: bf16 3e-5, fp16 3e-6
(plain FP32 SIMT GEMM shown beside for scale)."""
import numpy as np
import pytest

from tests import hipcall as H
from tomos_amd import _capi as C

pytestmark = pytest.mark.gpu


def run_split(a_type, A_planes, B_planes, M, N, K, lda, ldb, alpha=1.0, beta=0.0, C0=None):
    Cd = H.dev(np.zeros((M, N), np.float32) if C0 is None else C0)
    C.call("tomos_gemm_split", a_type, M, N, K, alpha, A_planes[0].ptr, A_planes[1].ptr, lda,
           B_planes[0].ptr, B_planes[1].ptr, ldb, beta, Cd.ptr, M, None, 0, None)
    return Cd.to_numpy(np.float32, (M, N))


@pytest.mark.parametrize("a_type,variant,dec,bound", [
    (C.FP32_BF16, C.PACK_FP32_BF16, 4, 3e-5),
    (C.FP32_16, C.PACK_FP32_16, 0, 3e-6),
])
@pytest.mark.parametrize("s", [-24, -12, 0, 12, 24])
@pytest.mark.parametrize("shape", [(512, 512, 512), (300, 200, 100), (129, 257, 1000)])
def test_split_gemm_accuracy(capi, orc, a_type, variant, dec, bound, s, shape):
    M, N, K = shape
    rng = np.random.default_rng(M + N + K + s)
    A = (rng.standard_normal((M, K)) * 2.0 ** s).astype(np.float32, order="F")
    B = (rng.standard_normal((K, N)) * 2.0 ** s).astype(np.float32, order="F")
    Ap, Bp = H.pack(variant, A), H.pack(variant, B)
    got = run_split(a_type, Ap, Bp, M, N, K, M, K)
    ea, ra = orc.pack(variant, A)
    eb, rb = orc.pack(variant, B)
    Ad = orc.unpack(dec, ea, ra).astype(np.float64)
    Bd = orc.unpack(dec, eb, rb).astype(np.float64)
    ref = Ad @ Bd
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("split gemm", a_type, s, shape, "L2 rel err %.3e" % err, flush=True)
    assert err < bound, err


def test_split_gemm_alpha_beta(capi):
    M, N, K = 200, 130, 70
    rng = np.random.default_rng(0)
    A = rng.standard_normal((M, K)).astype(np.float32, order="F")
    B = rng.standard_normal((K, N)).astype(np.float32, order="F")
    C0 = rng.standard_normal((M, N)).astype(np.float32, order="F")
    got = run_split(C.FP32_BF16, H.pack(C.PACK_FP32_BF16, A), H.pack(C.PACK_FP32_BF16, B), M, N, K,
                    M, K, alpha=-0.5, beta=2.0, C0=C0)
    ref = -0.5 * (A.astype(np.float64) @ B.astype(np.float64)) + 2.0 * C0
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 3e-5
