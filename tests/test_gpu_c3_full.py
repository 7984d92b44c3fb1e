"""Split-precision GEMM at the full GEMM configuration: C = A B with
FP32_16 and FP32_BF16 operands, M=N=K=8192 and M=16384, N=4096, K=16384,
entries N(0,1) * 2^s, s in {-24, -12, 0, 12, 24}, FP32 accumulate.  The
L2 relative error of sampled columns of C against FP64 products of the
exactly decoded operands (the exact-tier GEMV, FP64 accumulation) stays
under the split's error model: bf16 hi+lo keeps ~16 significand bits
(3e-5), fp16 hi+lo ~22 (3e-6)."""
import numpy as np
import pytest

from tomos_amd import _capi as C
from tomos_amd._capi import DeviceBuffer

pytestmark = pytest.mark.gpu

FORMATS = [(C.FP32_BF16, C.PACK_FP32_BF16, C.DECODE_FP32_BF16_EXACT, 3e-5),
           (C.FP32_16, C.PACK_FP32_16, C.DECODE_FP32_16_EXACT, 3e-6)]


@pytest.mark.parametrize("M,N,K", [(8192, 8192, 8192), (16384, 4096, 16384)])
def test_split_gemm_full_size(capi, M, N, K):
    A = DeviceBuffer(M * K * 4)
    B = DeviceBuffer(K * N * 4)
    Cm = DeviceBuffer(M * N * 4)
    pa = [DeviceBuffer(M * K * 2) for _ in range(2)]
    pb = [DeviceBuffer(K * N * 2) for _ in range(2)]
    sb = C.lib().tomos_gemm_split_scratch_size(M, N, K)
    scr = DeviceBuffer(sb)
    cols = np.random.default_rng(M + N + K).choice(N, 4, replace=False)
    bcol, ref = DeviceBuffer(K * 4), DeviceBuffer(M * 8)
    for s in (-24, -12, 0, 12, 24):
        C.call("tomos_fill_normal_f32", A.ptr, M, K, M, 31 + s, None, s, s, None)
        C.call("tomos_fill_normal_f32", B.ptr, K, N, K, 77 + s, None, s, s, None)
        for a_type, variant, dec, bound in FORMATS:
            C.call("tomos_pack", variant, M, K, A.ptr, M, pa[0].ptr, pa[1].ptr, M, None)
            C.call("tomos_pack", variant, K, N, B.ptr, K, pb[0].ptr, pb[1].ptr, K, None)
            C.call("tomos_gemm_split", a_type, M, N, K, 1.0, pa[0].ptr, pa[1].ptr, M, pb[0].ptr, pb[1].ptr, K,
                   0.0, Cm.ptr, M, scr.ptr, sb, None)
            num = den = 0.0
            for j in cols:
                C.call("tomos_unpack", dec, K, 1, pb[0].ptr + int(j) * K * 2, pb[1].ptr + int(j) * K * 2, K,
                       bcol.ptr, K, None)
                C.call("tomos_gemv", a_type, C.TIER_EXACT, 0, 1.0, M, K, pa[0].ptr, pa[1].ptr, M, C.F32, bcol.ptr,
                       0.0, C.F64, ref.ptr, None, 0, None)
                r = ref.to_numpy(np.float64, (M,))
                g = np.empty(M, np.float32)
                C.call("tomos_memcpy_dtoh", g.ctypes.data, Cm.ptr + int(j) * M * 4, M * 4, None)
                num += float(np.sum((g.astype(np.float64) - r) ** 2))
                den += float(np.sum(r ** 2))
            err = np.sqrt(num / den)
            print("C3 %dx%dx%d %s s=%d: L2 rel err %.3e" % (M, N, K, "FP32_BF16" if a_type == C.FP32_BF16 else
                                                            "FP32_16", s, err), flush=True)
            assert err < bound, (a_type, s, err)
