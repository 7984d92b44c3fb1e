"""The packed GEMV configuration at its full size: A of 1,000,000 x 1000
FP32 with entries N(0,1) * 2^k, k uniform in [-8, 8] (seed 4), x, y iid
N(0,1) (seeds 5, 6), stored as FP32_16 (full reconstruction) and as
FP32_BF16 (hi-only).  Pack bits are compared with the oracle's encoder
over the whole matrix, y = A x and z = A^T y with the oracle's FP64
products of the same decoded values, and the error vs an FP64 GEMV of the
original FP32 matrix is bounded (and printed)."""
import numpy as np
import pytest

from tests import hipcall as H
from tomos_amd import _capi as C

pytestmark = pytest.mark.gpu

M, N = 1_000_000, 1000


@pytest.fixture(scope="module")
def c2():
    rng = np.random.default_rng(4)
    z = rng.standard_normal((N, M), dtype=np.float32).T
    k = rng.integers(-8, 9, (N, M), dtype=np.int8).T
    A = np.ldexp(z, k).astype(np.float32)
    del z, k
    x = np.random.default_rng(5).standard_normal(N).astype(np.float32)
    y = np.random.default_rng(6).standard_normal(M).astype(np.float32)
    return np.asfortranarray(A), x, y


@pytest.mark.parametrize("a_type,variant,tier,dec", [
    (C.FP32_16, C.PACK_FP32_16, C.TIER_EXACT, C.DECODE_FP32_16_EXACT),
    (C.FP32_BF16, C.PACK_FP32_BF16, C.TIER_DEFAULT, C.DECODE_FP32_BF16_HI),
])
def test_c2_full_size(capi, orc, c2, a_type, variant, tier, dec):
    A, x, y = c2
    wh, wr = H.pack(variant, A)
    eh, er = orc.pack(variant, A)
    assert np.array_equal(H.from_dev16(wh, M, N), eh), "wh plane differs from the CPU encoder"
    assert np.array_equal(H.from_dev16(wr, M, N), er), "wr plane differs from the CPU encoder"
    for trans, vec in ((0, x), (1, y)):
        got = H.gemv(a_type, tier, trans, 1.0, M, N, wh, wr, M, vec)
        ref = orc.mv_packed(dec, trans, 1.0, eh, er, M, vec)  # FP64 accumulation, FP32 store
        absA = np.abs(A)
        sc = (absA.T @ np.abs(vec).astype(np.float64)) if trans else (absA @ np.abs(vec).astype(np.float64))
        del absA
        # exact tier: both accumulate in FP64 and round once to FP32; hi-only
        # tier: FP32 accumulation over row/column chunks
        tol = 1e-6 if tier == C.TIER_EXACT else 2e-5
        rel = np.max(np.abs(got.astype(np.float64) - ref) / np.maximum(sc, 1e-300))
        assert rel <= tol, (trans, rel)
        full = (A.T @ vec.astype(np.float64)) if trans else (A @ vec.astype(np.float64))
        l2 = np.linalg.norm(got - full) / np.linalg.norm(full)
        print("C2 %s %s: L2 error vs FP64 GEMV of the FP32 matrix %.3e" %
              ("FP32_16 exact" if a_type == C.FP32_16 else "FP32_BF16 hi-only", "T" if trans else "N", l2), flush=True)
        assert l2 < (1e-6 if tier == C.TIER_EXACT else 1e-2)
