"""Packed-A GEMM on matrix cores through tomos_gemm: tier DEFAULT (A's hi
word as the 16-bit tiers decode it times B rounded to nearest-even 16-bit,
FP32 accumulate) and tier MATERIALIZE (3-term split of the exact values).

DEFAULT is checked against the oracle's packed GEMM (orc_mm_packed, FP64
accumulation, mm_fp32_16_fast fp32_16_kernels.cpp:115-144 for the hi
decode) fed the same B rounded to the 16-bit grid -- numpy's float16 cast
for fp16, torch's bfloat16 cast for bf16 -- so the only difference left is
FP32 accumulation: |C - C_oracle| <= 1e-5 * (|alpha| |op(A)| |op(B)| + |beta C0|)
elementwise.  The unrounded FP64 product is held to the tier's own bound.
MATERIALIZE is checked against the FP64 GEMM of the exactly decoded A and
the FP32 B over the C3 exponent sweep (L2-relative bounds of the split:
bf16 3e-5, fp16 3e-6)."""
import numpy as np
import pytest
import torch

from tests import hipcall as H
from tomos_amd import _capi as C

pytestmark = pytest.mark.gpu

FORMATS = [  # a_type, pack variant, oracle hi decoder, oracle exact decoder
    (C.FP32_BF16, C.PACK_FP32_BF16, 5, 4),
    (C.FP32_16, C.PACK_FP32_16, 1, 0),
]


def round16(B, a_type):
    if a_type == C.FP32_16:
        return B.astype(np.float16).astype(np.float32)
    return torch.from_numpy(np.ascontiguousarray(B)).to(torch.bfloat16).to(torch.float32).numpy()


def gemm(a_type, tier, ta, tb, alpha, M, N, K, wh, wr, lda, B, ldb, beta, C0, ldc):
    Bd = H.dev(B)
    Cd = H.dev(C0)
    C.call("tomos_gemm", a_type, tier, ta, tb, alpha, M, N, K, wh.ptr, wr.ptr, lda, Bd.ptr, ldb, beta, Cd.ptr, ldc,
           None)
    return Cd.to_numpy(np.float32, (ldc, N))


def oracle_mm(orc, dec, ta, tb, alpha, M, N, K, eh, er, lda, B, ldb, beta, C0, ldc):
    out = np.array(C0, np.float32, order="F")
    B = np.asfortranarray(B, np.float32)
    orc.lib().orc_mm_packed(dec, ta, tb, alpha, M, N, K, eh.ctypes.data_as(orc.u16p), er.ctypes.data_as(orc.u16p),
                            lda, B.ctypes.data_as(orc.f32p), ldb, beta, out.ctypes.data_as(orc.f32p), ldc)
    return out


def operands(rng, M, N, K, ta, tb, pad, scale_exp=0):
    Ash = (K + pad, M) if ta else (M + pad, K)  # stored shape, leading dimension first
    Bsh = (N + pad, K) if tb else (K + pad, N)
    A = (rng.standard_normal(Ash) * 2.0 ** scale_exp).astype(np.float32, order="F")
    B = (rng.standard_normal(Bsh) * 2.0 ** scale_exp).astype(np.float32, order="F")
    C0 = rng.standard_normal((M + pad, N)).astype(np.float32, order="F")
    return A, B, C0


def views(A, B, M, N, K, ta, tb):
    Aop = (A[:K, :M].T if ta else A[:M, :K]).astype(np.float64)
    Bop = (B[:N, :K].T if tb else B[:K, :N]).astype(np.float64)
    return Aop, Bop


@pytest.mark.parametrize("a_type,variant,dec_hi,dec_ex", FORMATS)
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("shape,pad", [((300, 270, 130), 0), ((513, 257, 1000), 3), ((64, 8, 16), 5),
                                       ((1, 300, 7), 0)])
def test_default_tier_vs_oracle(capi, orc, a_type, variant, dec_hi, dec_ex, ta, tb, shape, pad):
    M, N, K = shape
    rng = np.random.default_rng(M * 7 + N + K + 2 * ta + tb)
    A, B, C0 = operands(rng, M, N, K, ta, tb, pad)
    lda, ldb, ldc = A.shape[0], B.shape[0], C0.shape[0]
    wh, wr = H.pack(variant, A)
    alpha, beta = 1.5, -0.75
    got = gemm(a_type, C.TIER_DEFAULT, ta, tb, alpha, M, N, K, wh, wr, lda, B, ldb, beta, C0, ldc)
    eh, er = orc.pack(variant, A)
    exp = oracle_mm(orc, dec_hi, ta, tb, alpha, M, N, K, eh, er, lda, round16(B, a_type), ldb, beta, C0, ldc)
    Ahi = orc.unpack(dec_hi, eh, er).astype(np.float64)
    Aop, Bop = views(Ahi, B, M, N, K, ta, tb)
    scale = abs(alpha) * (np.abs(Aop) @ np.abs(Bop)) + abs(beta) * np.abs(C0[:M])
    err = np.abs(got[:M] - exp[:M])
    assert np.all(err <= 1e-5 * scale + 1e-30), float(np.max(err / scale))
    # rows of C beyond M are untouched
    assert np.array_equal(got[M:], C0[M:])
    # the tier's own accuracy against the FP64 product of the FP32 inputs
    Af, Bf = views(A, B, M, N, K, ta, tb)
    ref = alpha * (Af @ Bf) + beta * C0[:M]
    rel = np.linalg.norm(got[:M] - ref) / np.linalg.norm(ref)
    assert rel < (2e-2 if a_type == C.FP32_BF16 else 2e-3), rel


@pytest.mark.parametrize("a_type,variant,dec_hi,dec_ex", FORMATS)
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
def test_default_tier_exactly_representable(capi, orc, a_type, variant, dec_hi, dec_ex, ta, tb):
    """Small-integer operands: every product and partial sum is exact in
    FP32, so the matrix-core result equals the oracle's bit for bit."""
    M, N, K = 260, 300, 96
    rng = np.random.default_rng(5)
    Ash = (K, M) if ta else (M, K)
    Bsh = (N, K) if tb else (K, N)
    A = rng.integers(-8, 9, Ash).astype(np.float32, order="F")
    B = rng.integers(-8, 9, Bsh).astype(np.float32, order="F")
    C0 = np.zeros((M, N), np.float32, order="F")
    wh, wr = H.pack(variant, A)
    got = gemm(a_type, C.TIER_DEFAULT, ta, tb, 1.0, M, N, K, wh, wr, Ash[0], B, Bsh[0], 0.0, C0, M)
    Aop, Bop = views(A, B, M, N, K, ta, tb)
    assert np.array_equal(got, (Aop @ Bop).astype(np.float32))


@pytest.mark.parametrize("a_type,variant,dec_hi,dec_ex", FORMATS)
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("s", [-24, -12, 0, 12, 24])
def test_materialize_tier_vs_fp64(capi, orc, a_type, variant, dec_hi, dec_ex, ta, tb, s):
    M, N, K = 300, 257, 400
    rng = np.random.default_rng(100 + s + 4 * ta + tb)
    A, B, C0 = operands(rng, M, N, K, ta, tb, 2, scale_exp=s)
    lda, ldb, ldc = A.shape[0], B.shape[0], C0.shape[0]
    wh, wr = H.pack(variant, A)
    got = gemm(a_type, C.TIER_MATERIALIZE, ta, tb, 1.0, M, N, K, wh, wr, lda, B, ldb, 0.0, C0, ldc)
    eh, er = orc.pack(variant, A)
    Aex = orc.unpack(dec_ex, eh, er).astype(np.float64)
    Aop, Bop = views(Aex, B, M, N, K, ta, tb)
    ref = Aop @ Bop
    rel = np.linalg.norm(got[:M] - ref) / np.linalg.norm(ref)
    print("materialize", a_type, ta, tb, s, "L2 rel err %.3e" % rel, flush=True)
    assert rel < (3e-5 if a_type == C.FP32_BF16 else 3e-6), rel
    assert np.array_equal(got[M:], C0[M:])


def test_materialize_alpha_beta_and_empty_k(capi):
    M, N, K = 200, 130, 70
    rng = np.random.default_rng(1)
    A = rng.standard_normal((M, K)).astype(np.float32, order="F")
    B = rng.standard_normal((K, N)).astype(np.float32, order="F")
    C0 = rng.standard_normal((M, N)).astype(np.float32, order="F")
    wh, wr = H.pack(C.PACK_FP32_BF16, A)
    got = gemm(C.FP32_BF16, C.TIER_MATERIALIZE, 0, 0, -0.5, M, N, K, wh, wr, M, B, K, 2.0, C0, M)
    ref = -0.5 * (A.astype(np.float64) @ B.astype(np.float64)) + 2.0 * C0
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 3e-5
    # K == 0: C = beta C
    got0 = gemm(C.FP32_BF16, C.TIER_DEFAULT, 0, 0, 1.0, M, N, 0, wh, wr, M, B, 1, 0.5, C0, M)
    assert np.allclose(got0, 0.5 * C0, rtol=1e-7, atol=0)


def test_tier_errors(capi):
    a = C.DeviceBuffer(64)
    with pytest.raises(C.TomosError):  # MATERIALIZE needs a packed A
        C.call("tomos_gemm", C.F32, C.TIER_MATERIALIZE, 0, 0, 1.0, 4, 4, 4, a.ptr, None, 4, a.ptr, 4, 0.0, a.ptr, 4,
               None)
    with pytest.raises(C.TomosError):  # MATERIALIZE needs the wr plane
        C.call("tomos_gemm", C.FP32_BF16, C.TIER_MATERIALIZE, 0, 0, 1.0, 4, 4, 4, a.ptr, None, 4, a.ptr, 4, 0.0,
               a.ptr, 4, None)
    with pytest.raises(C.TomosError):  # no such tier
        C.call("tomos_gemm", C.FP32_BF16, 7, 0, 0, 1.0, 4, 4, 4, a.ptr, a.ptr, 4, a.ptr, 4, 0.0, a.ptr, 4, None)
