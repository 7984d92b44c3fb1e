"""GPU GEMV parity against the CPU oracle tiers and an FP64 numpy reference.

Tolerances:
  * FAST / EXACT tiers and every F64 GEMV accumulate in FP64 over the same
    decoded values as the oracle: |gpu - oracle| <= 1e-12 * (|A||x|)_i + tiny.
  * DEFAULT tier (FP32 accumulation): |gpu - oracle_fp64acc| <= 2e-6 * (|A||x|)_i.
  * The reference's own tier tolerances vs the exact product
    (test/core/matrix/test_fp32_16.cpp:178-182, 1015-1016): pure_half
    0.15|ref|+0.15 elementwise on U[-8,8] data, L2-rel < 5e-2 at 140000 x 100;
    accurate 1e-3|ref|+1e-4 and L2-rel < 1e-3.
"""
import numpy as np
import pytest

from tests import hipcall as H
from tomos_amd import _capi as C

pytestmark = pytest.mark.gpu

FORMATS = {
    C.FP32_16: (C.PACK_FP32_16, {C.TIER_DEFAULT: 1, C.TIER_FAST: 1, C.TIER_EXACT: 0}),
    C.FP32_BF16: (C.PACK_FP32_BF16, {C.TIER_DEFAULT: 5, C.TIER_FAST: 5, C.TIER_EXACT: 4}),
}


def data(m, n, seed, wide=False):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n))
    if wide:
        A *= 2.0 ** rng.integers(-8, 9, (m, n))
    return np.asfortranarray(A.astype(np.float32))


def check_close(got, exp, scale, rtol):
    err = np.abs(got.astype(np.float64) - exp.astype(np.float64))
    bound = rtol * scale + 1e-30
    assert np.all(err <= bound), (np.max(err / bound), np.argmax(err / bound))


@pytest.mark.parametrize("a_type", [C.FP32_16, C.FP32_BF16])
@pytest.mark.parametrize("tier", [C.TIER_DEFAULT, C.TIER_FAST, C.TIER_EXACT])
@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("shape", [(3000, 70, 3000), (2999, 33, 2999), (1000, 257, 1008), (5, 5, 5)])
def test_packed_gemv_vs_oracle(capi, orc, a_type, tier, trans, shape):
    m, n, lda = shape
    print("gemv", a_type, tier, trans, shape, flush=True)
    A = data(m, n, seed=m + n + trans, wide=True)
    variant, decs = FORMATS[a_type]
    wh, wr = H.pack(variant, A, ld=lda)
    rng = np.random.default_rng(7)
    rows, cols = (n, m) if trans else (m, n)
    x = rng.standard_normal(cols).astype(np.float32)
    y0 = rng.standard_normal(rows).astype(np.float32)
    for alpha, beta in [(1.0, 0.0), (-0.75, 0.7)]:
        got = H.gemv(a_type, tier, trans, alpha, m, n, wh, wr, lda, x, beta, y0)
        ehh, err_ = orc.pack(variant, A)
        dec = decs[tier]
        exp = orc.mv_packed(dec, trans, alpha, ehh, err_, m, x, beta, y0)
        Adec = orc.unpack(dec, ehh, err_).astype(np.float64)
        scale = np.abs(alpha) * (np.abs(Adec.T if trans else Adec) @ np.abs(x.astype(np.float64)))
        scale += np.abs(beta) * np.abs(y0.astype(np.float64))
        rtol = 3e-5 if tier == C.TIER_DEFAULT else 1e-6  # y is stored in FP32
        check_close(got, exp, scale + 1e-30, rtol)


@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("xy", [(C.F32, C.F32), (C.F64, C.F32), (C.F32, C.F64), (C.F64, C.F64)])
def test_plain_gemv_types(capi, trans, xy):
    m, n = 4100, 130
    A = data(m, n, 11)
    x_type, y_type = xy
    rng = np.random.default_rng(5)
    rows, cols = (n, m) if trans else (m, n)
    x = rng.standard_normal(cols)
    y0 = rng.standard_normal(rows)
    Ad = H.dev(A)
    A64 = A.astype(np.float64)
    ref = 0.5 * ((A64.T if trans else A64) @ x.astype(np.float32 if x_type == C.F32 else np.float64)) + 2.0 * y0.astype(np.float32 if y_type == C.F32 else np.float64)
    scale = 0.5 * (np.abs(A64.T if trans else A64) @ np.abs(x)) + 2.0 * np.abs(y0)
    for tier, rtol in [(C.TIER_DEFAULT, 3e-5), (C.TIER_EXACT, 1e-12 if y_type == C.F64 else 1e-7)]:
        got = H.gemv(C.F32, tier, trans, 0.5, m, n, Ad, None, m, x, 2.0, y0, x_type, y_type)
        check_close(got, ref, scale, rtol)
    A2 = H.dev(A64)
    got = H.gemv(C.F64, C.TIER_DEFAULT, trans, 0.5, m, n, A2, None, m, x, 2.0, y0, x_type, y_type)
    check_close(got, ref, scale, 1e-12 if y_type == C.F64 else 1e-7)


def test_gemv_explicit_scratch_and_small_square(capi):
    # small square (R^-1 shape class): exercises the column-split path
    n = 1000
    rng = np.random.default_rng(1)
    R = np.triu(rng.standard_normal((n, n)))
    x = rng.standard_normal(n)
    Rd = H.dev(R)
    for trans in (0, 1):
        got = H.gemv(C.F64, C.TIER_DEFAULT, trans, 1.0, n, n, Rd, None, n, x, 0.0, None,
                     C.F64, C.F64, scratch="explicit")
        ref = (R.T if trans else R) @ x
        check_close(got, ref, np.abs(R.T if trans else R) @ np.abs(x), 1e-13)


def test_reference_tier_tolerances_small(capi, orc):
    """test_fp32_16.cpp:178-182 style: 5x5, U[-8,8], trans x beta grid."""
    rng = np.random.default_rng(101)
    for trial in range(3):
        A = rng.uniform(-8, 8, (5, 5)).astype(np.float32)
        x = rng.uniform(-8, 8, 5).astype(np.float32)
        y0 = rng.uniform(-8, 8, 5).astype(np.float32)
        wh, wr = H.pack(C.PACK_FP32_16, A)
        for trans in (0, 1):
            for beta in (0.0, 0.7):
                ref = (A.T.astype(np.float64) if trans else A.astype(np.float64)) @ x + beta * y0
                ph = H.gemv(C.FP32_16, C.TIER_DEFAULT, trans, 1.0, 5, 5, wh, wr, 5, x, beta, y0)
                fa = H.gemv(C.FP32_16, C.TIER_FAST, trans, 1.0, 5, 5, wh, wr, 5, x, beta, y0)
                ex = H.gemv(C.FP32_16, C.TIER_EXACT, trans, 1.0, 5, 5, wh, wr, 5, x, beta, y0)
                assert np.all(np.abs(ph - ref) <= 0.15 * np.abs(ref) + 0.15)
                assert np.all(np.abs(fa - ref) <= 0.05 * np.abs(ref) + 0.05)
                assert np.all(np.abs(ex - ref) <= 1e-3 * np.abs(ref) + 1e-4)


def test_reference_large_shape_l2(capi):
    """test_fp32_16.cpp:924-1017: 140000 x 100 and the ragged 140001 lead dim."""
    for m in (140000, 140001):
        n = 100
        rng = np.random.default_rng(m)
        A = rng.uniform(-8, 8, (m, n)).astype(np.float32, order="F")
        x = rng.uniform(-8, 8, n).astype(np.float32)
        u = rng.uniform(-8, 8, m).astype(np.float32)
        wh, wr = H.pack(C.PACK_FP32_16, A)
        A64 = A.astype(np.float64)
        for trans, vec in ((0, x), (1, u)):
            ref = (A64.T if trans else A64) @ vec.astype(np.float64)
            ph = H.gemv(C.FP32_16, C.TIER_DEFAULT, trans, 1.0, m, n, wh, wr, m, vec)
            ex = H.gemv(C.FP32_16, C.TIER_EXACT, trans, 1.0, m, n, wh, wr, m, vec)
            l2 = lambda g: np.linalg.norm(g - ref) / np.linalg.norm(ref)
            assert l2(ph) < 5e-2 and l2(ex) < 1e-3, (m, trans, l2(ph), l2(ex))


def bf16_bits_rne(A32):
    """Round-to-nearest-even FP32 -> bf16 bits (finite inputs)."""
    u = np.asarray(A32, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.astype(np.uint16)


@pytest.mark.parametrize("a_type", [C.F16, C.BF16])
@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("shape", [(3000, 70, 3000), (2999, 33, 2999), (1000, 257, 1008)])
def test_native_16bit_gemv(capi, a_type, trans, shape):
    """Native IEEE half / bf16 storage: the stored values times x, FP32
    accumulation (DEFAULT) or FP64 (FAST, EXACT), against FP64 numpy."""
    m, n, lda = shape
    A = data(m, n, 17 + m)
    if a_type == C.F16:
        bits = A.astype(np.float16).view(np.uint16)
        A_used = A.astype(np.float16).astype(np.float64)
    else:
        bits = bf16_bits_rne(A)
        A_used = (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    store = np.zeros((lda, n), np.uint16, order="F")
    store[:m] = bits
    Ad = H.dev(store)
    rng = np.random.default_rng(3)
    rows, cols = (n, m) if trans else (m, n)
    x = rng.standard_normal(cols)
    y0 = rng.standard_normal(rows)
    op = A_used.T if trans else A_used
    ref = 1.5 * (op @ x) - 0.5 * y0
    scale = 1.5 * (np.abs(op) @ np.abs(x)) + 0.5 * np.abs(y0)
    for tier, rtol in [(C.TIER_DEFAULT, 3e-6), (C.TIER_FAST, 1e-12), (C.TIER_EXACT, 1e-12)]:
        got = H.gemv(a_type, tier, trans, 1.5, m, n, Ad, None, lda, x, -0.5, y0, C.F64, C.F64)
        check_close(got, ref, scale, rtol)
