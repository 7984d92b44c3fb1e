"""GPU pack / decode: bit-exact against the CPU oracle encoders/decoders.

Inputs: a strided sweep of 2^26 fp32 bit patterns (every 64-pattern block of
the 2^32 space, every exponent, signs, zeros, subnormals, Inf, NaN) plus all
2^16 wh words for the wh-only decoders.
"""
import numpy as np
import pytest

from tests import hipcall as H
from tomos_amd import _capi as C

pytestmark = pytest.mark.gpu


def sweep_bits(n_log2=26, seed=0):
    i = np.arange(1 << n_log2, dtype=np.uint64)
    rng = np.random.default_rng(seed)
    low = rng.integers(0, 1 << (32 - n_log2), size=i.size, dtype=np.uint64)
    bits = ((i << np.uint64(32 - n_log2)) | low).astype(np.uint32)
    return bits.view(np.float32)


ENCODERS = [
    (C.PACK_FP32_16, 0), (C.PACK_FP32_16_CLAMPED, 1), (C.PACK_FP32_16_TRUNCATED, 2),
    (C.PACK_FP32_BF16, 3), (C.PACK_FP32_BF16_ROUNDED, 4),
]


@pytest.mark.parametrize("variant,orc_variant", ENCODERS)
def test_pack_bit_exact_sweep(capi, orc, variant, orc_variant):
    print("pack sweep variant", variant, flush=True)
    v = sweep_bits()
    m = 1 << 13
    n = v.size // m
    A = v.reshape(n, m).T  # (m, n) view
    wh, wr = H.pack(variant, A)
    gh = H.from_dev16(wh, m, n)
    gr = H.from_dev16(wr, m, n)
    eh, er = orc.pack(orc_variant, A)
    assert np.array_equal(gh, eh), np.argwhere(gh != eh)[:5]
    assert np.array_equal(gr, er), np.argwhere(gr != er)[:5]


@pytest.mark.parametrize("ld_pad", [0, 5, 8])
def test_pack_strided_and_ragged(capi, orc, ld_pad):
    rng = np.random.default_rng(3)
    m, n = 1001, 7
    A = (rng.standard_normal((m, n)) * 2.0 ** rng.integers(-30, 30, (m, n))).astype(np.float32)
    ld = m + ld_pad
    for variant in range(5):
        wh, wr = H.pack(variant, A, ld=ld)
        gh = H.from_dev16(wh, ld, n)[:m]
        gr = H.from_dev16(wr, ld, n)[:m]
        eh, er = orc.pack(variant, A)
        assert np.array_equal(gh, eh) and np.array_equal(gr, er), variant


DECODERS = [
    C.DECODE_FP32_16_EXACT, C.DECODE_FP32_16_HALF, C.DECODE_FP32_16_HALF_FAST,
    C.DECODE_FP32_16_BF16, C.DECODE_FP32_BF16_EXACT, C.DECODE_FP32_BF16_HI,
    C.DECODE_FP32_BF16_EXACT_ROUNDED, C.DECODE_FP32_16_EXACT_TRUNCATED,
]


@pytest.mark.parametrize("dec", DECODERS)
def test_unpack_bit_exact(capi, orc, dec):
    print("unpack decoder", dec, flush=True)
    v = sweep_bits(22, seed=dec)
    m = 1 << 11
    n = v.size // m
    A = v.reshape(n, m).T
    enc = {0: 0, 1: 0, 2: 0, 3: 0, 4: 3, 5: 3, 6: 4, 7: 2}[dec]
    eh, er = orc.pack(enc, A)
    dh, dr = H.dev(eh), H.dev(er)
    out = H.DeviceBuffer(m * n * 4)
    C.call("tomos_unpack", dec, m, n, dh.ptr, dr.ptr, m, out.ptr, m, None)
    got = out.to_numpy(np.float32, (m, n)).view(np.uint32)
    exp = orc.unpack(dec, eh, er).view(np.uint32)
    assert np.array_equal(got, exp), np.argwhere(got != exp)[:5]


@pytest.mark.parametrize("dec,fn", [
    (C.DECODE_FP32_16_HALF, "orc_fp32_16_half_only"),
    (C.DECODE_FP32_16_HALF_FAST, "orc_fp32_16_half_only_fast"),
    (C.DECODE_FP32_16_BF16, "orc_fp32_16_bf16_only"),
])
def test_decode_wh_all_words(capi, orc, dec, fn):
    words = np.arange(1 << 16, dtype=np.uint16)
    d = H.dev(words)
    out = H.DeviceBuffer(words.nbytes)
    C.call("tomos_decode_wh", dec, words.size, d.ptr, out.ptr, None)
    got = out.to_numpy(np.uint16, (words.size,))
    f = getattr(orc.lib(), fn)
    exp = np.array([f(int(w)) for w in words], dtype=np.uint16)
    if dec == C.DECODE_FP32_16_BF16:
        # control=0 words whose fp16 value is NaN (exponent field 31, never
        # produced by the encoders) may differ only in the NaN payload
        nan = ((words & 1) == 0) & (((words >> 10) & 0x1F) == 0x1F) & ((words & 0x3FF) != 0)
        assert np.array_equal(got[~nan], exp[~nan])
        assert np.all((got[nan] & 0x7F80) == 0x7F80) and np.all((got[nan] & 0x7F) != 0)
    else:
        assert np.array_equal(got, exp)
