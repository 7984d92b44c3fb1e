"""Pins the oracle's pack/decode restatement (oracle/tomos_oracle.c) to facts
that do not come from the oracle's own formulas:

* all 2^32 fp32 inputs: split->combine is exact on the normal/extended
  window, gives exactly +-Inf / +-0 on the saturate marker, NaN -> Inf with
  the NaN's sign, saturation only outside exponents [99,156];
* control=0 wh read as IEEE half == the value rounded to 9 mantissa bits
  (numpy rint, ties-to-even), for every exponent of the normal window;
* half_only: normal words pass through; overflow words -> +-Inf; underflow
  words -> the 10-bit significand truncated onto the fp16 subnormal grid
  (computed in float64 arithmetic);
* the IEEE half conversions the pure_half GEMV emulation uses == numpy
  float16 conversion; the pure_half GEMV == a numpy float16 loop;
* FP32_BF16 rounded split wh == torch's float->bfloat16 (RNE) conversion;
* fp32_16_bf16_only == torch bfloat16 rounding of the decoded value;
* the reference's alternative decoders agree with the primary ones.
"""
import ctypes

import numpy as np
import pytest

pytestmark = []


def _words():
    return np.arange(1 << 16, dtype=np.uint32)


@pytest.mark.parametrize("variant", [0, 1, 2], ids=["full_carry", "clamped", "truncated"])
def test_exhaustive_split_combine_contract(orc, variant):
    fb = ctypes.c_int64(0)
    bad = orc.lib().orc_exhaustive_split_combine(variant, 0, 1 << 32, ctypes.byref(fb))
    assert bad == 0, (bad, hex(fb.value))


def _split_many(orc, vals, variant=0):
    A = np.asarray(vals, np.float32).reshape(-1, 1)
    wh, wr = orc.pack(variant, A)
    return wh[:, 0].astype(np.uint32), wr[:, 0].astype(np.uint32)


def test_wh_is_value_rounded_to_9_bits(orc):
    rng = np.random.default_rng(7)
    n = 400000
    e = rng.integers(113, 143, n)  # the normal window (fp16 exponents)
    bits = (rng.integers(0, 2, n).astype(np.uint64) << 31) | (e.astype(np.uint64) << 23) | \
        rng.integers(0, 1 << 23, n).astype(np.uint64)
    v = bits.astype(np.uint32).view(np.float32)
    wh, _ = _split_many(orc, v)
    ctl0 = (wh & 1) == 0
    half = wh.astype(np.uint16).view(np.float16).astype(np.float64)
    x = v.astype(np.float64)
    ex = np.floor(np.log2(np.abs(x)))
    q = np.rint(x / 2.0 ** (ex - 9)) * 2.0 ** (ex - 9)  # 1 + 9 fraction bits, RNE
    fits = np.abs(q) < 65536.0  # rounding up past fp16's max leaves the window
    assert np.all(half[ctl0 & fits] == q[ctl0 & fits])
    assert ctl0[fits].mean() > 0.999


def test_half_only_math(orc):
    w = _words()
    L = orc.lib()
    got = np.array([L.orc_fp32_16_half_only(int(x)) for x in w], np.uint16)
    ctl = (w & 1) == 1
    direc = (w >> 14) & 1
    sign = np.where(w & 0x8000, -1.0, 1.0)
    gv = got.view(np.float16).astype(np.float64)
    # normal words pass through
    assert np.array_equal(got[~ctl], w[~ctl].astype(np.uint16))
    # overflow words: +-Inf
    ov = ctl & (direc == 1)
    assert np.all(np.isinf(gv[ov])) and np.all(np.sign(gv[ov]) == sign[ov])
    # underflow words: truncation of (1 + mant_hi/512) 2^(-15-sb) onto the 2^-24 grid
    un = ctl & (direc == 0)
    sb = ((w >> 10) & 0xF).astype(np.float64)
    mant = ((w >> 1) & 0x1FF).astype(np.float64)
    val = (1.0 + mant / 512.0) * 2.0 ** (-15.0 - sb)
    exp = sign * np.trunc(val / 2.0 ** -24) * 2.0 ** -24
    assert np.array_equal(gv[un], exp[un])


def test_half_only_vs_ieee_rounding(orc):
    """offset 1 (shift 0): half_only(split(v)) == fp16(v) exactly; deeper
    underflow double-rounds (RNE to 9 bits, then a truncating shift), so it
    can sit one fp16 subnormal step below IEEE rounding -- the header
    formula (fp32_16_pack.hpp:553-571) is the parity target."""
    rng = np.random.default_rng(11)
    L = orc.lib()
    for offset in range(1, 11):
        e32 = 113 - offset
        bits = (np.uint64(e32) << np.uint64(23)) | rng.integers(0, 1 << 23, 2000).astype(np.uint64)
        v = bits.astype(np.uint32).view(np.float32)
        wh, _ = _split_many(orc, v)
        got = np.array([L.orc_fp32_16_half_only(int(x)) for x in wh], np.uint16).view(np.float16)
        ieee = v.astype(np.float16)
        diff = got.astype(np.float64) - ieee.astype(np.float64)
        step = 2.0 ** -24
        if offset == 1:
            assert np.all(diff == 0)
        else:
            assert np.all((diff == 0) | (diff == -step)), offset
            assert np.any(diff != 0), offset


def test_decoder_variants_agree(orc):
    L = orc.lib()
    w = _words()
    ho = np.array([L.orc_fp32_16_half_only(int(x)) for x in w])
    for fn in ("orc_fp32_16_half_only_fast2", "orc_fp32_16_half_only_branchless"):
        f = getattr(L, fn)
        assert np.array_equal(ho, [f(int(x)) for x in w]), fn
    no_ovf = ~(((w & 1) == 1) & (((w >> 14) & 1) == 1))
    f = L.orc_fp32_16_half_only_no_overflow
    assert np.array_equal(ho[no_ovf], np.array([f(int(x)) for x in w[no_ovf]]))
    rng = np.random.default_rng(3)
    wr = rng.integers(0, 1 << 16, w.size)
    c = np.array([L.orc_fp32_16_combine(int(a), int(b)) for a, b in zip(w, wr)], np.float32)
    nh = np.array([L.orc_fp32_16_combine_branchless_nohw(int(a), int(b)) for a, b in zip(w, wr)],
                  np.float32)
    assert np.array_equal(c.view(np.uint32), nh.view(np.uint32))
    ctl0 = (w & 1) == 0
    for fn in ("orc_fp32_16_combine_fast_normal", "orc_fp32_16_combine_fast_normal_wrapborrow"):
        f = getattr(L, fn)
        g = np.array([f(int(a), int(b)) for a, b in zip(w[ctl0], wr[ctl0])], np.float32)
        assert np.array_equal(g.view(np.uint32), c[ctl0].view(np.uint32)), fn


def test_half_conversion_vs_numpy(orc):
    rng = np.random.default_rng(5)
    bits = rng.integers(0, 1 << 32, 200000, dtype=np.uint64).astype(np.uint32)
    # concentrate on the fp16 range and its subnormal/overflow edges
    e = rng.integers(95, 145, bits.size).astype(np.uint32)
    bits = (bits & np.uint32(0x807FFFFF)) | (e << np.uint32(23))
    v = bits.view(np.float32)
    L = orc.lib()
    got = np.array([L.orc_float_to_half(float(x)) for x in v], np.uint16)
    exp = v.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, exp)
    h = np.arange(1 << 16, dtype=np.uint32).astype(np.uint16)
    finite = np.isfinite(h.view(np.float16))
    back = np.array([L.orc_half_to_float(int(x)) for x in h[finite]], np.float32)
    assert np.array_equal(back, h[finite].view(np.float16).astype(np.float32))


def test_pure_half_gemv_vs_numpy_float16(orc):
    rng = np.random.default_rng(101)
    m, n = 9, 13
    A = rng.uniform(-8, 8, (m, n)).astype(np.float32)
    x = rng.uniform(-8, 8, n).astype(np.float32)
    wh, wr = orc.pack(0, A)
    L = orc.lib()
    Ah = np.array([[L.orc_fp32_16_half_only(int(wh[i, j])) for j in range(n)] for i in range(m)],
                  np.uint16).view(np.float16)
    for trans in (0, 1):
        got = orc.mv_pure_half(trans, 1.0, wh, m, x if not trans else rng.uniform(-8, 8, m).astype(np.float32))
    # direct emulation with numpy float16 arithmetic (each op rounds to fp16)
    xh = x.astype(np.float16)
    exp = np.zeros(m, np.float32)
    for i in range(m):
        acc = np.float16(0)
        for j in range(n):
            acc = np.float16(acc + np.float16(Ah[i, j] * xh[j]))
        exp[i] = np.float32(acc)
    got = orc.mv_pure_half(0, 1.0, wh, m, x)
    assert np.array_equal(got, exp)


def test_bf16_rounded_split_vs_torch(orc):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(9)
    bits = rng.integers(0, 1 << 32, 300000, dtype=np.uint64).astype(np.uint32)
    v = bits.view(np.float32)
    v = v[np.isfinite(v)]
    A = v.reshape(-1, 1)
    wh, wr = orc.pack(4, A)
    t = torch.from_numpy(v.copy()).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    ok = np.isfinite(torch.from_numpy(v.copy()).to(torch.bfloat16).float().numpy())
    assert np.array_equal(wh[:, 0][ok], t[ok])
    rec = orc.unpack(6, wh, wr)[:, 0]
    rb, vb = rec.view(np.uint32), v.view(np.uint32)
    assert np.all((rb & 1) == 0)
    assert np.array_equal(rb, vb & np.uint32(0xFFFFFFFE))
    wh0, wr0 = orc.pack(3, A)
    assert np.array_equal(wh0[:, 0], (vb >> 16).astype(np.uint16))
    assert np.array_equal(orc.unpack(4, wh0, wr0)[:, 0].view(np.uint32), vb)


def test_fp32_16_bf16_only_vs_torch(orc):
    torch = pytest.importorskip("torch")
    w = _words()
    L = orc.lib()
    got = np.array([L.orc_fp32_16_bf16_only(int(x)) for x in w], np.uint16)
    ctl = (w & 1) == 1
    sb = (w >> 10) & 0xF
    direc = (w >> 14) & 1
    sign = np.where(w & 0x8000, -1.0, 1.0)
    mant = ((w >> 1) & 0x1FF).astype(np.float64)
    e = np.where(direc == 1, 143 + sb, 112 - sb).astype(np.float64) - 127.0
    val = np.where(ctl, sign * (1.0 + mant / 512.0) * 2.0 ** e,
                   w.astype(np.uint16).view(np.float16).astype(np.float64))
    sat = ctl & (sb == 14)
    val = np.where(sat, np.where(direc == 1, sign * np.inf, sign * 0.0), val)
    fin = np.isfinite(val) | sat
    ref = torch.from_numpy(val.astype(np.float32)).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    keep = fin & ~np.isnan(val)
    assert np.array_equal(got[keep], ref[keep])


def _wh_value(w):
    """Value of FP32_16 hi words by the header's closed-form layout
    (fp32_16_pack.hpp:11-67): control bit 0 -> the word is fp16; control 1 ->
    sign | dir | shift_by | mant9 with exponent 143 + shift_by (dir 1) or
    112 - shift_by (dir 0), shift_by 14 saturating."""
    w = w.astype(np.int64)
    ctl = (w & 1) == 1
    sb = (w >> 10) & 0xF
    direc = (w >> 14) & 1
    sign = np.where(w & 0x8000, -1.0, 1.0)
    mant = ((w >> 1) & 0x1FF).astype(np.float64)
    e = np.where(direc == 1, 143 + sb, 112 - sb).astype(np.float64) - 127.0
    val = np.where(ctl, sign * (1.0 + mant / 512.0) * 2.0 ** e, w.astype(np.uint16).view(np.float16).astype(np.float64))
    sat = ctl & (sb == 14)
    return np.where(sat, np.where(direc == 1, sign * np.inf, sign * 0.0), val), ctl


@pytest.mark.parametrize("band", ["overflow", "underflow"])
def test_fp32_16_split_extended_bands_vs_rne9(orc, band):
    """The split encoder's hi word in the extended bands (fp32 exponents
    143..156 above the fp16 range, 99..112 below it): its value by the
    closed-form layout equals v rounded to nearest-even at 9 fraction bits
    (numpy, FP64), and fp32_16_bf16_only of it equals torch's bf16 rounding
    of that value.  Independent of the oracle's own combine."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(21 if band == "overflow" else 22)
    lo, hi = (143, 156) if band == "overflow" else (99, 112)
    e32 = rng.integers(lo, hi + 1, 200000).astype(np.uint32)
    mant = rng.integers(0, 1 << 23, 200000).astype(np.uint32)
    sgn = rng.integers(0, 2, 200000).astype(np.uint32)
    bits = (sgn << 31) | (e32 << 23) | mant
    v = bits.view(np.float32)
    wh, _ = orc.pack(0, v.reshape(-1, 1))
    w = wh[:, 0]
    got, ctl = _wh_value(w)
    ex = np.floor(np.log2(np.abs(v.astype(np.float64))))
    q = 2.0 ** (ex - 9)
    v9 = np.round(v.astype(np.float64) / q) * q  # half to even
    # a carry out of the 9 bits can leave the band (it saturates there)
    e9 = np.floor(np.log2(np.abs(v9)))
    inband = (e9 + 127 >= lo) & (e9 + 127 <= hi)
    assert np.all(ctl[inband]), "extended-band words carry the control bit"
    assert np.array_equal(got[inband], v9[inband])
    L = orc.lib()
    b16 = np.array([L.orc_fp32_16_bf16_only(int(x)) for x in w[inband][:20000]], np.uint16)
    ref = torch.from_numpy(v9[inband][:20000].astype(np.float32)).to(torch.bfloat16).view(torch.int16).numpy()
    assert np.array_equal(b16, ref.view(np.uint16))


def test_pack_compare_range_detects_mismatch(orc):
    """orc_pack_compare_range (the exhaustive GPU sweep's checker) counts
    every element that differs from the oracle's encoding: zero on the
    oracle's own output, and a single flipped bit is found with its pattern."""
    import ctypes
    start, count = 0x3F7FFF00, 4096
    bits = (np.arange(count, dtype=np.uint64) + start).astype(np.uint32)
    for variant in range(5):
        wh, wr = orc.pack(variant, bits.view(np.float32).reshape(-1, 1))
        wh, wr = np.ascontiguousarray(wh[:, 0]), np.ascontiguousarray(wr[:, 0])
        fb = ctypes.c_int64(0)
        L = orc.lib()
        assert L.orc_pack_compare_range(variant, start, count, wh.ctypes.data_as(orc.u16p),
                                        wr.ctypes.data_as(orc.u16p), ctypes.byref(fb)) == 0
        assert fb.value == -1
        wr[1234] ^= 4
        assert L.orc_pack_compare_range(variant, start, count, wh.ctypes.data_as(orc.u16p),
                                        wr.ctypes.data_as(orc.u16p), ctypes.byref(fb)) == 1
        assert fb.value == start + 1234
