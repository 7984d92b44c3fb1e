// Bit codecs for the two packed FP32 storage formats, usable on host and
// device.  Formats (element = one FP32 value stored as two uint16 words):
//
// FP32_16 (src/core/matrix/fp32_16_pack.hpp:11-67):
//   wh = sign(1) | field(5) | mant_hi(9) | control(1)
//   wr = mant_lo(14) | reserved(1) | round_flag(1)
//   control=0: field = fp32 exponent - 112 (fp32 exponent in [113,142]);
//              wh is then a valid fp16 whose mantissa is mant_hi RNE'd.
//   control=1: field = dir(1) | shift(4).  dir=1 overflow (exp = 143+shift),
//              dir=0 underflow (exp = 112-shift), shift in [0,13]; shift=14
//              marks saturation (value decodes to +-Inf / +-0).
//   mant_hi is the top 9 of the 23 mantissa bits rounded to nearest-even;
//   round_flag records a round-up so the exact value can be rebuilt.
//
// FP32_BF16 (src/core/matrix/fp32_bf16_pack.hpp:9-45):
//   wh = top 16 bits of the fp32 pattern (a bf16), wr = low 16 bits;
//   the "rounded" variant rounds wh to nearest-even and keeps the round-up
//   flag in wr bit 0 (losing the last mantissa bit).
//
// The encode/decode results are bit-identical to the reference header
// functions named at each definition; the formulations are this library's
// own (mask-select form, no data-dependent branches, so a wavefront never
// diverges on the element class).
#pragma once

#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TOMOS_HD __host__ __device__ __forceinline__
#else
#define TOMOS_HD inline
#endif

namespace tomos_hip {
namespace codec {

TOMOS_HD uint32_t f2u(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_uint(v);
#else
    uint32_t u;
    std::memcpy(&u, &v, 4);
    return u;
#endif
}
TOMOS_HD float u2f(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(u);
#else
    float v;
    std::memcpy(&v, &u, 4);
    return v;
#endif
}

// Branch-free select: c ? a : b with c in {0,1}.
TOMOS_HD uint32_t sel(uint32_t c, uint32_t a, uint32_t b) {
    uint32_t m = 0u - c;
    return (a & m) | (b & ~m);
}

// ---------------------------------------------------------------- FP32_16
// Encoders.  mode 0 = full-carry (fp32_16_split, fp32_16_pack.hpp:129-190),
// 1 = clamped (fp32_16_split_clamped :200-248), 2 = truncated
// (fp32_16_split_truncated :263-299).
template <int MODE>
TOMOS_HD void fp32_16_encode(float value, uint16_t& wh, uint16_t& wr) {
    const uint32_t bits = f2u(value);
    const uint32_t sign = (bits >> 16) & 0x8000u;
    const uint32_t e = (bits >> 23) & 0xFFu;
    const uint32_t mant = bits & 0x7FFFFFu;
    const uint32_t hi = mant >> 14;
    const uint32_t lo = mant & 0x3FFFu;

    // Round-to-nearest-even decision for dropping the 14 low bits.
    uint32_t up = 0u;
    if (MODE != 2) {
        const uint32_t rbit = (lo >> 13) & 1u;
        const uint32_t sticky = (lo & 0x1FFFu) != 0u;
        up = rbit & (sticky | (hi & 1u));
    }
    uint32_t hi_w, rf, e_eff, in_window;
    if (MODE == 0) {
        // Eligible exponents [99,156]; a carry out of mant_hi bumps e.
        in_window = (e >= 99u) & (e <= 156u);
        const uint32_t hr = hi + up;
        const uint32_t carry = hr >> 9;
        hi_w = hr & 0x1FFu;
        rf = up;
        e_eff = e + carry;
    } else {
        // Clamped: a would-be carry keeps mant_hi = 0x1FF, no flag.
        in_window = 1u;
        const uint32_t blocked = up & (hi == 0x1FFu);
        hi_w = hi + (up & ~blocked);
        rf = up & ~blocked;
        e_eff = e;
    }
    const uint32_t dir = e_eff > 142u;
    const uint32_t normal = (e_eff >= 113u) & (e_eff <= 142u);
    const uint32_t off = dir ? (e_eff - 142u) : (113u - e_eff);  // >= 1 when !normal
    const uint32_t exact_ext = in_window & (off <= 14u);

    const uint32_t wh_norm = sign | ((e_eff - 112u) << 10) | (hi_w << 1);
    const uint32_t wh_ext = sign | (dir << 14) | (((off - 1u) & 0xFu) << 10) | (hi_w << 1) | 1u;
    // Saturation direction follows the original exponent, except when the
    // carry itself pushed past the window edge (e_eff decides then).
    const uint32_t sat_dir = (MODE == 0) ? sel(in_window, dir, e > 142u) : (e > 142u);
    const uint32_t wh_sat = sign | (sat_dir << 14) | (14u << 10) | 1u;
    const uint32_t wr_val = (lo << 2) | rf;

    const uint32_t is_norm = in_window & normal;
    const uint32_t is_ext = (1u - normal) & exact_ext;
    wh = static_cast<uint16_t>(sel(is_norm, wh_norm, sel(is_ext, wh_ext, wh_sat)));
    wr = static_cast<uint16_t>(sel(is_norm | is_ext, wr_val, 0u));
}

// Exact reconstruction: fp32_16_combine (fp32_16_pack.hpp:301-359); also
// bit-identical to fp32_16_combine_branchless_nohw (:440-482).
TOMOS_HD float fp32_16_decode_exact(uint32_t h, uint32_t r) {
    const uint32_t sign = (h & 0x8000u) << 16;
    const uint32_t ctl = h & 1u;
    const uint32_t hi = (h >> 1) & 0x1FFu;
    const uint32_t lo = (r >> 2) & 0x3FFFu;
    const uint32_t rf = r & 1u;
    // Undo the rounding: hi - rf, where a borrow out of 9 bits means the
    // encoder carried into the exponent.
    const uint32_t d = hi - rf;
    const uint32_t borrow = (d >> 9) & 1u;
    const uint32_t m32 = ((d & 0x1FFu) << 14) | lo;
    const uint32_t sh = (h >> 10) & 0xFu;
    const uint32_t dir = (h >> 14) & 1u;
    const uint32_t e_norm = ((h >> 10) & 0x1Fu) + 112u;
    const uint32_t e_ext = dir ? (143u + sh) : (112u - sh);
    const uint32_t e = sel(ctl, e_ext, e_norm) - borrow;
    const uint32_t exact = sign | (e << 23) | m32;
    const uint32_t sat = sign | (dir ? 0x7F800000u : 0u);
    return u2f(sel(ctl & (sh == 14u), sat, exact));
}

// fp32_16_combine_truncated (fp32_16_pack.hpp:498-522): no rounding undo.
TOMOS_HD float fp32_16_decode_exact_truncated(uint32_t h, uint32_t r) {
    const uint32_t sign = (h & 0x8000u) << 16;
    const uint32_t ctl = h & 1u;
    const uint32_t m32 = (((h >> 1) & 0x1FFu) << 14) | ((r >> 2) & 0x3FFFu);
    const uint32_t sh = (h >> 10) & 0xFu;
    const uint32_t dir = (h >> 14) & 1u;
    const uint32_t e = sel(ctl, dir ? (143u + sh) : (112u - sh), ((h >> 10) & 0x1Fu) + 112u);
    const uint32_t sat = sign | (dir ? 0x7F800000u : 0u);
    return u2f(sel(ctl & (sh == 14u), sat, sign | (e << 23) | m32));
}

// wh -> fp16 bits, fp32_16_half_only (fp32_16_pack.hpp:553-571):
// normal words pass through; overflow -> +-Inf; underflow -> the fp16
// subnormal (0x200 | mant_hi) >> shift (0 once shift >= 10).
TOMOS_HD uint32_t fp32_16_half_bits(uint32_t h) {
    const uint32_t sign = h & 0x8000u;
    const uint32_t ctl = h & 1u;
    const uint32_t dir = (h >> 14) & 1u;
    const uint32_t sh = (h >> 10) & 0xFu;
    const uint32_t sub = sign | ((0x200u | ((h >> 1) & 0x1FFu)) >> sh);
    const uint32_t special = sel(dir, sign | 0x7C00u, sub);
    return sel(ctl, special, h & 0xFFFFu);
}

// fp32_16_half_only_fast (fp32_16_pack.hpp:586-595): every control=1 word
// saturates (+-Inf overflow, +-0 underflow).
TOMOS_HD uint32_t fp32_16_half_bits_fast(uint32_t h) {
    const uint32_t sign = h & 0x8000u;
    const uint32_t dir = (h >> 14) & 1u;
    return sel(h & 1u, sign | (dir ? 0x7C00u : 0u), h & 0xFFFFu);
}

// Round-to-nearest-even fp32 -> bf16 bits (NaN stays NaN).
TOMOS_HD uint32_t f32_to_bf16_bits(float v) {
    uint32_t b = f2u(v);
    const uint32_t is_nan = ((b & 0x7F800000u) == 0x7F800000u) & ((b & 0x7FFFFFu) != 0u);
    const uint32_t rounded = (b + 0x7FFFu + ((b >> 16) & 1u)) >> 16;
    return sel(is_nan, (b >> 16) | 0x40u, rounded);
}

// IEEE fp16 bits -> fp32 (exact), integer-only so host and device agree.
TOMOS_HD float half_bits_to_f32(uint32_t h) {
    const uint32_t sign = (h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1Fu;
    const uint32_t m = h & 0x3FFu;
    if (e == 0x1Fu) return u2f(sign | 0x7F800000u | (m << 13));
    if (e == 0u) {
        // subnormal m * 2^-24: exact in fp32
        const float v = static_cast<float>(m) * 5.9604644775390625e-08f;
        return u2f(f2u(v) | sign);
    }
    return u2f(sign | ((e + 112u) << 23) | (m << 13));
}

// wh -> bf16 bits, fp32_16_bf16_only (fp32_16_pack.hpp:812-849).
TOMOS_HD uint32_t fp32_16_bf16_bits(uint32_t h) {
    const uint32_t s16 = h & 0x8000u;
    if (!(h & 1u)) return f32_to_bf16_bits(half_bits_to_f32(h));
    const uint32_t sh = (h >> 10) & 0xFu;
    const uint32_t dir = (h >> 14) & 1u;
    if (sh == 14u) return s16 | (dir ? 0x7F80u : 0u);
    const uint32_t hi = (h >> 1) & 0x1FFu;
    const uint32_t e = dir ? (143u + sh) : (112u - sh);
    // 9 -> 7 mantissa bits, nearest-even; a carry ripples into e.
    const uint32_t up = ((hi >> 1) & 1u) & ((hi & 1u) | ((hi >> 2) & 1u));
    return s16 + ((e << 7) | (hi >> 2)) + up;
}

// --------------------------------------------------------------- FP32_BF16
// fp32_bf16_split (fp32_bf16_pack.hpp:68-73) and split_rounded (:117-145).
template <bool ROUNDED>
TOMOS_HD void fp32_bf16_encode(float value, uint16_t& wh, uint16_t& wr) {
    const uint32_t b = f2u(value);
    if (!ROUNDED) {
        wh = static_cast<uint16_t>(b >> 16);
        wr = static_cast<uint16_t>(b & 0xFFFFu);
    } else {
        const uint32_t lo = b & 0xFFFFu;
        const uint32_t up = ((lo >> 15) & 1u) & (((lo & 0x7FFFu) != 0u) | ((b >> 16) & 1u));
        wh = static_cast<uint16_t>((b >> 16) + up);
        wr = static_cast<uint16_t>((lo & 0xFFFEu) | up);
    }
}
TOMOS_HD float fp32_bf16_decode_exact(uint32_t h, uint32_t r) { return u2f((h << 16) | r); }
TOMOS_HD float fp32_bf16_decode_hi(uint32_t h) { return u2f(h << 16); }
TOMOS_HD float fp32_bf16_decode_exact_rounded(uint32_t h, uint32_t r) {
    return u2f((((h - (r & 1u)) & 0xFFFFu) << 16) | (r & 0xFFFEu));
}

}  // namespace codec
}  // namespace tomos_hip
