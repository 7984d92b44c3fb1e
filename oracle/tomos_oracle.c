/*
 * tomos_oracle.c -- CPU ORACLE (test infrastructure only).
 *
 * A plain-C restatement of the reference TOMOS CPU algorithms on the LSQR
 * hot path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker or as the
 * timed CPU baseline -- never as part of the product path.
 *
 * Every function cites the reference file:line it follows (paths relative
 * to the reference checkout).  How each group is pinned independently of
 * this file (exhaustive enumeration, numpy/torch/scipy library routines,
 * closed forms, known-answer systems) is listed in tests/test_oracle_*.py.
 *
 * Parity status per group:
 *   pack / decode (FP32_16, FP32_BF16) ........ pinned (exhaustive + numpy/torch)
 *   CPU GEMV tiers (accurate/fast/pure_half) .. pinned (numpy fp64 / numpy fp16)
 *   hashed CountSketch ........................ pinned (pure-Python splitmix64)
 *   CSR products (A x, A^T x) ................. pinned (scipy.sparse, dense numpy)
 *   dense Gaussian sketch values .............. pinned (pure-Python SplitMix64
 *                                                stream, KS test against the
 *                                                normal CDF, moments, pair and
 *                                                neighbour correlations)
 *   Householder QR / triangular inverse ....... pinned (numpy.linalg.qr, R R^-1 = I)
 *   restarted LSQR ............................ pinned (known answers, numpy lstsq,
 *                                                scipy.sparse.linalg.lsqr)
 */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* bit helpers                                                               */
/* ------------------------------------------------------------------------ */
static inline uint32_t f2u(float v) { uint32_t u; memcpy(&u, &v, 4); return u; }
static inline float u2f(uint32_t u) { float v; memcpy(&v, &u, 4); return v; }

/* IEEE binary16 -> binary32 (exact). */
ORC_API float orc_half_to_float(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = ((uint32_t)h >> 10) & 0x1Fu;
    uint32_t mnt = (uint32_t)h & 0x3FFu;
    if (e == 0x1Fu) return u2f(sign | 0x7F800000u | (mnt << 13));
    if (e == 0) {
        if (mnt == 0) return u2f(sign);
        /* subnormal: mnt * 2^-24 */
        float v = (float)mnt * 5.9604644775390625e-08f;
        return (sign ? -v : v);
    }
    return u2f(sign | ((e + 112u) << 23) | (mnt << 13));
}

/* IEEE binary32 -> binary16, round to nearest even, overflow -> Inf. */
ORC_API uint16_t orc_float_to_half(float v) {
    uint32_t b = f2u(v);
    uint32_t sign = (b >> 16) & 0x8000u;
    uint32_t e = (b >> 23) & 0xFFu;
    uint32_t mnt = b & 0x7FFFFFu;
    if (e == 0xFFu) return (uint16_t)(sign | 0x7C00u | (mnt ? 0x200u : 0u));
    int32_t ue = (int32_t)e - 127;
    if (ue > 15) return (uint16_t)(sign | 0x7C00u);
    if (ue >= -14) {
        /* normal half: keep 10 bits of mantissa, RNE on the 13 dropped */
        uint32_t hm = mnt >> 13, rem = mnt & 0x1FFFu;
        uint32_t out = ((uint32_t)(ue + 15) << 10) | hm;
        if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) out += 1u; /* carry may bump exp (ok) */
        return (uint16_t)(sign | out);
    }
    /* subnormal half or zero: value = sig24 * 2^(ue-23); target grid 2^-24 */
    if (ue < -25) return (uint16_t)sign; /* below half of the smallest subnormal */
    uint32_t sig = (e ? (mnt | 0x800000u) : mnt);
    int shift = (e ? (-14 - ue) + 13 : 0); /* bits to drop to land on 2^-24 grid */
    if (!e) return (uint16_t)sign;         /* fp32 subnormals are far below 2^-25 */
    uint32_t q = sig >> shift;
    uint32_t rem = sig & ((1u << shift) - 1u);
    uint32_t halfway = 1u << (shift - 1);
    if (rem > halfway || (rem == halfway && (q & 1u))) q += 1u;
    return (uint16_t)(sign | q);
}

/* binary32 -> bfloat16, round to nearest even (NaN kept quiet). */
ORC_API uint16_t orc_float_to_bf16(float v) {
    uint32_t b = f2u(v);
    if ((b & 0x7F800000u) == 0x7F800000u && (b & 0x7FFFFFu))
        return (uint16_t)((b >> 16) | 0x40u);
    uint32_t lsb = (b >> 16) & 1u;
    b += 0x7FFFu + lsb;
    return (uint16_t)(b >> 16);
}

static inline float h2f(uint16_t h) { return orc_half_to_float(h); }

/* ------------------------------------------------------------------------ */
/* FP32_16 pack / decode  -- src/core/matrix/fp32_16_pack.hpp               */
/* ------------------------------------------------------------------------ */
enum {
    kSignMask = 0x80000000u, kExpMask = 0x7F800000u, kMantMask = 0x007FFFFFu,
    kExpMin = 113, kExpMax = 142, kRebias = 112, kOffSat = 15, kOffMax = 14,
    kOffSatStored = 14
};

/* fp32_16_pack.hpp:115-119 */
static inline int round_up_9(uint32_t mant_lo, int hi_odd) {
    int round_bit = (mant_lo >> 13) & 1u;
    int sticky = (mant_lo & 0x1FFFu) != 0u;
    return round_bit && (sticky || hi_odd);
}

/* fp32_16_split (full-carry encoder), fp32_16_pack.hpp:129-190 */
ORC_API void orc_fp32_16_split(float value, uint16_t* wh, uint16_t* wr) {
    uint32_t bits = f2u(value);
    uint32_t sign = (bits & kSignMask) >> 16;
    uint32_t e = (bits & kExpMask) >> 23;
    uint32_t mant = bits & kMantMask;
    if (e >= kExpMin - kOffMax && e <= kExpMax + kOffMax) {
        uint32_t hi_pre = mant >> 14, lo = mant & 0x3FFFu;
        int up = round_up_9(lo, hi_pre & 1u);
        uint32_t hi_r = hi_pre + (uint32_t)up;
        int ovf = hi_r > 0x1FFu;
        uint32_t e_eff = e + (uint32_t)ovf;
        uint32_t hi_w = ovf ? 0u : hi_r;
        uint32_t rf = (uint32_t)up;
        if (e_eff >= kExpMin && e_eff <= kExpMax) {
            *wh = (uint16_t)(sign | ((e_eff - kRebias) << 10) | (hi_w << 1));
            *wr = (uint16_t)((lo << 2) | rf);
            return;
        }
        uint32_t dir = e_eff > kExpMax;
        uint32_t off = dir ? (e_eff - kExpMax) : (kExpMin - e_eff);
        if (off >= 1u && off <= kOffMax) {
            *wh = (uint16_t)(sign | (dir << 14) | ((off - 1u) << 10) | (hi_w << 1) | 1u);
            *wr = (uint16_t)((lo << 2) | rf);
            return;
        }
        *wh = (uint16_t)(sign | (dir << 14) | (kOffSatStored << 10) | 1u);
        *wr = 0;
        return;
    }
    uint32_t dir = e > kExpMax;
    *wh = (uint16_t)(sign | (dir << 14) | (kOffSatStored << 10) | 1u);
    *wr = 0;
}

/* fp32_16_split_clamped, fp32_16_pack.hpp:200-248 */
ORC_API void orc_fp32_16_split_clamped(float value, uint16_t* wh, uint16_t* wr) {
    uint32_t bits = f2u(value);
    uint32_t sign = (bits & kSignMask) >> 16;
    uint32_t e = (bits & kExpMask) >> 23;
    uint32_t mant = bits & kMantMask;
    uint32_t hi_pre = mant >> 14, lo = mant & 0x3FFFu;
    int up = round_up_9(lo, hi_pre & 1u);
    int ovf = up && hi_pre == 0x1FFu;
    uint32_t hi_w = ovf ? hi_pre : hi_pre + (uint32_t)up;
    uint32_t rf = ovf ? 0u : (uint32_t)up;
    if (e >= kExpMin && e <= kExpMax) {
        *wh = (uint16_t)(sign | ((e - kRebias) << 10) | (hi_w << 1));
        *wr = (uint16_t)((lo << 2) | rf);
        return;
    }
    uint32_t dir = e > kExpMax;
    uint32_t off = dir ? (e - kExpMax) : (kExpMin - e);
    if (off <= kOffMax) {
        *wh = (uint16_t)(sign | (dir << 14) | ((off - 1u) << 10) | (hi_w << 1) | 1u);
        *wr = (uint16_t)((lo << 2) | rf);
        return;
    }
    *wh = (uint16_t)(sign | (dir << 14) | (kOffSatStored << 10) | 1u);
    *wr = 0;
}

/* fp32_16_split_truncated, fp32_16_pack.hpp:263-299 */
ORC_API void orc_fp32_16_split_truncated(float value, uint16_t* wh, uint16_t* wr) {
    uint32_t bits = f2u(value);
    uint32_t sign = (bits & kSignMask) >> 16;
    uint32_t e = (bits & kExpMask) >> 23;
    uint32_t mant = bits & kMantMask;
    if (e >= kExpMin && e <= kExpMax) {
        *wh = (uint16_t)(sign | ((e - kRebias) << 10) | ((mant >> 14) << 1));
        *wr = (uint16_t)((mant & 0x3FFFu) << 2);
        return;
    }
    uint32_t dir = e > kExpMax;
    uint32_t off = dir ? (e - kExpMax) : (kExpMin - e);
    if (off <= kOffMax) {
        *wh = (uint16_t)(sign | (dir << 14) | ((off - 1u) << 10) | ((mant >> 14) << 1) | 1u);
        *wr = (uint16_t)((mant & 0x3FFFu) << 2);
        return;
    }
    *wh = (uint16_t)(sign | (dir << 14) | (kOffSatStored << 10) | 1u);
    *wr = 0;
}

/* fp32_16_combine, fp32_16_pack.hpp:301-359 */
ORC_API float orc_fp32_16_combine(uint16_t wh, uint16_t wr) {
    uint32_t h = wh, r = wr;
    uint32_t sign = (h & 0x8000u) << 16;
    uint32_t control = h & 1u;
    uint32_t hi = (h >> 1) & 0x1FFu;
    uint32_t lo = (r >> 2) & 0x3FFFu;
    uint32_t rf = r & 1u;
    if (!control) {
        uint32_t e16 = (h >> 10) & 0x1Fu;
        int carry = rf && hi == 0u;
        uint32_t hi_pre = carry ? 0x1FFu : hi - rf;
        uint32_t e = e16 + kRebias - (uint32_t)carry;
        return u2f(sign | (e << 23) | (hi_pre << 14) | lo);
    }
    uint32_t sb = (h >> 10) & 0xFu, dir = (h >> 14) & 1u;
    if (sb != kOffSatStored) {
        int carry = rf && hi == 0u;
        uint32_t hi_pre = carry ? 0x1FFu : hi - rf;
        uint32_t e = (dir ? (kExpMax + 1u + sb) : (kExpMin - 1u - sb)) - (uint32_t)carry;
        return u2f(sign | (e << 23) | (hi_pre << 14) | lo);
    }
    return u2f(dir ? (sign | 0x7F800000u) : sign);
}

/* fp32_16_combine_fast_normal, fp32_16_pack.hpp:374-386 (control must be 0) */
ORC_API float orc_fp32_16_combine_fast_normal(uint16_t wh, uint16_t wr) {
    uint32_t h = wh, r = wr;
    uint32_t sign = (h & 0x8000u) << 16;
    uint32_t hi = (h >> 1) & 0x1FFu, lo = (r >> 2) & 0x3FFFu, rf = r & 1u;
    uint32_t e16 = (h >> 10) & 0x1Fu;
    int carry = rf && hi == 0u;
    uint32_t hi_pre = carry ? 0x1FFu : hi - rf;
    return u2f(sign | ((e16 + kRebias - (uint32_t)carry) << 23) | (hi_pre << 14) | lo);
}

/* fp32_16_combine_fast_normal_wrapborrow, fp32_16_pack.hpp:409-422 */
ORC_API float orc_fp32_16_combine_fast_normal_wrapborrow(uint16_t wh, uint16_t wr) {
    uint32_t h = wh, r = wr;
    uint32_t sign = (h & 0x8000u) << 16;
    uint32_t hi = (h >> 1) & 0x1FFu, lo = (r >> 2) & 0x3FFFu, rf = r & 1u;
    uint32_t e16 = (h >> 10) & 0x1Fu;
    uint32_t sub = hi - rf;
    uint32_t borrow = (sub >> 9) & 1u;
    return u2f(sign | ((e16 + kRebias - borrow) << 23) | ((sub & 0x1FFu) << 14) | lo);
}

/* fp32_16_combine_branchless_nohw, fp32_16_pack.hpp:440-482 */
ORC_API float orc_fp32_16_combine_branchless_nohw(uint16_t wh, uint16_t wr) {
    uint32_t h = wh, r = wr;
    uint32_t sign = (h & 0x8000u) << 16;
    uint32_t control = h & 1u;
    uint32_t hi = (h >> 1) & 0x1FFu, lo = (r >> 2) & 0x3FFFu, rf = r & 1u;
    uint32_t carry = (uint32_t)(rf != 0u && hi == 0u);
    uint32_t mc = 0u - carry;
    uint32_t hi_pre = (mc & 0x1FFu) | (~mc & (hi - rf));
    uint32_t m32 = (hi_pre << 14) | lo;
    uint32_t e16 = (h >> 10) & 0x1Fu;
    uint32_t b_norm = sign | (((e16 + kRebias) - carry) << 23) | m32;
    uint32_t sb = (h >> 10) & 0xFu, dir = (h >> 14) & 1u;
    uint32_t e_ext = dir ? (kExpMax + 1u + sb) : (kExpMin - 1u - sb);
    uint32_t b_ext = sign | ((e_ext - carry) << 23) | m32;
    uint32_t b_sat = dir ? (sign | 0x7F800000u) : sign;
    uint32_t ms = 0u - (uint32_t)(sb == kOffSatStored);
    uint32_t mctl = 0u - control;
    uint32_t b1 = (ms & b_sat) | (~ms & b_ext);
    return u2f((mctl & b1) | (~mctl & b_norm));
}

/* fp32_16_combine_truncated, fp32_16_pack.hpp:498-522 */
ORC_API float orc_fp32_16_combine_truncated(uint16_t wh, uint16_t wr) {
    uint32_t h = wh, r = wr;
    uint32_t sign = (h & 0x8000u) << 16;
    uint32_t hi = (h >> 1) & 0x1FFu, lo = (r >> 2) & 0x3FFFu;
    if (!(h & 1u)) return u2f(sign | ((((h >> 10) & 0x1Fu) + kRebias) << 23) | (hi << 14) | lo);
    uint32_t sb = (h >> 10) & 0xFu, dir = (h >> 14) & 1u;
    if (sb != kOffSatStored) {
        uint32_t e = dir ? (kExpMax + 1u + sb) : (kExpMin - 1u - sb);
        return u2f(sign | (e << 23) | (hi << 14) | lo);
    }
    return u2f(dir ? (sign | 0x7F800000u) : sign);
}

/* fp32_16_half_only, fp32_16_pack.hpp:553-571 (returns half bits) */
ORC_API uint16_t orc_fp32_16_half_only(uint16_t wh) {
    uint32_t h = wh, sign = h & 0x8000u;
    if (!(h & 1u)) return wh;
    if ((h >> 14) & 1u) return (uint16_t)(sign | 0x7C00u);
    uint32_t sb = (h >> 10) & 0xFu, hi = (h >> 1) & 0x1FFu;
    return (uint16_t)(sign | (((0x200u | hi) >> sb) & 0x3FFu));
}

/* fp32_16_half_only_fast, fp32_16_pack.hpp:586-595 */
ORC_API uint16_t orc_fp32_16_half_only_fast(uint16_t wh) {
    uint32_t h = wh, sign = h & 0x8000u;
    if (!(h & 1u)) return wh;
    return (uint16_t)(sign | (((h >> 14) & 1u) ? 0x7C00u : 0u));
}

/* fp32_16_half_only_fast2, fp32_16_pack.hpp:619-634 */
ORC_API uint16_t orc_fp32_16_half_only_fast2(uint16_t wh) {
    uint32_t h = wh, sign = h & 0x8000u;
    uint32_t control = h & 1u, dir = (h >> 14) & 1u;
    if (control & (1u - dir)) {
        uint32_t sb = (h >> 10) & 0xFu, hi = (h >> 1) & 0x1FFu;
        return (uint16_t)(sign | (((0x200u | hi) >> sb) & 0x3FFu));
    }
    return control ? (uint16_t)(sign | 0x7C00u) : wh;
}

/* fp32_16_half_only_branchless, fp32_16_pack.hpp:661-678 */
ORC_API uint16_t orc_fp32_16_half_only_branchless(uint16_t wh) {
    uint32_t h = wh, sign = h & 0x8000u;
    uint32_t control = h & 1u, dir = (h >> 14) & 1u;
    uint32_t und = control & (1u - dir);
    uint32_t sb = (h >> 10) & 0xFu, hi = (h >> 1) & 0x1FFu;
    uint32_t ub = sign | (((0x200u | hi) >> sb) & 0x3FFu);
    uint32_t cheap = control ? (sign | 0x7C00u) : h;
    uint32_t mk = 0u - und;
    return (uint16_t)((ub & mk) | (cheap & ~mk));
}

/* fp32_16_half_only_branchless_no_overflow, fp32_16_pack.hpp:705-718.
 * Caller contract: no overflow-direction words.  (_nomask :738-751 and
 * _shiftenc :761-774 are bit-identical; the mask is provably redundant.) */
ORC_API uint16_t orc_fp32_16_half_only_no_overflow(uint16_t wh) {
    uint32_t h = wh, sign = h & 0x8000u;
    uint32_t sb = (h >> 10) & 0xFu, hi = (h >> 1) & 0x1FFu;
    uint32_t ub = sign | (((0x200u | hi) >> sb) & 0x3FFu);
    uint32_t mk = 0u - (h & 1u);
    return (uint16_t)((ub & mk) | (h & ~mk));
}

/* fp32_16_bf16_only, fp32_16_pack.hpp:812-849 (returns bf16 bits) */
ORC_API uint16_t orc_fp32_16_bf16_only(uint16_t wh) {
    uint32_t h = wh, control = h & 1u, s16 = (h >> 15) & 1u;
    if (!control) return orc_float_to_bf16(h2f(wh));
    uint32_t sb = (h >> 10) & 0xFu, dir = (h >> 14) & 1u;
    if (sb == kOffSatStored) return (uint16_t)((s16 << 15) | (dir ? 0x7F80u : 0u));
    uint32_t e = dir ? (kExpMax + 1u + sb) : (kExpMin - 1u - sb);
    uint32_t hi = (h >> 1) & 0x1FFu;
    int rb = (hi & 2u) != 0, st = (hi & 1u) != 0, odd = ((hi >> 2) & 1u) != 0;
    uint32_t bm = (hi >> 2) + (uint32_t)(rb && (st || odd));
    if (bm > 0x7Fu) { bm = 0; e += 1u; }
    return (uint16_t)((s16 << 15) | (e << 7) | bm);
}

/* ------------------------------------------------------------------------ */
/* FP32_BF16 pack / decode -- src/core/matrix/fp32_bf16_pack.hpp            */
/* ------------------------------------------------------------------------ */
/* fp32_bf16_split, fp32_bf16_pack.hpp:68-73 */
ORC_API void orc_fp32_bf16_split(float v, uint16_t* wh, uint16_t* wr) {
    uint32_t b = f2u(v);
    *wh = (uint16_t)(b >> 16);
    *wr = (uint16_t)(b & 0xFFFFu);
}
/* fp32_bf16_combine, fp32_bf16_pack.hpp:78-81 */
ORC_API float orc_fp32_bf16_combine(uint16_t wh, uint16_t wr) {
    return u2f(((uint32_t)wh << 16) | wr);
}
/* fp32_bf16_split_rounded, fp32_bf16_pack.hpp:117-145 */
ORC_API void orc_fp32_bf16_split_rounded(float v, uint16_t* wh, uint16_t* wr) {
    uint32_t b = f2u(v);
    uint16_t wt = (uint16_t)(b >> 16);
    uint32_t lo = b & 0xFFFFu, hp = wt & 0x7Fu;
    int rb = (lo >> 15) & 1u, st = (lo & 0x7FFFu) != 0u;
    int up = rb && (st || (hp & 1u));
    *wh = (uint16_t)(wt + (uint16_t)up);
    *wr = (uint16_t)((lo & 0xFFFEu) | (uint32_t)up);
}
/* fp32_bf16_combine_rounded, fp32_bf16_pack.hpp:180-186 */
ORC_API float orc_fp32_bf16_combine_rounded(uint16_t wh, uint16_t wr) {
    uint32_t rf = wr & 1u;
    uint32_t pre = ((uint32_t)wh - rf) & 0xFFFFu;
    return u2f((pre << 16) | ((uint32_t)wr & 0xFFFEu));
}

/* ------------------------------------------------------------------------ */
/* array-level pack/unpack (planar, column-major)                           */
/* cpu::split_dense_fp32_16 / combine_dense_fp32_16                        */
/* src/kernels/cpu/matrix/fp32_16_kernels.cpp:8-26                          */
/* ------------------------------------------------------------------------ */
enum { ORC_SPLIT16 = 0, ORC_SPLIT16_CLAMPED = 1, ORC_SPLIT16_TRUNC = 2,
       ORC_SPLITBF = 3, ORC_SPLITBF_ROUNDED = 4 };

ORC_API void orc_pack(int variant, int64_t m, int64_t n, const float* src, int64_t ld_src,
                      uint16_t* wh, uint16_t* wr, int64_t ld_dst) {
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            float v = src[i + j * ld_src];
            uint16_t* h = &wh[i + j * ld_dst];
            uint16_t* r = &wr[i + j * ld_dst];
            switch (variant) {
                case ORC_SPLIT16: orc_fp32_16_split(v, h, r); break;
                case ORC_SPLIT16_CLAMPED: orc_fp32_16_split_clamped(v, h, r); break;
                case ORC_SPLIT16_TRUNC: orc_fp32_16_split_truncated(v, h, r); break;
                case ORC_SPLITBF: orc_fp32_bf16_split(v, h, r); break;
                default: orc_fp32_bf16_split_rounded(v, h, r); break;
            }
        }
}

/* Exhaustive encoder check: wh/wr hold the encoding (variant as orc_pack)  */
/* of the fp32 bit patterns start .. start + count - 1, in order.  Returns   */
/* the number of elements that differ from the oracle's encoding and, in    */
/* *first_bad, the smallest such bit pattern (-1 if none).  OpenMP.          */
ORC_API int64_t orc_pack_compare_range(int variant, uint64_t start, int64_t count, const uint16_t* wh,
                                       const uint16_t* wr, int64_t* first_bad) {
    int64_t bad = 0, fb = INT64_MAX;
#pragma omp parallel for reduction(+ : bad) reduction(min : fb) schedule(static)
    for (int64_t i = 0; i < count; ++i) {
        const uint32_t bits = (uint32_t)(start + (uint64_t)i);
        uint16_t h, r;
        orc_pack(variant, 1, 1, (const float*)&bits, 1, &h, &r, 1);
        if (h != wh[i] || r != wr[i]) {
            bad++;
            if ((int64_t)bits < fb) fb = (int64_t)bits;
        }
    }
    if (first_bad) *first_bad = bad ? fb : -1;
    return bad;
}

/* decoders selectable by id, for array decode / GEMV operand decode */
enum { ORC_DEC16_EXACT = 0, ORC_DEC16_HALF = 1, ORC_DEC16_HALF_FAST = 2,
       ORC_DEC16_BF16 = 3, ORC_DECBF_EXACT = 4, ORC_DECBF_HI = 5,
       ORC_DECBF_EXACT_ROUNDED = 6, ORC_DEC16_EXACT_TRUNC = 7 };

static inline float decode_elem(int dec, uint16_t h, uint16_t r) {
    switch (dec) {
        case ORC_DEC16_EXACT: return orc_fp32_16_combine(h, r);
        case ORC_DEC16_HALF: return h2f(orc_fp32_16_half_only(h));
        case ORC_DEC16_HALF_FAST: return h2f(orc_fp32_16_half_only_fast(h));
        case ORC_DEC16_BF16: return u2f((uint32_t)orc_fp32_16_bf16_only(h) << 16);
        case ORC_DECBF_EXACT: return orc_fp32_bf16_combine(h, r);
        case ORC_DECBF_HI: return u2f((uint32_t)h << 16);
        case ORC_DECBF_EXACT_ROUNDED: return orc_fp32_bf16_combine_rounded(h, r);
        default: return orc_fp32_16_combine_truncated(h, r);
    }
}

ORC_API void orc_unpack(int dec, int64_t m, int64_t n, const uint16_t* wh, const uint16_t* wr,
                        int64_t ld_src, float* dst, int64_t ld_dst) {
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i)
            dst[i + j * ld_dst] = decode_elem(dec, wh[i + j * ld_src],
                                              wr ? wr[i + j * ld_src] : 0);
}

/* ------------------------------------------------------------------------ */
/* CPU GEMV tiers -- src/kernels/cpu/matrix/fp32_16_kernels.cpp             */
/*   accurate  :28-55   (decode exact, double accumulate)                   */
/*   fast      :90-113  (half_only decode, double accumulate)               */
/*   pure_half :146-172 (half_only decode, __half multiply-then-add,         */
/*                       each op = float op then RNE to half)               */
/* y = alpha*op(A)*x + beta*y; y not read when beta == 0.                   */
/* ------------------------------------------------------------------------ */
ORC_API void orc_mv_packed(int dec, int trans, double alpha, int64_t m, int64_t n,
                           const uint16_t* wh, const uint16_t* wr, int64_t lda,
                           const float* x, double beta, float* y) {
    int64_t rows = trans ? n : m, cols = trans ? m : n;
    for (int64_t i = 0; i < rows; ++i) {
        double acc = 0.0;
        for (int64_t j = 0; j < cols; ++j) {
            int64_t idx = trans ? (j + i * lda) : (i + j * lda);
            acc += (double)decode_elem(dec, wh[idx], wr ? wr[idx] : 0) * (double)x[j];
        }
        y[i] = (beta != 0.0) ? (float)(alpha * acc + beta * (double)y[i]) : (float)(alpha * acc);
    }
}

ORC_API void orc_mv_fp32_16_pure_half(int trans, double alpha, int64_t m, int64_t n,
                                      const uint16_t* wh, int64_t lda, const float* x,
                                      double beta, float* y) {
    int64_t rows = trans ? n : m, cols = trans ? m : n;
    for (int64_t i = 0; i < rows; ++i) {
        uint16_t acc = 0;
        for (int64_t j = 0; j < cols; ++j) {
            int64_t idx = trans ? (j + i * lda) : (i + j * lda);
            uint16_t a = orc_fp32_16_half_only(wh[idx]);
            uint16_t xv = orc_float_to_half(x[j]);
            uint16_t prod = orc_float_to_half(h2f(a) * h2f(xv));
            acc = orc_float_to_half(h2f(acc) + h2f(prod));
        }
        double accd = (double)h2f(acc);
        y[i] = (beta != 0.0) ? (float)(alpha * accd + beta * (double)y[i]) : (float)(alpha * accd);
    }
}

/* packed GEMM, accurate/fast tiers: mm_fp32_16_accurate :57-88, fast :115-144.
 * C = alpha*op(A)*op(B) + beta*C, double accumulation. */
ORC_API void orc_mm_packed(int dec, int transA, int transB, double alpha,
                           int64_t M, int64_t N, int64_t K,
                           const uint16_t* wh, const uint16_t* wr, int64_t lda,
                           const float* B, int64_t ldb, double beta, float* C, int64_t ldc) {
    for (int64_t j = 0; j < N; ++j)
        for (int64_t i = 0; i < M; ++i) {
            double acc = 0.0;
            for (int64_t l = 0; l < K; ++l) {
                int64_t ia = transA ? (l + i * lda) : (i + l * lda);
                double a = decode_elem(dec, wh[ia], wr ? wr[ia] : 0);
                double b = transB ? B[j + l * ldb] : B[l + j * ldb];
                acc += a * b;
            }
            C[i + j * ldc] = (float)(alpha * acc + beta * (double)C[i + j * ldc]);
        }
}

/* ------------------------------------------------------------------------ */
/* hashed CountSketch -- src/kernels/cpu/sketch/countsketch_kernels.cpp:59-85 */
/* ------------------------------------------------------------------------ */
static const uint64_t CS_BUCKET_SEED = 0x12345678abcdef00ull;
static const uint64_t CS_SIGN_SEED = 0xf00dd00d42babe55ull;

ORC_API uint64_t orc_splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

ORC_API void orc_countsketch_indices(int64_t m, int64_t s, int32_t* bucket, int8_t* sign) {
    for (int64_t i = 0; i < m; ++i) {
        uint64_t hb = orc_splitmix64((uint64_t)i ^ CS_BUCKET_SEED);
        uint64_t hs = orc_splitmix64((uint64_t)i ^ CS_SIGN_SEED);
        bucket[i] = (int32_t)(hb % (uint64_t)s);
        sign[i] = (hs & 1ull) ? 1 : -1;
    }
}

/* Y[h(i), j] += s(i) * X[i, j]  (Y accumulated, caller zeroes it) */
ORC_API void orc_countsketch_apply_f64(int64_t m, int64_t n, const double* X, int64_t ldx,
                                       int64_t s, double* Y, int64_t ldy) {
    for (int64_t i = 0; i < m; ++i) {
        uint64_t hb = orc_splitmix64((uint64_t)i ^ CS_BUCKET_SEED);
        uint64_t hs = orc_splitmix64((uint64_t)i ^ CS_SIGN_SEED);
        int64_t b = (int64_t)(hb % (uint64_t)s);
        double sg = (hs & 1ull) ? 1.0 : -1.0;
        for (int64_t j = 0; j < n; ++j) Y[b + j * ldy] += sg * X[i + j * ldx];
    }
}

/* ------------------------------------------------------------------------ */
/* CSR y = alpha op(A) x + beta y -- cpu::mv_csr,                            */
/* src/kernels/cpu/matrix/matrix_cpu_kernels.cpp:92-131, computed in FP64    */
/* (the reference's compute_t for any FP64 operand); beta == 0 leaves y     */
/* unread.  int64 indices, FP64 values: callers convert.                    */
/* ------------------------------------------------------------------------ */
ORC_API void orc_csr_mv(int trans, double alpha, int64_t m, int64_t n, const double* vals,
                        const int64_t* row_ptrs, const int64_t* col_idxs, const double* x,
                        double beta, double* y) {
    if (!trans) {
        for (int64_t i = 0; i < m; ++i) {
            double acc = 0.0;
            for (int64_t k = row_ptrs[i]; k < row_ptrs[i + 1]; ++k) acc += vals[k] * x[col_idxs[k]];
            y[i] = beta != 0.0 ? alpha * acc + beta * y[i] : alpha * acc;
        }
    } else {
        for (int64_t j = 0; j < n; ++j) y[j] = beta != 0.0 ? beta * y[j] : 0.0;
        for (int64_t i = 0; i < m; ++i)
            for (int64_t k = row_ptrs[i]; k < row_ptrs[i + 1]; ++k)
                y[col_idxs[k]] += alpha * vals[k] * x[i];
    }
}

/* ------------------------------------------------------------------------ */
/* dense Gaussian sketch -- create_gaussian_sketch(Dense&),                  */
/* src/dispatch/gaussian_sketch_registries.cpp:16-30: i.i.d. N(0,1) entries. */
/* The reference draws them from a clock-seeded generator                    */
/* (src/kernels/cpu/sketch/gaussian_sketch_kernels.cpp:10-18), so only the   */
/* distribution is shared; this restates the library's seeded counter       */
/* layout in FP64: entries (2q, j), (2q+1, j) are the Box-Muller pair of    */
/* z = the p-th SplitMix64 output from `seed`, p = j*ceil(rows/2) + q,      */
/* u1 = ((z >> 41) + 1/2) 2^-23, u2 = ((z >> 8) & 0xFFFFFF) 2^-24.          */
/* ------------------------------------------------------------------------ */
ORC_API void orc_gaussian_fill_f64(int64_t rows, int64_t cols, double* out, int64_t ld, uint64_t seed) {
    const int64_t half = (rows + 1) / 2;
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t q = 0; q < half; ++q) {
            const uint64_t p = (uint64_t)j * (uint64_t)half + (uint64_t)q;
            /* SplitMix64: state_p = seed + (p+1) gamma; orc_splitmix64 adds
             * one gamma itself */
            const uint64_t z = orc_splitmix64(seed + p * 0x9e3779b97f4a7c15ull);
            const double u1 = ((double)(z >> 41) + 0.5) * 0x1p-23;
            const double u2 = (double)((z >> 8) & 0xFFFFFFull) * 0x1p-24;
            const double r = sqrt(-2.0 * log(u1));
            out[2 * q + j * ld] = r * cos(two_pi * u2);
            if (2 * q + 1 < rows) out[2 * q + 1 + j * ld] = r * sin(two_pi * u2);
        }
}

/* ------------------------------------------------------------------------ */
/* Householder QR + triangular inverse                                      */
/* src/core/precond/sketch_qr.cpp:38-64 (LAPACKE dgeqrf, zero reflectors,   */
/* dtrtri('U','N')).  Restated as textbook Householder (Golub & Van Loan    */
/* Alg. 5.2.1) and column-oriented back substitution.                       */
/* ------------------------------------------------------------------------ */
/* A (k x n, col-major, ld) overwritten: upper triangle = R; returns 0 or
 * -1 if k < n. */
ORC_API int orc_householder_qr(int64_t k, int64_t n, double* A, int64_t ld) {
    if (k < n) return -1;
    for (int64_t c = 0; c < n; ++c) {
        double* col = A + c * ld;
        double norm2 = 0.0;
        for (int64_t i = c; i < k; ++i) norm2 += col[i] * col[i];
        double nrm = sqrt(norm2);
        if (nrm == 0.0) continue;
        double x0 = col[c];
        double beta = (x0 >= 0.0) ? -nrm : nrm;   /* R_cc = beta */
        double v0 = x0 - beta;
        /* v = [1, col[c+1:]/v0], tau = (beta - x0)/beta */
        double tau = (beta - x0) / beta;
        for (int64_t j = c + 1; j < n; ++j) {
            double* cj = A + j * ld;
            double w = cj[c];
            for (int64_t i = c + 1; i < k; ++i) w += (col[i] / v0) * cj[i];
            w *= tau;
            cj[c] -= w;
            for (int64_t i = c + 1; i < k; ++i) cj[i] -= w * (col[i] / v0);
        }
        col[c] = beta;
        for (int64_t i = c + 1; i < k; ++i) col[i] = 0.0;
    }
    return 0;
}

/* In-place inverse of the n x n upper-triangular R (ld).  Returns 0, or
 * j+1 if R[j,j] == 0 (singular, as LAPACK's info). */
ORC_API int orc_trtri_upper(int64_t n, double* R, int64_t ld) {
    for (int64_t j = 0; j < n; ++j)
        if (R[j + j * ld] == 0.0) return (int)(j + 1);
    double* X = (double*)calloc((size_t)(n * n), sizeof(double));
    double* b = (double*)malloc((size_t)n * sizeof(double));
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = 0; i <= j; ++i) b[i] = (i == j) ? 1.0 : 0.0;
        for (int64_t i = j; i >= 0; --i) {
            double xi = b[i] / R[i + i * ld];
            X[i + j * n] = xi;
            for (int64_t r = 0; r < i; ++r) b[r] -= xi * R[r + i * ld];
        }
    }
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) R[i + j * ld] = (i <= j) ? X[i + j * n] : 0.0;
    free(X);
    free(b);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Restarted, right-preconditioned LSQR, FP64 throughout.                   */
/* src/core/solver/lsqr.cpp:417-747 (check_stopping_criteria :417-425,      */
/* initialize :431-513, step_1_orth_basis :557-626, step_2_update_sol       */
/* :632-679, fused_lsqr_impl :685-747), lsqr_registries.cpp:148-158.        */
/* Rinv (n x n, col-major) is the preconditioner P = R^-1 (nullable);       */
/* A (m x n, col-major, lda) is used for every matvec.                      */
/* ------------------------------------------------------------------------ */
static void dgemv_n(int64_t m, int64_t n, double alpha, const double* A, int64_t lda,
                    const double* x, double beta, double* y) {
    for (int64_t i = 0; i < m; ++i) y[i] = (beta == 0.0) ? 0.0 : beta * y[i];
    for (int64_t j = 0; j < n; ++j) {
        double s = alpha * x[j];
        const double* c = A + j * lda;
        for (int64_t i = 0; i < m; ++i) y[i] += s * c[i];
    }
}
static void dgemv_t(int64_t m, int64_t n, double alpha, const double* A, int64_t lda,
                    const double* x, double beta, double* y) {
    for (int64_t j = 0; j < n; ++j) {
        const double* c = A + j * lda;
        double acc = 0.0;
        for (int64_t i = 0; i < m; ++i) acc += c[i] * x[i];
        y[j] = alpha * acc + ((beta == 0.0) ? 0.0 : beta * y[j]);
    }
}
static double dnrm2(int64_t n, const double* x) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += x[i] * x[i];
    return sqrt(s);
}

ORC_API int orc_lsqr_f64(int64_t m, int64_t n, const double* A, int64_t lda,
                         const double* b, double* x, const double* Rinv,
                         double tol, int max_iters, double max_relres, int restarts,
                         int* iterations, double* history, int history_cap,
                         int* history_len) {
    double* u = (double*)calloc((size_t)m, sizeof(double));
    double* v = (double*)calloc((size_t)n, sizeof(double));
    double* w = (double*)calloc((size_t)n, sizeof(double));
    double* t = (double*)calloc((size_t)n, sizeof(double));
    double* pv = (double*)calloc((size_t)n, sizeof(double));
    int completed = 0, hl = 0;
    double alpha = 0, beta = 0, rho_bar = 0, phi_bar = 0;
    for (int64_t j = 0; j < n; ++j) x[j] = 0.0;
    double rhsnorm = dnrm2(m, b);
    for (int r = 0; r < restarts; ++r) {
        /* check_stopping_criteria at the top of each restart (:708) */
        if (completed > 0) {
            if (rhsnorm == 0.0) goto done;
            double rr = fabs(phi_bar) / rhsnorm;
            if (history && hl < history_cap) history[hl] = rr;
            hl++;
            if (rr < tol || rr < max_relres) goto done;
        }
        /* initialize (:431-513): u = b - A x; beta = |u|; u /= beta */
        for (int64_t i = 0; i < m; ++i) u[i] = b[i];
        dgemv_n(m, n, -1.0, A, lda, x, 1.0, u);
        beta = dnrm2(m, u);
        for (int64_t i = 0; i < m; ++i) u[i] *= 1.0 / beta;
        if (Rinv) {
            dgemv_t(m, n, 1.0, A, lda, u, 0.0, t);
            dgemv_t(n, n, 1.0, Rinv, n, t, 0.0, v);
        } else {
            dgemv_t(m, n, 1.0, A, lda, u, 0.0, v);
        }
        alpha = dnrm2(n, v);
        for (int64_t j = 0; j < n; ++j) v[j] *= 1.0 / alpha;
        for (int64_t j = 0; j < n; ++j) w[j] = v[j];
        rho_bar = alpha;
        phi_bar = beta;
        if (rhsnorm == 0.0) goto done;
        {
            double true_rr = beta / rhsnorm; /* :723-729 */
            if (true_rr < tol || true_rr < max_relres) goto done;
        }
        for (int it = 0; it < max_iters; ++it) {
            /* step_1_orth_basis */
            if (Rinv) {
                dgemv_n(n, n, 1.0, Rinv, n, v, 0.0, pv);
                dgemv_n(m, n, 1.0, A, lda, pv, -alpha, u);
                beta = dnrm2(m, u);
                for (int64_t i = 0; i < m; ++i) u[i] *= 1.0 / beta;
                dgemv_t(m, n, 1.0, A, lda, u, 0.0, t);
                dgemv_t(n, n, 1.0, Rinv, n, t, 0.0, pv);
                for (int64_t j = 0; j < n; ++j) pv[j] += -beta * v[j];
                alpha = dnrm2(n, pv);
                for (int64_t j = 0; j < n; ++j) pv[j] *= 1.0 / alpha;
                for (int64_t j = 0; j < n; ++j) v[j] = pv[j];
            } else {
                for (int64_t j = 0; j < n; ++j) t[j] = v[j];
                dgemv_n(m, n, 1.0, A, lda, t, -alpha, u);
                beta = dnrm2(m, u);
                for (int64_t i = 0; i < m; ++i) u[i] *= 1.0 / beta;
                dgemv_t(m, n, 1.0, A, lda, u, 0.0, t);
                for (int64_t j = 0; j < n; ++j) t[j] += -beta * v[j];
                alpha = dnrm2(n, t);
                for (int64_t j = 0; j < n; ++j) t[j] *= 1.0 / alpha;
                for (int64_t j = 0; j < n; ++j) v[j] = t[j];
            }
            /* step_2_update_sol */
            double rho = hypot(rho_bar, beta);
            double c = rho_bar / rho, s = beta / rho;
            double theta = s * alpha;
            rho_bar = -c * alpha;
            double phi = c * phi_bar;
            phi_bar = s * phi_bar;
            if (Rinv) {
                dgemv_n(n, n, 1.0, Rinv, n, w, 0.0, pv);
                for (int64_t j = 0; j < n; ++j) x[j] += (phi / rho) * pv[j];
            } else {
                for (int64_t j = 0; j < n; ++j) x[j] += (phi / rho) * w[j];
            }
            for (int64_t j = 0; j < n; ++j) w[j] = v[j] + (-theta / rho) * w[j];
            completed++;
            if (rhsnorm == 0.0) goto done;
            double rr = fabs(phi_bar) / rhsnorm;
            if (history && hl < history_cap) history[hl] = rr;
            hl++;
            if (rr < tol || rr < max_relres) goto done;
        }
    }
done:
    if (iterations) *iterations = completed;
    if (history_len) *history_len = hl;
    free(u); free(v); free(w); free(t); free(pv);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FP32 LSQR with OpenBLAS (the CPU baseline).                              */
/* Follows the lsqr.cpp control flow with value_t = float: vectors in float, */
/* BLAS-1 through cblas_snrm2/sscal/saxpy/scopy (lsqr.cpp:194-319), matvecs  */
/* through cblas_sgemv (src/kernels/cpu/matrix/matrix_cpu_kernels.cpp:      */
/* 44-54), scalars carried in double between steps; with Rinv (n x n, FP32) */
/* the right preconditioner is applied by sgemv on Rinv (precond_apply,     */
/* blas.cpp:88-96), as orc_lsqr_f64 does.  The OpenBLAS shared object is the */
/* one numpy bundles (ILP64 symbols with the scipy_ prefix); its path is     */
/* passed in.  Falls back to plain loops when no library is given.          */
/* ------------------------------------------------------------------------ */
typedef void (*sgemv_fn)(int, int, int64_t, int64_t, float, const float*, int64_t,
                         const float*, int64_t, float, float*, int64_t);
typedef float (*snrm2_fn)(int64_t, const float*, int64_t);
typedef void (*sscal_fn)(int64_t, float, float*, int64_t);
typedef void (*saxpy_fn)(int64_t, float, const float*, int64_t, float*, int64_t);
typedef void (*scopy_fn)(int64_t, const float*, int64_t, float*, int64_t);
typedef void (*set_threads_fn)(int);
static sgemv_fn p_sgemv = 0;
static snrm2_fn p_snrm2 = 0;
static sscal_fn p_sscal = 0;
static saxpy_fn p_saxpy = 0;
static scopy_fn p_scopy = 0;
static set_threads_fn p_set_threads = 0;

ORC_API int orc_blas_load(const char* path, int threads) {
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    p_sgemv = (sgemv_fn)dlsym(h, "scipy_cblas_sgemv64_");
    p_snrm2 = (snrm2_fn)dlsym(h, "scipy_cblas_snrm264_");
    p_sscal = (sscal_fn)dlsym(h, "scipy_cblas_sscal64_");
    p_saxpy = (saxpy_fn)dlsym(h, "scipy_cblas_saxpy64_");
    p_scopy = (scopy_fn)dlsym(h, "scipy_cblas_scopy64_");
    p_set_threads = (set_threads_fn)dlsym(h, "scipy_openblas_set_num_threads64_");
    if (!p_sgemv || !p_snrm2 || !p_sscal || !p_saxpy || !p_scopy) {
        p_sgemv = 0; p_snrm2 = 0; p_sscal = 0; p_saxpy = 0; p_scopy = 0;
        return -2;
    }
    if (p_set_threads && threads > 0) p_set_threads(threads);
    return 0;
}

/* 1 when the OpenBLAS routines are bound (orc_blas_load succeeded) */
ORC_API int orc_blas_loaded(void) { return p_sgemv != 0; }

static void sgemv(int trans, int64_t m, int64_t n, float alpha, const float* A, int64_t lda,
                  const float* x, float beta, float* y) {
    if (p_sgemv) { p_sgemv(102, trans ? 112 : 111, m, n, alpha, A, lda, x, 1, beta, y, 1); return; }
    if (!trans) {
        for (int64_t i = 0; i < m; ++i) y[i] = (beta == 0.0f) ? 0.0f : beta * y[i];
        for (int64_t j = 0; j < n; ++j) {
            float s = alpha * x[j];
            for (int64_t i = 0; i < m; ++i) y[i] += s * A[i + j * lda];
        }
    } else {
        for (int64_t j = 0; j < n; ++j) {
            float acc = 0.0f;
            for (int64_t i = 0; i < m; ++i) acc += A[i + j * lda] * x[i];
            y[j] = alpha * acc + ((beta == 0.0f) ? 0.0f : beta * y[j]);
        }
    }
}
static double snrm2(int64_t n, const float* x) {
    if (p_snrm2) return (double)p_snrm2(n, x, 1);
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += (double)x[i] * (double)x[i];
    return sqrt(s);
}
static void sscal(int64_t n, float a, float* x) {
    if (p_sscal) { p_sscal(n, a, x, 1); return; }
    for (int64_t i = 0; i < n; ++i) x[i] *= a;
}
static void saxpy(int64_t n, float a, const float* x, float* y) {
    if (p_saxpy) { p_saxpy(n, a, x, 1, y, 1); return; }
    for (int64_t i = 0; i < n; ++i) y[i] += a * x[i];
}
static void scopy(int64_t n, const float* x, float* y) {
    if (p_scopy) { p_scopy(n, x, 1, y, 1); return; }
    memcpy(y, x, (size_t)n * sizeof(float));
}

/* y = alpha op(A) x + beta y through the bound cblas_sgemv (the CPU      */
/* baseline's GEMV timing).                                                */
ORC_API void orc_sgemv(int trans, int64_t m, int64_t n, float alpha, const float* A, int64_t lda,
                       const float* x, float beta, float* y) {
    sgemv(trans, m, n, alpha, A, lda, x, beta, y);
}

ORC_API int orc_lsqr_f32(int64_t m, int64_t n, const float* A, int64_t lda,
                         const float* b, float* x, const float* Rinv, double tol, int max_iters,
                         double max_relres, int restarts, int* iterations,
                         double* history, int history_cap, int* history_len) {
    float* u = (float*)calloc((size_t)m, sizeof(float));
    float* v = (float*)calloc((size_t)n, sizeof(float));
    float* w = (float*)calloc((size_t)n, sizeof(float));
    float* t = (float*)calloc((size_t)n, sizeof(float));
    float* pv = (float*)calloc((size_t)n, sizeof(float));
    int completed = 0, hl = 0;
    double alpha = 0, beta = 0, rho_bar = 0, phi_bar = 0;
    for (int64_t j = 0; j < n; ++j) x[j] = 0.0f;
    double rhsnorm = snrm2(m, b);
    for (int r = 0; r < restarts; ++r) {
        if (completed > 0) {
            if (rhsnorm == 0.0) goto done;
            double rr = fabs(phi_bar) / rhsnorm;
            if (history && hl < history_cap) history[hl] = rr;
            hl++;
            if (rr < tol || rr < max_relres) goto done;
        }
        scopy(m, b, u);
        sgemv(0, m, n, -1.0f, A, lda, x, 1.0f, u);
        beta = snrm2(m, u);
        sscal(m, (float)(1.0 / beta), u);
        if (Rinv) {
            sgemv(1, m, n, 1.0f, A, lda, u, 0.0f, t);
            sgemv(1, n, n, 1.0f, Rinv, n, t, 0.0f, v);
        } else {
            sgemv(1, m, n, 1.0f, A, lda, u, 0.0f, v);
        }
        alpha = snrm2(n, v);
        sscal(n, (float)(1.0 / alpha), v);
        scopy(n, v, w);
        rho_bar = alpha;
        phi_bar = beta;
        if (rhsnorm == 0.0) goto done;
        if (beta / rhsnorm < tol || beta / rhsnorm < max_relres) goto done;
        for (int it = 0; it < max_iters; ++it) {
            if (Rinv) sgemv(0, n, n, 1.0f, Rinv, n, v, 0.0f, t);
            else scopy(n, v, t);
            sgemv(0, m, n, 1.0f, A, lda, t, (float)(-alpha), u);
            beta = snrm2(m, u);
            sscal(m, (float)(1.0 / beta), u);
            if (Rinv) {
                sgemv(1, m, n, 1.0f, A, lda, u, 0.0f, pv);
                sgemv(1, n, n, 1.0f, Rinv, n, pv, 0.0f, t);
            } else {
                sgemv(1, m, n, 1.0f, A, lda, u, 0.0f, t);
            }
            saxpy(n, (float)(-beta), v, t);
            alpha = snrm2(n, t);
            sscal(n, (float)(1.0 / alpha), t);
            scopy(n, t, v);
            double rho = hypot(rho_bar, beta);
            double c = rho_bar / rho, s = beta / rho;
            double theta = s * alpha;
            rho_bar = -c * alpha;
            double phi = c * phi_bar;
            phi_bar = s * phi_bar;
            if (Rinv) {
                sgemv(0, n, n, 1.0f, Rinv, n, w, 0.0f, pv);
                saxpy(n, (float)(phi / rho), pv, x);
            } else {
                saxpy(n, (float)(phi / rho), w, x);
            }
            /* w = v - (theta / rho) w */
            sscal(n, (float)(-theta / rho), w);
            saxpy(n, 1.0f, v, w);
            completed++;
            if (rhsnorm == 0.0) goto done;
            double rr = fabs(phi_bar) / rhsnorm;
            if (history && hl < history_cap) history[hl] = rr;
            hl++;
            if (rr < tol || rr < max_relres) goto done;
        }
    }
done:
    if (iterations) *iterations = completed;
    if (history_len) *history_len = hl;
    free(u); free(v); free(w); free(t); free(pv);
    return 0;
}

/* Hashed CountSketch of an FP32 matrix on the host cores (the CPU baseline's */
/* preconditioner build): Y[h(i), j] += s(i) * X[i, j] in FP64, the same     */
/* hash and sign as orc_countsketch_apply_f64 (countsketch_kernels.cpp:      */
/* 59-85); columns are independent, so they are split over OpenMP threads.  */
ORC_API void orc_countsketch_apply_f32_omp(int64_t m, int64_t n, const float* X, int64_t ldx,
                                           int64_t s, double* Y, int64_t ldy) {
    int32_t* bucket = (int32_t*)malloc((size_t)(m > 0 ? m : 1) * sizeof(int32_t));
    int8_t* sign = (int8_t*)malloc((size_t)(m > 0 ? m : 1));
    orc_countsketch_indices(m, s, bucket, sign);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t j = 0; j < n; ++j) {
        const float* xc = X + j * ldx;
        double* yc = Y + j * ldy;
        for (int64_t i = 0; i < m; ++i) yc[bucket[i]] += (double)sign[i] * (double)xc[i];
    }
    free(bucket);
    free(sign);
}

/* Host memory bandwidth probe (STREAM "copy"): dst[i] = src[i] over OpenMP */
/* threads; the caller times it and counts 2 * 8 * n bytes.                 */
ORC_API void orc_stream_copy(int64_t n, const double* src, double* dst) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) dst[i] = src[i];
}

/* ------------------------------------------------------------------------ */
/* Exhaustive checks over all 2^32 fp32 bit patterns (run by the CPU tests). */
/* Returns the number of violations of the split->combine contract:          */
/*  exact where wh is normal/extended, exactly +-Inf / +-0 on the saturate   */
/*  marker, NaN -> +Inf; and the saturate marker only where the exponent     */
/*  really is outside [99,156] (or pushed outside by the mantissa carry).    */
/* ------------------------------------------------------------------------ */
ORC_API int64_t orc_exhaustive_split_combine(int variant, uint64_t start, uint64_t stop,
                                             int64_t* first_bad) {
    int64_t bad = 0;
    int64_t fb = -1;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t bi = (int64_t)start; bi < (int64_t)stop; ++bi) {
        uint32_t bits = (uint32_t)bi;
        float v = u2f(bits);
        uint16_t h, r;
        if (variant == ORC_SPLIT16) orc_fp32_16_split(v, &h, &r);
        else if (variant == ORC_SPLIT16_CLAMPED) orc_fp32_16_split_clamped(v, &h, &r);
        else orc_fp32_16_split_truncated(v, &h, &r);
        float rec = orc_fp32_16_combine(h, r);
        float nohw = orc_fp32_16_combine_branchless_nohw(h, r);
        uint32_t e = (bits >> 23) & 0xFFu;
        int sat = (h & 1u) && (((h >> 10) & 0xFu) == 14u);
        int ok = f2u(nohw) == f2u(rec);
        if (variant == ORC_SPLIT16_TRUNC) {
            float rt = orc_fp32_16_combine_truncated(h, r);
            ok = ok && f2u(rt) == f2u(rec);
        }
        if (isnan(v)) {
            /* NaN saturates to Inf carrying the NaN's sign bit */
            ok = ok && f2u(rec) == ((bits & 0x80000000u) | 0x7F800000u);
        } else if (!sat) {
            ok = ok && f2u(rec) == bits && e >= 99u && e <= 156u;
        } else {
            uint32_t dir = (h >> 14) & 1u;
            uint32_t expect = (bits & 0x80000000u) | (dir ? 0x7F800000u : 0u);
            ok = ok && f2u(rec) == expect;
            /* saturation only outside the exact window (or full carry at its edge) */
            int outside = e < 99u || e > 156u;
            int carry_edge = (variant == ORC_SPLIT16) && e == 156u && ((bits >> 14) & 0x1FFu) == 0x1FFu;
            ok = ok && (outside || carry_edge);
            ok = ok && (dir == (e > 142u));
        }
        if (!ok) {
            bad += 1;
#pragma omp critical
            if (fb < 0 || bi < fb) fb = bi;
        }
    }
    if (first_bad) *first_bad = fb;
    return bad;
}
