// Packed GEMMs on CDNA4 matrix cores (gfx950), FP32 accumulate:
//   hi-only   C = A_hi B_hi                      (tomos_gemm, tier DEFAULT)
//   split     C ~= A_hi B_hi + A_hi B_lo + A_lo B_hi
//             (tomos_gemm_split with both operands packed; tomos_gemm, tier
//             MATERIALIZE, with FP32 B).
// The hi-only form is the reference's 16-bit tensor-core GEMM on wh
// (mm_fp32_bf16_pure_bf16_gemm_direct_fastk_precomp2_serp,
// mm_cuda_kernels.cu:35372; mm_fp32_16_pure_half_wide_bigreg_sq_*, :37826);
// the split has no reference counterpart.
//
// Two steps:
//  1. stage: each operand is written as 16-bit planes in a K-contiguous
//     layout (A as [M][K], B as [N][K]) so every MFMA operand is one 16-byte
//     load.  Hi-only: A's hi word as the 16-bit tiers decode it (FP32_BF16:
//     bf16 as stored; FP32_16: fp32_16_half_only, fp16) and FP32 B rounded
//     to nearest-even bf16 / fp16.  Split: the exact value a, hi = rn16(a),
//     lo = rn16(a - hi); FP32_16 uses fp16 planes with a power-of-two scale
//     2^s per operand, chosen from the operand's largest magnitude, so the hi
//     and lo terms stay clear of fp16 overflow and underflow (C is rescaled
//     by 2^-(sA+sB) in the epilogue); FP32_BF16 uses bf16 planes.
//  2. gemm: 512 threads, 256 x 256 C tile, K slice of 32 staged through LDS
//     (rows padded against bank conflicts); each wave owns 128 x 64 = 4 x 2
//     MFMA 32x32x16 tiles and issues 1 (hi-only) or 3 (split) MFMAs per
//     tile per 16-wide k step into one FP32 accumulator.
#include <algorithm>

#include "codec.hpp"
#include "common.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 256, BN = 256, BK = 32;
constexpr int LDK = BK + 8;  // padded K stride (elements) of an LDS row
constexpr int kGemmThreads = 512;                     // 8 waves: 2 (M) x 4 (N)
constexpr int kPlaneElems = BM * LDK;                 // one plane's K slice in LDS (BM == BN)

// ---------------------------------------------------------------- operands
// A GEMM operand as the prep kernels read it: X column-major with leading
// dimension ld, either packed planes (wh/wr of FP32_16 or FP32_BF16) or a
// plain FP32 array (f != nullptr).
struct Src {
    int fmt;                 // TOMOS_FP32_16 | TOMOS_FP32_BF16 (packed); ignored for FP32
    const uint16_t* wh;
    const uint16_t* wr;
    const float* f;
    int64_t ld;
};

__device__ __forceinline__ float decode_exact(int fmt, uint32_t h, uint32_t r) {
    return fmt == TOMOS_FP32_16 ? codec::fp32_16_decode_exact(h, r) : __uint_as_float((h << 16) | r);
}

// exact value of element idx (split mode)
__device__ __forceinline__ float src_value(const Src& s, int64_t idx) {
    return s.f ? s.f[idx] : decode_exact(s.fmt, s.wh[idx], s.wr[idx]);
}

__device__ __forceinline__ uint16_t bf16_rn(float v) {
    return static_cast<uint16_t>(codec::f32_to_bf16_bits(v));
}
__device__ __forceinline__ uint16_t f16_rn(float v) {
    return __builtin_bit_cast(uint16_t, static_cast<_Float16>(v));
}
__device__ __forceinline__ float bf16_val(uint16_t b) { return __uint_as_float(static_cast<uint32_t>(b) << 16); }

// 16-bit MFMA operand bits of element idx (hi-only mode): the packed hi word
// decoded as the reference's 16-bit tiers read it (FP32_BF16: wh is the
// bf16 value, fp32_bf16_bf16_only; FP32_16: fp32_16_half_only), or an FP32
// value rounded to nearest-even bf16 / fp16.
__device__ __forceinline__ uint16_t src_hi_bits(const Src& s, int64_t idx, bool f16) {
    if (s.f) return f16 ? f16_rn(s.f[idx]) : bf16_rn(s.f[idx]);
    const uint32_t h = s.wh[idx];
    return static_cast<uint16_t>(s.fmt == TOMOS_FP32_16 ? codec::fp32_16_half_bits(h) : h);
}

// max |x| over an operand (exact values), for the fp16 split scale
__global__ __launch_bounds__(256) void absmax_kernel(Src s, int64_t rows, int64_t cols,
                                                     unsigned* __restrict__ out_bits) {
    float mx = 0.0f;
    const int64_t total = rows * cols;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; t < total;
         t += static_cast<int64_t>(gridDim.x) * 256) {
        const int64_t j = t / rows, i = t - j * rows;
        const float v = fabsf(src_value(s, i + j * s.ld));
        if (v <= 3.0e38f) mx = fmaxf(mx, v);  // ignore Inf/NaN
    }
    for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out_bits, __float_as_uint(mx));
}

__device__ __forceinline__ int fp16_scale_exp(bool f16, const unsigned* absmax_bits) {
    // power-of-two scale bringing max |x| to about 2^13 (fp16 max is 65504)
    if (!f16) return 0;
    const float mx = __uint_as_float(*absmax_bits);
    int e = 0;
    frexpf(mx > 0.0f ? mx : 1.0f, &e);
    return 13 - e;
}

__device__ __forceinline__ void split16(bool f16, float v, int sc, uint16_t& h16, uint16_t& l16) {
    if (f16) {
        const float x = ldexpf(v, sc);
        const _Float16 h = static_cast<_Float16>(x);
        const _Float16 l = static_cast<_Float16>(x - static_cast<float>(h));
        h16 = __builtin_bit_cast(uint16_t, h);
        l16 = __builtin_bit_cast(uint16_t, l);
    } else {
        h16 = bf16_rn(v);
        l16 = bf16_rn(v - bf16_val(h16));
    }
}

enum { PREP_HI = 1, PREP_SPLIT = 3 };

// hi (and, for PREP_SPLIT, lo) 16-bit bits of element idx
template <int MODE>
__device__ __forceinline__ void prep_elem(const Src& s, int64_t idx, bool valid, bool f16, int sc,
                                          uint16_t& h16, uint16_t& l16) {
    if constexpr (MODE == PREP_HI) {
        h16 = valid ? src_hi_bits(s, idx, f16) : uint16_t(0);
        l16 = 0;
    } else {
        split16(f16, valid ? src_value(s, idx) : 0.0f, sc, h16, l16);
    }
}

// K-contiguous operand (X is K x R column-major): out[r][k] = X(k, r) for
// r < pad_r, k < pad_k, zero outside X.  hi (and lo) planes have row
// stride out_ld.
template <int MODE>
__global__ __launch_bounds__(256) void prep_k_kernel(Src s, int64_t rows, int64_t cols, int64_t out_ld,
                                                     uint16_t* __restrict__ hi, uint16_t* __restrict__ lo,
                                                     const unsigned* __restrict__ absmax_bits, bool f16,
                                                     int64_t pad_r, int64_t pad_k) {
    const int sc = MODE == PREP_SPLIT ? fp16_scale_exp(f16, absmax_bits) : 0;
    const int64_t total = pad_r * pad_k;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; t < total;
         t += static_cast<int64_t>(gridDim.x) * 256) {
        const int64_t r = t / pad_k, k = t - r * pad_k;
        uint16_t h16, l16;
        prep_elem<MODE>(s, k + r * s.ld, k < rows && r < cols, f16, sc, h16, l16);
        hi[r * out_ld + k] = h16;
        if constexpr (MODE == PREP_SPLIT) lo[r * out_ld + k] = l16;
    }
}

// Vector form of prep_k_kernel for operands whose columns are whole 16-byte
// runs of 16-bit words (rows % 8 == 0, ld % 8 == 0, 16-byte aligned planes)
// or 32-byte runs of FP32: 8 elements per thread, 16-byte stores.
template <int MODE>
__global__ __launch_bounds__(256) void prep_k_vec_kernel(Src s, int64_t rows, int64_t cols, int64_t out_ld,
                                                         uint16_t* __restrict__ hi, uint16_t* __restrict__ lo,
                                                         const unsigned* __restrict__ absmax_bits, bool f16,
                                                         int64_t pad_r, int64_t pad_k) {
    const int sc = MODE == PREP_SPLIT ? fp16_scale_exp(f16, absmax_bits) : 0;
    const int64_t cpr = pad_k / 8, total = pad_r * cpr;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; t < total;
         t += static_cast<int64_t>(gridDim.x) * 256) {
        const int64_t r = t / cpr, k = (t - r * cpr) * 8;  // out[r][k..k+7] = X(k..k+7, r)
        const bool valid = k < rows && r < cols;
        const int64_t base = k + r * s.ld;
        uint16_t h[8], l[8];
        if (s.f) {
            float v[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            if (valid) {
                const float4 v0 = *reinterpret_cast<const float4*>(s.f + base);
                const float4 v1 = *reinterpret_cast<const float4*>(s.f + base + 4);
                v[0] = v0.x; v[1] = v0.y; v[2] = v0.z; v[3] = v0.w;
                v[4] = v1.x; v[5] = v1.y; v[6] = v1.z; v[7] = v1.w;
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                if constexpr (MODE == PREP_HI) {
                    h[e] = f16 ? f16_rn(v[e]) : bf16_rn(v[e]);
                    l[e] = 0;
                } else {
                    split16(f16, v[e], sc, h[e], l[e]);
                }
            }
        } else {
            u32x4 wh8 = {0u, 0u, 0u, 0u}, wr8 = {0u, 0u, 0u, 0u};
            if (valid) {
                wh8 = *reinterpret_cast<const u32x4*>(s.wh + base);
                if constexpr (MODE == PREP_SPLIT) wr8 = *reinterpret_cast<const u32x4*>(s.wr + base);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const uint32_t hw = b ? (wh8[q] >> 16) : (wh8[q] & 0xFFFFu);
                    const int e = 2 * q + b;
                    if constexpr (MODE == PREP_HI) {
                        h[e] = valid ? static_cast<uint16_t>(s.fmt == TOMOS_FP32_16 ? codec::fp32_16_half_bits(hw)
                                                                                    : hw)
                                     : uint16_t(0);
                        l[e] = 0;
                    } else {
                        const uint32_t rw = b ? (wr8[q] >> 16) : (wr8[q] & 0xFFFFu);
                        split16(f16, valid ? decode_exact(s.fmt, hw, rw) : 0.0f, sc, h[e], l[e]);
                    }
                }
        }
        u32x4 h8, l8;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            h8[q] = static_cast<uint32_t>(h[2 * q]) | (static_cast<uint32_t>(h[2 * q + 1]) << 16);
            l8[q] = static_cast<uint32_t>(l[2 * q]) | (static_cast<uint32_t>(l[2 * q + 1]) << 16);
        }
        *reinterpret_cast<u32x4*>(hi + r * out_ld + k) = h8;
        if constexpr (MODE == PREP_SPLIT) *reinterpret_cast<u32x4*>(lo + r * out_ld + k) = l8;
    }
}

// R-contiguous operand (X is R x K column-major): out[r][k] = X(r, k) needs
// a transpose, done per 64 x 64 tile through LDS so both the reads (r
// fastest) and the writes (k fastest) are contiguous.
// grid = (ceil(pad_k/64), ceil(pad_r/64)).
template <int MODE>
__global__ __launch_bounds__(256) void prep_t_kernel(Src s, int64_t rows, int64_t cols, int64_t out_ld,
                                                     uint16_t* __restrict__ hi, uint16_t* __restrict__ lo,
                                                     const unsigned* __restrict__ absmax_bits, bool f16,
                                                     int64_t pad_r, int64_t pad_k) {
    __shared__ uint16_t th[64][65];
    __shared__ uint16_t tl[64][65];
    const int sc = MODE == PREP_SPLIT ? fp16_scale_exp(f16, absmax_bits) : 0;
    const int64_t k0 = static_cast<int64_t>(blockIdx.x) * 64, r0 = static_cast<int64_t>(blockIdx.y) * 64;
    const int tid = threadIdx.x;
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {
        const int idx = tid + it * 256, i = idx & 63, j = idx >> 6;  // i: r, j: k
        const int64_t r = r0 + i, k = k0 + j;
        uint16_t h16, l16;
        prep_elem<MODE>(s, r + k * s.ld, r < rows && k < cols, f16, sc, h16, l16);
        th[j][i] = h16;
        if constexpr (MODE == PREP_SPLIT) tl[j][i] = l16;
    }
    __syncthreads();
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {
        const int idx = tid + it * 256, j = idx & 63, i = idx >> 6;  // j: k, i: r
        const int64_t r = r0 + i, k = k0 + j;
        if (r < pad_r && k < pad_k) {
            hi[r * out_ld + k] = th[j][i];
            if constexpr (MODE == PREP_SPLIT) lo[r * out_ld + k] = tl[j][i];
        }
    }
}

// ------------------------------------------------------------------- gemm
template <bool F16>
__device__ __forceinline__ void mfma32(const u32x4& a, const u32x4& b, f32x16& c) {
    if constexpr (F16) {
        c = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
    } else {
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    }
}

// C tile -> (tm, tn).  Block ids go round-robin over the 8 XCDs; the remap
// gives each XCD a contiguous run of ids, and inside a run the tiles are
// ordered in groups of kGroupN tile columns walked down M, so the ~32 blocks
// an XCD holds at once cover an 8 x 4 patch of C (8 A row panels and 4 B
// column panels in that XCD's L2) instead of one whole tile column.
constexpr int kGroupN = 4;
__device__ __forceinline__ void tile_of(int64_t tiles_m, int64_t tiles_n, int64_t& tm, int64_t& tn) {
    const int64_t nb = tiles_m * tiles_n;
    int64_t bid = blockIdx.x;
    if (nb % 8 == 0) bid = (bid % 8) * (nb / 8) + bid / 8;
    const int64_t per_group = tiles_m * kGroupN;
    const int64_t g = bid / per_group, w = bid - g * per_group;
    const int64_t gn = min(static_cast<int64_t>(kGroupN), tiles_n - g * kGroupN);
    tn = g * kGroupN + w % gn;
    tm = w / gn;
}

// Ah/Al: [Mp][Kp], Bh/Bl: [Np][Kp] (K contiguous, padded to tile multiples).
// NT threads = NT/64 waves as 2 (M) x WN (N); a wave owns 128 x (256/WN)
// of the 256 x 256 C tile: 4 x TJ MFMA 32x32x16 tiles, 3 MFMAs (lo*hi,
// hi*lo, hi*hi) per tile per 16-wide k step into one FP32 accumulator.  The
// B fragment is the MFMA's first operand, so the accumulator is C^T per
// tile and each lane holds C rows (lane % 32): the column-major stores of a
// 32-lane half are 128 contiguous bytes.  The K slice (32) of all four
// planes is staged through LDS (rows padded to 80 bytes: conflict-free
// 16-byte fragment reads); the next slice is loaded into registers while
// the current one is multiplied.
template <bool F16, int NT, int TERMS>
__global__ __launch_bounds__(NT) void gemm_mfma_kernel(
    int64_t M, int64_t N, int64_t Kp, const uint16_t* __restrict__ Ah, const uint16_t* __restrict__ Al,
    const uint16_t* __restrict__ Bh, const uint16_t* __restrict__ Bl, int64_t ldk, double alpha,
    double beta, float* __restrict__ C, int64_t ldc, const unsigned* __restrict__ absA,
    const unsigned* __restrict__ absB) {
    constexpr int WN = NT / 128;      // waves along N
    constexpr int TJ = 8 / WN;        // 32-wide n tiles per wave
    constexpr int CH = 1024 / NT;     // 16-byte chunks per plane slice per thread
    static_assert(TERMS == 1 || TERMS == 3, "hi-only or 3-term split");
    extern __shared__ __attribute__((aligned(16))) uint16_t gemm_lds[];
    uint16_t* sAh = gemm_lds;
    uint16_t* sBh = sAh + kPlaneElems;
    uint16_t* sAl = sBh + kPlaneElems;  // TERMS == 3 only
    uint16_t* sBl = sAl + kPlaneElems;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int64_t tm, tn;
    tile_of((M + BM - 1) / BM, (N + BN - 1) / BN, tm, tn);
    const int64_t m0 = tm * BM, n0 = tn * BN;

    f32x16 acc[TJ][4];  // [n tile][m tile]
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[j][i][v] = 0.0f;

    // loader: a plane slice is 256 rows x 32 k = 1024 16-byte chunks, CH per thread
    int so[CH];
    int64_t oa[CH], ob[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
        const int ch = tid + q * NT, row = ch >> 2, kc = (ch & 3) * 8;
        so[q] = row * LDK + kc;
        oa[q] = (m0 + row) * ldk + kc;
        ob[q] = (n0 + row) * ldk + kc;
    }
    // next K slice in flight in registers: A hi/lo and B hi/lo
    u32x4 pah[CH], pal[CH], pbh[CH], pbl[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
        pah[q] = *reinterpret_cast<const u32x4*>(Ah + oa[q]);
        pbh[q] = *reinterpret_cast<const u32x4*>(Bh + ob[q]);
        if constexpr (TERMS == 3) {
            pal[q] = *reinterpret_cast<const u32x4*>(Al + oa[q]);
            pbl[q] = *reinterpret_cast<const u32x4*>(Bl + ob[q]);
        }
    }
    for (int64_t k0 = 0; k0 < Kp; k0 += BK) {
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            *reinterpret_cast<u32x4*>(&sAh[so[q]]) = pah[q];
            *reinterpret_cast<u32x4*>(&sBh[so[q]]) = pbh[q];
            if constexpr (TERMS == 3) {
                *reinterpret_cast<u32x4*>(&sAl[so[q]]) = pal[q];
                *reinterpret_cast<u32x4*>(&sBl[so[q]]) = pbl[q];
            }
        }
        __syncthreads();
        if (k0 + BK < Kp) {
            const int64_t k1 = k0 + BK;
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                pah[q] = *reinterpret_cast<const u32x4*>(Ah + oa[q] + k1);
                pbh[q] = *reinterpret_cast<const u32x4*>(Bh + ob[q] + k1);
                if constexpr (TERMS == 3) {
                    pal[q] = *reinterpret_cast<const u32x4*>(Al + oa[q] + k1);
                    pbl[q] = *reinterpret_cast<const u32x4*>(Bl + ob[q] + k1);
                }
            }
        }
#pragma unroll
        for (int ks = 0; ks < BK; ks += 16) {
            const int kk = ks + 8 * (lane >> 5);
            u32x4 ah[4], al[4], bh[TJ], bl[TJ];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = (wm * 128 + i * 32 + (lane & 31)) * LDK + kk;
                ah[i] = *reinterpret_cast<const u32x4*>(&sAh[r]);
                if constexpr (TERMS == 3) al[i] = *reinterpret_cast<const u32x4*>(&sAl[r]);
            }
#pragma unroll
            for (int j = 0; j < TJ; ++j) {
                const int r = (wn * (32 * TJ) + j * 32 + (lane & 31)) * LDK + kk;
                bh[j] = *reinterpret_cast<const u32x4*>(&sBh[r]);
                if constexpr (TERMS == 3) bl[j] = *reinterpret_cast<const u32x4*>(&sBl[r]);
            }
#pragma unroll
            for (int j = 0; j < TJ; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if constexpr (TERMS == 3) {
                        mfma32<F16>(bh[j], al[i], acc[j][i]);
                        mfma32<F16>(bl[j], ah[i], acc[j][i]);
                    }
                    mfma32<F16>(bh[j], ah[i], acc[j][i]);
                }
        }
        __syncthreads();
    }
    // epilogue: acc[j][i][v] is C(row, col) with row = m tile i, lane % 32 and
    // col = n tile j, 8 (v / 4) + 4 (lane / 32) + v % 4
    double scale = alpha;
    if (F16 && absA) {  // fp16 split planes carry the scale 2^(13 - e) per operand
        int ea = 0, eb = 0;
        frexpf(fmaxf(__uint_as_float(*absA), 1e-38f), &ea);
        frexpf(fmaxf(__uint_as_float(*absB), 1e-38f), &eb);
        scale = ldexp(alpha, -((13 - ea) + (13 - eb)));
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t row = m0 + wm * 128 + i * 32 + (lane & 31);
        if (row >= M) continue;
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int64_t col = n0 + wn * (32 * TJ) + j * 32 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
                if (col < N) {
                    float* p = C + row + col * ldc;
                    double r = scale * static_cast<double>(acc[j][i][v]);
                    if (beta != 0.0) r += beta * static_cast<double>(*p);
                    *p = static_cast<float>(r);
                }
            }
    }
}

// Staged operand planes: absmax words, then A hi [Mp][Kp], B hi [Np][Kp]
// and, for the split, A lo and B lo.
struct Planes {
    int64_t Mp, Np, Kp;
    unsigned* absmax;
    uint16_t *Ah, *Bh, *Al, *Bl;
};

Planes planes_of(int terms, int64_t M, int64_t N, int64_t K, void* base_ptr) {
    Planes p;
    p.Mp = ceil_div(M, BM) * BM;
    p.Np = ceil_div(N, BN) * BN;
    p.Kp = std::max<int64_t>(ceil_div(K, BK) * BK, BK);
    char* base = static_cast<char*>(base_ptr);
    p.absmax = reinterpret_cast<unsigned*>(base);
    p.Ah = reinterpret_cast<uint16_t*>(base + 256);
    p.Bh = p.Ah + p.Mp * p.Kp;
    p.Al = terms == 3 ? p.Bh + p.Np * p.Kp : nullptr;
    p.Bl = terms == 3 ? p.Al + p.Mp * p.Kp : nullptr;
    return p;
}

// Stage one operand: X is rows x cols column-major; kcontig means the GEMM's
// K runs down X's columns (X is K x R), otherwise along its rows (X is R x K).
void stage(int mode, bool f16, const Src& x, bool kcontig, int64_t rows, int64_t cols, int64_t pad_r,
           int64_t Kp, uint16_t* hi, uint16_t* lo, const unsigned* absmax, hipStream_t s) {
    if (kcontig) {
        const bool vec = rows % 8 == 0 && x.ld % 8 == 0 &&
                         (x.f ? reinterpret_cast<uintptr_t>(x.f) % 16 == 0
                              : reinterpret_cast<uintptr_t>(x.wh) % 16 == 0 &&
                                    (mode == PREP_HI || reinterpret_cast<uintptr_t>(x.wr) % 16 == 0));
        if (vec && rows > 0) {
            const unsigned blocks = static_cast<unsigned>(std::min<int64_t>(ceil_div(pad_r * Kp / 8, 256), 65536));
            if (mode == PREP_HI)
                prep_k_vec_kernel<PREP_HI><<<blocks, 256, 0, s>>>(x, rows, cols, Kp, hi, lo, absmax, f16, pad_r, Kp);
            else
                prep_k_vec_kernel<PREP_SPLIT><<<blocks, 256, 0, s>>>(x, rows, cols, Kp, hi, lo, absmax, f16, pad_r,
                                                                     Kp);
        } else {
            const unsigned blocks = static_cast<unsigned>(std::min<int64_t>(ceil_div(pad_r * Kp, 256), 65536));
            if (mode == PREP_HI)
                prep_k_kernel<PREP_HI><<<blocks, 256, 0, s>>>(x, rows, cols, Kp, hi, lo, absmax, f16, pad_r, Kp);
            else
                prep_k_kernel<PREP_SPLIT><<<blocks, 256, 0, s>>>(x, rows, cols, Kp, hi, lo, absmax, f16, pad_r, Kp);
        }
    } else {
        require(ceil_div(pad_r, 64) < 65536, "gemm: operand too large");
        const dim3 grid(static_cast<unsigned>(ceil_div(Kp, 64)), static_cast<unsigned>(ceil_div(pad_r, 64)));
        if (mode == PREP_HI)
            prep_t_kernel<PREP_HI><<<grid, 256, 0, s>>>(x, rows, cols, Kp, hi, lo, absmax, f16, pad_r, Kp);
        else
            prep_t_kernel<PREP_SPLIT><<<grid, 256, 0, s>>>(x, rows, cols, Kp, hi, lo, absmax, f16, pad_r, Kp);
    }
    TOMOS_LAUNCH_CHECK("gemm operand staging");
}

}  // namespace

size_t gemm_mfma_scratch_bytes(int terms, int64_t M, int64_t N, int64_t K) {
    const int64_t Mp = ceil_div(M, BM) * BM, Np = ceil_div(N, BN) * BN;
    const int64_t Kp = std::max<int64_t>(ceil_div(K, BK) * BK, BK);
    return static_cast<size_t>((terms == 3 ? 2 : 1) * (Mp + Np) * Kp) * 2 + 256;
}

size_t gemm_split_scratch_bytes(int64_t M, int64_t N, int64_t K) { return gemm_mfma_scratch_bytes(3, M, N, K); }

// C = alpha op(A) op(B) + beta C on 16-bit MFMA with FP32 accumulation:
// terms == 1 multiplies the hi words (A's packed hi word as the 16-bit
// tiers decode it, B rounded to nearest-even bf16/fp16), terms == 3 the
// split A_hi B_hi + A_hi B_lo + A_lo B_hi of the exact values.  f16 picks
// fp16 planes (FP32_16) over bf16 ones (FP32_BF16).  A is M x K (a_kcontig:
// stored K x M, i.e. transposed), B is K x N (b_kcontig: stored K x N, the
// plain layout; otherwise stored N x K).
void gemm_mfma(int terms, bool f16, const Src& a, bool a_kcontig, const Src& b, bool b_kcontig, int64_t M,
               int64_t N, int64_t K, double alpha, double beta, float* c, int64_t ldc, void* scratch,
               size_t scratch_bytes, hipStream_t s) {
    require(terms == 1 || terms == 3, "gemm_mfma: terms must be 1 or 3");
    if (M == 0 || N == 0) return;
    const size_t need = gemm_mfma_scratch_bytes(terms, M, N, K);
    if (!scratch) scratch = fallback_scratch(need);
    else require(scratch_bytes >= need, "gemm: scratch too small (need " + std::to_string(need) + " bytes)");
    const Planes p = planes_of(terms, M, N, K, scratch);
    const int mode = terms == 3 ? PREP_SPLIT : PREP_HI;
    const bool scaled = terms == 3 && f16;
    if (scaled) {
        TOMOS_HIP_CHECK(hipMemsetAsync(p.absmax, 0, 8, s));
        if (K > 0) {
            absmax_kernel<<<2048, 256, 0, s>>>(a, a_kcontig ? K : M, a_kcontig ? M : K, p.absmax);
            absmax_kernel<<<2048, 256, 0, s>>>(b, b_kcontig ? K : N, b_kcontig ? N : K, p.absmax + 1);
            TOMOS_LAUNCH_CHECK("absmax_kernel");
        }
    }
    stage(mode, f16, a, a_kcontig, a_kcontig ? K : M, a_kcontig ? M : K, p.Mp, p.Kp, p.Ah, p.Al, p.absmax, s);
    stage(mode, f16, b, b_kcontig, b_kcontig ? K : N, b_kcontig ? N : K, p.Np, p.Kp, p.Bh, p.Bl, p.absmax + 1, s);
    const int64_t nblocks = (p.Mp / BM) * (p.Np / BN);
    require(nblocks < (int64_t(1) << 31), "gemm: grid too large");
    const unsigned* absA = scaled ? p.absmax : nullptr;
    const unsigned* absB = scaled ? p.absmax + 1 : nullptr;
    const int lds = (terms == 3 ? 4 : 2) * kPlaneElems * 2;
    auto launch = [&](auto* k) {
        TOMOS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        k<<<static_cast<unsigned>(nblocks), kGemmThreads, lds, s>>>(M, N, p.Kp, p.Ah, p.Al, p.Bh, p.Bl, p.Kp, alpha,
                                                                    beta, c, ldc, absA, absB);
    };
    if (terms == 3) {
        if (f16) launch(gemm_mfma_kernel<true, kGemmThreads, 3>);
        else launch(gemm_mfma_kernel<false, kGemmThreads, 3>);
    } else {
        if (f16) launch(gemm_mfma_kernel<true, kGemmThreads, 1>);
        else launch(gemm_mfma_kernel<false, kGemmThreads, 1>);
    }
    TOMOS_LAUNCH_CHECK("gemm_mfma_kernel");
}

void gemm_split(int a_type, int64_t M, int64_t N, int64_t K, double alpha, const uint16_t* awh,
                const uint16_t* awr, int64_t lda, const uint16_t* bwh, const uint16_t* bwr, int64_t ldb,
                double beta, float* c, int64_t ldc, void* scratch, size_t scratch_bytes, hipStream_t s) {
    require(a_type == TOMOS_FP32_16 || a_type == TOMOS_FP32_BF16, "gemm_split: packed operands only");
    require(M >= 0 && N >= 0 && K >= 0, "gemm_split: negative size");
    require(lda >= std::max<int64_t>(M, 1) && ldb >= std::max<int64_t>(K, 1) && ldc >= std::max<int64_t>(M, 1),
            "gemm_split: bad leading dimension");
    if (M == 0 || N == 0) return;
    require(c != nullptr && (K == 0 || (awh && awr && bwh && bwr)), "gemm_split: null pointer");
    const Src a{a_type, awh, awr, nullptr, lda}, b{a_type, bwh, bwr, nullptr, ldb};
    gemm_mfma(3, a_type == TOMOS_FP32_16, a, false, b, true, M, N, K, alpha, beta, c, ldc, scratch, scratch_bytes,
              s);
}

// tomos_gemm's packed-A tiers on matrix cores: DEFAULT (hi words) and
// MATERIALIZE (3-term split of the exact values), FP32 B.
void gemm_packed_mfma(int a_type, int tier, bool ta, bool tb, double alpha, int64_t M, int64_t N, int64_t K,
                      const uint16_t* wh, const uint16_t* wr, int64_t lda, const float* b, int64_t ldb,
                      double beta, float* c, int64_t ldc, hipStream_t s) {
    const int terms = tier == TOMOS_TIER_MATERIALIZE ? 3 : 1;
    const Src a{a_type, wh, wr, nullptr, lda}, bs{a_type, nullptr, nullptr, b, ldb};
    gemm_mfma(terms, a_type == TOMOS_FP32_16, a, ta, bs, !tb, M, N, K, alpha, beta, c, ldc, nullptr, 0, s);
}

}  // namespace tomos_hip

using namespace tomos_hip;

extern "C" {

size_t tomos_gemm_split_scratch_size(int64_t M, int64_t N, int64_t K) {
    if (M < 0 || N < 0 || K < 0) return 0;
    return gemm_split_scratch_bytes(M, N, K);
}

int tomos_gemm_split(int a_type, int64_t M, int64_t N, int64_t K, double alpha, const uint16_t* a_wh,
                     const uint16_t* a_wr, int64_t lda, const uint16_t* b_wh, const uint16_t* b_wr,
                     int64_t ldb, double beta, float* c, int64_t ldc, void* scratch,
                     size_t scratch_bytes, tomos_stream_t s) {
    return guarded([&] {
        gemm_split(a_type, M, N, K, alpha, a_wh, a_wr, lda, b_wh, b_wr, ldb, beta, c, ldc, scratch,
                   scratch_bytes, s);
    });
}

}  // extern "C"
