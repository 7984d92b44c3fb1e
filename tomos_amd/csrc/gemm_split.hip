// Split-precision packed GEMM on CDNA4 matrix cores (gfx950):
//   C = A B  ~=  A_hi B_hi + A_hi B_lo + A_lo B_hi,   FP32 accumulate.
// (No reference counterpart: the reference's 16-bit packed GEMMs use the hi
// word only, mm_cuda_kernels.cu:35372 / :37826.)
//
// Two steps:
//  1. split: each packed operand (FP32_16 or FP32_BF16 planes, column-major)
//     is decoded exactly to FP32 and written as hi/lo 16-bit planes in a
//     K-contiguous layout (A as [M][K], B as [N][K]) so every MFMA operand is
//     one 16-byte load:  hi = rn16(a), lo = rn16(a - hi).  FP32_BF16 uses
//     bf16 planes; FP32_16 uses fp16 planes with a power-of-two scale 2^s
//     per operand, chosen from the operand's largest magnitude, so the hi and
//     lo terms stay clear of fp16 overflow and underflow (C is rescaled by
//     2^-(sA+sB) in the epilogue).
//  2. gemm: 512 threads, 256 x 256 C tile, K slice of 32 staged through LDS
//     (hi and lo of both operands, rows padded against bank conflicts);
//     each wave owns 128 x 64 = 4 x 2 MFMA 32x32x16 tiles and issues 3
//     MFMAs per tile per 16-wide k step into one FP32 accumulator.
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
constexpr int kGemmLds = 4 * kPlaneElems * 2;          // bytes: hi/lo of A and B, 80 KiB

// ------------------------------------------------------------------ split
__device__ __forceinline__ float decode_exact(int fmt, uint32_t h, uint32_t r) {
    return fmt == TOMOS_FP32_16 ? codec::fp32_16_decode_exact(h, r) : __uint_as_float((h << 16) | r);
}

__device__ __forceinline__ uint16_t bf16_rn(float v) {
    return static_cast<uint16_t>(codec::f32_to_bf16_bits(v));
}
__device__ __forceinline__ float bf16_val(uint16_t b) { return __uint_as_float(static_cast<uint32_t>(b) << 16); }

// max |a| over a packed column-major operand (exact decode), for the fp16 scale
__global__ __launch_bounds__(256) void absmax_kernel(int fmt, int64_t rows, int64_t cols,
                                                     const uint16_t* __restrict__ wh,
                                                     const uint16_t* __restrict__ wr, int64_t ld,
                                                     unsigned* __restrict__ out_bits) {
    float mx = 0.0f;
    const int64_t total = rows * cols;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; t < total;
         t += static_cast<int64_t>(gridDim.x) * 256) {
        const int64_t j = t / rows, i = t - j * rows;
        const float v = fabsf(decode_exact(fmt, wh[i + j * ld], wr[i + j * ld]));
        if (v <= 3.0e38f) mx = fmaxf(mx, v);  // ignore Inf/NaN
    }
    for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out_bits, __float_as_uint(mx));
}

__device__ __forceinline__ int fp16_scale_exp(int fmt, const unsigned* absmax_bits) {
    // power-of-two scale bringing max |x| to about 2^13 (fp16 max is 65504)
    if (fmt != TOMOS_FP32_16) return 0;
    const float mx = __uint_as_float(*absmax_bits);
    int e = 0;
    frexpf(mx > 0.0f ? mx : 1.0f, &e);
    return 13 - e;
}

__device__ __forceinline__ void split16(int fmt, float v, int sc, uint16_t& h16, uint16_t& l16) {
    if (fmt == TOMOS_FP32_16) {
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

// B operand (K x N column-major, K contiguous already): out[n][k], n < Np,
// k < Kp, zero outside the operand.  hi/lo are bf16 (FP32_BF16) or fp16 of
// x * 2^scale (FP32_16).
__global__ __launch_bounds__(256) void split_planes_kernel(
    int fmt, int64_t rows, int64_t cols, const uint16_t* __restrict__ wh,
    const uint16_t* __restrict__ wr, int64_t ld, int64_t out_ld, uint16_t* __restrict__ hi,
    uint16_t* __restrict__ lo, const unsigned* __restrict__ absmax_bits, int64_t pad_cols,
    int64_t pad_k) {
    const int sc = fp16_scale_exp(fmt, absmax_bits);
    const int64_t total = pad_cols * pad_k;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; t < total;
         t += static_cast<int64_t>(gridDim.x) * 256) {
        const int64_t r = t / pad_k, k = t - r * pad_k;  // out[r][k] = X(k, r)
        float v = 0.0f;
        if (k < rows && r < cols) v = decode_exact(fmt, wh[k + r * ld], wr[k + r * ld]);
        uint16_t h16, l16;
        split16(fmt, v, sc, h16, l16);
        hi[r * out_ld + k] = h16;
        lo[r * out_ld + k] = l16;
    }
}

// Vector forms of absmax_kernel and split_planes_kernel for operands whose
// columns are whole 16-byte runs (rows % 8 == 0, ld % 8 == 0, 16-byte
// aligned planes): 8 elements per thread, 16-byte loads and stores.
__device__ __forceinline__ void decode8(int fmt, const u32x4& h, const u32x4& r, float* v) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[2 * q] = decode_exact(fmt, h[q] & 0xFFFFu, r[q] & 0xFFFFu);
        v[2 * q + 1] = decode_exact(fmt, h[q] >> 16, r[q] >> 16);
    }
}

__global__ __launch_bounds__(256) void absmax_vec_kernel(int fmt, int64_t rows, int64_t cols,
                                                         const uint16_t* __restrict__ wh,
                                                         const uint16_t* __restrict__ wr, int64_t ld,
                                                         unsigned* __restrict__ out_bits) {
    float mx = 0.0f;
    const int64_t cpc = rows / 8, total = cpc * cols;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; t < total;
         t += static_cast<int64_t>(gridDim.x) * 256) {
        const int64_t j = t / cpc, i = (t - j * cpc) * 8;
        const u32x4 h = *reinterpret_cast<const u32x4*>(wh + i + j * ld);
        const u32x4 r = *reinterpret_cast<const u32x4*>(wr + i + j * ld);
        float v[8];
        decode8(fmt, h, r, v);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float a = fabsf(v[e]);
            if (a <= 3.0e38f) mx = fmaxf(mx, a);  // ignore Inf/NaN
        }
    }
    for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out_bits, __float_as_uint(mx));
}

__global__ __launch_bounds__(256) void split_planes_vec_kernel(
    int fmt, int64_t rows, int64_t cols, const uint16_t* __restrict__ wh,
    const uint16_t* __restrict__ wr, int64_t ld, int64_t out_ld, uint16_t* __restrict__ hi,
    uint16_t* __restrict__ lo, const unsigned* __restrict__ absmax_bits, int64_t pad_cols,
    int64_t pad_k) {
    const int sc = fp16_scale_exp(fmt, absmax_bits);
    const int64_t cpr = pad_k / 8, total = pad_cols * cpr;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; t < total;
         t += static_cast<int64_t>(gridDim.x) * 256) {
        const int64_t r = t / cpr, k = (t - r * cpr) * 8;  // out[r][k..k+7] = X(k..k+7, r)
        float v[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (k < rows && r < cols)
            decode8(fmt, *reinterpret_cast<const u32x4*>(wh + k + r * ld),
                    *reinterpret_cast<const u32x4*>(wr + k + r * ld), v);
        u32x4 h8, l8;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint16_t h0, l0, h1, l1;
            split16(fmt, v[2 * q], sc, h0, l0);
            split16(fmt, v[2 * q + 1], sc, h1, l1);
            h8[q] = static_cast<uint32_t>(h0) | (static_cast<uint32_t>(h1) << 16);
            l8[q] = static_cast<uint32_t>(l0) | (static_cast<uint32_t>(l1) << 16);
        }
        *reinterpret_cast<u32x4*>(hi + r * out_ld + k) = h8;
        *reinterpret_cast<u32x4*>(lo + r * out_ld + k) = l8;
    }
}

// A operand (M x K column-major, M contiguous): out[m][k] needs a transpose,
// done per 64 x 64 tile through LDS so both the reads (m fastest) and the
// writes (k fastest) are contiguous.  grid = (ceil(pad_k/64), ceil(pad_m/64)).
__global__ __launch_bounds__(256) void split_planes_t_kernel(
    int fmt, int64_t rows, int64_t cols, const uint16_t* __restrict__ wh,
    const uint16_t* __restrict__ wr, int64_t ld, int64_t out_ld, uint16_t* __restrict__ hi,
    uint16_t* __restrict__ lo, const unsigned* __restrict__ absmax_bits, int64_t pad_m,
    int64_t pad_k) {
    __shared__ uint16_t th[64][65];
    __shared__ uint16_t tl[64][65];
    const int sc = fp16_scale_exp(fmt, absmax_bits);
    const int64_t k0 = static_cast<int64_t>(blockIdx.x) * 64, m0 = static_cast<int64_t>(blockIdx.y) * 64;
    const int tid = threadIdx.x;
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {
        const int idx = tid + it * 256, i = idx & 63, j = idx >> 6;  // i: m, j: k
        const int64_t m = m0 + i, k = k0 + j;
        float v = 0.0f;
        if (m < rows && k < cols) v = decode_exact(fmt, wh[m + k * ld], wr[m + k * ld]);
        uint16_t h16, l16;
        split16(fmt, v, sc, h16, l16);
        th[j][i] = h16;
        tl[j][i] = l16;
    }
    __syncthreads();
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {
        const int idx = tid + it * 256, j = idx & 63, i = idx >> 6;  // j: k, i: m
        const int64_t m = m0 + i, k = k0 + j;
        if (m < pad_m && k < pad_k) {
            hi[m * out_ld + k] = th[j][i];
            lo[m * out_ld + k] = tl[j][i];
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
template <bool F16, int NT>
__global__ __launch_bounds__(NT) void gemm_split_kernel(
    int64_t M, int64_t N, int64_t Kp, const uint16_t* __restrict__ Ah, const uint16_t* __restrict__ Al,
    const uint16_t* __restrict__ Bh, const uint16_t* __restrict__ Bl, int64_t ldk, double alpha,
    double beta, float* __restrict__ C, int64_t ldc, const unsigned* __restrict__ absA,
    const unsigned* __restrict__ absB) {
    constexpr int WN = NT / 128;      // waves along N
    constexpr int TJ = 8 / WN;        // 32-wide n tiles per wave
    constexpr int CH = 1024 / NT;     // 16-byte chunks per plane slice per thread
    extern __shared__ __attribute__((aligned(16))) uint16_t gemm_lds[];
    uint16_t* sAh = gemm_lds;
    uint16_t* sAl = sAh + kPlaneElems;
    uint16_t* sBh = sAl + kPlaneElems;
    uint16_t* sBl = sBh + kPlaneElems;
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
        pal[q] = *reinterpret_cast<const u32x4*>(Al + oa[q]);
        pbh[q] = *reinterpret_cast<const u32x4*>(Bh + ob[q]);
        pbl[q] = *reinterpret_cast<const u32x4*>(Bl + ob[q]);
    }
    for (int64_t k0 = 0; k0 < Kp; k0 += BK) {
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            *reinterpret_cast<u32x4*>(&sAh[so[q]]) = pah[q];
            *reinterpret_cast<u32x4*>(&sAl[so[q]]) = pal[q];
            *reinterpret_cast<u32x4*>(&sBh[so[q]]) = pbh[q];
            *reinterpret_cast<u32x4*>(&sBl[so[q]]) = pbl[q];
        }
        __syncthreads();
        if (k0 + BK < Kp) {
            const int64_t k1 = k0 + BK;
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                pah[q] = *reinterpret_cast<const u32x4*>(Ah + oa[q] + k1);
                pal[q] = *reinterpret_cast<const u32x4*>(Al + oa[q] + k1);
                pbh[q] = *reinterpret_cast<const u32x4*>(Bh + ob[q] + k1);
                pbl[q] = *reinterpret_cast<const u32x4*>(Bl + ob[q] + k1);
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
                al[i] = *reinterpret_cast<const u32x4*>(&sAl[r]);
            }
#pragma unroll
            for (int j = 0; j < TJ; ++j) {
                const int r = (wn * (32 * TJ) + j * 32 + (lane & 31)) * LDK + kk;
                bh[j] = *reinterpret_cast<const u32x4*>(&sBh[r]);
                bl[j] = *reinterpret_cast<const u32x4*>(&sBl[r]);
            }
#pragma unroll
            for (int j = 0; j < TJ; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    mfma32<F16>(bh[j], al[i], acc[j][i]);
                    mfma32<F16>(bl[j], ah[i], acc[j][i]);
                    mfma32<F16>(bh[j], ah[i], acc[j][i]);
                }
        }
        __syncthreads();
    }
    // epilogue: acc[j][i][v] is C(row, col) with row = m tile i, lane % 32 and
    // col = n tile j, 8 (v / 4) + 4 (lane / 32) + v % 4
    double scale = alpha;
    if (F16) {
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

}  // namespace

size_t gemm_split_scratch_bytes(int64_t M, int64_t N, int64_t K) {
    const int64_t Mp = ceil_div(M, BM) * BM, Np = ceil_div(N, BN) * BN;
    const int64_t Kp = std::max<int64_t>(ceil_div(K, BK) * BK, BK);
    return static_cast<size_t>(2 * (Mp + Np) * Kp) * 2 + 256;
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
    const int64_t Mp = ceil_div(M, BM) * BM, Np = ceil_div(N, BN) * BN, Kp = std::max<int64_t>(ceil_div(K, BK) * BK, BK);
    const size_t need = gemm_split_scratch_bytes(M, N, K);
    if (!scratch) scratch = fallback_scratch(need);
    else require(scratch_bytes >= need, "gemm_split: scratch too small");
    char* base = static_cast<char*>(scratch);
    unsigned* absmax = reinterpret_cast<unsigned*>(base);
    uint16_t* Ah = reinterpret_cast<uint16_t*>(base + 256);
    uint16_t* Al = Ah + Mp * Kp;
    uint16_t* Bh = Al + Mp * Kp;
    uint16_t* Bl = Bh + Np * Kp;
    TOMOS_HIP_CHECK(hipMemsetAsync(absmax, 0, 8, s));
    const bool f16 = a_type == TOMOS_FP32_16;
    // whole 16-byte column runs: the vector kernels apply
    auto runs16 = [](int64_t rows, int64_t ld, const uint16_t* h, const uint16_t* r) {
        return rows % 8 == 0 && ld % 8 == 0 && reinterpret_cast<uintptr_t>(h) % 16 == 0 &&
               reinterpret_cast<uintptr_t>(r) % 16 == 0;
    };
    const bool vecA = runs16(M, lda, awh, awr), vecB = runs16(K, ldb, bwh, bwr);
    if (f16 && K > 0) {
        if (vecA) absmax_vec_kernel<<<2048, 256, 0, s>>>(a_type, M, K, awh, awr, lda, absmax);
        else absmax_kernel<<<1024, 256, 0, s>>>(a_type, M, K, awh, awr, lda, absmax);
        if (vecB) absmax_vec_kernel<<<2048, 256, 0, s>>>(a_type, K, N, bwh, bwr, ldb, absmax + 1);
        else absmax_kernel<<<1024, 256, 0, s>>>(a_type, K, N, bwh, bwr, ldb, absmax + 1);
        TOMOS_LAUNCH_CHECK("absmax_kernel");
    }
    require(ceil_div(Mp, 64) < 65536, "gemm_split: M too large");
    const dim3 gridA(static_cast<unsigned>(ceil_div(Kp, 64)), static_cast<unsigned>(ceil_div(Mp, 64)));
    const unsigned blocksB = static_cast<unsigned>(std::min<int64_t>(ceil_div(Np * Kp, 256), 65536));
    split_planes_t_kernel<<<gridA, 256, 0, s>>>(a_type, M, K, awh, awr, lda, Kp, Ah, Al, absmax, Mp, Kp);
    if (vecB && K > 0)
        split_planes_vec_kernel<<<static_cast<unsigned>(std::min<int64_t>(ceil_div(Np * Kp / 8, 256), 65536)), 256,
                                  0, s>>>(a_type, K, N, bwh, bwr, ldb, Kp, Bh, Bl, absmax + 1, Np, Kp);
    else
        split_planes_kernel<<<blocksB, 256, 0, s>>>(a_type, K, N, bwh, bwr, ldb, Kp, Bh, Bl, absmax + 1, Np, Kp);
    TOMOS_LAUNCH_CHECK("split_planes_kernel");
    const int64_t nblocks = (Mp / BM) * (Np / BN);
    require(nblocks < (int64_t(1) << 31), "gemm_split: grid too large");
    auto launch = [&](auto* k, int nt) {
        TOMOS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, kGemmLds));
        k<<<static_cast<unsigned>(nblocks), nt, kGemmLds, s>>>(M, N, Kp, Ah, Al, Bh, Bl, Kp, alpha, beta, c,
                                                             ldc, absmax, absmax + 1);
    };
    if (f16) launch(gemm_split_kernel<true, kGemmThreads>, kGemmThreads);
    else launch(gemm_split_kernel<false, kGemmThreads>, kGemmThreads);
    TOMOS_LAUNCH_CHECK("gemm_split_kernel");
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
