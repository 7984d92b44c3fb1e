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
//  2. gemm: 256 threads, 128 x 128 C tile, K slice of 32 staged through LDS
//     (hi and lo of both operands, rows padded against bank conflicts);
//     each wave owns a 64 x 64 quadrant = 2 x 2 MFMA 32x32x16 tiles and
//     issues 3 MFMAs per tile per 16-wide k step into one FP32 accumulator.
#include <algorithm>

#include "codec.hpp"
#include "common.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int BM = 128, BN = 128, BK = 32;
constexpr int LDK = BK + 8;  // padded K stride (elements) of an LDS row

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
__device__ __forceinline__ void mfma32(const uint4& a, const uint4& b, f32x16& c) {
    if constexpr (F16) {
        c = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
    } else {
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    }
}

// Ah/Al: [Mp][Kp], Bh/Bl: [Np][Kp] (K contiguous, padded to tile multiples)
template <bool F16>
__global__ __launch_bounds__(256) void gemm_split_kernel(
    int64_t M, int64_t N, int64_t Kp, const uint16_t* __restrict__ Ah, const uint16_t* __restrict__ Al,
    const uint16_t* __restrict__ Bh, const uint16_t* __restrict__ Bl, int64_t ldk, float alpha_pow2_scale,
    double alpha, double beta, float* __restrict__ C, int64_t ldc, const unsigned* __restrict__ absA,
    const unsigned* __restrict__ absB) {
    __shared__ __attribute__((aligned(16))) uint16_t sAh[BM * LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sAl[BM * LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBh[BN * LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBl[BN * LDK];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;  // 2 x 2 waves, 64 x 64 each
    // XCD-aware tile order: consecutive block ids go round-robin to the 8
    // XCDs; remap so that each XCD walks a contiguous band of C tiles
    // (sharing A rows / B columns in its own L2).
    const int64_t tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
    const int64_t nb = tiles_m * tiles_n;
    int64_t bid = blockIdx.x;
    if (nb % 8 == 0) bid = (bid % 8) * (nb / 8) + bid / 8;
    const int64_t tm = bid % tiles_m, tn = bid / tiles_m;
    const int64_t m0 = tm * BM, n0 = tn * BN;

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

    // global -> LDS: each tile is 128 rows x 32 k = 512 16-byte chunks per plane
    for (int64_t k0 = 0; k0 < Kp; k0 += BK) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int ch = tid + q * 256;       // 0..511
            const int row = ch >> 2, kc = (ch & 3) * 8;
            const int64_t ga = (m0 + row) * ldk + k0 + kc;
            const int64_t gb = (n0 + row) * ldk + k0 + kc;
            *reinterpret_cast<uint4*>(&sAh[row * LDK + kc]) = *reinterpret_cast<const uint4*>(Ah + ga);
            *reinterpret_cast<uint4*>(&sAl[row * LDK + kc]) = *reinterpret_cast<const uint4*>(Al + ga);
            *reinterpret_cast<uint4*>(&sBh[row * LDK + kc]) = *reinterpret_cast<const uint4*>(Bh + gb);
            *reinterpret_cast<uint4*>(&sBl[row * LDK + kc]) = *reinterpret_cast<const uint4*>(Bl + gb);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 16) {
            const int kk = ks + 8 * (lane >> 5);
            uint4 ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int r = wm * 64 + i * 32 + (lane & 31);
                ah[i] = *reinterpret_cast<const uint4*>(&sAh[r * LDK + kk]);
                al[i] = *reinterpret_cast<const uint4*>(&sAl[r * LDK + kk]);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int c = wn * 64 + j * 32 + (lane & 31);
                bh[j] = *reinterpret_cast<const uint4*>(&sBh[c * LDK + kk]);
                bl[j] = *reinterpret_cast<const uint4*>(&sBl[c * LDK + kk]);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    mfma32<F16>(al[i], bh[j], acc[i][j]);
                    mfma32<F16>(ah[i], bl[j], acc[i][j]);
                    mfma32<F16>(ah[i], bh[j], acc[i][j]);
                }
        }
        __syncthreads();
    }
    // epilogue: lane holds C rows 8*(v/4) + 4*(lane/32) + v%4, column lane%32
    double scale = alpha;
    if (F16) {
        int ea = 0, eb = 0;
        frexpf(fmaxf(__uint_as_float(*absA), 1e-38f), &ea);
        frexpf(fmaxf(__uint_as_float(*absB), 1e-38f), &eb);
        scale = ldexp(alpha, -((13 - ea) + (13 - eb)));
    }
    (void)alpha_pow2_scale;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t col = n0 + wn * 64 + j * 32 + (lane & 31);
            if (col >= N) continue;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int64_t row = m0 + wm * 64 + i * 32 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
                if (row < M) {
                    float* p = C + row + col * ldc;
                    double r = scale * static_cast<double>(acc[i][j][v]);
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
    if (f16 && K > 0) {
        absmax_kernel<<<1024, 256, 0, s>>>(a_type, M, K, awh, awr, lda, absmax);
        absmax_kernel<<<1024, 256, 0, s>>>(a_type, K, N, bwh, bwr, ldb, absmax + 1);
        TOMOS_LAUNCH_CHECK("absmax_kernel");
    }
    require(ceil_div(Mp, 64) < 65536, "gemm_split: M too large");
    const dim3 gridA(static_cast<unsigned>(ceil_div(Kp, 64)), static_cast<unsigned>(ceil_div(Mp, 64)));
    const unsigned blocksB = static_cast<unsigned>(std::min<int64_t>(ceil_div(Np * Kp, 256), 65536));
    split_planes_t_kernel<<<gridA, 256, 0, s>>>(a_type, M, K, awh, awr, lda, Kp, Ah, Al, absmax, Mp, Kp);
    split_planes_kernel<<<blocksB, 256, 0, s>>>(a_type, K, N, bwh, bwr, ldb, Kp, Bh, Bl, absmax + 1, Np, Kp);
    TOMOS_LAUNCH_CHECK("split_planes_kernel");
    const int64_t nblocks = (Mp / BM) * (Np / BN);
    require(nblocks < (int64_t(1) << 31), "gemm_split: grid too large");
    if (f16)
        gemm_split_kernel<true><<<static_cast<unsigned>(nblocks), 256, 0, s>>>(
            M, N, Kp, Ah, Al, Bh, Bl, Kp, 1.0f, alpha, beta, c, ldc, absmax, absmax + 1);
    else
        gemm_split_kernel<false><<<static_cast<unsigned>(nblocks), 256, 0, s>>>(
            M, N, Kp, Ah, Al, Bh, Bl, Kp, 1.0f, alpha, beta, c, ldc, absmax, absmax + 1);
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
