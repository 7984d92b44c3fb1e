// Hashed CountSketch Y = S * A on the GPU.
//
// S is the s x m sketch with one +-1 per column: row i of A goes to bucket
// h(i) with sign g(i), both from splitmix64 of i xor a fixed seed
// (src/kernels/cpu/sketch/countsketch_kernels.cpp:59-85).
//
// Apply: A (m x n, column-major, any storage format/tier) is streamed once
// in its natural order.  A block owns (row split r, column group of CG
// columns); its lanes read 16-byte row vectors of each column (coalesced),
// look up the packed bucket|sign word of each row (shared by every column
// group of the same row split, so it is L2-resident), and add
// sign*A[i,j] into an LDS copy of its s x CG slice of Y.  The LDS slice is
// written to a per-split partial, and a second pass sums the splits in a
// fixed order in FP64.  Bucket/sign indices are bit-exact; the values are
// FP32 sums within a split (order depends on LDS atomic order), FP64 across
// splits.
#include <algorithm>

#include "codec.hpp"
#include "common.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

constexpr uint64_t kBucketSeed = 0x12345678abcdef00ull;
constexpr uint64_t kSignSeed = 0xf00dd00d42babe55ull;

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// packed word: bucket in bits 0..30, bit 31 set for sign -1
__global__ __launch_bounds__(256) void cs_index_kernel(int64_t m, int64_t s,
                                                       int32_t* __restrict__ bucket,
                                                       int8_t* __restrict__ sign,
                                                       uint32_t* __restrict__ packed) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= m) return;
    const uint64_t hb = splitmix64(static_cast<uint64_t>(i) ^ kBucketSeed);
    const uint64_t hs = splitmix64(static_cast<uint64_t>(i) ^ kSignSeed);
    const uint32_t b = static_cast<uint32_t>(hb % static_cast<uint64_t>(s));
    const bool pos = (hs & 1ull) != 0;
    if (bucket) bucket[i] = static_cast<int32_t>(b);
    if (sign) sign[i] = pos ? 1 : -1;
    if (packed) packed[i] = b | (pos ? 0u : 0x80000000u);
}

enum { FMT_F32 = 0, FMT_F64 = 1, FMT_P16 = 2, FMT_PBF = 3 };

template <int FMT, bool EXACT>
__device__ __forceinline__ float load_elem(const void* __restrict__ a,
                                           const uint16_t* __restrict__ wr, int64_t idx) {
    if constexpr (FMT == FMT_F32) return static_cast<const float*>(a)[idx];
    else if constexpr (FMT == FMT_F64) return static_cast<float>(static_cast<const double*>(a)[idx]);
    else {
        const uint32_t h = static_cast<const uint16_t*>(a)[idx];
        if constexpr (FMT == FMT_P16) {
            if constexpr (EXACT) return codec::fp32_16_decode_exact(h, wr[idx]);
            else return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(codec::fp32_16_half_bits(h))));
        } else {
            if constexpr (EXACT) return __uint_as_float((h << 16) | wr[idx]);
            else return __uint_as_float(h << 16);
        }
    }
}

template <int FMT, bool EXACT, int CG>
__global__ __launch_bounds__(256) void cs_apply_kernel(int64_t m, int64_t n,
                                                       const void* __restrict__ a,
                                                       const uint16_t* __restrict__ wr,
                                                       int64_t lda, int64_t s,
                                                       const uint32_t* __restrict__ packed,
                                                       int64_t rows_per_split, int ncg,
                                                       float* __restrict__ partial) {
    extern __shared__ float ylds[];  // [s][CG]
    const int split = blockIdx.x / ncg;
    const int cg = blockIdx.x - split * ncg;
    const int64_t c0 = static_cast<int64_t>(cg) * CG;
    const int ncols = static_cast<int>(min<int64_t>(CG, n - c0));
    for (int64_t t = threadIdx.x; t < s * CG; t += 256) ylds[t] = 0.0f;
    __syncthreads();

    const int64_t r0 = static_cast<int64_t>(split) * rows_per_split;
    const int64_t r1 = min(m, r0 + rows_per_split);
    for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) {
        const uint32_t w = packed[i];
        const int b = static_cast<int>(w & 0x7FFFFFFFu);
        const float sg = (w & 0x80000000u) ? -1.0f : 1.0f;
#pragma unroll
        for (int c = 0; c < CG; ++c) {
            if (c < ncols) {
                const float v = load_elem<FMT, EXACT>(a, wr, i + (c0 + c) * lda);
                atomicAdd(&ylds[b * CG + c], sg * v);
            }
        }
    }
    __syncthreads();
    float* out = partial + static_cast<int64_t>(split) * s * n;
    for (int64_t t = threadIdx.x; t < s * ncols; t += 256) {
        const int64_t c = t / s;
        const int64_t b = t - c * s;
        out[(c0 + c) * s + b] = ylds[b * CG + c];
    }
}

__global__ __launch_bounds__(256) void cs_finish_kernel(int64_t s, int64_t n, int splits,
                                                        const float* __restrict__ partial,
                                                        double* __restrict__ y, int64_t ldy) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= s * n) return;
    const int64_t c = t / s;
    const int64_t b = t - c * s;
    double acc = 0.0;
    for (int r = 0; r < splits; ++r) acc += static_cast<double>(partial[static_cast<int64_t>(r) * s * n + t]);
    y[b + c * ldy] = acc;
}

constexpr int64_t kLdsBytes = 128 * 1024;

int cg_for(int64_t s) {
    int cg = 8;
    while (cg > 1 && s * cg * 4 > kLdsBytes) cg >>= 1;
    return cg;
}

struct CsPlan {
    int cg, ncg, splits;
    int64_t rows_per_split;
};

CsPlan cs_plan(int64_t m, int64_t n, int64_t s) {
    CsPlan p;
    p.cg = cg_for(s);
    p.ncg = static_cast<int>(ceil_div(std::max<int64_t>(n, 1), p.cg));
    const int64_t want = std::max<int64_t>(1, ceil_div(1024, p.ncg));
    const int64_t max_splits = std::max<int64_t>(1, ceil_div(m, 4096));
    p.splits = static_cast<int>(std::min(want, max_splits));
    p.rows_per_split = ceil_div(std::max<int64_t>(m, 1), p.splits);
    p.splits = static_cast<int>(ceil_div(std::max<int64_t>(m, 1), p.rows_per_split));
    return p;
}

size_t cs_scratch(int64_t m, int64_t n, int64_t s) {
    const CsPlan p = cs_plan(m, n, s);
    const size_t idx = ((static_cast<size_t>(m) * 4 + 255) / 256) * 256;
    return idx + static_cast<size_t>(p.splits) * s * n * 4;
}

template <int FMT, bool EXACT, int CG>
void launch_apply(const CsPlan& p, int64_t m, int64_t n, const void* a, const uint16_t* wr,
                  int64_t lda, int64_t s, const uint32_t* packed, float* partial,
                  hipStream_t st) {
    const size_t lds = static_cast<size_t>(s) * CG * 4;
    auto* k = cs_apply_kernel<FMT, EXACT, CG>;
    TOMOS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
    const int64_t blocks = static_cast<int64_t>(p.splits) * p.ncg;
    k<<<dim3(static_cast<unsigned>(blocks)), 256, lds, st>>>(m, n, a, wr, lda, s, packed,
                                                            p.rows_per_split, p.ncg, partial);
    TOMOS_LAUNCH_CHECK("cs_apply_kernel");
}

template <int FMT, bool EXACT>
void launch_apply_cg(const CsPlan& p, int64_t m, int64_t n, const void* a, const uint16_t* wr,
                     int64_t lda, int64_t s, const uint32_t* packed, float* partial,
                     hipStream_t st) {
    switch (p.cg) {
        case 8: launch_apply<FMT, EXACT, 8>(p, m, n, a, wr, lda, s, packed, partial, st); break;
        case 4: launch_apply<FMT, EXACT, 4>(p, m, n, a, wr, lda, s, packed, partial, st); break;
        case 2: launch_apply<FMT, EXACT, 2>(p, m, n, a, wr, lda, s, packed, partial, st); break;
        default: launch_apply<FMT, EXACT, 1>(p, m, n, a, wr, lda, s, packed, partial, st); break;
    }
}

}  // namespace

size_t countsketch_scratch_bytes(int64_t m, int64_t n, int64_t s) { return cs_scratch(m, n, s); }

void countsketch_apply(int a_type, int tier, int64_t m, int64_t n, const void* a,
                       const uint16_t* wr, int64_t lda, int64_t s, double* y, int64_t ldy,
                       void* scratch, size_t scratch_bytes, hipStream_t st) {
    require(m >= 0 && n >= 0 && s >= 1, "countsketch_apply: bad sizes");
    require(s <= kLdsBytes / 4, "countsketch_apply: s too large (max 32768)");
    require(m < (int64_t(1) << 31), "countsketch_apply: m must be < 2^31");
    require(lda >= std::max<int64_t>(m, 1) && ldy >= s, "countsketch_apply: bad leading dimension");
    require(tier == TOMOS_TIER_DEFAULT || tier == TOMOS_TIER_FAST || tier == TOMOS_TIER_EXACT,
            "countsketch_apply: bad tier");
    if (n == 0) return;
    require(y != nullptr, "countsketch_apply: null y");
    const size_t need = cs_scratch(m, n, s);
    if (scratch == nullptr) scratch = fallback_scratch(need);
    else require(scratch_bytes >= need, "countsketch_apply: scratch too small");
    const CsPlan p = cs_plan(m, n, s);
    uint32_t* packed = static_cast<uint32_t*>(scratch);
    float* partial = reinterpret_cast<float*>(static_cast<char*>(scratch) +
                                              ((static_cast<size_t>(m) * 4 + 255) / 256) * 256);
    if (m > 0) {
        require(a != nullptr, "countsketch_apply: null A");
        cs_index_kernel<<<dim3(static_cast<unsigned>(ceil_div(m, 256))), 256, 0, st>>>(
            m, s, nullptr, nullptr, packed);
        TOMOS_LAUNCH_CHECK("cs_index_kernel");
        const bool exact = tier == TOMOS_TIER_EXACT;
        switch (a_type) {
            case TOMOS_F32: launch_apply_cg<FMT_F32, false>(p, m, n, a, wr, lda, s, packed, partial, st); break;
            case TOMOS_F64: launch_apply_cg<FMT_F64, false>(p, m, n, a, wr, lda, s, packed, partial, st); break;
            case TOMOS_FP32_16:
                require(!exact || wr, "countsketch_apply: EXACT needs wr");
                if (exact) launch_apply_cg<FMT_P16, true>(p, m, n, a, wr, lda, s, packed, partial, st);
                else launch_apply_cg<FMT_P16, false>(p, m, n, a, wr, lda, s, packed, partial, st);
                break;
            case TOMOS_FP32_BF16:
                require(!exact || wr, "countsketch_apply: EXACT needs wr");
                if (exact) launch_apply_cg<FMT_PBF, true>(p, m, n, a, wr, lda, s, packed, partial, st);
                else launch_apply_cg<FMT_PBF, false>(p, m, n, a, wr, lda, s, packed, partial, st);
                break;
            default: throw Error(kErrUnsupported, "countsketch_apply: unsupported a_type");
        }
    }
    const int64_t total = s * n;
    cs_finish_kernel<<<dim3(static_cast<unsigned>(ceil_div(total, 256))), 256, 0, st>>>(
        s, n, m > 0 ? p.splits : 0, partial, y, ldy);
    TOMOS_LAUNCH_CHECK("cs_finish_kernel");
}

}  // namespace tomos_hip

using namespace tomos_hip;

extern "C" {

int tomos_countsketch_indices(int64_t m, int64_t s, int32_t* bucket, int8_t* sign,
                              tomos_stream_t st) {
    return guarded([&] {
        require(m >= 0 && s >= 1, "tomos_countsketch_indices: bad sizes");
        require(s <= 0x7FFFFFFF, "tomos_countsketch_indices: s too large");
        if (m == 0) return;
        require(bucket || sign, "tomos_countsketch_indices: no output");
        cs_index_kernel<<<dim3(static_cast<unsigned>(ceil_div(m, 256))), 256, 0, st>>>(
            m, s, bucket, sign, nullptr);
        TOMOS_LAUNCH_CHECK("cs_index_kernel");
    });
}

size_t tomos_countsketch_scratch_size(int64_t m, int64_t n, int64_t s) {
    if (m < 0 || n < 0 || s < 1) return 0;
    return countsketch_scratch_bytes(m, n, s);
}

int tomos_countsketch_apply(int a_type, int tier, int64_t m, int64_t n, const void* a,
                            const uint16_t* wr, int64_t lda, int64_t s, double* y, int64_t ldy,
                            void* scratch, size_t scratch_bytes, tomos_stream_t st) {
    return guarded([&] {
        countsketch_apply(a_type, tier, m, n, a, wr, lda, s, y, ldy, scratch, scratch_bytes, st);
    });
}

}  // extern "C"
