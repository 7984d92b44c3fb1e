// Packed GEMM, C = alpha*op(A)*op(B) + beta*C, FP32 B and C, A plain or
// packed (decoded while staging the A tile into LDS).
//
// Tiers follow dispatch::Fp32_16's mm slots (fp32_16_registries.cpp:370-453):
// DEFAULT decodes wh only and accumulates in FP32 (pure_half / hi-only; on
// matrix cores for packed A), FAST decodes wh only and accumulates in FP64,
// EXACT rebuilds FP32 from wh+wr and accumulates in FP64 ("accurate"),
// MATERIALIZE (packed A only) is the 3-term split on matrix cores.
//
// This file holds the SIMT tile kernel (64x64 C tile per 256-thread block,
// 4x4 outputs per thread, K staged through LDS in slices of 16) that runs
// the FP64-accumulating tiers and plain F32/F64 A.  Packed A's DEFAULT and
// MATERIALIZE tiers run on matrix cores (gemm_split.hip).
#include <algorithm>

#include "codec.hpp"
#include "common.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

enum { FMT_F32 = 0, FMT_F64 = 1, FMT_P16 = 2, FMT_PBF = 3 };
constexpr int BM = 64, BN = 64, BK = 16;

template <int FMT, bool EXACT>
__device__ __forceinline__ float a_elem(const void* __restrict__ a, const uint16_t* __restrict__ wr,
                                        int64_t idx) {
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

template <int FMT, bool EXACT, class ACC>
__global__ __launch_bounds__(256) void gemm_tile_kernel(int64_t M, int64_t N, int64_t K,
                                                        const void* __restrict__ a,
                                                        const uint16_t* __restrict__ wr,
                                                        int64_t lda, bool ta,
                                                        const float* __restrict__ b, int64_t ldb,
                                                        bool tb, double alpha, double beta,
                                                        float* __restrict__ c, int64_t ldc) {
    __shared__ float As[BK][BM + 4];
    __shared__ float Bs[BK][BN + 4];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int64_t m0 = static_cast<int64_t>(blockIdx.x) * BM;
    const int64_t n0 = static_cast<int64_t>(blockIdx.y) * BN;
    ACC acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = ACC(0);

    for (int64_t k0 = 0; k0 < K; k0 += BK) {
        // A tile: BM x BK elements, 4 per thread
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = threadIdx.x + r * 256;
            const int ii = e % BM, kk = e / BM;
            const int64_t gi = m0 + ii, gk = k0 + kk;
            float v = 0.0f;
            if (gi < M && gk < K) v = a_elem<FMT, EXACT>(a, wr, ta ? (gk + gi * lda) : (gi + gk * lda));
            As[kk][ii] = v;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = threadIdx.x + r * 256;
            const int kk = e % BK, jj = e / BK;
            const int64_t gj = n0 + jj, gk = k0 + kk;
            float v = 0.0f;
            if (gj < N && gk < K) v = tb ? b[gj + gk * ldb] : b[gk + gj * ldb];
            Bs[kk][jj] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; ++kk) {
            ACC av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = static_cast<ACC>(As[kk][ty * 4 + i]);
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = static_cast<ACC>(Bs[kk][tx * 4 + j]);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(av[i], bv[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t gi = m0 + ty * 4 + i, gj = n0 + tx * 4 + j;
            if (gi < M && gj < N) {
                float* p = c + gi + gj * ldc;
                double r = alpha * static_cast<double>(acc[i][j]);
                if (beta != 0.0) r += beta * static_cast<double>(*p);
                *p = static_cast<float>(r);
            }
        }
}

template <int FMT, bool EXACT, class ACC>
void launch(int64_t M, int64_t N, int64_t K, const void* a, const uint16_t* wr, int64_t lda,
            bool ta, const float* b, int64_t ldb, bool tb, double alpha, double beta, float* c,
            int64_t ldc, hipStream_t s) {
    const dim3 grid(static_cast<unsigned>(ceil_div(M, BM)), static_cast<unsigned>(ceil_div(N, BN)));
    gemm_tile_kernel<FMT, EXACT, ACC><<<grid, 256, 0, s>>>(M, N, K, a, wr, lda, ta, b, ldb, tb,
                                                           alpha, beta, c, ldc);
    TOMOS_LAUNCH_CHECK("gemm_tile_kernel");
}

}  // namespace

// packed-A DEFAULT and MATERIALIZE tiers on matrix cores (gemm_split.hip)
void gemm_packed_mfma(int a_type, int tier, bool ta, bool tb, double alpha, int64_t M, int64_t N, int64_t K,
                      const uint16_t* wh, const uint16_t* wr, int64_t lda, const float* b, int64_t ldb,
                      double beta, float* c, int64_t ldc, hipStream_t s);

void gemm(int a_type, int tier, int trans_a, int trans_b, double alpha, int64_t M, int64_t N,
          int64_t K, const void* a, const uint16_t* wr, int64_t lda, const float* b, int64_t ldb,
          double beta, float* c, int64_t ldc, hipStream_t s) {
    require(M >= 0 && N >= 0 && K >= 0, "gemm: negative size");
    require(tier == TOMOS_TIER_DEFAULT || tier == TOMOS_TIER_FAST || tier == TOMOS_TIER_EXACT ||
                tier == TOMOS_TIER_MATERIALIZE,
            "gemm: bad tier");
    const bool packed = a_type == TOMOS_FP32_16 || a_type == TOMOS_FP32_BF16;
    require(tier != TOMOS_TIER_MATERIALIZE || packed, "gemm: MATERIALIZE needs a packed A");
    const bool ta = trans_a == TOMOS_TRANSPOSE, tb = trans_b == TOMOS_TRANSPOSE;
    require(lda >= std::max<int64_t>(ta ? K : M, 1), "gemm: lda too small");
    require(ldb >= std::max<int64_t>(tb ? N : K, 1), "gemm: ldb too small");
    require(ldc >= std::max<int64_t>(M, 1), "gemm: ldc too small");
    require(ceil_div(N, BN) <= 65535, "gemm: N too large");
    if (M == 0 || N == 0) return;
    require(c != nullptr, "gemm: null C");
    if (K > 0) require(a && b, "gemm: null A or B");
    if (packed && (tier == TOMOS_TIER_DEFAULT || tier == TOMOS_TIER_MATERIALIZE)) {
        require(tier == TOMOS_TIER_DEFAULT || wr || K == 0, "gemm: MATERIALIZE needs wr");
        gemm_packed_mfma(a_type, tier, ta, tb, alpha, M, N, K, static_cast<const uint16_t*>(a), wr, lda, b, ldb,
                         beta, c, ldc, s);
        return;
    }
    const bool exact = tier == TOMOS_TIER_EXACT;
    const bool acc64 = tier != TOMOS_TIER_DEFAULT || a_type == TOMOS_F64;
    switch (a_type) {
        case TOMOS_F32:
            if (acc64) launch<FMT_F32, false, double>(M, N, K, a, wr, lda, ta, b, ldb, tb, alpha, beta, c, ldc, s);
            else launch<FMT_F32, false, float>(M, N, K, a, wr, lda, ta, b, ldb, tb, alpha, beta, c, ldc, s);
            break;
        case TOMOS_F64:
            launch<FMT_F64, false, double>(M, N, K, a, wr, lda, ta, b, ldb, tb, alpha, beta, c, ldc, s);
            break;
        case TOMOS_FP32_16:
            require(!exact || wr || K == 0, "gemm: EXACT needs wr");
            if (exact) launch<FMT_P16, true, double>(M, N, K, a, wr, lda, ta, b, ldb, tb, alpha, beta, c, ldc, s);
            else if (acc64) launch<FMT_P16, false, double>(M, N, K, a, wr, lda, ta, b, ldb, tb, alpha, beta, c, ldc, s);
            else launch<FMT_P16, false, float>(M, N, K, a, wr, lda, ta, b, ldb, tb, alpha, beta, c, ldc, s);
            break;
        case TOMOS_FP32_BF16:
            require(!exact || wr || K == 0, "gemm: EXACT needs wr");
            if (exact) launch<FMT_PBF, true, double>(M, N, K, a, wr, lda, ta, b, ldb, tb, alpha, beta, c, ldc, s);
            else if (acc64) launch<FMT_PBF, false, double>(M, N, K, a, wr, lda, ta, b, ldb, tb, alpha, beta, c, ldc, s);
            else launch<FMT_PBF, false, float>(M, N, K, a, wr, lda, ta, b, ldb, tb, alpha, beta, c, ldc, s);
            break;
        default:
            throw Error(kErrUnsupported, "gemm: unsupported a_type");
    }
}

}  // namespace tomos_hip

using namespace tomos_hip;

extern "C" int tomos_gemm(int a_type, int tier, int trans_a, int trans_b, double alpha, int64_t M,
                          int64_t N, int64_t K, const void* a, const uint16_t* wr, int64_t lda,
                          const float* b, int64_t ldb, double beta, float* c, int64_t ldc,
                          tomos_stream_t s) {
    return guarded([&] {
        gemm(a_type, tier, trans_a, trans_b, alpha, M, N, K, a, wr, lda, b, ldb, beta, c, ldc, s);
    });
}
