// One pass over A per LSQR iteration (gfx950).
//
// LSQR's step 1 needs u_new = A p - alpha u followed by A^T u_new / beta,
// where beta = |u_new| is a global reduction.  Because
// A^T (u_raw / beta) = (A^T u_raw) / beta, the transpose product can use the
// unnormalised u_raw of the same rows and the 1/beta scale is applied when
// the partials are summed.  So a block that owns a tile of rows computes
// u_raw for them (phase 1) and immediately the tile's contribution to
// A^T u_raw (phase 2) from the SAME register-resident tile: A is streamed
// from HBM once per iteration instead of twice.
//
// Tile: 8 lanes per column, each lane one 16-byte vector (8 packed rows or
// 4 fp32 rows), so a tile is TR = 64 (packed) or 32 (fp32) rows by all n
// columns; column slot c = tid/8 takes columns r*32 + c for rounds
// r < ROUNDS (n <= 32*ROUNDS).  The whole tile sits in VGPRs (4*ROUNDS per
// lane).
//   phase 1: per-lane row sums over its columns, then a 3-stage transpose
//            reduction across the 8 column slots of a wave and an LDS sum
//            over the 4 waves give u_raw for the tile's rows;
//   phase 2: per-lane column sums over its 8 rows, then a 3-stage transpose
//            reduction across the 8 row groups; every column ends in exactly
//            one lane, which keeps it in a VGPR across all tiles of the
//            block and writes one partial per block at the end.
// Deterministic: fixed reduction order, no atomics.
#include <algorithm>

#include "codec.hpp"
#include "common.hpp"
#include "lsqr_fused.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

constexpr int kThreads = 256;
enum { FMT_F32 = 0, FMT_PBF = 1, FMT_P16 = 2 };

template <int FMT>
constexpr int rows_per_lane() {
    return FMT == FMT_F32 ? 4 : 8;
}

template <int FMT>
__device__ __forceinline__ void decode_vec(const uint4& t, float (&out)[rows_per_lane<FMT>()]) {
    if constexpr (FMT == FMT_F32) {
        out[0] = __uint_as_float(t.x);
        out[1] = __uint_as_float(t.y);
        out[2] = __uint_as_float(t.z);
        out[3] = __uint_as_float(t.w);
    } else {
        const uint32_t w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if constexpr (FMT == FMT_PBF) {
                out[2 * k] = __uint_as_float(w[k] << 16);
                out[2 * k + 1] = __uint_as_float(w[k] & 0xFFFF0000u);
            } else {
                out[2 * k] = static_cast<float>(__builtin_bit_cast(
                    _Float16, static_cast<uint16_t>(codec::fp32_16_half_bits(w[k] & 0xFFFFu))));
                out[2 * k + 1] = static_cast<float>(__builtin_bit_cast(
                    _Float16, static_cast<uint16_t>(codec::fp32_16_half_bits(w[k] >> 16))));
            }
        }
    }
}

// Transposing reduction over the lane bits OFF, OFF/2, ..., LO of L values:
// at each stage a lane keeps the half of its values selected by its bit and
// adds the partner's copy of that half.  Stages run while L > 1; any
// remaining offsets down to LO reduce the single value.
template <int L, int OFF, int LO>
struct LaneTR {
    static __device__ __forceinline__ void run(float* v, int lane) {
        if constexpr (OFF < LO) {
            return;
        } else if constexpr (L == 1) {
            v[0] += __shfl_xor(v[0], OFF);
            LaneTR<1, OFF / 2, LO>::run(v, lane);
        } else {
            const bool upper = (lane & OFF) != 0;
#pragma unroll
            for (int t = 0; t < L / 2; ++t) {
                const float send = upper ? v[t] : v[t + L / 2];
                const float keep = upper ? v[t + L / 2] : v[t];
                v[t] = keep + __shfl_xor(send, OFF);
            }
            LaneTR<L / 2, OFF / 2, LO>::run(v, lane);
        }
    }
};
template <int L, int LO>
struct LaneTR<L, 0, LO> {
    static __device__ __forceinline__ void run(float*, int) {}
};

template <int FMT, int ROUNDS>
__global__ __launch_bounds__(kThreads) void lsqr_fused_kernel(
    int64_t m, int64_t n, const void* __restrict__ a, int64_t lda, const void* __restrict__ p,
    bool p64, void* __restrict__ u, bool u64, const double* __restrict__ u_coef, int64_t tiles_per_block,
    double* __restrict__ sumsq, float* __restrict__ tpart, const int* __restrict__ skip) {
    if (skip && *skip) return;
    constexpr int RPL = rows_per_lane<FMT>();
    constexpr int TR = 8 * RPL;
    constexpr int ESZ = FMT == FMT_F32 ? 4 : 2;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int g = tid & 7;    // row group inside the tile
    const int c = tid >> 3;   // column slot 0..31
    const char* abase = static_cast<const char*>(a);

    // p for this lane's columns, kept for the whole block
    float pv[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int64_t j = static_cast<int64_t>(r) * 32 + c;
        pv[r] = j < n ? (p64 ? static_cast<float>(static_cast<const double*>(p)[j])
                             : static_cast<const float*>(p)[j])
                      : 0.0f;
    }
    const double cu = *u_coef;
    constexpr int TOUT = ROUNDS / 8;  // column partials owned by this lane
    float tacc[TOUT];
#pragma unroll
    for (int i = 0; i < TOUT; ++i) tacc[i] = 0.0f;

    __shared__ float red[4][64];
    __shared__ float ush[TR];
    double sq = 0.0;

    const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * tiles_per_block;
    const int64_t ntiles = (m + TR - 1) / TR;
    const int64_t tile1 = min(ntiles, tile0 + tiles_per_block);
    for (int64_t tile = tile0; tile < tile1; ++tile) {
        const int64_t row0 = tile * TR;
        const int64_t my_row = row0 + g * RPL;
        const bool in = my_row < m;
        const int nvalid = in ? static_cast<int>(min<int64_t>(RPL, m - my_row)) : 0;
        uint4 t[ROUNDS];
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            const int64_t j = static_cast<int64_t>(r) * 32 + c;
            if (in && j < n)
                t[r] = *reinterpret_cast<const uint4*>(abase + (my_row + j * lda) * ESZ);
            else
                t[r] = make_uint4(0u, 0u, 0u, 0u);
        }
        // phase 1: row sums over this lane's columns
        float acc[RPL];
#pragma unroll
        for (int k = 0; k < RPL; ++k) acc[k] = 0.0f;
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            float av[RPL];
            decode_vec<FMT>(t[r], av);
#pragma unroll
            for (int k = 0; k < RPL; ++k) acc[k] = fmaf(av[k], pv[r], acc[k]);
        }
        // reduce over the 8 column slots of this wave (lane bits 3..5)
        LaneTR<RPL, 32, 8>::run(acc, lane);
        // lane holds row k of group g: k = lane bits (5,4,3) for RPL 8,
        // (5,4) for RPL 4
        {
            const int k = RPL == 8 ? ((lane >> 3) & 7) : ((lane >> 4) & 3);
            const bool writer = RPL == 8 ? true : ((lane & 8) == 0);
            if (writer) red[wave][g * RPL + k] = acc[0];
        }
        __syncthreads();
        if (tid < TR) {
            const int64_t row = row0 + tid;
            float un = 0.0f;
            if (row < m) {
                const float s = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
                double uo = u64 ? static_cast<double*>(u)[row] : static_cast<double>(static_cast<float*>(u)[row]);
                const double r = static_cast<double>(s) + cu * uo;
                if (u64) {
                    static_cast<double*>(u)[row] = r;
                    un = static_cast<float>(r);
                    sq += r * r;
                } else {
                    un = static_cast<float>(r);
                    static_cast<float*>(u)[row] = un;
                    sq += static_cast<double>(un) * un;
                }
            }
            ush[tid] = un;
        }
        __syncthreads();
        // phase 2: column sums over this lane's rows with u_raw
        float ur[RPL];
#pragma unroll
        for (int k = 0; k < RPL; ++k) ur[k] = k < nvalid ? ush[g * RPL + k] : 0.0f;
        float cs[ROUNDS];
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            float av[RPL];
            decode_vec<FMT>(t[r], av);
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < RPL; ++k) s = fmaf(k < nvalid ? av[k] : 0.0f, ur[k], s);
            cs[r] = s;
        }
        // reduce over the 8 row groups (lane bits 0..2); lane keeps TOUT columns
        LaneTR<ROUNDS, 4, 1>::run(cs, lane);
#pragma unroll
        for (int i = 0; i < TOUT; ++i) tacc[i] += cs[i];
        __syncthreads();  // red/ush reuse in the next tile
    }
    // column ids of tacc: r = base + i, base = b2*(R/2) + b1*(R/4) + b0*(R/8)
    const int base = ((g >> 2) & 1) * (ROUNDS / 2) + ((g >> 1) & 1) * (ROUNDS / 4) + (g & 1) * (ROUNDS / 8);
#pragma unroll
    for (int i = 0; i < TOUT; ++i) {
        const int64_t j = static_cast<int64_t>(base + i) * 32 + c;
        if (j < n) tpart[static_cast<int64_t>(blockIdx.x) * n + j] = tacc[i];
    }
    // sum of squares of this block's u_raw rows
    for (int o = 32; o >= 1; o >>= 1) sq += __shfl_xor(sq, o);
    __shared__ double sqw[4];
    if (lane == 0) sqw[wave] = sq;
    __syncthreads();
    if (tid == 0) sumsq[blockIdx.x] = (sqw[0] + sqw[1]) + (sqw[2] + sqw[3]);
}

// level-1 reduction of the block partials: out[s][j] = sum over the blocks
// of split s (fixed order)
__global__ __launch_bounds__(kThreads) void fused_tsum_kernel(int64_t n, int nblocks, int splits,
                                                              const float* __restrict__ tpart,
                                                              double* __restrict__ out,
                                                              const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int64_t j = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    const int sp = blockIdx.y;
    if (j >= n) return;
    const int per = (nblocks + splits - 1) / splits;
    const int b0 = sp * per, b1 = min(nblocks, b0 + per);
    double acc = 0.0;
    for (int b = b0; b < b1; ++b) acc += static_cast<double>(tpart[static_cast<int64_t>(b) * n + j]);
    out[static_cast<int64_t>(sp) * n + j] = acc;
}

template <int FMT>
int rounds_for(int64_t n) {
    (void)FMT;
    if (n <= 256) return 8;
    if (n <= 512) return 16;
    if (n <= 1024) return 32;
    if (n <= 1536) return 48;
    return 0;
}

}  // namespace

bool lsqr_fused_supported(int a_type, int tier, int64_t m, int64_t n, int64_t lda, const void* a) {
    if (tier != TOMOS_TIER_DEFAULT) return false;
    if (a_type != TOMOS_F32 && a_type != TOMOS_FP32_BF16 && a_type != TOMOS_FP32_16) return false;
    const int vec = a_type == TOMOS_F32 ? 4 : 8;
    if (n < 1 || n > 1536 || m < 1) return false;
    return lda % vec == 0 && reinterpret_cast<uintptr_t>(a) % 16 == 0;
}

FusedPlan lsqr_fused_plan(int a_type, int64_t m, int64_t n) {
    FusedPlan p;
    const int64_t tr = a_type == TOMOS_F32 ? 32 : 64;
    const int64_t ntiles = ceil_div(std::max<int64_t>(m, 1), tr);
    const int64_t want = 512;  // two blocks per CU on a 256-CU MI355X
    p.tiles_per_block = std::max<int64_t>(1, ceil_div(ntiles, want));
    p.nblocks = static_cast<int>(ceil_div(ntiles, p.tiles_per_block));
    p.splits = std::min(16, std::max(1, p.nblocks / 16));
    p.tpart_bytes = static_cast<size_t>(p.nblocks) * n * 4;
    p.tsum_bytes = static_cast<size_t>(p.splits) * n * 8;
    (void)n;
    return p;
}

void lsqr_fused_launch(int a_type, int64_t m, int64_t n, const void* a, int64_t lda, const void* p,
                       bool p64, void* u, bool u64, const double* u_coef, double* sumsq,
                       float* tpart, double* tsum, const FusedPlan& plan, const int* skip,
                       hipStream_t s) {
    const dim3 grid(static_cast<unsigned>(plan.nblocks));
#define FUSED_CASE(F, R)                                                                        \
    lsqr_fused_kernel<F, R><<<grid, kThreads, 0, s>>>(m, n, a, lda, p, p64, u, u64, u_coef,      \
                                                      plan.tiles_per_block, sumsq, tpart, skip)
    const int rounds = rounds_for<0>(n);
    require(rounds > 0, "lsqr_fused: n too large");
    if (a_type == TOMOS_FP32_BF16) {
        if (rounds == 8) FUSED_CASE(FMT_PBF, 8);
        else if (rounds == 16) FUSED_CASE(FMT_PBF, 16);
        else if (rounds == 32) FUSED_CASE(FMT_PBF, 32);
        else FUSED_CASE(FMT_PBF, 48);
    } else if (a_type == TOMOS_FP32_16) {
        if (rounds == 8) FUSED_CASE(FMT_P16, 8);
        else if (rounds == 16) FUSED_CASE(FMT_P16, 16);
        else if (rounds == 32) FUSED_CASE(FMT_P16, 32);
        else FUSED_CASE(FMT_P16, 48);
    } else {
        if (rounds == 8) FUSED_CASE(FMT_F32, 8);
        else if (rounds == 16) FUSED_CASE(FMT_F32, 16);
        else if (rounds == 32) FUSED_CASE(FMT_F32, 32);
        else FUSED_CASE(FMT_F32, 48);
    }
#undef FUSED_CASE
    TOMOS_LAUNCH_CHECK("lsqr_fused_kernel");
    fused_tsum_kernel<<<dim3(static_cast<unsigned>(ceil_div(n, kThreads)), static_cast<unsigned>(plan.splits)),
                        kThreads, 0, s>>>(n, plan.nblocks, plan.splits, tpart, tsum, skip);
    TOMOS_LAUNCH_CHECK("fused_tsum_kernel");
}

}  // namespace tomos_hip
