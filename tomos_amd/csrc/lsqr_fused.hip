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
#include <map>
#include <mutex>
#include <vector>

#include "codec.hpp"
#include "common.hpp"
#include "lsqr_fused.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

constexpr int kThreads = 256;
constexpr int64_t kFusedMaxN = 2048;
constexpr int kMaxSplits = 16;
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
template <int L, int OFF, int LO, class T = float>
struct LaneTR {
    static __device__ __forceinline__ void run(T* v, int lane) {
        if constexpr (OFF < LO) {
            return;
        } else if constexpr (L == 1) {
            v[0] += __shfl_xor(v[0], OFF);
            LaneTR<1, OFF / 2, LO, T>::run(v, lane);
        } else {
            const bool upper = (lane & OFF) != 0;
#pragma unroll
            for (int t = 0; t < L / 2; ++t) {
                const T send = upper ? v[t] : v[t + L / 2];
                const T keep = upper ? v[t + L / 2] : v[t];
                v[t] = keep + __shfl_xor(send, OFF);
            }
            LaneTR<L / 2, OFF / 2, LO, T>::run(v, lane);
        }
    }
};
template <int L, int LO, class T>
struct LaneTR<L, 0, LO, T> {
    static __device__ __forceinline__ void run(T*, int) {}
};

// LaneTR for float values without the LDS crossbar (gfx950): the stage on
// lane bit 5 / 4 with v_permlane32_swap / v_permlane16_swap (after the swap
// each lane holds its kept half and the partner's copy of it, and their sum
// is the stage's result), the stage on bit 3 with DPP row_ror:8 (lane ^ 8
// exactly), bits 1 and 0 with quad permutations (lane ^ 2, lane ^ 1).  Bit 2
// has no exact single-instruction DPP partner: with MIRROR (every lower lane
// bit is reduced by the later stages, so the partner only has to differ in
// bit 2 and agree above it) row_half_mirror (lane ^ 7) serves, otherwise the
// crossbar shuffle does.  Same kept halves and per-stage sums as LaneTR.
template <bool ROWS16>
__device__ __forceinline__ float swap_add_f(float lo, float hi) {
    if constexpr (ROWS16) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(lo), __float_as_uint(hi), false, false);
        return __uint_as_float(r[0]) + __uint_as_float(r[1]);
    } else {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(lo), __float_as_uint(hi), false, false);
        return __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
}
template <int OFF, bool MIRROR>
__device__ __forceinline__ float partner_f(float x) {
    constexpr int kCtrl = OFF == 8 ? 0x128 : (OFF == 4 ? 0x141 : (OFF == 2 ? 0x4E : 0xB1));
    if constexpr (OFF == 4 && !MIRROR) {
        return __shfl_xor(x, 4);
    } else {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), kCtrl, 0xF, 0xF, true));
    }
}
template <int L, int OFF, int LO, bool MIRROR>
struct LaneTRf {
    static __device__ __forceinline__ void run(float* v, int lane) {
        if constexpr (OFF < LO || OFF == 0) {
            return;
        } else if constexpr (L == 1) {
            if constexpr (OFF >= 16) v[0] += __shfl_xor(v[0], OFF);
            else v[0] += partner_f<OFF, MIRROR>(v[0]);
            LaneTRf<1, OFF / 2, LO, MIRROR>::run(v, lane);
        } else {
            if constexpr (OFF >= 16) {
#pragma unroll
                for (int t = 0; t < L / 2; ++t) v[t] = swap_add_f<OFF == 16>(v[t], v[t + L / 2]);
            } else {
                const bool upper = (lane & OFF) != 0;
#pragma unroll
                for (int t = 0; t < L / 2; ++t) {
                    const float send = upper ? v[t] : v[t + L / 2];
                    const float keep = upper ? v[t + L / 2] : v[t];
                    v[t] = keep + partner_f<OFF, MIRROR>(send);
                }
            }
            LaneTRf<L / 2, OFF / 2, LO, MIRROR>::run(v, lane);
        }
    }
};

template <int G>
constexpr int log2_small() {
    return G >= 16 ? 4 : (G >= 8 ? 3 : (G >= 4 ? 2 : (G >= 2 ? 1 : 0)));
}

// Block barrier for the LDS exchanges between the waves (red, ush) that
// leaves the next tile's DMA in flight.  A workgroup fence would also wait
// for the DMA (its LDS writes count as outstanding vector-memory work), so
// instead: this wave's own DS operations are waited for (lgkmcnt(0)), the
// compiler is kept from moving memory operations across the barrier, and
// the hardware barrier synchronises the waves.  The DMA writes only the
// slots of the next tile, which nothing reads before its own wait.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
template <bool STAGE>
__device__ __forceinline__ void tile_barrier() {
    if constexpr (STAGE) lds_barrier();
    else __syncthreads();
}

// One block's share of the single-pass product: tiles [tile0, tile1) of an
// m x n matrix; writes the block's sum of u_raw^2 to *sumsq_out and its
// A^T u_raw partial (n floats) to tpart_out.  NTL: non-temporal loads of A.
// STAGE: the tile goes through LDS by DMA (global_load_lds_dwordx4, each
// lane's 16 bytes at its own slot: base + 16 * lane), and the next tile's
// loads are issued before this tile's arithmetic and barriers, so HBM
// requests stay in flight throughout (one tile buffer: NW x ROUNDS KB).
// GROUPED only tells the grouped kernel's instances apart from the
// one-problem kernel's (the HIP host pass rejects one staged instance
// reached from both kinds of kernel); it changes nothing in the code.
template <int FMT, int G, int ROUNDS, int NT, bool NTL, bool STAGE = false, bool GROUPED = false>
__device__ __forceinline__ void fused_body(int64_t m, int64_t n, const void* __restrict__ a,
                                           int64_t lda, const void* __restrict__ p, bool p64,
                                           void* __restrict__ u, bool u64, double cu, int64_t tile0,
                                           int64_t tile1, double* __restrict__ sumsq_out,
                                           float* __restrict__ tpart_out) {
    constexpr int RPL = rows_per_lane<FMT>();
    constexpr int LOG_RPL = log2_small<RPL>();
    constexpr int LOG_G = log2_small<G>();
    constexpr int TR = G * RPL;          // rows per tile
    constexpr int CS = NT / G;           // column slots per round
    constexpr int NW = NT / 64;          // waves per block
    constexpr int ESZ = FMT == FMT_F32 ? 4 : 2;
    static_assert(ROUNDS % G == 0, "ROUNDS must be a multiple of G");
    static_assert(LOG_G + LOG_RPL <= 6, "row groups and row bits must fit a wave");
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int g = tid & (G - 1);  // row group inside the tile
    const int c = tid >> LOG_G;   // column slot
    const char* abase = static_cast<const char*>(a);

    // p for this lane's columns, kept for the whole block
    float pv[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int64_t j = static_cast<int64_t>(r) * CS + c;
        pv[r] = j < n ? (p64 ? static_cast<float>(static_cast<const double*>(p)[j])
                             : static_cast<const float*>(p)[j])
                      : 0.0f;
    }
    constexpr int TOUT = ROUNDS / G;  // column partials owned by this lane
    float tacc[TOUT];
#pragma unroll
    for (int i = 0; i < TOUT; ++i) tacc[i] = 0.0f;

    // row partials and u rows in two copies, by tile parity: a wave may
    // write the next tile's partials while slower waves still read this
    // tile's u rows
    __shared__ float red[2][NW][TR];
    __shared__ float ush[2][TR];
    __shared__ uint4 stage[STAGE ? NW : 1][STAGE ? ROUNDS : 1][64];
    double sq = 0.0;
    // LDS-DMA of tile `tile` into this wave's slots; lanes outside A load
    // the first vector instead (in bounds) and are masked when read back
    auto issue = [&](int64_t tile) {
        const int64_t row = tile * TR + g * RPL;
        const bool ok_row = tile < tile1 && row < m;
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            const int64_t j = static_cast<int64_t>(r) * CS + c;
            const char* src = (ok_row && j < n) ? abase + (row + j * lda) * ESZ : abase;
            __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint4*>(src), &stage[wave][r][0], 16, 0,
                                             NTL ? 2 : 0);
        }
    };
    if constexpr (STAGE) issue(tile0);
    double upre = 0.0;  // STAGE: this tile's u_raw row (threads < TR)

    int par = 0;
    for (int64_t tile = tile0; tile < tile1; ++tile, par ^= 1) {
        const int64_t row0 = tile * TR;
        const int64_t my_row = row0 + g * RPL;
        const bool in = my_row < m;
        const int nvalid = in ? static_cast<int>(min<int64_t>(RPL, m - my_row)) : 0;
        const bool tile_full = row0 + TR <= m;  // every lane has RPL rows
        uint4 t[ROUNDS];
        if constexpr (STAGE) {
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's DMA has landed
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) t[r] = stage[wave][r][lane];
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): slots read before they are refilled
            // this tile's u_raw rows, loaded before the DMA so that waiting
            // for them does not wait for the next tile
            if (tid < TR && row0 + tid < m && cu != 0.0)
                upre = u64 ? static_cast<double*>(u)[row0 + tid]
                           : static_cast<double>(static_cast<float*>(u)[row0 + tid]);
            issue(tile + 1);
            // lanes outside A hold the stand-in vector: clear them (bit masks;
            // only the last tile and the rounds that reach past column n have
            // such lanes, a block-uniform test)
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) {
                if (!tile_full || static_cast<int64_t>(r + 1) * CS > n) {
                    const int64_t j = static_cast<int64_t>(r) * CS + c;
                    const uint32_t keep = (in && j < n) ? 0xFFFFFFFFu : 0u;
                    t[r].x &= keep;
                    t[r].y &= keep;
                    t[r].z &= keep;
                    t[r].w &= keep;
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) {
                const int64_t j = static_cast<int64_t>(r) * CS + c;
                if (in && j < n)
                    t[r] = load_a16<NTL>(abase + (my_row + j * lda) * ESZ);
                else
                    t[r] = make_uint4(0u, 0u, 0u, 0u);
            }
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
        // reduce over the column slots of this wave (lane bits LOG_G..5):
        // halving stages on bits 5.. leave row k = lane bits (5..6-LOG_RPL),
        // plain stages on the bits below make the lanes of a row agree
        LaneTRf<RPL, 32, G, false>::run(acc, lane);
        {
            const int k = (lane >> (6 - LOG_RPL)) & (RPL - 1);
            constexpr int kDupMask = ((1 << (6 - LOG_RPL)) - 1) & ~(G - 1);
            if ((lane & kDupMask) == 0) red[par][wave][g * RPL + k] = acc[0];
        }
        tile_barrier<STAGE>();
        if (tid < TR) {
            const int64_t row = row0 + tid;
            float un = 0.0f;
            if (row < m) {
                float s = red[par][0][tid];
#pragma unroll
                for (int w = 1; w < NW; ++w) s += red[par][w][tid];
                double r = static_cast<double>(s);
                if (cu != 0.0) {  // u is not read for a zero coefficient
                    if constexpr (STAGE) r += cu * upre;
                    else r += cu * (u64 ? static_cast<double*>(u)[row]
                                        : static_cast<double>(static_cast<float*>(u)[row]));
                }
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
            ush[par][tid] = un;
        }
        tile_barrier<STAGE>();
        // phase 2: column sums over this lane's rows with u_raw
        float ur[RPL];
#pragma unroll
        for (int k = 0; k < RPL; ++k) ur[k] = k < nvalid ? ush[par][g * RPL + k] : 0.0f;
        float cs[ROUNDS];
        if (tile_full) {
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) {
                float av[RPL];
                decode_vec<FMT>(t[r], av);
                float s = 0.0f;
#pragma unroll
                for (int k = 0; k < RPL; ++k) s = fmaf(av[k], ur[k], s);
                cs[r] = s;
            }
        } else {
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) {
                float av[RPL];
                decode_vec<FMT>(t[r], av);
                float s = 0.0f;
#pragma unroll
                for (int k = 0; k < RPL; ++k) s = fmaf(k < nvalid ? av[k] : 0.0f, ur[k], s);
                cs[r] = s;
            }
        }
        // reduce over the G row groups (lane bits 0..LOG_G-1); lane keeps TOUT columns
        LaneTRf<ROUNDS, G / 2, 1, true>::run(cs, lane);
#pragma unroll
        for (int i = 0; i < TOUT; ++i) tacc[i] += cs[i];
        // no barrier here: the next tile writes the other copy, and this
        // copy is rewritten only after every wave has passed the next
        // tile's first barrier, that is, finished reading it
    }
    // column ids of tacc: r = base + i, the halving stage on lane bit b kept
    // the upper half for lanes with bit b set
    int base = 0;
#pragma unroll
    for (int b = 0; b < LOG_G; ++b) base += ((g >> b) & 1) * (ROUNDS >> (LOG_G - b));
#pragma unroll
    for (int i = 0; i < TOUT; ++i) {
        const int64_t j = static_cast<int64_t>(base + i) * CS + c;
        if (j < n) tpart_out[j] = tacc[i];
    }
    // sum of squares of this block's u_raw rows
    for (int o = 32; o >= 1; o >>= 1) sq += __shfl_xor(sq, o);
    __shared__ double sqw[NW];
    if (lane == 0) sqw[wave] = sq;
    __syncthreads();
    if (tid == 0) {
        double t = sqw[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) t += sqw[w];
        *sumsq_out = t;
    }
}

// Variant with long column runs.  A tile is RL * RPL rows (32 row lanes x
// 8 packed rows = 256 rows: 512 contiguous bytes per column) by all n
// columns; 32 column slots of 32 row lanes each (1024 threads); slot s
// owns columns s, s + 32, s + 64, ...  The tile is too large for
// registers, so phase 2 reads it a second time right after phase 1 (the
// re-read is meant to be served by L2 / the MALL, not HBM):
//   phase 1: each lane sums its 8 rows over its slot's columns, CPB column
//            loads in flight per batch; the 32 slots combine in LDS -> u_raw;
//   phase 2: per batch of CPB columns each lane forms a . u over its 8
//            rows, a transposing reduction over the 32 row lanes leaves one
//            column per lane, added into an LDS column accumulator that
//            the column's lane owns (no atomics, fixed order).
// The loops over column batches are not unrolled: registers stay bounded
// whatever n (up to 2048).
template <int FMT, bool NTL>
__device__ __forceinline__ void fused2_body(int64_t m, int64_t n, const void* __restrict__ a,
                                            int64_t lda, const void* __restrict__ p, bool p64,
                                            void* __restrict__ u, bool u64, double cu, int64_t tile0,
                                            int64_t tile1, double* __restrict__ sumsq_out,
                                            float* __restrict__ tpart_out) {
    constexpr int NT = 1024, RL = 32, NS = NT / RL, CPB = 8;
    constexpr int RPL = rows_per_lane<FMT>();
    constexpr int TR = RL * RPL;
    constexpr int ESZ = FMT == FMT_F32 ? 4 : 2;
    constexpr int kMaxN = 2048;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int rl = tid & (RL - 1);  // row lane
    const int slot = tid / RL;      // column slot
    const char* abase = static_cast<const char*>(a);
    const int64_t step = static_cast<int64_t>(NS) * lda * ESZ;  // bytes between a slot's columns
    const int nbatch = static_cast<int>((n + NS * CPB - 1) / (NS * CPB));
    __shared__ float pvs[kMaxN];
    __shared__ float colacc[kMaxN];
    __shared__ float red[NS][TR];
    __shared__ float ush[TR];
    for (int64_t j = tid; j < n; j += NT) {
        pvs[j] = p64 ? static_cast<float>(static_cast<const double*>(p)[j]) : static_cast<const float*>(p)[j];
        colacc[j] = 0.0f;
    }
    double sq = 0.0;
    // after the row-lane reduction, lane bits 4..2 select the batch column
    // (halving stages on bits 4, 3, 2); bits 1..0 hold duplicates
    const int sub = (((rl >> 4) & 1) << 2) | (((rl >> 3) & 1) << 1) | ((rl >> 2) & 1);
    const bool owner = (rl & 3) == 0;
    __syncthreads();
    for (int64_t tile = tile0; tile < tile1; ++tile) {
        const int64_t row0 = tile * TR;
        const int64_t my_row = row0 + rl * RPL;
        const bool in = my_row < m;
        const int nvalid = in ? static_cast<int>(min<int64_t>(RPL, m - my_row)) : 0;
        const char* colp = abase + (static_cast<int64_t>(slot) * lda + my_row) * ESZ;
        float acc[RPL];
#pragma unroll
        for (int k = 0; k < RPL; ++k) acc[k] = 0.0f;
        for (int bt = 0; bt < nbatch; ++bt) {
            uint4 t[CPB];
#pragma unroll
            for (int q = 0; q < CPB; ++q) {
                const int r = bt * CPB + q;
                const int64_t j = slot + static_cast<int64_t>(NS) * r;
                t[q] = (in && j < n) ? load_a16<NTL>(colp + r * step)
                                     : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int q = 0; q < CPB; ++q) {
                const int64_t j = slot + static_cast<int64_t>(NS) * (bt * CPB + q);
                const float pj = j < n ? pvs[j] : 0.0f;
                float av[RPL];
                decode_vec<FMT>(t[q], av);
#pragma unroll
                for (int k = 0; k < RPL; ++k) acc[k] = fmaf(av[k], pj, acc[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < RPL; ++k) red[slot][rl * RPL + k] = acc[k];
        __syncthreads();
        if (tid < TR) {
            const int64_t row = row0 + tid;
            float un = 0.0f;
            if (row < m) {
                float sacc = red[0][tid];
                for (int w = 1; w < NS; ++w) sacc += red[w][tid];
                double r = static_cast<double>(sacc);
                if (cu != 0.0)  // u is not read for a zero coefficient
                    r += cu * (u64 ? static_cast<double*>(u)[row]
                                   : static_cast<double>(static_cast<float*>(u)[row]));
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
        float ur[RPL];
#pragma unroll
        for (int k = 0; k < RPL; ++k) ur[k] = k < nvalid ? ush[rl * RPL + k] : 0.0f;
        for (int bt = 0; bt < nbatch; ++bt) {
            uint4 t[CPB];
#pragma unroll
            for (int q = 0; q < CPB; ++q) {
                const int r = bt * CPB + q;
                const int64_t j = slot + static_cast<int64_t>(NS) * r;
                t[q] = (in && j < n) ? load_a16<NTL>(colp + r * step)
                                     : make_uint4(0u, 0u, 0u, 0u);
            }
            float cs[CPB];
#pragma unroll
            for (int q = 0; q < CPB; ++q) {
                float av[RPL];
                decode_vec<FMT>(t[q], av);
                float sacc = 0.0f;
#pragma unroll
                for (int k = 0; k < RPL; ++k) sacc = fmaf(av[k], ur[k], sacc);
                cs[q] = sacc;
            }
            LaneTR<CPB, 16, 1>::run(cs, lane);
            const int64_t j = slot + static_cast<int64_t>(NS) * (bt * CPB + sub);
            if (owner && j < n) colacc[j] += cs[0];
        }
        // red / ush are rewritten only after the next tile's first barrier
    }
    __syncthreads();
    for (int64_t j = tid; j < n; j += NT) tpart_out[j] = colacc[j];
    for (int o = 32; o >= 1; o >>= 1) sq += __shfl_xor(sq, o);
    __shared__ double sqw[NT / 64];
    if (lane == 0) sqw[wave] = sq;
    __syncthreads();
    if (tid == 0) {
        double tt = sqw[0];
        for (int w = 1; w < NT / 64; ++w) tt += sqw[w];
        *sumsq_out = tt;
    }
}

template <int FMT, bool NTL>
__global__ __launch_bounds__(1024) void lsqr_fused2_kernel(
    int64_t m, int64_t n, const void* __restrict__ a, int64_t lda, const void* __restrict__ p,
    bool p64, void* __restrict__ u, bool u64, const double* __restrict__ u_coef, int64_t tiles_per_block,
    double* __restrict__ sumsq, float* __restrict__ tpart, const int* __restrict__ skip) {
    if (skip && *skip) return;
    constexpr int TR = 32 * rows_per_lane<FMT>();
    const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * tiles_per_block;
    const int64_t ntiles = (m + TR - 1) / TR;
    const int64_t tile1 = min(ntiles, tile0 + tiles_per_block);
    fused2_body<FMT, NTL>(m, n, a, lda, p, p64, u, u64, *u_coef, tile0, tile1, sumsq + blockIdx.x,
                        tpart + static_cast<int64_t>(blockIdx.x) * n);
}

template <int FMT>
__global__ __launch_bounds__(1024) void grouped_fused2_kernel(const GroupProb* __restrict__ probs,
                                                              const GroupItem* __restrict__ items) {
    const GroupItem it = items[blockIdx.x];
    const GroupProb& P = probs[it.prob];
    if (P.skip && *P.skip) return;
    fused2_body<FMT, false>(P.m, P.n, P.a, P.lda, P.p, true, P.u, false, *P.u_coef, it.tile0, it.tile1,
                        P.sumsq + it.slot, P.tpart + static_cast<int64_t>(it.slot) * P.n);
}

template <int FMT, int G, int ROUNDS, int NT, bool NTL, bool STAGE>
__global__ __launch_bounds__(NT) void lsqr_fused_kernel(
    int64_t m, int64_t n, const void* __restrict__ a, int64_t lda, const void* __restrict__ p,
    bool p64, void* __restrict__ u, bool u64, const double* __restrict__ u_coef, int64_t tiles_per_block,
    double* __restrict__ sumsq, float* __restrict__ tpart, const int* __restrict__ skip) {
    if (skip && *skip) return;
    constexpr int TR = G * rows_per_lane<FMT>();
    const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * tiles_per_block;
    const int64_t ntiles = (m + TR - 1) / TR;
    const int64_t tile1 = min(ntiles, tile0 + tiles_per_block);
    fused_body<FMT, G, ROUNDS, NT, NTL, STAGE>(m, n, a, lda, p, p64, u, u64, *u_coef, tile0, tile1, sumsq + blockIdx.x,
                               tpart + static_cast<int64_t>(blockIdx.x) * n);
}

// Body of a grouped work item with the fewest column rounds R (G, 2G, 4G,
// ... up to RMAX) that cover the item's problem (P.rounds): a narrow
// problem does not pay for the widest problem's rounds.  The widest form
// stages its tiles through LDS by DMA when G == 8 (as the one-problem
// kernel does); only it does, because each form's stage buffer is LDS of
// its own.
template <int FMT, int G, int R, int RMAX, int NT, bool NTL>
__device__ __forceinline__ void grouped_item(const GroupProb& P, const GroupItem& it) {
    if constexpr (R < RMAX) {
        if (P.rounds <= R) {
            fused_body<FMT, G, R, NT, false>(P.m, P.n, P.a, P.lda, P.p, true, P.u, false, *P.u_coef, it.tile0,
                                             it.tile1, P.sumsq + it.slot,
                                             P.tpart + static_cast<int64_t>(it.slot) * P.n);
            return;
        }
        grouped_item<FMT, G, (2 * R < RMAX ? 2 * R : RMAX), RMAX, NT, NTL>(P, it);
    } else {
        fused_body<FMT, G, RMAX, NT, NTL, G == 8, true>(P.m, P.n, P.a, P.lda, P.p, true, P.u, false, *P.u_coef,
                                                  it.tile0, it.tile1, P.sumsq + it.slot,
                                                  P.tpart + static_cast<int64_t>(it.slot) * P.n);
    }
}

// Many problems in one launch: block b runs work item items[b] (a tile
// range of one problem).  A problem whose skip flag is set is a no-op.
// NTL: non-temporal loads in the widest (LDS-staged) instance, for a batch
// far larger than the MALL (C5 with 128-byte columns: uniform wide batches
// 812 -> 748 us a pass, the C5 solve 30.2 -> 30.0 ms); a batch that fits is
// left cacheable and is reread from the MALL at the next iteration.
template <int FMT, int G, int ROUNDS, int NT, bool NTL>
__global__ __launch_bounds__(NT) void grouped_fused_kernel(const GroupProb* __restrict__ probs,
                                                                 const GroupItem* __restrict__ items) {
    const GroupItem it = items[blockIdx.x];
    const GroupProb& P = probs[it.prob];
    if (P.skip && *P.skip) return;
    grouped_item<FMT, G, G, ROUNDS, NT, NTL>(P, it);
}

// Exact values of the RPL rows of one 16-byte vector: F32 as stored, packed
// formats from the hi word h and the lo word l (wh + wr, exact FP32).
template <int FMT>
__device__ __forceinline__ void decode_exact_vec(const uint4& h, const uint4& l,
                                                 float (&out)[rows_per_lane<FMT>()]) {
    if constexpr (FMT == FMT_F32) {
        decode_vec<FMT>(h, out);
    } else {
        const uint32_t hw[4] = {h.x, h.y, h.z, h.w}, lw[4] = {l.x, l.y, l.z, l.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if constexpr (FMT == FMT_PBF) {
                out[2 * k] = __uint_as_float((hw[k] << 16) | (lw[k] & 0xFFFFu));
                out[2 * k + 1] = __uint_as_float((hw[k] & 0xFFFF0000u) | (lw[k] >> 16));
            } else {
                out[2 * k] = codec::fp32_16_decode_exact(hw[k] & 0xFFFFu, lw[k] & 0xFFFFu);
                out[2 * k + 1] = codec::fp32_16_decode_exact(hw[k] >> 16, lw[k] >> 16);
            }
        }
    }
}

// LSQR restart in one pass over A (lsqr.cpp:441-469 initialize + the first
// A^T u): u_raw = u - A x with the exact values (wh + wr) and FP64 row sums
// (u holds b on entry; x in FP64), |u_raw|^2 per block, and the block
// partials of A_hi^T u_raw from the hi words of the same tile (the
// per-iteration tier).  Tile: 8 row groups x 512 threads x 16 column rounds
// (n <= 1024), both planes in registers, loaded directly.  XZERO (x == 0,
// the first restart of a solve): A x is the zero vector for finite A, so the
// lo plane is not read and u_raw = u.
constexpr int kRsG = 8, kRsNT = 512, kRsR = 16;
constexpr int64_t kRestartMaxN = static_cast<int64_t>(kRsNT / kRsG) * kRsR;

template <int FMT, bool NTL, bool XZERO = false>
__global__ __launch_bounds__(kRsNT) void lsqr_restart_kernel(
    int64_t m, int64_t n, const void* __restrict__ a, const uint16_t* __restrict__ wr, int64_t lda,
    const double* __restrict__ x, void* __restrict__ u, bool u64, int64_t tiles_per_block,
    double* __restrict__ sumsq, float* __restrict__ tpart, const int* __restrict__ skip) {
    if (skip && *skip) return;
    constexpr int G = kRsG, NT = kRsNT, R = kRsR;
    constexpr int RPL = rows_per_lane<FMT>();
    constexpr int LOG_RPL = log2_small<RPL>();
    constexpr int LOG_G = log2_small<G>();
    constexpr int TR = G * RPL, CS = NT / G, NW = NT / 64, TOUT = R / G;
    constexpr int ESZ = FMT == FMT_F32 ? 4 : 2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = tid & (G - 1), c = tid >> LOG_G;
    const char* abase = static_cast<const char*>(a);
    const int64_t ntiles = (m + TR - 1) / TR;
    const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * tiles_per_block;
    const int64_t tile1 = min(ntiles, tile0 + tiles_per_block);
    __shared__ double xs[kRestartMaxN];  // x in LDS (registers go to the two planes)
    for (int j = tid; j < kRestartMaxN; j += NT) xs[j] = j < n ? x[j] : 0.0;
    __syncthreads();
    float tacc[TOUT];
#pragma unroll
    for (int i = 0; i < TOUT; ++i) tacc[i] = 0.0f;
    // row partials and u rows in two copies by tile parity (the next
    // tile writes the other copy)
    __shared__ double red[2][NW][TR];
    __shared__ float ush[2][TR];
    double sq = 0.0;
    int par = 0;
    for (int64_t tile = tile0; tile < tile1; ++tile, par ^= 1) {
        const int64_t row0 = tile * TR;
        const int64_t my_row = row0 + g * RPL;
        const bool in = my_row < m;
        const int nvalid = in ? static_cast<int>(min<int64_t>(RPL, m - my_row)) : 0;
        uint4 th[R], tl[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t j = static_cast<int64_t>(r) * CS + c;
            const bool ok = in && j < n;
            th[r] = ok ? load_a16<NTL>(abase + (my_row + j * lda) * ESZ) : make_uint4(0u, 0u, 0u, 0u);
            if constexpr (FMT != FMT_F32 && !XZERO)
                tl[r] = ok ? load_a16<NTL>(wr + my_row + j * lda) : make_uint4(0u, 0u, 0u, 0u);
            else
                tl[r] = th[r];
        }
        // (A x) rows in FP64
        double acc[RPL];
#pragma unroll
        for (int k = 0; k < RPL; ++k) acc[k] = 0.0;
        if constexpr (!XZERO) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float av[RPL];
                decode_exact_vec<FMT>(th[r], tl[r], av);
                const double xr = xs[r * CS + c];
#pragma unroll
                for (int k = 0; k < RPL; ++k) acc[k] = fma(static_cast<double>(av[k]), xr, acc[k]);
            }
        }
        LaneTR<RPL, 32, G, double>::run(acc, lane);
        {
            const int k = (lane >> (6 - LOG_RPL)) & (RPL - 1);
            constexpr int kDupMask = ((1 << (6 - LOG_RPL)) - 1) & ~(G - 1);
            if ((lane & kDupMask) == 0) red[par][wave][g * RPL + k] = acc[0];
        }
        __syncthreads();
        if (tid < TR) {
            const int64_t row = row0 + tid;
            float un = 0.0f;
            if (row < m) {
                double sx = red[par][0][tid];
#pragma unroll
                for (int w = 1; w < NW; ++w) sx += red[par][w][tid];
                if (u64) {
                    const double r = static_cast<double*>(u)[row] - sx;
                    static_cast<double*>(u)[row] = r;
                    un = static_cast<float>(r);
                    sq += r * r;
                } else {
                    un = static_cast<float>(static_cast<double>(static_cast<float*>(u)[row]) - sx);
                    static_cast<float*>(u)[row] = un;
                    sq += static_cast<double>(un) * un;
                }
            }
            ush[par][tid] = un;
        }
        __syncthreads();
        // A_hi^T u_raw over this lane's rows
        float ur[RPL];
#pragma unroll
        for (int k = 0; k < RPL; ++k) ur[k] = k < nvalid ? ush[par][g * RPL + k] : 0.0f;
        float cs[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float av[RPL];
            decode_vec<FMT>(th[r], av);
            float sacc = 0.0f;
#pragma unroll
            for (int k = 0; k < RPL; ++k) sacc = fmaf(k < nvalid ? av[k] : 0.0f, ur[k], sacc);
            cs[r] = sacc;
        }
        LaneTRf<R, G / 2, 1, true>::run(cs, lane);
#pragma unroll
        for (int i = 0; i < TOUT; ++i) tacc[i] += cs[i];
        // no barrier here: this copy is rewritten only after every wave has
        // passed the next tile's first barrier, that is, finished reading it
    }
    int base = 0;
#pragma unroll
    for (int bb = 0; bb < LOG_G; ++bb) base += ((g >> bb) & 1) * (R >> (LOG_G - bb));
    float* tp = tpart + static_cast<int64_t>(blockIdx.x) * n;
#pragma unroll
    for (int i = 0; i < TOUT; ++i) {
        const int64_t j = static_cast<int64_t>(base + i) * CS + c;
        if (j < n) tp[j] = tacc[i];
    }
    for (int o = 32; o >= 1; o >>= 1) sq += __shfl_xor(sq, o);
    __shared__ double sqw[NW];
    if (lane == 0) sqw[wave] = sq;
    __syncthreads();
    if (tid == 0) {
        double tt = sqw[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) tt += sqw[w];
        sumsq[blockIdx.x] = tt;
    }
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

// Instantiated launch shapes: (row groups G, threads NT, column rounds R);
// a lane covers R columns, a round CS = NT / G columns.
struct Shape {
    int g, nt, r;
};
constexpr Shape kShapes[] = {
    {2, 256, 2}, {2, 256, 4}, {2, 256, 6}, {2, 256, 8}, {2, 256, 12}, {2, 256, 16},
    {4, 256, 4}, {4, 256, 8}, {4, 256, 12}, {4, 256, 16},
    {8, 256, 8}, {8, 256, 16}, {8, 256, 24}, {8, 256, 32},
    {4, 512, 8}, {4, 512, 16},
    {8, 512, 8}, {8, 512, 16},
    {32, 1024, 64},  // long column runs, tile read twice (up to 64 x 32 = 2048 columns)
};

// Smallest instantiated R covering n for (g, nt); 0 when none.
int rounds_for(int g, int nt, int64_t n) {
    int best = 0;
    for (const Shape& sh : kShapes) {
        if (sh.g != g || sh.nt != nt) continue;
        if (static_cast<int64_t>(sh.r) * (nt / g) >= n && (best == 0 || sh.r < best)) best = sh.r;
    }
    return best;
}

#define TOMOS_FUSED_SHAPES(X)                                                                         \
    X(2, 256, 2) X(2, 256, 4) X(2, 256, 6) X(2, 256, 8) X(2, 256, 12) X(2, 256, 16)                   \
    X(4, 256, 4) X(4, 256, 8) X(4, 256, 12) X(4, 256, 16)                                               \
    X(8, 256, 8) X(8, 256, 16) X(8, 256, 24) X(8, 256, 32)                                              \
    X(4, 512, 8) X(4, 512, 16)                                                                          \
    X(8, 512, 8) X(8, 512, 16)

// The kernel instance for (format, shape), or nullptr.
template <int FMT, bool NTL>
const void* fused_kernel_ptr(int g, int nt, int rounds) {
    if (g == 32 && nt == 1024)
        return rounds > 0 ? reinterpret_cast<const void*>(lsqr_fused2_kernel<FMT, NTL>) : nullptr;
// 8 row groups (128-byte column runs) stage tiles through LDS by DMA
#define FUSED_CASE(G, NT, R)                                                                      \
    if (g == G && nt == NT && rounds == R)                                                        \
        return reinterpret_cast<const void*>(lsqr_fused_kernel<FMT, G, R, NT, NTL, G == 8>);
    TOMOS_FUSED_SHAPES(FUSED_CASE)
#undef FUSED_CASE
    return nullptr;
}

template <int FMT, bool NTL>
const void* grouped_kernel_ptr(int g, int nt, int rounds) {
    if (g == 32 && nt == 1024) return rounds > 0 ? reinterpret_cast<const void*>(grouped_fused2_kernel<FMT>) : nullptr;
#define FUSED_CASE(G, NT, R)                           \
    if (g == G && nt == NT && rounds == R)             \
        return reinterpret_cast<const void*>(grouped_fused_kernel<FMT, G, R, NT, NTL>);
    TOMOS_FUSED_SHAPES(FUSED_CASE)
#undef FUSED_CASE
    return nullptr;
}

const void* grouped_kernel_for(int a_type, int g, int nt, int rounds, bool ntl) {
    if (ntl) {
        if (a_type == TOMOS_FP32_BF16) return grouped_kernel_ptr<FMT_PBF, true>(g, nt, rounds);
        if (a_type == TOMOS_FP32_16) return grouped_kernel_ptr<FMT_P16, true>(g, nt, rounds);
        return grouped_kernel_ptr<FMT_F32, true>(g, nt, rounds);
    }
    if (a_type == TOMOS_FP32_BF16) return grouped_kernel_ptr<FMT_PBF, false>(g, nt, rounds);
    if (a_type == TOMOS_FP32_16) return grouped_kernel_ptr<FMT_P16, false>(g, nt, rounds);
    return grouped_kernel_ptr<FMT_F32, false>(g, nt, rounds);
}

// Native BF16 storage runs the FP32_BF16 kernels: the array is a hi plane.
const void* fused_kernel_for(int a_type, int g, int nt, int rounds, bool ntl) {
    if (a_type == TOMOS_BF16) a_type = TOMOS_FP32_BF16;
    if (ntl) {
        if (a_type == TOMOS_FP32_BF16) return fused_kernel_ptr<FMT_PBF, true>(g, nt, rounds);
        if (a_type == TOMOS_FP32_16) return fused_kernel_ptr<FMT_P16, true>(g, nt, rounds);
        return fused_kernel_ptr<FMT_F32, true>(g, nt, rounds);
    }
    if (a_type == TOMOS_FP32_BF16) return fused_kernel_ptr<FMT_PBF, false>(g, nt, rounds);
    if (a_type == TOMOS_FP32_16) return fused_kernel_ptr<FMT_P16, false>(g, nt, rounds);
    return fused_kernel_ptr<FMT_F32, false>(g, nt, rounds);
}

// Kernel timer: HIP events around every single-pass kernel launch made by
// this host thread between tomos_kernel_timer_begin and _end (a fixed pool
// of event pairs, created at begin; launches beyond it are not timed).
struct KernelTimer {
    bool on = false;
    std::vector<hipEvent_t> ev;  // 2 per launch
    size_t used = 0;             // pairs recorded
};
thread_local KernelTimer g_timer;

void timer_destroy() {
    for (hipEvent_t e : g_timer.ev) (void)hipEventDestroy(e);
    g_timer.ev.clear();
    g_timer.used = 0;
    g_timer.on = false;
}

}  // namespace

FusedShape lsqr_fused_default_shape(int64_t n) {
    // 128 contiguous bytes per column per tile (G = 8): the fetch granule,
    // so no over-fetch.  Above 512 columns, 512 threads with the tiles staged
    // through LDS by DMA (the shape the autotuner picks at 2e6 x 1000:
    // 0.64 ms per launch against 0.80 ms for 256 threads); up to 1024
    // columns the whole tile fits the block.  Beyond, 64-byte runs (G = 4,
    // twice the fetch).
    if (n > 512 && n <= 1024) return FusedShape{8, 512};
    if (n <= 1024) return FusedShape{8, 256};
    return FusedShape{4, 512};
}

int64_t fused_tile_rows(int a_type, int g) { return static_cast<int64_t>(g) * (a_type == TOMOS_F32 ? 4 : 8); }

int fused_rounds(int g, int threads, int64_t n) { return rounds_for(g, threads, n); }

void grouped_fused_launch(int a_type, int g, int threads, int rounds, const GroupProb* probs,
                          const GroupItem* items, int nitems, hipStream_t s, bool ntl) {
    require(a_type == TOMOS_F32 || a_type == TOMOS_FP32_BF16 || a_type == TOMOS_FP32_16,
            "grouped_fused: unsupported a_type");
    const void* k = grouped_kernel_for(a_type, g, threads, rounds, ntl);
    if (!k) throw Error(kErrInternal, "grouped_fused: no kernel for this shape");
    if (nitems <= 0) return;
    void* args[] = {const_cast<GroupProb**>(&probs), const_cast<GroupItem**>(&items)};
    TOMOS_HIP_CHECK(hipLaunchKernel(k, dim3(static_cast<unsigned>(nitems)), dim3(threads), args, 0, s));
}

int grouped_fused_occupancy(int a_type, int g, int threads, int rounds) {
    const void* k = grouped_kernel_for(a_type, g, threads, rounds, false);
    int occ = 0;
    if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, threads, 0) != hipSuccess) occ = 0;
    return std::max(occ, 1);
}

bool lsqr_fused_supported(int a_type, int tier, int64_t m, int64_t n, int64_t lda, const void* a) {
    if (tier != TOMOS_TIER_DEFAULT) return false;
    if (a_type != TOMOS_F32 && a_type != TOMOS_FP32_BF16 && a_type != TOMOS_FP32_16 && a_type != TOMOS_BF16)
        return false;
    const int vec = a_type == TOMOS_F32 ? 4 : 8;
    if (n < 1 || n > kFusedMaxN || m < 1) return false;
    return lda % vec == 0 && reinterpret_cast<uintptr_t>(a) % 16 == 0;
}

FusedPlan lsqr_fused_plan(int a_type, int64_t m, int64_t n, int g, int blocks_per_cu, int threads) {
    FusedPlan p;
    if (g <= 0 && blocks_per_cu <= 0 && threads <= 0) lsqr_fused_config_get(a_type, n, &g, &blocks_per_cu, &threads);
    if (threads <= 0) threads = 256;
    if (g <= 0 || rounds_for(g, threads, n) == 0) {
        const FusedShape d = lsqr_fused_default_shape(n);
        g = d.g;
        threads = d.threads;
    }
    p.g = g;
    p.threads = threads;
    p.rounds = rounds_for(g, threads, n);
    p.stream = stream_loads(std::max<int64_t>(m, 1) * n * (a_type == TOMOS_F32 ? 4 : 2));
    if (blocks_per_cu <= 0) {
        // as many blocks as fit on the device at once: one wave of blocks
        int occ = 0;
        const void* k = p.rounds > 0 ? fused_kernel_for(a_type, g, threads, p.rounds, p.stream) : nullptr;
        if (k && hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, threads, 0) != hipSuccess) occ = 0;
        blocks_per_cu = occ > 0 ? occ : 2;
    }
    blocks_per_cu = std::min(blocks_per_cu, 8);
    const int64_t tr = static_cast<int64_t>(g) * (a_type == TOMOS_F32 ? 4 : 8);
    const int64_t ntiles = ceil_div(std::max<int64_t>(m, 1), tr);
    const int64_t want = static_cast<int64_t>(blocks_per_cu) * cu_count();
    p.tiles_per_block = std::max<int64_t>(1, ceil_div(ntiles, want));
    p.nblocks = static_cast<int>(ceil_div(ntiles, p.tiles_per_block));
    p.splits = std::min(kMaxSplits, std::max(1, p.nblocks / 16));
    p.tpart_bytes = static_cast<size_t>(p.nblocks) * n * 4;
    p.tsum_bytes = static_cast<size_t>(p.splits) * n * 8;
    return p;
}

int64_t lsqr_fused_max_blocks(int64_t m) {
    // at most 8 blocks per CU, and never more blocks than 8-row tiles
    return std::max<int64_t>(1, std::min<int64_t>(8LL * cu_count(), ceil_div(std::max<int64_t>(m, 1), 8)));
}

size_t lsqr_fused_max_tpart_bytes(int64_t m, int64_t n) {
    return static_cast<size_t>(lsqr_fused_max_blocks(m)) * n * 4;
}

size_t lsqr_fused_max_tsum_bytes(int64_t n) { return static_cast<size_t>(kMaxSplits) * n * 8; }

namespace {
struct Cfg {
    int g, bpc, threads;
};
std::mutex g_cfg_mu;
std::map<std::pair<int, int64_t>, Cfg> g_cfg;  // (a_type, n) -> shape
}  // namespace

void lsqr_fused_config_set(int a_type, int64_t n, int g, int blocks_per_cu, int threads) {
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    if (g <= 0 && blocks_per_cu <= 0 && threads <= 0) g_cfg.erase({a_type, n});
    else g_cfg[{a_type, n}] = Cfg{g, blocks_per_cu, threads};
}

void lsqr_fused_config_clear() {
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    g_cfg.clear();
}

bool lsqr_fused_config_get(int a_type, int64_t n, int* g, int* blocks_per_cu, int* threads) {
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    auto it = g_cfg.find({a_type, n});
    if (it == g_cfg.end()) return false;
    *g = it->second.g;
    *blocks_per_cu = it->second.bpc;
    *threads = it->second.threads;
    return true;
}

void lsqr_fused_launch(int a_type, int64_t m, int64_t n, const void* a, int64_t lda, const void* p,
                       bool p64, void* u, bool u64, const double* u_coef, double* sumsq,
                       float* tpart, double* tsum, const FusedPlan& plan, const int* skip,
                       hipStream_t s) {
    require(plan.rounds > 0 && rounds_for(plan.g, plan.threads, n) == plan.rounds,
            "lsqr_fused: plan does not fit n");
    const void* k = fused_kernel_for(a_type, plan.g, plan.threads, plan.rounds, plan.stream);
    if (!k) throw Error(kErrInternal, "lsqr_fused: no kernel for this plan");
    int64_t tpb = plan.tiles_per_block;
    void* args[] = {&m, &n, const_cast<void**>(&a), &lda, const_cast<void**>(&p), &p64, &u, &u64,
                    const_cast<double**>(&u_coef), &tpb, &sumsq, &tpart, const_cast<int**>(&skip)};
    const bool timed = g_timer.on && 2 * (g_timer.used + 1) <= g_timer.ev.size();
    if (timed) TOMOS_HIP_CHECK(hipEventRecord(g_timer.ev[2 * g_timer.used], s));
    TOMOS_HIP_CHECK(hipLaunchKernel(k, dim3(static_cast<unsigned>(plan.nblocks)), dim3(plan.threads), args, 0, s));
    TOMOS_LAUNCH_CHECK("lsqr_fused_kernel");
    if (timed) TOMOS_HIP_CHECK(hipEventRecord(g_timer.ev[2 * g_timer.used++ + 1], s));
    fused_tsum_kernel<<<dim3(static_cast<unsigned>(ceil_div(n, kThreads)), static_cast<unsigned>(plan.splits)),
                        kThreads, 0, s>>>(n, plan.nblocks, plan.splits, tpart, tsum, skip);
    TOMOS_LAUNCH_CHECK("fused_tsum_kernel");
}

bool lsqr_restart_supported(int a_type, int tier, int64_t m, int64_t n, int64_t lda, const void* a,
                            const uint16_t* wr) {
    if (tier != TOMOS_TIER_EXACT || m < 1 || n < 1 || n > kRestartMaxN) return false;
    // FP32_16's exact decode does not fit the register budget next to both
    // planes: its restarts keep the two-kernel path
    if (a_type != TOMOS_F32 && a_type != TOMOS_FP32_BF16) return false;
    const int vec = a_type == TOMOS_F32 ? 4 : 8;
    if (lda % vec != 0 || reinterpret_cast<uintptr_t>(a) % 16 != 0) return false;
    return a_type == TOMOS_F32 || (wr != nullptr && reinterpret_cast<uintptr_t>(wr) % 16 == 0);
}

namespace {
const void* restart_kernel_for(int a_type, bool ntl, bool xzero = false) {
    if (a_type == TOMOS_FP32_BF16) {
        if (xzero)
            return ntl ? reinterpret_cast<const void*>(lsqr_restart_kernel<FMT_PBF, true, true>)
                       : reinterpret_cast<const void*>(lsqr_restart_kernel<FMT_PBF, false, true>);
        return ntl ? reinterpret_cast<const void*>(lsqr_restart_kernel<FMT_PBF, true>)
                   : reinterpret_cast<const void*>(lsqr_restart_kernel<FMT_PBF, false>);
    }
    // F32: the one plane is needed for A^T u anyway
    return ntl ? reinterpret_cast<const void*>(lsqr_restart_kernel<FMT_F32, true>)
               : reinterpret_cast<const void*>(lsqr_restart_kernel<FMT_F32, false>);
}
}  // namespace

FusedPlan lsqr_restart_plan(int a_type, int64_t m, int64_t n) {
    FusedPlan p;
    p.g = kRsG;
    p.threads = kRsNT;
    p.rounds = kRsR;
    p.stream = stream_loads(std::max<int64_t>(m, 1) * n * 4);  // hi + lo, or F32
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, restart_kernel_for(a_type, p.stream), kRsNT, 0) !=
        hipSuccess)
        occ = 0;
    const int64_t tr = static_cast<int64_t>(kRsG) * (a_type == TOMOS_F32 ? 4 : 8);
    const int64_t ntiles = ceil_div(std::max<int64_t>(m, 1), tr);
    const int64_t want = std::min<int64_t>(static_cast<int64_t>(std::max(occ, 1)) * cu_count(),
                                           lsqr_fused_max_blocks(m));
    p.tiles_per_block = std::max<int64_t>(1, ceil_div(ntiles, want));
    p.nblocks = static_cast<int>(ceil_div(ntiles, p.tiles_per_block));
    p.splits = std::min(kMaxSplits, std::max(1, p.nblocks / 16));
    p.tpart_bytes = static_cast<size_t>(p.nblocks) * n * 4;
    p.tsum_bytes = static_cast<size_t>(p.splits) * n * 8;
    return p;
}

void lsqr_restart_launch(int a_type, int64_t m, int64_t n, const void* a, const uint16_t* wr, int64_t lda,
                         const double* x, bool xzero, void* u, bool u64, double* sumsq, float* tpart,
                         double* tsum, const FusedPlan& plan, const int* skip, hipStream_t s) {
    require(n <= kRestartMaxN && plan.nblocks <= lsqr_fused_max_blocks(m), "lsqr_restart: plan does not fit");
    const void* k = restart_kernel_for(a_type, plan.stream, xzero);
    int64_t tpb = plan.tiles_per_block;
    void* args[] = {&m, &n, const_cast<void**>(&a), const_cast<uint16_t**>(&wr), &lda, const_cast<double**>(&x),
                    &u, &u64, &tpb, &sumsq, &tpart, const_cast<int**>(&skip)};
    TOMOS_HIP_CHECK(hipLaunchKernel(k, dim3(static_cast<unsigned>(plan.nblocks)), dim3(kRsNT), args, 0, s));
    TOMOS_LAUNCH_CHECK("lsqr_restart_kernel");
    fused_tsum_kernel<<<dim3(static_cast<unsigned>(ceil_div(n, kThreads)), static_cast<unsigned>(plan.splits)),
                        kThreads, 0, s>>>(n, plan.nblocks, plan.splits, tpart, tsum, skip);
    TOMOS_LAUNCH_CHECK("fused_tsum_kernel");
}

}  // namespace tomos_hip

namespace tomos_hip {
namespace {

__global__ void pair_setup_kernel(double* __restrict__ coef, double value) { *coef = value; }

// t[j] = sum over splits of tsum[s][j]; block 0 also sums the block sumsq
// partials into *unorm2 (fixed order).
__global__ __launch_bounds__(kThreads) void pair_finish_kernel(int64_t n, int splits,
                                                               const double* __restrict__ tsum,
                                                               double* __restrict__ t, int nb,
                                                               const double* __restrict__ sumsq,
                                                               double* __restrict__ unorm2) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (j < n) {
        double acc = 0.0;
        for (int sp = 0; sp < splits; ++sp) acc += tsum[static_cast<int64_t>(sp) * n + j];
        t[j] = acc;
    }
    if (blockIdx.x == 0 && unorm2) {
        __shared__ double red[kThreads];
        double acc = 0.0;
        for (int b = threadIdx.x; b < nb; b += kThreads) acc += sumsq[b];
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int w = kThreads / 2; w >= 1; w >>= 1) {
            if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) *unorm2 = red[0];
    }
}

struct PairLayout {
    size_t coef, sumsq, tpart, tsum, total;
};

// Sized for every plan of (m, n), so a tuned plan never outgrows scratch
// that was sized before tuning.
PairLayout pair_layout(int64_t m, int64_t n) {
    PairLayout L;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = (off + bytes + 255) / 256 * 256;
        return at;
    };
    L.coef = take(8);
    L.sumsq = take(static_cast<size_t>(lsqr_fused_max_blocks(m)) * 8);
    L.tpart = take(lsqr_fused_max_tpart_bytes(m, n));
    L.tsum = take(lsqr_fused_max_tsum_bytes(n));
    L.total = off;
    return L;
}

}  // namespace
}  // namespace tomos_hip

using namespace tomos_hip;

extern "C" {

int tomos_fused_config_set(int a_type, int64_t n, int row_groups, int blocks_per_cu, int threads) {
    return guarded([&] {
        require(a_type == TOMOS_F32 || a_type == TOMOS_FP32_16 || a_type == TOMOS_FP32_BF16,
                "tomos_fused_config_set: bad a_type");
        require(n >= 1 && n <= kFusedMaxN, "tomos_fused_config_set: bad n");
        require(row_groups == 0 || row_groups == 2 || row_groups == 4 || row_groups == 8 || row_groups == 32,
                "tomos_fused_config_set: row_groups must be 0, 2, 4, 8 or 32");
        require(threads == 0 || threads == 256 || threads == 512 || threads == 1024,
                "tomos_fused_config_set: threads must be 0, 256, 512 or 1024");
        require(row_groups == 0 || rounds_for(row_groups, threads ? threads : 256, n) > 0,
                "tomos_fused_config_set: no kernel for this shape at this n");
        require(blocks_per_cu >= 0 && blocks_per_cu <= 8, "tomos_fused_config_set: blocks_per_cu in 0..8");
        lsqr_fused_config_set(a_type, n, row_groups, blocks_per_cu, threads);
    });
}

int tomos_fused_config_clear(void) {
    return guarded([&] { lsqr_fused_config_clear(); });
}

int tomos_fused_config_get(int a_type, int64_t n, int* row_groups, int* blocks_per_cu, int* threads) {
    return guarded([&] {
        require(row_groups && blocks_per_cu && threads, "tomos_fused_config_get: null pointer");
        if (!lsqr_fused_config_get(a_type, n, row_groups, blocks_per_cu, threads))
            *row_groups = *blocks_per_cu = *threads = 0;
    });
}

int tomos_kernel_timer_begin(int capacity) {
    return guarded([&] {
        require(capacity >= 1 && capacity <= 65536, "tomos_kernel_timer_begin: capacity out of range");
        timer_destroy();
        g_timer.ev.resize(2 * static_cast<size_t>(capacity), nullptr);
        for (hipEvent_t& e : g_timer.ev) TOMOS_HIP_CHECK(hipEventCreate(&e));
        g_timer.on = true;
    });
}

int tomos_kernel_timer_end(float* durations_ms, int capacity, int* launches) {
    return guarded([&] {
        require(launches && (durations_ms || capacity == 0), "tomos_kernel_timer_end: null pointer");
        require(g_timer.on, "tomos_kernel_timer_end: no timer running on this thread");
        const size_t k = g_timer.used;
        for (size_t i = 0; i < k; ++i) {
            TOMOS_HIP_CHECK(hipEventSynchronize(g_timer.ev[2 * i + 1]));
            float ms = 0.0f;
            TOMOS_HIP_CHECK(hipEventElapsedTime(&ms, g_timer.ev[2 * i], g_timer.ev[2 * i + 1]));
            if (static_cast<int64_t>(i) < capacity) durations_ms[i] = ms;
        }
        *launches = static_cast<int>(k);
        timer_destroy();
    });
}

size_t tomos_matvec_pair_scratch_size(int a_type, int64_t m, int64_t n) {
    if (m < 1 || n < 1 || n > kFusedMaxN) return 0;
    (void)a_type;
    return pair_layout(m, n).total;
}

int tomos_matvec_pair(int a_type, int64_t m, int64_t n, const void* a, int64_t lda,
                      const double* p, double u_coef, float* u, double* t, double* unorm2,
                      void* scratch, size_t scratch_bytes, tomos_stream_t s) {
    return guarded([&] {
        require(m >= 1 && n >= 1 && lda >= m, "tomos_matvec_pair: bad sizes");
        require(a && p && u && t, "tomos_matvec_pair: null pointer");
        if (!lsqr_fused_supported(a_type, TOMOS_TIER_DEFAULT, m, n, lda, a))
            throw Error(kErrUnsupported,
                        "tomos_matvec_pair: needs F32/FP32_16/FP32_BF16 A, n <= 2048, 16-byte aligned "
                        "A and lda a multiple of the 16-byte row vector");
        const FusedPlan plan = lsqr_fused_plan(a_type, m, n);
        const PairLayout L = pair_layout(m, n);
        if (!scratch) scratch = fallback_scratch(L.total);
        else require(scratch_bytes >= L.total, "tomos_matvec_pair: scratch too small");
        char* base = static_cast<char*>(scratch);
        double* coef = reinterpret_cast<double*>(base + L.coef);
        double* sumsq = reinterpret_cast<double*>(base + L.sumsq);
        float* tpart = reinterpret_cast<float*>(base + L.tpart);
        double* tsum = reinterpret_cast<double*>(base + L.tsum);
        pair_setup_kernel<<<1, 1, 0, s>>>(coef, u_coef);
        TOMOS_LAUNCH_CHECK("pair_setup_kernel");
        lsqr_fused_launch(a_type, m, n, a, lda, p, true, u, false, coef, sumsq, tpart, tsum, plan,
                          nullptr, s);
        pair_finish_kernel<<<static_cast<unsigned>(ceil_div(n, kThreads)), kThreads, 0, s>>>(
            n, plan.splits, tsum, t, plan.nblocks, sumsq, unorm2);
        TOMOS_LAUNCH_CHECK("pair_finish_kernel");
    });
}

}  // extern "C"
