// The adjoint of many least-squares problems on the device, all problems in
// the same launches (tomos_lsq_adjoint_batched with workers == 0).
//
// Every problem runs the single-problem adjoint's restarted,
// R-preconditioned CG on the normal equations (adjoint.hip): restarts form
// the exact residual res = g - A^T (A w) at the restart tier (FP64
// accumulation) and stop when |res|/|g| < tol; inner iterations form
// q = A d and A^T q from one read of A at the per-iteration tier.  What is
// grouped is the launch:
//   iteration: the grouped single-pass kernel (the batched LSQR solver's,
//              u = A d with a zero u coefficient, A^T u per tile range, |u|^2
//              per tile range) for all iterating problems, then one
//              1024-thread block per problem for the n-sized work: alpha =
//              rho/|q|^2, w += alpha d, res -= alpha A^T q, the stop test;
//              then, over many blocks per problem, y = R^-T res (a wave per
//              column) and z = R^-1 y (64 rows a block), and one block per
//              problem for rho' = res.z, d = z + (rho'/rho) d;
//   restart:   A^T (A w) at the exact tier for every restarting problem
//              (one grouped launch over row items; a per-problem GEMV pair
//              when some n > 1024), then one block per problem: res,
//              |res|/|g|, the stop test; z, rho and d = z as in a step.
// Each problem carries a small state machine on the device (restart
// pending -> iterating -> restart pending ... -> done), so the host never
// looks at a problem: it replays a captured segment of `kSegIters`
// iterations (one graph launch) and, when some problem is waiting for a
// restart, the captured restart sequence, checking after each launch
// whether every problem is done.  The segment graphs are instantiated once
// per launch configuration and kept.
//
// The results agree with the one-problem adjoint to rounding (the partial
// sums of A^T q and |q|^2 are taken per tile range rather than per block),
// not bit for bit.  No reference counterpart: the reference has no adjoint.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "adjoint.hpp"
#include "codec.hpp"
#include "common.hpp"
#include "gemv.hpp"
#include "lsqr_fused.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

constexpr int kTail = 1024;
constexpr int kSegIters = 16;  // inner iterations per captured segment

enum : int { kPhaseRestart = 0, kPhaseIterate = 1, kPhaseDone = 2 };

struct GState {
    double rho, gnorm, resnorm;
    double res0;        // |res| at the start of the current restart
    double zero;        // 0.0: u coefficient of the single-pass kernel (u = A d)
    int phase;          // kPhase*
    int restarts_done;  // exact residuals formed so far
    int iters;          // inner iterations of the current restart
    int completed;      // inner iterations in total
    int skip_restart;   // restart GEMVs are a no-op (not restarting, or w == 0)
    int skip_iter;      // grouped single-pass kernel is a no-op (not iterating)
    int pc;             // a step continues: z = R^-1 R^-T res and d pending
    int pad;
};

struct GProb {
    int64_t m, n, ldr, lda;
    const void* a;        // A (hi plane when packed)
    const uint16_t* wr;   // lo plane (packed)
    const double* g;
    double* w;
    const double* rinv;
    double *res, *z, *d, *rfull;  // n each
    double* y;                    // R^-T res (n)
    double* rzp;                  // res.z partial per 64-row chunk of z
    double* rpart;                // exact restart: A^T A w per row item [nrx][n]
    int nrx;                      // row items of the exact restart (0: rfull from the GEMV pair)
    GState* st;
    int nitems;
    double* sumsq;  // [nitems]
    float* tpart;   // [nitems][n]
};

__global__ void k_gadj_init(const GProb* __restrict__ P, int nprob) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= nprob) return;
    GState* s = P[i].st;
    s->rho = s->gnorm = s->resnorm = s->zero = s->res0 = 0.0;
    s->phase = P[i].n > 0 ? kPhaseRestart : kPhaseDone;
    s->restarts_done = s->iters = s->completed = 0;
    s->skip_restart = 1;  // the first restart has w = 0: res = g, no products
    s->skip_iter = 1;
    s->pc = 0;
}

__global__ __launch_bounds__(kTail) void k_gadj_zero_w(const GProb* __restrict__ P) {
    const GProb& p = P[blockIdx.x];
    for (int64_t j = threadIdx.x; j < p.n; j += kTail) p.w[j] = 0.0;
}

__device__ __forceinline__ double gsum(double v, double* sh) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < kTail / 64; ++i) t += sh[i];  // same order in every thread
    return t;
}

// Exact restart products of every restarting problem in one launch.  An
// item is kRxRows rows of one problem: each lane takes one 16-byte row
// vector (8 packed / 4 FP32 rows) of both planes, forms its rows of
// q = A w with the exact values and FP64 sums over all n columns (phase 1,
// q stays in registers), then reads the same vectors again (from the
// caches) for the item's share of A^T q (phase 2): per batch of 8 columns
// a transposing reduction over the wave's lanes leaves each column's wave
// sum in one lane group, the 4 waves' sums are added in a fixed order and
// written to rpart[item] (n FP64).  The restart kernel adds the items of a
// problem in order.
constexpr int kRxThreads = 256;
constexpr int64_t kRxMaxN = 1024;  // LDS: 4 waves x n FP64 sums
struct RxItem {
    int prob, slot;
    int64_t row0, row1;
};

template <int FMT>  // 0: F32, 1: FP32_BF16, 2: FP32_16 (wh + wr)
__device__ __forceinline__ void rx_decode(const uint4& h, const uint4& l, float (&v)[FMT == 0 ? 4 : 8]) {
    if constexpr (FMT == 0) {
        v[0] = __uint_as_float(h.x);
        v[1] = __uint_as_float(h.y);
        v[2] = __uint_as_float(h.z);
        v[3] = __uint_as_float(h.w);
    } else {
        const uint32_t hw[4] = {h.x, h.y, h.z, h.w}, lw[4] = {l.x, l.y, l.z, l.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if constexpr (FMT == 1) {
                v[2 * k] = __uint_as_float((hw[k] << 16) | (lw[k] & 0xFFFFu));
                v[2 * k + 1] = __uint_as_float((hw[k] & 0xFFFF0000u) | (lw[k] >> 16));
            } else {
                v[2 * k] = codec::fp32_16_decode_exact(hw[k] & 0xFFFFu, lw[k] & 0xFFFFu);
                v[2 * k + 1] = codec::fp32_16_decode_exact(hw[k] >> 16, lw[k] >> 16);
            }
        }
    }
}

template <int FMT>
__device__ __forceinline__ void rx_load(const GProb& p, int64_t r, int nv, int64_t j, float (&v)[FMT == 0 ? 4 : 8]) {
    constexpr int RPL = FMT == 0 ? 4 : 8;
    constexpr int ESZ = FMT == 0 ? 4 : 2;
    const int64_t e = r + j * p.lda;
    if (nv == RPL) {
        const uint4 h = *reinterpret_cast<const uint4*>(static_cast<const char*>(p.a) + e * ESZ);
        uint4 l = make_uint4(0u, 0u, 0u, 0u);
        if constexpr (FMT != 0) l = *reinterpret_cast<const uint4*>(p.wr + e);
        rx_decode<FMT>(h, l, v);
    } else {
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
            if (k < nv) {
                if constexpr (FMT == 0) {
                    v[k] = static_cast<const float*>(p.a)[e + k];
                } else {
                    const uint32_t h = static_cast<const uint16_t*>(p.a)[e + k], l = p.wr[e + k];
                    v[k] = FMT == 1 ? __uint_as_float((h << 16) | l) : codec::fp32_16_decode_exact(h, l);
                }
            } else {
                v[k] = 0.0f;
            }
        }
    }
}

template <int FMT>
__global__ __launch_bounds__(kRxThreads) void k_gadj_rx(const GProb* __restrict__ P, const RxItem* __restrict__ items) {
    constexpr int RPL = FMT == 0 ? 4 : 8;
    constexpr int CB = 8;  // columns per transposing reduction
    const RxItem it = items[blockIdx.x];
    const GProb& p = P[it.prob];
    if (p.st->skip_restart) return;
    __shared__ double wsum[kRxThreads / 64][kRxMaxN];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t n = p.n;
    const int64_t r = it.row0 + static_cast<int64_t>(threadIdx.x) * RPL;
    const int nv = r < it.row1 ? static_cast<int>(min<int64_t>(RPL, it.row1 - r)) : 0;
    // phase 1: q = A w for this lane's rows
    double q[RPL];
#pragma unroll
    for (int k = 0; k < RPL; ++k) q[k] = 0.0;
    if (nv > 0) {
        int64_t j = 0;
        for (; j + 4 <= n; j += 4) {
            float v[4][RPL];
#pragma unroll
            for (int c = 0; c < 4; ++c) rx_load<FMT>(p, r, nv, j + c, v[c]);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double wj = p.w[j + c];
#pragma unroll
                for (int k = 0; k < RPL; ++k) q[k] = fma(static_cast<double>(v[c][k]), wj, q[k]);
            }
        }
        for (; j < n; ++j) {
            float v[RPL];
            rx_load<FMT>(p, r, nv, j, v);
            const double wj = p.w[j];
#pragma unroll
            for (int k = 0; k < RPL; ++k) q[k] = fma(static_cast<double>(v[k]), wj, q[k]);
        }
    }
    // phase 2: A^T q over this item's rows, CB columns at a time
    for (int64_t j0 = 0; j0 < n; j0 += CB) {
        double cs[CB];
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            double t = 0.0;
            if (nv > 0 && j0 + c < n) {
                float v[RPL];
                rx_load<FMT>(p, r, nv, j0 + c, v);
#pragma unroll
                for (int k = 0; k < RPL; ++k) t = fma(static_cast<double>(v[k]), q[k], t);
            }
            cs[c] = t;
        }
        // halving stages on lane bits 5, 4, 3 (8 -> 1 values), then bits 2..0
#pragma unroll
        for (int st = 0; st < 3; ++st) {
            const int off = 32 >> st, half = CB >> (st + 1);
            const bool upper = (lane & off) != 0;
#pragma unroll
            for (int t2 = 0; t2 < 4; ++t2) {
                if (t2 < half) {
                    const double send = upper ? cs[t2] : cs[t2 + half];
                    const double keep = upper ? cs[t2 + half] : cs[t2];
                    cs[t2] = keep + __shfl_xor(send, off);
                }
            }
        }
        for (int off = 4; off >= 1; off >>= 1) cs[0] += __shfl_xor(cs[0], off);
        // the column this lane group holds: bit 5 -> +4, bit 4 -> +2, bit 3 -> +1
        const int sub = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
        if ((lane & 7) == 0 && j0 + sub < n) wsum[wave][j0 + sub] = cs[0];
    }
    __syncthreads();
    double* out = p.rpart + static_cast<int64_t>(it.slot) * n;
    for (int64_t j = threadIdx.x; j < n; j += kRxThreads) {
        double t = wsum[0][j];
#pragma unroll
        for (int w2 = 1; w2 < kRxThreads / 64; ++w2) t += wsum[w2][j];
        out[j] = t;
    }
}

// The same products with the tile held in registers (packed formats,
// n <= kRx2MaxN): a tile is 64 rows (8 row groups x 8 rows) by 512 columns
// (64 column slots x 8 rounds), both planes, 16 vectors a lane; a problem
// uses only the first R rounds, the fewest of 1, 2, 4, 8 that cover its n.
// The tile comes through LDS by DMA, and the next tile's DMA is issued as
// soon as this one is in registers, so it stays in flight through this
// tile's arithmetic and LDS-only barriers.  Phase 1: a lane's 8 row sums
// over its columns (FP64), a halving reduction over the wave's 8 column
// slots (lane bits 5..3) leaves each lane one row, the 8 waves add through
// LDS in a fixed order -> q for the tile's rows.  Phase 2 reuses the
// registers: a lane's column sums over its 8 rows with q, a halving
// reduction over the 8 row groups (lane bits 2..0) leaves lane (g, c)
// column g * 64 + c, kept in a register across the item's tiles and
// written once to rpart[item].  Items are row ranges of whole tiles.
constexpr int kRx2Threads = 512;
constexpr int64_t kRx2MaxN = 512;
constexpr int64_t kRx2Rows = 64;  // rows per tile

// Workgroup barrier for LDS exchanges only: waits for this wave's LDS
// reads and writes (lgkmcnt) but not for its outstanding LDS-DMA loads,
// which __syncthreads() (a fence over all memory) would drain; the
// "memory" clobber keeps the compiler from moving memory accesses across.
__device__ __forceinline__ void lds_exchange_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <int FMT>  // 1: FP32_BF16, 2: FP32_16
__global__ __launch_bounds__(kRx2Threads) void k_gadj_rx2(const GProb* __restrict__ P,
                                                          const RxItem* __restrict__ items) {
    const RxItem it = items[blockIdx.x];
    const GProb& p = P[it.prob];
    if (p.st->skip_restart) return;
    // row partials and q in two copies, by tile parity: a wave may write
    // the next tile's partials while slower waves still read this tile's q
    __shared__ double red[2][kRx2Threads / 64][kRx2Rows];
    __shared__ double qsh[2][kRx2Rows];
    __shared__ uint4 stage[kRx2Threads / 64][16][64];  // a wave's tile: 8 hi, 8 lo vectors a lane
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = tid & 7, c = tid >> 3;  // row group, column slot
    const int64_t n = p.n, m = p.m;
    const char* hb = static_cast<const char*>(p.a);
    double wv[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int64_t j = static_cast<int64_t>(r) * 64 + c;
        wv[r] = j < n ? p.w[j] : 0.0;
    }
    double tacc = 0.0;
    auto run = [&](auto rounds) {
        constexpr int R = decltype(rounds)::value;
        // LDS-DMA of the tile at t0 into this wave's stage slots (each
        // lane's 16 bytes at slot base + 16 * lane); lanes outside the item
        // or past column n load the first vector of the planes instead (in
        // bounds) and are cleared when read back
        auto issue = [&](int64_t t0) {
            const int64_t row = t0 + g * 8;
            const bool ok = row < it.row1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int64_t j = static_cast<int64_t>(r) * 64 + c;
                const int64_t e = (ok && j < n) ? row + j * p.lda : 0;
                __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint4*>(hb + e * 2), &stage[wave][r][0], 16,
                                                 0, 0);
                __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint4*>(p.wr + e), &stage[wave][8 + r][0],
                                                 16, 0, 0);
            }
        };
        issue(it.row0);
        int par = 0;
        for (int64_t t0 = it.row0; t0 < it.row1; t0 += kRx2Rows, par ^= 1) {
            uint4 h[R], l[R];
            const int64_t row = t0 + g * 8;
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's DMA has landed
#pragma unroll
            for (int r = 0; r < R; ++r) {
                h[r] = stage[wave][r][lane];
                l[r] = stage[wave][8 + r][lane];
            }
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): slots read before they are refilled
            if (t0 + kRx2Rows < it.row1) issue(t0 + kRx2Rows);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int64_t j = static_cast<int64_t>(r) * 64 + c;
                if (!(row < it.row1 && j < n)) {
                    h[r] = make_uint4(0u, 0u, 0u, 0u);
                    l[r] = make_uint4(0u, 0u, 0u, 0u);
                }
            }
            const int nv = row < m ? static_cast<int>(min<int64_t>(8, m - row)) : 0;
            // phase 1: row sums
            double acc[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) acc[k] = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float v[8];
                rx_decode<FMT>(h[r], l[r], v);
#pragma unroll
                for (int k = 0; k < 8; ++k) acc[k] = fma(static_cast<double>(v[k]), wv[r], acc[k]);
            }
#pragma unroll
            for (int st = 0; st < 3; ++st) {
                const int off = 32 >> st, half = 4 >> st;
                const bool upper = (lane & off) != 0;
#pragma unroll
                for (int t2 = 0; t2 < 4; ++t2) {
                    if (t2 < half) {
                        const double send = upper ? acc[t2] : acc[t2 + half];
                        const double keep = upper ? acc[t2 + half] : acc[t2];
                        acc[t2] = keep + __shfl_xor(send, off);
                    }
                }
            }
            const int kk = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
            red[par][wave][g * 8 + kk] = acc[0];
            lds_exchange_barrier();
            if (tid < kRx2Rows) {
                double qv = 0.0;
#pragma unroll
                for (int w2 = 0; w2 < kRx2Threads / 64; ++w2) qv += red[par][w2][tid];
                qsh[par][tid] = t0 + tid < m ? qv : 0.0;
            }
            lds_exchange_barrier();
            // phase 2: column sums with q over this lane's rows
            double qk[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) qk[k] = qsh[par][g * 8 + k];
            double cs[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float v[8];
                rx_decode<FMT>(h[r], l[r], v);
                double sacc = 0.0;  // rows past m (padding of any value) are left out
#pragma unroll
                for (int k = 0; k < 8; ++k) sacc = k < nv ? fma(static_cast<double>(v[k]), qk[k], sacc) : sacc;
                cs[r] = sacc;
            }
#pragma unroll
            for (int st = 0; st < 3; ++st) {
                const int off = 4 >> st, half = 4 >> st;
                const bool upper = (lane & off) != 0;
#pragma unroll
                for (int t2 = 0; t2 < 4; ++t2) {
                    if (t2 < half) {
                        const double send = upper ? cs[t2] : cs[t2 + half];
                        const double keep = upper ? cs[t2 + half] : cs[t2];
                        cs[t2] = keep + __shfl_xor(send, off);
                    }
                }
            }
            tacc += cs[0];
            // no barrier here: the next tile writes the other copy, and this
            // copy is rewritten only after every wave has passed the next
            // tile's first barrier, that is, finished reading it
        }
    };
    if (n <= 64) run(std::integral_constant<int, 1>{});
    else if (n <= 128) run(std::integral_constant<int, 2>{});
    else if (n <= 256) run(std::integral_constant<int, 4>{});
    else run(std::integral_constant<int, 8>{});
    const int64_t j = static_cast<int64_t>(g) * 64 + c;  // this lane's column after the row-group reduction
    if (j < n) p.rpart[static_cast<int64_t>(it.slot) * n + j] = tacc;
}

// The gradients of every problem in one launch over the restart's row
// items (packed formats, n <= kRx2MaxN, lsq_grad's products in one pass
// over A): per 64-row tile, the exact row sums A w and A x in FP64 (the
// first phase of k_gadj_rx2 with two vectors), then grad_b = A w,
// r = b - A x and the tile's rows of grad_A = r w^T - (A w) x^T (FP32,
// column-major, ldg), so A is read once and grad_A written once.
struct GGrad {
    int64_t m, n, lda, ldg;
    const void* a;
    const uint16_t* wr;
    const double *b, *x, *w;
    double* grad_b;
    float* grad_a;  // null: grad_b only
};

template <int FMT>  // 1: FP32_BF16, 2: FP32_16
__global__ __launch_bounds__(kRx2Threads) void k_gadj_grad(const GGrad* __restrict__ T,
                                                           const RxItem* __restrict__ items) {
    const RxItem it = items[blockIdx.x];
    const GGrad& p = T[it.prob];
    // row sums and (A w, r) rows in two copies by tile parity, as in k_gadj_rx2
    __shared__ double redw[2][kRx2Threads / 64][kRx2Rows];
    __shared__ double redx[2][kRx2Threads / 64][kRx2Rows];
    __shared__ double gsh[2][kRx2Rows];
    __shared__ double rsh[2][kRx2Rows];
    __shared__ uint4 stage[kRx2Threads / 64][16][64];  // a wave's tile: 8 hi, 8 lo vectors a lane
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = tid & 7, c = tid >> 3;  // row group, column slot
    const int64_t n = p.n;
    const char* hb = static_cast<const char*>(p.a);
    double wv[8], xv[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int64_t j = static_cast<int64_t>(r) * 64 + c;
        wv[r] = j < n ? p.w[j] : 0.0;
        xv[r] = j < n ? p.x[j] : 0.0;
    }
    auto run = [&](auto rounds) {
        constexpr int R = decltype(rounds)::value;
        auto issue = [&](int64_t t0) {
            const int64_t row = t0 + g * 8;
            const bool ok = row < it.row1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int64_t j = static_cast<int64_t>(r) * 64 + c;
                const int64_t e = (ok && j < n) ? row + j * p.lda : 0;
                __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint4*>(hb + e * 2), &stage[wave][r][0], 16,
                                                 0, 0);
                __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint4*>(p.wr + e), &stage[wave][8 + r][0],
                                                 16, 0, 0);
            }
        };
        issue(it.row0);
        int par = 0;
        for (int64_t t0 = it.row0; t0 < it.row1; t0 += kRx2Rows, par ^= 1) {
            uint4 h[R], l[R];
            const int64_t row = t0 + g * 8;
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's DMA has landed
#pragma unroll
            for (int r = 0; r < R; ++r) {
                h[r] = stage[wave][r][lane];
                l[r] = stage[wave][8 + r][lane];
            }
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): slots read before they are refilled
            if (t0 + kRx2Rows < it.row1) issue(t0 + kRx2Rows);
            double aw[8], ax[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) aw[k] = ax[k] = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int64_t j = static_cast<int64_t>(r) * 64 + c;
                if (!(row < it.row1 && j < n)) {
                    h[r] = make_uint4(0u, 0u, 0u, 0u);
                    l[r] = make_uint4(0u, 0u, 0u, 0u);
                }
                float v[8];
                rx_decode<FMT>(h[r], l[r], v);
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    aw[k] = fma(static_cast<double>(v[k]), wv[r], aw[k]);
                    ax[k] = fma(static_cast<double>(v[k]), xv[r], ax[k]);
                }
            }
#pragma unroll
            for (int st = 0; st < 3; ++st) {
                const int off = 32 >> st, half = 4 >> st;
                const bool upper = (lane & off) != 0;
#pragma unroll
                for (int t2 = 0; t2 < 4; ++t2) {
                    if (t2 < half) {
                        const double sw = upper ? aw[t2] : aw[t2 + half];
                        const double kw = upper ? aw[t2 + half] : aw[t2];
                        aw[t2] = kw + __shfl_xor(sw, off);
                        const double sx = upper ? ax[t2] : ax[t2 + half];
                        const double kx = upper ? ax[t2 + half] : ax[t2];
                        ax[t2] = kx + __shfl_xor(sx, off);
                    }
                }
            }
            const int kk = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
            redw[par][wave][g * 8 + kk] = aw[0];
            redx[par][wave][g * 8 + kk] = ax[0];
            lds_exchange_barrier();
            if (tid < kRx2Rows) {
                double sw = 0.0, sx = 0.0;
#pragma unroll
                for (int w2 = 0; w2 < kRx2Threads / 64; ++w2) {
                    sw += redw[par][w2][tid];
                    sx += redx[par][w2][tid];
                }
                const int64_t i = t0 + tid;
                double ri = 0.0;
                if (i < it.row1) {
                    ri = p.b[i] - sx;
                    p.grad_b[i] = sw;
                }
                gsh[par][tid] = sw;
                rsh[par][tid] = ri;
            }
            lds_exchange_barrier();
            // grad_A rows of this tile from the (A w, r) rows in LDS: the 8
            // lanes of column slot c write 128 contiguous bytes of column
            // r * 64 + c per store (lane g: rows 32 h + 4 g .. + 3, h = 0, 1)
            if (p.grad_a) {
                const bool al = p.ldg % 4 == 0 && reinterpret_cast<uintptr_t>(p.grad_a) % 16 == 0;
#pragma unroll
                for (int hh = 0; hh < 2; ++hh) {
                    const int lr = hh * 32 + g * 4;  // first of this lane's 4 rows in the tile
                    const int64_t r0 = t0 + lr;
                    const int nv = r0 < it.row1 ? static_cast<int>(min<int64_t>(4, it.row1 - r0)) : 0;
                    if (nv == 0) continue;
                    double rk[4], gk[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        rk[k] = rsh[par][lr + k];
                        gk[k] = gsh[par][lr + k];
                    }
                    const bool vec = al && nv == 4;  // r0 is a multiple of 4: 16-byte aligned
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int64_t j = static_cast<int64_t>(r) * 64 + c;
                        if (j >= n) continue;
                        float* out = p.grad_a + r0 + j * p.ldg;
                        float o[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) o[k] = static_cast<float>(rk[k] * wv[r] - gk[k] * xv[r]);
                        if (vec) {
                            *reinterpret_cast<float4*>(out) = make_float4(o[0], o[1], o[2], o[3]);
                        } else {
#pragma unroll
                            for (int k = 0; k < 4; ++k)
                                if (k < nv) out[k] = o[k];
                        }
                    }
                }
            }
            // no third barrier: the next tile writes the other copies
        }
    };
    if (n <= 64) run(std::integral_constant<int, 1>{});
    else if (n <= 128) run(std::integral_constant<int, 2>{});
    else if (n <= 256) run(std::integral_constant<int, 4>{});
    else run(std::integral_constant<int, 8>{});
}

// Restart of every problem whose phase is kPhaseRestart: res = g - rfull
// (rfull = A^T A w, or nothing on the first restart), |g| on the first, the
// exact-residual stop test (the restart after `max_restarts` earlier ones is
// final whatever its residual), then z = M res, rho = res.z, d = z.
__global__ __launch_bounds__(kTail) void k_gadj_restart(const GProb* __restrict__ P, double tol, int max_restarts) {
    const GProb& p = P[blockIdx.x];
    GState* S = p.st;
    if (S->phase != kPhaseRestart) return;
    __shared__ double sh[kTail / 64];
    const int k = S->restarts_done;  // read before the barriers: thread 0 rewrites it
    const bool first = k == 0;
    double gg = 0.0, rr = 0.0;
    for (int64_t j = threadIdx.x; j < p.n; j += kTail) {
        double rf = 0.0;
        if (!first) {
            if (p.nrx > 0) {
                for (int k2 = 0; k2 < p.nrx; ++k2) rf += p.rpart[static_cast<int64_t>(k2) * p.n + j];
            } else {
                rf = p.rfull[j];
            }
        }
        const double rj = p.g[j] - rf;
        p.res[j] = rj;
        gg += p.g[j] * p.g[j];
        rr += rj * rj;
    }
    gg = gsum(gg, sh);
    rr = gsum(rr, sh);
    if (threadIdx.x == 0) {
        if (first) S->gnorm = sqrt(gg);
        S->resnorm = sqrt(rr);
        S->res0 = S->resnorm;
        S->restarts_done = k + 1;
        S->skip_restart = 1;
        const double rel = S->gnorm > 0.0 ? S->resnorm / S->gnorm : 0.0;
        const bool stop = S->gnorm == 0.0 || rel < tol || k >= max_restarts;
        S->phase = stop ? kPhaseDone : kPhaseIterate;
        S->skip_iter = stop ? 1 : 0;
        S->iters = 0;
        // z = R^-1 R^-T res, rho = res.z and d = z follow in k_gadj_pre1 / _pre2 / _fin
        S->pc = stop ? 0 : 2;
    }
}

// The n-sized part of one CG step of every iterating problem, after the
// grouped single-pass kernel left |q|^2 and A^T q as per-item partials.
// The restart's inner loop ends when |res|/|g| < tol, |q| = 0 or after
// max_iters steps; the problem then waits for its next restart.
__global__ __launch_bounds__(kTail) void k_gadj_step(const GProb* __restrict__ P, double tol, int max_iters,
                                                    double drop) {
    const GProb& p = P[blockIdx.x];
    GState* S = p.st;
    if (S->phase != kPhaseIterate) return;
    __shared__ double sh[kTail / 64];
    const double rho = S->rho;
    const int iters = S->iters;
    const int64_t n = p.n;
    double q2 = 0.0;
    for (int k = threadIdx.x; k < p.nitems; k += kTail) q2 += p.sumsq[k];
    q2 = gsum(q2, sh);
    const double alpha = q2 > 0.0 ? rho / q2 : 0.0;
    double rr = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += kTail) {
        double t = 0.0;
        for (int k = 0; k < p.nitems; ++k) t += static_cast<double>(p.tpart[static_cast<int64_t>(k) * n + j]);
        p.w[j] += alpha * p.d[j];
        const double rj = p.res[j] - alpha * t;
        p.res[j] = rj;
        rr += rj * rj;
    }
    rr = gsum(rr, sh);
    if (threadIdx.x == 0) {
        S->completed += 1;
        S->iters = iters + 1;
        S->resnorm = sqrt(rr);
        const double rel = S->resnorm / S->gnorm;
        const bool end = rel < tol || q2 <= 0.0 || iters + 1 >= max_iters || S->resnorm < drop * S->res0;
        if (end) {
            S->phase = kPhaseRestart;
            S->skip_iter = 1;
            S->skip_restart = 0;
        }
        S->pc = end ? 0 : 1;  // k_gadj_pre1 / _pre2 / _fin finish a continuing step
    }
}

// The step's z = R^-1 (R^-T res), rho' = res.z and d = z + (rho'/rho) d,
// spread over many blocks per problem (one block per problem reads all of
// R^-1 twice through one CU).  pre1: y = R^-T res, one wave per column j
// (column j of R^-1, rows 0..j, dotted with res).
constexpr int kPreRows = 64;                 // rows of z per k_gadj_pre2 block
constexpr int kPreSlices = kTail / kPreRows;  // column slices per row in k_gadj_pre2
__global__ __launch_bounds__(kTail) void k_gadj_pre1(const GProb* __restrict__ P) {
    const GProb& p = P[blockIdx.y];
    if (!p.st->pc) return;
    const int lane = threadIdx.x & 63;
    const int64_t j = static_cast<int64_t>(blockIdx.x) * (kTail / 64) + (threadIdx.x >> 6);
    if (j >= p.n) return;
    const double* col = p.rinv + j * p.ldr;
    double a = 0.0;
    for (int64_t i = lane; i <= j; i += 64) a += col[i] * p.res[i];
    for (int o = 32; o >= 1; o >>= 1) a += __shfl_xor(a, o);
    if (lane == 0) p.y[j] = a;
}

// pre2: z = R^-1 y for kPreRows rows per block, kPreSlices column slices
// (slice k takes columns i0 + k, i0 + k + kPreSlices, ...; a wave reads one
// column's kPreRows entries per step), the slices added in a fixed order;
// res.z of the block's rows into rzp[blockIdx.x]
__global__ __launch_bounds__(kTail) void k_gadj_pre2(const GProb* __restrict__ P) {
    const GProb& p = P[blockIdx.y];
    if (!p.st->pc) return;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kPreRows;
    if (i0 >= p.n) return;
    __shared__ double part[kPreSlices][kPreRows];
    const int r = threadIdx.x & (kPreRows - 1), sl = threadIdx.x / kPreRows;
    const int64_t i = i0 + r;
    double a = 0.0;
    if (i < p.n) {
#pragma unroll 4
        for (int64_t j = i0 + sl; j < p.n; j += kPreSlices)
            if (j >= i) a += p.rinv[i + j * p.ldr] * p.y[j];
    }
    part[sl][r] = a;
    __syncthreads();
    if (threadIdx.x < kPreRows) {
        double z = 0.0;
#pragma unroll
        for (int k = 0; k < kPreSlices; ++k) z += part[k][r];
        double rz = 0.0;
        if (i < p.n) {
            p.z[i] = z;
            rz = p.res[i] * z;
        }
        for (int o = 32; o >= 1; o >>= 1) rz += __shfl_xor(rz, o);
        if (r == 0) p.rzp[blockIdx.x] = rz;
    }
}

// fin: rho' = sum of the chunks' res.z (fixed order), d = z + (rho'/rho) d
__global__ __launch_bounds__(256) void k_gadj_fin(const GProb* __restrict__ P) {
    const GProb& p = P[blockIdx.x];
    GState* S = p.st;
    if (!S->pc) return;
    const int64_t nch = (p.n + kPreRows - 1) / kPreRows;
    double rz = 0.0;
    for (int64_t c = 0; c < nch; ++c) rz += p.rzp[c];
    const double rho = S->rho;
    const bool restart = S->pc == 2;  // a restart's first direction: d = z
    const double beta = rho != 0.0 ? rz / rho : 0.0;
    for (int64_t j = threadIdx.x; j < p.n; j += 256) p.d[j] = restart ? p.z[j] : p.z[j] + beta * p.d[j];
    __syncthreads();
    if (threadIdx.x == 0) {
        S->rho = rz;
        S->pc = 0;
    }
}

size_t a256(size_t v) { return (v + 255) / 256 * 256; }
constexpr int kNVecs = 6;  // n-vectors per problem in the scratch

size_t rzp_bytes(int64_t n) { return a256(static_cast<size_t>(ceil_div(std::max<int64_t>(n, 1), kPreRows)) * 8); }

struct GPlan {
    int g = 0, threads = 256, rounds = 0;
    int64_t nmax = 0, tiles_per_item = 1;
    int nitems = 0;
    std::vector<GroupItem> items;
    std::vector<int> prob_items;
    size_t probs = 0, gprobs = 0, items_off = 0, states = 0, gemv_sc = 0, gemv_bytes = 0, grad_sc = 0,
           grad_bytes = 0, total = 0;
    std::vector<size_t> vec_at, sumsq_at, tpart_at;
    // exact restart products in one grouped launch (k_gadj_rx) when the
    // restart tier decodes exactly, n <= kRxMaxN and the planes are aligned;
    // otherwise the per-problem GEMV pair
    bool stream = false;  // non-temporal loads in the grouped pass (stream_loads of its bytes)
    bool use_rx = false;
    bool rx2 = false;  // the register-tile form (packed, n <= kRx2MaxN)
    int nrx = 0;
    std::vector<RxItem> rx;
    std::vector<int> prob_rx;
    std::vector<size_t> rpart_at;
    size_t rx_off = 0;
    size_t grad_tab = 0;  // rx2: the gradients' table (GGrad per problem)
};

void grouped_validate(const tomos_lsqr_options& o, int nprob, const tomos_lsq_problem* pr) {
    require(o.a_type == TOMOS_F32 || o.a_type == TOMOS_FP32_BF16 || o.a_type == TOMOS_FP32_16,
            "adjoint (grouped): a_type must be F32, FP32_16 or FP32_BF16");
    require(o.mv_tier == TOMOS_TIER_DEFAULT, "adjoint (grouped): iterations run at the DEFAULT tier");
    require(o.precond_type == TOMOS_F64, "adjoint (grouped): precond_type must be F64 (R^-1 required)");
    require(o.max_iters >= 1 && o.restarts >= 1, "adjoint (grouped): need max_iters >= 1 and restarts >= 1");
    const int vec = o.a_type == TOMOS_F32 ? 4 : 8;
    const bool packed = o.a_type != TOMOS_F32;
    for (int i = 0; i < nprob; ++i) {
        const tomos_lsq_problem& q = pr[i];
        require(q.m >= q.n && q.n >= 0 && q.n <= 2048, "adjoint (grouped): need m >= n and n <= 2048");
        if (q.n == 0) continue;
        require(q.rinv != nullptr && q.ldr >= q.n, "adjoint (grouped): R^-1 missing or ldr < n");
        require(q.lda >= q.m && q.lda % vec == 0,
                "adjoint (grouped): lda must be >= m and a multiple of the 16-byte row vector");
        require(reinterpret_cast<uintptr_t>(q.a) % 16 == 0, "adjoint (grouped): A must be 16-byte aligned");
        if (packed && o.restart_tier == TOMOS_TIER_EXACT) require(q.wr != nullptr, "adjoint (grouped): EXACT needs wr");
    }
}

GPlan grouped_plan(const tomos_lsqr_options& o, int nprob, const tomos_lsq_problem* pr) {
    GPlan L;
    for (int i = 0; i < nprob; ++i) L.nmax = std::max(L.nmax, pr[i].n);
    int g = 0, nt = 0, bpc = 0;
    lsqr_fused_config_get(o.a_type, L.nmax, &g, &bpc, &nt);
    if (nt <= 0) nt = 256;
    if (g <= 0 || fused_rounds(g, nt, std::max<int64_t>(L.nmax, 1)) == 0) {
        const FusedShape d = lsqr_fused_default_shape(std::max<int64_t>(L.nmax, 1));
        g = d.g;
        nt = d.threads;
    }
    L.g = g;
    L.threads = nt;
    L.rounds = fused_rounds(g, nt, std::max<int64_t>(L.nmax, 1));
    require(L.rounds > 0, "adjoint (grouped): n too large for the single-pass kernel (max 2048)");
    const int64_t tr = fused_tile_rows(o.a_type, g);
    int64_t tiles_total = 0;
    for (int i = 0; i < nprob; ++i) tiles_total += pr[i].n > 0 ? ceil_div(std::max<int64_t>(pr[i].m, 1), tr) : 0;
    const int64_t resident = static_cast<int64_t>(grouped_fused_occupancy(o.a_type, g, nt, L.rounds)) * cu_count();
    L.tiles_per_item = std::max<int64_t>(1, ceil_div(std::max<int64_t>(tiles_total, 1), 4 * resident));
    L.prob_items.resize(static_cast<size_t>(nprob));
    for (int i = 0; i < nprob; ++i) {
        const int64_t tiles = pr[i].n > 0 ? ceil_div(std::max<int64_t>(pr[i].m, 1), tr) : 0;
        int k = 0;
        for (int64_t t = 0; t < tiles; t += L.tiles_per_item, ++k)
            L.items.push_back(GroupItem{i, k, t, std::min(tiles, t + L.tiles_per_item)});
        L.prob_items[static_cast<size_t>(i)] = k;
    }
    L.nitems = static_cast<int>(L.items.size());
    int64_t pass_bytes = 0;
    for (int i = 0; i < nprob; ++i) pass_bytes += pr[i].m * pr[i].n * (o.a_type == TOMOS_F32 ? 4 : 2);
    L.stream = stream_loads(pass_bytes);
    L.use_rx = (o.a_type == TOMOS_F32 || o.restart_tier == TOMOS_TIER_EXACT) && L.nmax <= kRxMaxN;
    for (int i = 0; i < nprob && L.use_rx; ++i)
        if (pr[i].n > 0 && o.a_type != TOMOS_F32 && reinterpret_cast<uintptr_t>(pr[i].wr) % 16 != 0) L.use_rx = false;
    L.prob_rx.assign(static_cast<size_t>(nprob), 0);
    L.rx2 = L.use_rx && o.a_type != TOMOS_F32 && L.nmax <= kRx2MaxN;
    if (L.use_rx) {
        // rx2: 2048-row items (32 tiles; 1024-row items measured 0.3 ms slower over
        // a 64-problem backward); rx: one 16-byte row vector a lane
        const int64_t rows = L.rx2 ? 32 * kRx2Rows : static_cast<int64_t>(kRxThreads) * (o.a_type == TOMOS_F32 ? 4 : 8);
        for (int i = 0; i < nprob; ++i) {
            if (pr[i].n == 0) continue;
            int k = 0;
            for (int64_t r0 = 0; r0 < pr[i].m; r0 += rows, ++k)
                L.rx.push_back(RxItem{i, k, r0, std::min(pr[i].m, r0 + rows)});
            L.prob_rx[static_cast<size_t>(i)] = k;
        }
        L.nrx = static_cast<int>(L.rx.size());
    }
    size_t off = 0;
    auto take = [&](size_t b) {
        const size_t at = off;
        off = a256(off + b);
        return at;
    };
    L.probs = take(sizeof(GProb) * static_cast<size_t>(std::max(nprob, 1)));
    L.gprobs = take(sizeof(GroupProb) * static_cast<size_t>(std::max(nprob, 1)));
    L.items_off = take(sizeof(GroupItem) * static_cast<size_t>(std::max(L.nitems, 1)));
    L.states = take(sizeof(GState) * static_cast<size_t>(std::max(nprob, 1)));
    L.rx_off = take(sizeof(RxItem) * static_cast<size_t>(std::max(L.nrx, 1)));
    L.grad_tab = take(sizeof(GGrad) * static_cast<size_t>(std::max(nprob, 1)));
    for (int i = 0; i < nprob; ++i) {
        L.rpart_at.push_back(take(static_cast<size_t>(L.prob_rx[static_cast<size_t>(i)]) *
                                  static_cast<size_t>(std::max<int64_t>(pr[i].n, 1)) * 8 + 8));
        const size_t nb = a256(static_cast<size_t>(std::max<int64_t>(pr[i].n, 1)) * 8);
        // res, z, d, rfull, w, y (FP64, n), rzp (FP64, one per 64 rows of n),
        // q (FP32, m), q64 (FP64, m)
        L.vec_at.push_back(take(kNVecs * nb + rzp_bytes(pr[i].n) +
                                a256(static_cast<size_t>(std::max<int64_t>(pr[i].m, 1)) * 4) +
                                a256(static_cast<size_t>(std::max<int64_t>(pr[i].m, 1)) * 8)));
        L.sumsq_at.push_back(take(static_cast<size_t>(std::max(L.prob_items[static_cast<size_t>(i)], 1)) * 8));
        L.tpart_at.push_back(take(static_cast<size_t>(std::max(L.prob_items[static_cast<size_t>(i)], 1)) *
                                  static_cast<size_t>(std::max<int64_t>(pr[i].n, 1)) * 4));
        for (int tr2 = 0; tr2 < 2; ++tr2)
            L.gemv_bytes = std::max(L.gemv_bytes, gemv_scratch_bytes(o.a_type, o.restart_tier, tr2, pr[i].m, pr[i].n));
        L.grad_bytes = std::max(L.grad_bytes, grad_scratch_bytes(o.a_type, o.restart_tier, pr[i].m, pr[i].n));
    }
    L.gemv_sc = take(std::max<size_t>(L.gemv_bytes, 1));
    L.grad_sc = take(std::max<size_t>(L.grad_bytes, 1));
    L.total = off;
    return L;
}

// Everything a captured segment depends on (grids, LDS, scalars, scratch).
// The per-problem GEMV pair of the restart (no grouped exact pass) takes
// each problem's matrix, shape and vectors as launch arguments, so its
// graph also depends on those (`prob`).
struct SegKey {
    int device, kind, nprob, nitems, nrx, stream, g, threads, rounds, a_type, restart_tier, max_iters, restarts;
    size_t lds;
    double tol;
    const void* scratch;
    std::vector<int64_t> prob;
    bool operator==(const SegKey& k) const {
        return device == k.device && kind == k.kind && nprob == k.nprob && nitems == k.nitems && nrx == k.nrx &&
               stream == k.stream &&
               g == k.g && threads == k.threads && rounds == k.rounds && a_type == k.a_type &&
               restart_tier == k.restart_tier && max_iters == k.max_iters && restarts == k.restarts && lds == k.lds &&
               tol == k.tol && scratch == k.scratch && prob == k.prob;
    }
};

struct SegCache {
    struct Entry {
        SegKey key;
        hipGraphExec_t exec;
    };
    std::mutex mu;
    std::vector<Entry> v;
    static constexpr int kMax = 8;
    hipGraphExec_t find(const SegKey& k) {
        std::lock_guard<std::mutex> lock(mu);
        for (const Entry& e : v)
            if (e.key == k) return e.exec;
        return nullptr;
    }
    // Every call synchronizes before it returns, so an evicted graph is idle.
    void insert(const SegKey& k, hipGraphExec_t exec) {
        std::lock_guard<std::mutex> lock(mu);
        if (static_cast<int>(v.size()) >= kMax) {
            (void)hipGraphExecDestroy(v.front().exec);
            v.erase(v.begin());
        }
        v.push_back(Entry{k, exec});
    }
};
SegCache& seg_cache() {
    static SegCache c;
    return c;
}

struct Grouped {
    tomos_lsqr_options o{};
    int nprob = 0;
    const tomos_lsq_problem* pr = nullptr;
    GPlan L;
    char* sc = nullptr;
    hipStream_t s = nullptr;
    std::vector<double*> q64;    // per problem: A w of the restart (FP64, m)
    std::vector<double*> wbuf;   // per problem: w, accumulated in the scratch (FP64, n)
    std::vector<double*> rfull;  // per problem: A^T A w of the restart (FP64, n)

    GProb* dprobs() const { return reinterpret_cast<GProb*>(sc + L.probs); }
    GroupProb* dgprobs() const { return reinterpret_cast<GroupProb*>(sc + L.gprobs); }
    GroupItem* ditems() const { return reinterpret_cast<GroupItem*>(sc + L.items_off); }
    GState* states() const { return reinterpret_cast<GState*>(sc + L.states); }
    size_t lds() const { return static_cast<size_t>(std::max<int64_t>(L.nmax, 1)) * 8; }

    void upload(const double* const* g) {
        std::vector<GProb> bp(static_cast<size_t>(nprob));
        std::vector<GroupProb> gp(static_cast<size_t>(nprob));
        q64.assign(static_cast<size_t>(nprob), nullptr);
        wbuf.assign(static_cast<size_t>(nprob), nullptr);
        rfull.assign(static_cast<size_t>(nprob), nullptr);
        for (int i = 0; i < nprob; ++i) {
            const tomos_lsq_problem& p = pr[i];
            GProb& q = bp[static_cast<size_t>(i)];
            const size_t nb = a256(static_cast<size_t>(std::max<int64_t>(p.n, 1)) * 8);
            char* vb = sc + L.vec_at[static_cast<size_t>(i)];
            q.m = p.m;
            q.n = p.n;
            q.ldr = p.ldr;
            q.lda = p.lda;
            q.a = p.a;
            q.wr = p.wr;
            q.g = g[i];
            q.rinv = p.rinv;
            q.res = reinterpret_cast<double*>(vb);
            q.z = reinterpret_cast<double*>(vb + nb);
            q.d = reinterpret_cast<double*>(vb + 2 * nb);
            q.rfull = reinterpret_cast<double*>(vb + 3 * nb);
            q.w = reinterpret_cast<double*>(vb + 4 * nb);
            q.y = reinterpret_cast<double*>(vb + 5 * nb);
            q.rzp = reinterpret_cast<double*>(vb + kNVecs * nb);
            char* qb = vb + kNVecs * nb + rzp_bytes(p.n);
            float* qf = reinterpret_cast<float*>(qb);
            q64[static_cast<size_t>(i)] =
                reinterpret_cast<double*>(qb + a256(static_cast<size_t>(std::max<int64_t>(p.m, 1)) * 4));
            wbuf[static_cast<size_t>(i)] = q.w;
            rfull[static_cast<size_t>(i)] = q.rfull;
            q.st = states() + i;
            q.rpart = reinterpret_cast<double*>(sc + L.rpart_at[static_cast<size_t>(i)]);
            q.nrx = L.use_rx ? L.prob_rx[static_cast<size_t>(i)] : 0;
            q.nitems = L.prob_items[static_cast<size_t>(i)];
            q.sumsq = reinterpret_cast<double*>(sc + L.sumsq_at[static_cast<size_t>(i)]);
            q.tpart = reinterpret_cast<float*>(sc + L.tpart_at[static_cast<size_t>(i)]);
            GroupProb& r = gp[static_cast<size_t>(i)];
            r.m = p.m;
            r.n = p.n;
            r.lda = p.lda;
            r.a = p.a;
            r.p = q.d;
            r.u = qf;
            r.u_coef = &q.st->zero;
            r.skip = &q.st->skip_iter;
            r.sumsq = q.sumsq;
            r.tpart = q.tpart;
            r.rounds = fused_rounds(L.g, L.threads, std::max<int64_t>(p.n, 1));
        }
        TOMOS_HIP_CHECK(hipMemcpyAsync(dprobs(), bp.data(), sizeof(GProb) * bp.size(), hipMemcpyHostToDevice, s));
        TOMOS_HIP_CHECK(hipMemcpyAsync(dgprobs(), gp.data(), sizeof(GroupProb) * gp.size(), hipMemcpyHostToDevice, s));
        if (L.nitems)
            TOMOS_HIP_CHECK(hipMemcpyAsync(ditems(), L.items.data(), sizeof(GroupItem) * L.items.size(),
                                           hipMemcpyHostToDevice, s));
        if (L.nrx)
            TOMOS_HIP_CHECK(hipMemcpyAsync(sc + L.rx_off, L.rx.data(), sizeof(RxItem) * L.rx.size(),
                                           hipMemcpyHostToDevice, s));
        TOMOS_HIP_CHECK(hipStreamSynchronize(s));  // the host vectors must outlive the copies
    }

    // Restart sequence: per problem the exact-tier pair q64 = A w,
    // rfull = A^T q64 (no-ops unless that problem restarts with w != 0),
    // then the restart block of every problem.
    void restart_seq() {
        if (L.use_rx) {
            if (L.nrx > 0) {
                const RxItem* d = reinterpret_cast<const RxItem*>(sc + L.rx_off);
                const unsigned nb = static_cast<unsigned>(L.nrx);
                if (L.rx2 && o.a_type == TOMOS_FP32_BF16) k_gadj_rx2<1><<<nb, kRx2Threads, 0, s>>>(dprobs(), d);
                else if (L.rx2) k_gadj_rx2<2><<<nb, kRx2Threads, 0, s>>>(dprobs(), d);
                else if (o.a_type == TOMOS_F32) k_gadj_rx<0><<<nb, kRxThreads, 0, s>>>(dprobs(), d);
                else if (o.a_type == TOMOS_FP32_BF16) k_gadj_rx<1><<<nb, kRxThreads, 0, s>>>(dprobs(), d);
                else k_gadj_rx<2><<<nb, kRxThreads, 0, s>>>(dprobs(), d);
                TOMOS_LAUNCH_CHECK("k_gadj_rx");
            }
            k_gadj_restart<<<static_cast<unsigned>(nprob), kTail, 0, s>>>(dprobs(), o.tol, o.restarts);
            TOMOS_LAUNCH_CHECK("k_gadj_restart");
            precond_seq();
            return;
        }
        void* gsc = sc + L.gemv_sc;
        for (int i = 0; i < nprob; ++i) {
            const tomos_lsq_problem& p = pr[i];
            if (p.n == 0) continue;
            GState* st = states() + i;
            Coef c;
            c.alpha = 1.0;
            c.beta = 0.0;
            c.skip = &st->skip_restart;
            const size_t k = static_cast<size_t>(i);
            gemv(o.a_type, o.restart_tier, TOMOS_NO_TRANSPOSE, c, p.m, p.n, p.a, p.wr, p.lda, TOMOS_F64, wbuf[k],
                 TOMOS_F64, q64[k], gsc, L.gemv_bytes, s);
            gemv(o.a_type, o.restart_tier, TOMOS_TRANSPOSE, c, p.m, p.n, p.a, p.wr, p.lda, TOMOS_F64, q64[k],
                 TOMOS_F64, rfull[k], gsc, L.gemv_bytes, s);
        }
        k_gadj_restart<<<static_cast<unsigned>(nprob), kTail, 0, s>>>(dprobs(), o.tol, o.restarts);
        TOMOS_LAUNCH_CHECK("k_gadj_restart");
        precond_seq();
    }
    // z = R^-1 R^-T res, rho and d for the problems whose pc flag is set
    void precond_seq() {
        k_gadj_pre1<<<dim3(static_cast<unsigned>(ceil_div(L.nmax, static_cast<int64_t>(kTail / 64))),
                           static_cast<unsigned>(nprob)),
                      kTail, 0, s>>>(dprobs());
        k_gadj_pre2<<<dim3(static_cast<unsigned>(ceil_div(L.nmax, static_cast<int64_t>(kPreRows))),
                           static_cast<unsigned>(nprob)),
                      kTail, 0, s>>>(dprobs());
        k_gadj_fin<<<static_cast<unsigned>(nprob), 256, 0, s>>>(dprobs());
        TOMOS_LAUNCH_CHECK("k_gadj_pre");
    }
    void iter_seq() {
        for (int it = 0; it < kSegIters; ++it) {
            if (L.nitems)
                grouped_fused_launch(o.a_type, L.g, L.threads, L.rounds, dgprobs(), ditems(), L.nitems, s, L.stream);
            k_gadj_step<<<static_cast<unsigned>(nprob), kTail, 0, s>>>(dprobs(), o.tol, o.max_iters,
                                                                      adjoint_inner_drop(o));
            precond_seq();
        }
        TOMOS_LAUNCH_CHECK("k_gadj_step");
    }

    SegKey key(int kind) const {
        SegKey k;
        TOMOS_HIP_CHECK(hipGetDevice(&k.device));
        k.kind = kind;
        k.nprob = nprob;
        k.nitems = L.nitems;
        k.g = L.g;
        k.threads = L.threads;
        k.rounds = L.rounds;
        k.a_type = o.a_type;
        k.restart_tier = o.restart_tier;
        k.max_iters = o.max_iters;
        k.restarts = o.restarts;
        k.lds = lds();
        k.tol = o.tol;
        k.scratch = sc;
        k.nrx = L.nrx;
        k.stream = L.stream ? 1 : 0;
        if (kind == 0 && !L.use_rx) {
            for (int i = 0; i < nprob; ++i) {
                const tomos_lsq_problem& p = pr[i];
                for (int64_t v : {p.m, p.n, p.lda, static_cast<int64_t>(reinterpret_cast<uintptr_t>(p.a)),
                                  static_cast<int64_t>(reinterpret_cast<uintptr_t>(p.wr))})
                    k.prob.push_back(v);
            }
        }
        return k;
    }
    hipGraphExec_t graph(int kind) {
        const SegKey k = key(kind);
        hipGraphExec_t exec = seg_cache().find(k);
        if (exec) return exec;
        hipGraph_t gr = nullptr;
        TOMOS_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        try {
            if (kind == 0) restart_seq();
            else iter_seq();
        } catch (...) {
            hipGraph_t g2 = nullptr;
            (void)hipStreamEndCapture(s, &g2);
            if (g2) (void)hipGraphDestroy(g2);
            throw;
        }
        TOMOS_HIP_CHECK(hipStreamEndCapture(s, &gr));
        const hipError_t e = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        TOMOS_HIP_CHECK(e);
        seg_cache().insert(k, exec);
        return exec;
    }

    void states_to_host(std::vector<GState>& h) {
        TOMOS_HIP_CHECK(hipMemcpyAsync(h.data(), states(), sizeof(GState) * h.size(), hipMemcpyDeviceToHost, s));
        TOMOS_HIP_CHECK(hipStreamSynchronize(s));
    }

    void run(const double* const* g, double* const* w, int* iterations) {
        upload(g);
        k_gadj_init<<<static_cast<unsigned>(ceil_div(nprob, 64)), 64, 0, s>>>(dprobs(), nprob);
        k_gadj_zero_w<<<static_cast<unsigned>(nprob), kTail, 0, s>>>(dprobs());
        TOMOS_LAUNCH_CHECK("adjoint (grouped) set-up");
        hipGraphExec_t rexec = graph(0);
        hipGraphExec_t iexec = graph(1);
        std::vector<GState> h(static_cast<size_t>(nprob));
        // a restart cycle takes one restart launch and at most
        // ceil(max_iters / kSegIters) + 1 rounds: a bound on the rounds
        const int64_t max_rounds =
            static_cast<int64_t>(o.restarts + 1) * (ceil_div(static_cast<int64_t>(o.max_iters), kSegIters) + 1) + 2;
        bool restart_pending = true;  // every problem starts with its first restart
        bool all_done = false;
        for (int64_t round = 0; round < max_rounds && !all_done; ++round) {
            if (restart_pending) TOMOS_HIP_CHECK(hipGraphLaunch(rexec, s));
            TOMOS_HIP_CHECK(hipGraphLaunch(iexec, s));
            states_to_host(h);
            all_done = true;
            restart_pending = false;
            for (const GState& b : h) {
                all_done = all_done && b.phase == kPhaseDone;
                restart_pending = restart_pending || b.phase == kPhaseRestart;
            }
        }
        if (!all_done) throw Error(kErrInternal, "adjoint (grouped): a problem did not finish within the round bound");
        for (int i = 0; i < nprob; ++i) {
            if (pr[i].n > 0)
                TOMOS_HIP_CHECK(hipMemcpyAsync(w[i], wbuf[static_cast<size_t>(i)], static_cast<size_t>(pr[i].n) * 8,
                                               hipMemcpyDeviceToDevice, s));
            if (iterations) iterations[i] = h[static_cast<size_t>(i)].completed;
        }
    }
};

}  // namespace

size_t adjoint_grouped_scratch_bytes(const tomos_lsqr_options& o, int nprob, const tomos_lsq_problem* probs) {
    grouped_validate(o, nprob, probs);
    return grouped_plan(o, nprob, probs).total;
}

void adjoint_grouped(const tomos_lsqr_options& o, int nprob, const tomos_lsq_problem* probs, const double* const* g,
                     double* const* w, double* const* grad_b, float* const* grad_a, const int64_t* ldg,
                     void* scratch, size_t scratch_bytes, int* iterations, hipStream_t stream) {
    grouped_validate(o, nprob, probs);
    if (nprob == 0) return;
    Grouped G;
    G.o = o;
    G.nprob = nprob;
    G.pr = probs;
    G.L = grouped_plan(o, nprob, probs);
    require(scratch != nullptr && scratch_bytes >= G.L.total, "tomos_lsq_adjoint_batched: scratch missing or too small");
    G.sc = static_cast<char*>(scratch);
    // a graph cannot be captured from the legacy NULL stream: a private
    // stream (default priority, ordered after earlier NULL-stream work)
    hipStream_t own = nullptr;
    if (stream == nullptr) TOMOS_HIP_CHECK(hipStreamCreate(&own));
    G.s = own ? own : stream;
    try {
        G.run(g, w, iterations);
        std::vector<GGrad> tab;  // outlives its copy: the stream is synchronized below
        if (G.L.rx2) {
            // one launch for every problem with columns (rows of n = 0
            // problems: lsq_grad's zero fill)
            tab.resize(static_cast<size_t>(nprob));
            for (int i = 0; i < nprob; ++i) {
                const tomos_lsq_problem& q = probs[i];
                const int64_t lg = ldg ? ldg[i] : q.m;
                float* ga = grad_a ? grad_a[i] : nullptr;
                if (q.m > 0 && q.n > 0) {
                    require(q.b && q.x && w[i] && grad_b[i], "tomos_lsq_grad: null pointer");
                    require(!ga || lg >= q.m, "tomos_lsq_grad: ldg < m");
                }
                tab[static_cast<size_t>(i)] = GGrad{q.m, q.n, q.lda, lg, q.a, q.wr, q.b, q.x, w[i], grad_b[i], ga};
                if (q.n == 0)
                    lsq_grad(o.a_type, o.restart_tier, q.m, q.n, q.a, q.wr, q.lda, q.b, q.x, w[i], grad_b[i], ga, lg,
                             G.sc + G.L.grad_sc, G.L.grad_bytes, G.s);
            }
            if (G.L.nrx > 0) {
                GGrad* dt = reinterpret_cast<GGrad*>(G.sc + G.L.grad_tab);
                TOMOS_HIP_CHECK(hipMemcpyAsync(dt, tab.data(), sizeof(GGrad) * tab.size(), hipMemcpyHostToDevice, G.s));
                const RxItem* d = reinterpret_cast<const RxItem*>(G.sc + G.L.rx_off);
                const unsigned nb = static_cast<unsigned>(G.L.nrx);
                if (o.a_type == TOMOS_FP32_BF16) k_gadj_grad<1><<<nb, kRx2Threads, 0, G.s>>>(dt, d);
                else k_gadj_grad<2><<<nb, kRx2Threads, 0, G.s>>>(dt, d);
                TOMOS_LAUNCH_CHECK("k_gadj_grad");
            }
        } else {
            for (int i = 0; i < nprob; ++i) {
                const tomos_lsq_problem& q = probs[i];
                lsq_grad(o.a_type, o.restart_tier, q.m, q.n, q.a, q.wr, q.lda, q.b, q.x, w[i], grad_b[i],
                         grad_a ? grad_a[i] : nullptr, ldg ? ldg[i] : q.m, G.sc + G.L.grad_sc, G.L.grad_bytes, G.s);
            }
        }
        TOMOS_HIP_CHECK(hipStreamSynchronize(G.s));
    } catch (...) {
        if (own) {
            (void)hipStreamSynchronize(own);
            (void)hipStreamDestroy(own);
        }
        throw;
    }
    if (own) TOMOS_HIP_CHECK(hipStreamDestroy(own));
}

}  // namespace tomos_hip
