// Many independent, uneven LSQR problems solved together on one GPU.
//
// Each problem i has its own A_i (m_i x n_i, packed planes or FP32), b_i,
// the sketch-QR preconditioner R_i^-1 and output x_i; every problem runs
// the restarted, right-preconditioned LSQR of the single-problem solver
// (the control flow of fused_lsqr_impl, src/core/solver/lsqr.cpp:685-747,
// per problem: restart check, exact-residual initialize, inner
// iterations, per-iteration stop test) with its own FP64 scalars and stop
// flag.  What is batched is the launch: one step of ALL problems is three
// kernels whatever their number and shapes,
//   restart:   grouped exact residual  u_i = b_i - A_i x_i (wh+wr, FP64
//              accumulation), then the grouped single-pass kernel with
//              p = 0, coefficient 1 (A_i^T u_i and |u_i|^2), then the tail;
//   iteration: the grouped single-pass kernel (u_i <- A_i p_i - (alpha/beta)
//              u_i and A_i^T u_i from one read of A_i) and the tail,
// where the tail is one 1024-thread block per problem doing the n-sized
// work: beta, A^T u / beta from the partial sums, R^-T (A^T u), alpha and
// v, the Givens rotation, z = P w and x updates, the stop test and the
// next p = R^-1 v.  A converged problem's blocks exit at once.
//
// The whole solve (state set-up and `restarts` restarts of 1 + max_iters
// steps, every launch) is captured as one hipGraph, instantiated once per
// launch configuration and kept, and submitted as ONE graph launch; a
// stopped problem's blocks exit at once in the launches that follow, and
// the host synchronizes once at the end.  The grouped kernels' work lists
// (tile ranges per block) are built on the host per call and read by the
// graph from the scratch; their shape (row groups, tiles per block) is a
// tunable.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "codec.hpp"
#include "common.hpp"
#include "lsqr_fused.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

constexpr int kT = 256;
constexpr int kTail = 1024;

struct BState {
    double alpha, beta, rho_bar, phi_bar, rhsnorm;
    double w_coef;    // -theta / rho of the previous step (z recurrence)
    double u_coef;    // -alpha / beta: next u_raw = A p + u_coef u_raw
    double one;       // 1.0: u coefficient of the restart's single-pass call
    int done, completed, hist_len, z_fresh;
    int tail_on;  // the current step's tail runs (set by k_batch_tail_a)
    int pad[3];
};

// Per-problem device descriptor (the tail and residual kernels).
struct BProb {
    int64_t m, n, lda, ldr;
    const void* a;
    const uint16_t* wr;
    const double* b;
    double* x;
    const double* rinv;
    float* u;
    double* v;
    double* z;
    double* pv;
    double* temp;  // A^T u / beta (n)
    double* src;   // R^-T temp (n)
    BState* st;
    double* hist;
    int hist_cap;
    int nitems;        // single-pass partial slots of this problem
    double* sumsq;     // [nitems]
    float* tpart;      // [nitems][n]
};

struct RowItem {
    int prob;
    int64_t row0, row1;
};

__global__ void k_bstate_init(BProb* P, int nprob) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= nprob) return;
    BState* s = P[i].st;
    s->alpha = s->beta = s->rho_bar = s->phi_bar = s->rhsnorm = 0.0;
    s->w_coef = s->u_coef = 0.0;
    s->one = 1.0;
    s->done = s->completed = s->hist_len = 0;
    s->z_fresh = 1;
    s->tail_on = 0;
}

// |b_i| (FP64, fixed order), x_i = 0
__global__ __launch_bounds__(kTail) void k_batch_setup(BProb* P) {
    const BProb& p = P[blockIdx.x];
    __shared__ double sh[kTail / 64];
    double a = 0.0;
    for (int64_t i = threadIdx.x; i < p.m; i += kTail) a += p.b[i] * p.b[i];
    for (int64_t j = threadIdx.x; j < p.n; j += kTail) p.x[j] = 0.0;
    for (int o = 32; o >= 1; o >>= 1) a += __shfl_xor(a, o);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kTail / 64; ++w) t += sh[w];
        p.st->rhsnorm = sqrt(t);
        if (p.st->rhsnorm == 0.0 || p.n == 0) p.st->done = 1;
    }
}

// check_stopping_criteria at the top of a restart (lsqr.cpp:417-425, :708).
// confirm: a problem whose estimate fell below tol (done == 2) runs this
// restart again, and only its true residual (the tail's initialize test)
// stops it.
__global__ void k_batch_restart_check(BProb* P, int nprob, double tol, double maxrr, int confirm) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= nprob) return;
    BState* s = P[i].st;
    if (s->done == 2) s->done = 0;
    if (s->done || s->completed == 0) return;
    const double rr = fabs(s->phi_bar) / s->rhsnorm;
    if (s->hist_len < P[i].hist_cap) P[i].hist[s->hist_len] = rr;
    s->hist_len += 1;
    if (!confirm && (rr < tol || rr < maxrr)) s->done = 1;
}

// Exact value of the RPL rows in one 16-byte vector (hi h, lo l).
template <int FMT>
__device__ __forceinline__ void residual_decode(const uint4& h, const uint4& l, float* av) {
    if constexpr (FMT == 0) {
        av[0] = __uint_as_float(h.x);
        av[1] = __uint_as_float(h.y);
        av[2] = __uint_as_float(h.z);
        av[3] = __uint_as_float(h.w);
    } else {
        const uint32_t hw[4] = {h.x, h.y, h.z, h.w}, lw[4] = {l.x, l.y, l.z, l.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const uint32_t hh = half ? hw[k] >> 16 : hw[k] & 0xFFFFu;
                const uint32_t ll = half ? lw[k] >> 16 : lw[k] & 0xFFFFu;
                av[2 * k + half] = FMT == 1 ? __uint_as_float((hh << 16) | ll) : codec::fp32_16_decode_exact(hh, ll);
            }
        }
    }
}

// u_i = b_i - A_i x_i for a range of rows: exact decode (wh+wr for packed),
// FP64 accumulation, FP32 store.  One lane per 16-byte row vector, kU
// columns' loads in flight per lane.  Non-temporal loads: a batch worth
// grouping is far larger than the MALL, and the residual reads both planes
// (measured at C5: the forward solve 34.3 -> 33.3 ms against cacheable loads).
// XZERO (the first restart of a solve, x = 0): A x is the zero vector for
// finite A, so u = b and A is not read.
template <int FMT, bool XZERO>
__global__ __launch_bounds__(kT) void k_batch_residual(const BProb* __restrict__ P,
                                                       const RowItem* __restrict__ items) {
    constexpr int RPL = FMT == 0 ? 4 : 8;  // FP32 / packed rows per 16-byte vector
    constexpr int ESZ = FMT == 0 ? 4 : 2;
    constexpr int kU = 8;
    const RowItem it = items[blockIdx.x];
    const BProb& p = P[it.prob];
    if (p.st->done) return;
    const int64_t r = it.row0 + static_cast<int64_t>(threadIdx.x) * RPL;
    if (r >= it.row1) return;
    const int nv = static_cast<int>(min<int64_t>(RPL, p.m - r));
    if constexpr (XZERO) {
        for (int k = 0; k < nv; ++k) p.u[r + k] = static_cast<float>(p.b[r + k]);
        return;
    }
    const char* hb = static_cast<const char*>(p.a) + r * ESZ;
    const uint16_t* lb = FMT != 0 ? p.wr + r : nullptr;
    double acc[RPL];
#pragma unroll
    for (int k = 0; k < RPL; ++k) acc[k] = 0.0;
    int64_t j = 0;
    for (; j + kU <= p.n; j += kU) {
        uint4 h[kU], l[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            h[u] = load_stream16(hb + (j + u) * p.lda * ESZ);
            if constexpr (FMT != 0) l[u] = load_stream16(lb + (j + u) * p.lda);
            else l[u] = h[u];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            float av[RPL];
            residual_decode<FMT>(h[u], l[u], av);
            const double xj = p.x[j + u];
#pragma unroll
            for (int k = 0; k < RPL; ++k) acc[k] = fma(static_cast<double>(av[k]), xj, acc[k]);
        }
    }
    for (; j < p.n; ++j) {
        const uint4 h = *reinterpret_cast<const uint4*>(hb + j * p.lda * ESZ);
        const uint4 l = FMT != 0 ? *reinterpret_cast<const uint4*>(lb + j * p.lda) : h;
        float av[RPL];
        residual_decode<FMT>(h, l, av);
        const double xj = p.x[j];
#pragma unroll
        for (int k = 0; k < RPL; ++k) acc[k] = fma(static_cast<double>(av[k]), xj, acc[k]);
    }
    for (int k = 0; k < nv; ++k) p.u[r + k] = static_cast<float>(p.b[r + k] - acc[k]);
}

__device__ __forceinline__ double tail_sum(double v, double* sh) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < kTail / 64; ++i) t += sh[i];  // same order in every thread
    return t;
}

// The n-sized part of one LSQR step (lsqr.cpp:557-590 after A^T u, and
// :632-679), with P = R^-1, in four launches so that the two products with
// R^-1 (n^2 / 2 entries each) spread over many blocks per problem instead
// of one block streaming R^-1 twice:
//   a (one block per problem): beta = |u_raw|, temp = A^T u_raw / beta;
//   b (a wave per column):     src = R^-T temp;
//   c (one block per problem): q = src - beta v (iteration) or src
//        (initialize), alpha = |q|, v = q/alpha; iteration: Givens update,
//        z = P v_old + w_coef_prev z (z = P w), x += (phi/rho) z;
//        initialize: rho_bar = alpha, phi_bar = beta, true-residual stop
//        test; u_coef = -alpha/beta, stop test + history;
//   d (64 rows a block):       next p = R^-1 v.
// tail_on (written by a) gates b, c and d: a problem that has stopped skips
// the whole tail.
__global__ __launch_bounds__(kTail) void k_batch_tail_a(const BProb* __restrict__ P) {
    const BProb& p = P[blockIdx.x];
    BState* S = p.st;
    __shared__ double sh[kTail / 64];
    if (S->done) {
        if (threadIdx.x == 0) S->tail_on = 0;
        return;
    }
    const int64_t n = p.n;
    double ss = 0.0;
    for (int k = threadIdx.x; k < p.nitems; k += kTail) ss += p.sumsq[k];
    const double beta = sqrt(tail_sum(ss, sh));
    const double inv_beta = 1.0 / beta;
    for (int64_t j = threadIdx.x; j < n; j += kTail) {
        double a = 0.0;
        for (int k = 0; k < p.nitems; ++k) a += static_cast<double>(p.tpart[static_cast<int64_t>(k) * n + j]);
        p.temp[j] = a * inv_beta;
    }
    if (threadIdx.x == 0) {
        S->beta = beta;
        S->tail_on = 1;
    }
}

// src = R^-T temp: column j of R^-1 (upper: rows 0..j) dotted with temp,
// one wave per column, lanes along the column; grid (ceil(nmax / 16), problems)
__global__ __launch_bounds__(kTail) void k_batch_tail_b(const BProb* __restrict__ P) {
    const BProb& p = P[blockIdx.y];
    if (!p.st->tail_on) return;
    const int lane = threadIdx.x & 63;
    const int64_t j = static_cast<int64_t>(blockIdx.x) * (kTail / 64) + (threadIdx.x >> 6);
    if (j >= p.n) return;
    const double* col = p.rinv + j * p.ldr;
    double a = 0.0;
    for (int64_t i = lane; i <= j; i += 64) a += col[i] * p.temp[i];
    for (int o = 32; o >= 1; o >>= 1) a += __shfl_xor(a, o);
    if (lane == 0) p.src[j] = a;
}

__global__ __launch_bounds__(kTail) void k_batch_tail_c(const BProb* __restrict__ P, int init, double tol,
                                                        double maxrr, int confirm) {
    const BProb& p = P[blockIdx.x];
    BState* S = p.st;
    if (!S->tail_on) return;
    __shared__ double sh[kTail / 64];
    __shared__ double bc[2];
    const int64_t n = p.n;
    // state read before any barrier (thread 0 rewrites it at the end)
    const double rho_bar = S->rho_bar, phi_bar = S->phi_bar, wc_prev = S->w_coef, beta = S->beta;
    const bool fresh = S->z_fresh != 0;
    double qq = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += kTail) {
        const double q = init ? p.src[j] : p.src[j] - beta * p.v[j];
        qq += q * q;
    }
    const double alpha = sqrt(tail_sum(qq, sh));
    if (threadIdx.x == 0) {
        if (init) {
            bc[0] = 0.0;
            bc[1] = 0.0;
        } else {
            const double rho = hypot(rho_bar, beta);
            const double c = rho_bar / rho, sn = beta / rho;
            const double theta = sn * alpha;
            const double phi = c * phi_bar;
            S->rho_bar = -c * alpha;
            S->phi_bar = sn * phi_bar;
            bc[0] = phi / rho;     // x coefficient
            bc[1] = -theta / rho;  // w coefficient
        }
    }
    __syncthreads();
    const double xc = bc[0], wc = bc[1];
    const double inv_alpha = 1.0 / alpha;
    for (int64_t j = threadIdx.x; j < n; j += kTail) {
        const double q = init ? p.src[j] : p.src[j] - beta * p.v[j];
        if (!init) {
            const double pj = p.pv[j];
            const double zj = fresh ? pj : pj + wc_prev * p.z[j];
            p.z[j] = zj;
            p.x[j] += xc * zj;
        }
        p.v[j] = q * inv_alpha;
    }
    if (threadIdx.x != 0) return;
    S->alpha = alpha;
    S->u_coef = -alpha / beta;
    if (init) {
        S->rho_bar = alpha;
        S->phi_bar = beta;
        S->z_fresh = 1;
        S->w_coef = 0.0;
        const double rr = beta / S->rhsnorm;
        if (rr < tol || rr < maxrr) S->done = 1;
        return;
    }
    S->z_fresh = 0;
    S->w_coef = wc;
    S->completed += 1;
    const double rr = fabs(S->phi_bar) / S->rhsnorm;
    if (S->hist_len < p.hist_cap) p.hist[S->hist_len] = rr;
    S->hist_len += 1;
    if (rr < tol || rr < maxrr) S->done = confirm ? 2 : 1;
}

// next p = R^-1 v (upper: row i uses columns i..n-1), kTailRows rows a
// block, kTailSlices column slices (slice k takes columns i0 + k,
// i0 + k + kTailSlices, ...; a wave reads one column's kTailRows entries
// per step) added in a fixed order; grid (ceil(nmax / 64), problems)
constexpr int kTailRows = 64;
constexpr int kTailSlices = kTail / kTailRows;
__global__ __launch_bounds__(kTail) void k_batch_tail_d(const BProb* __restrict__ P) {
    const BProb& p = P[blockIdx.y];
    if (!p.st->tail_on) return;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kTailRows;
    if (i0 >= p.n) return;
    __shared__ double part[kTailSlices][kTailRows];
    const int r = threadIdx.x & (kTailRows - 1), sl = threadIdx.x / kTailRows;
    const int64_t i = i0 + r;
    double a = 0.0;
    if (i < p.n) {
#pragma unroll 4
        for (int64_t j = i0 + sl; j < p.n; j += kTailSlices)
            if (j >= i) a += p.rinv[i + j * p.ldr] * p.v[j];
    }
    part[sl][r] = a;
    __syncthreads();
    if (threadIdx.x < kTailRows && i < p.n) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < kTailSlices; ++k) t += part[k][r];
        p.pv[i] = t;
    }
}

size_t a256(size_t v) { return (v + 255) / 256 * 256; }

struct Plan {
    int g = 0, threads = 256, rounds = 0;
    int64_t tiles_per_item = 1;
    int nitems = 0, nrow_items = 0;
    int64_t nmax = 0;
    std::vector<GroupItem> items;
    std::vector<RowItem> row_items;
    std::vector<int> prob_items;  // items per problem
    bool stream = false;          // non-temporal loads in the grouped pass (stream_loads of its bytes)
    // scratch layout
    size_t probs, gprobs_iter, gprobs_init, items_off, row_items_off, zeros, states, hist, sumsq,
        tpart, vecs, total;
    std::vector<size_t> sumsq_at, tpart_at, vec_at, hist_at;
    int hist_cap = 0;
};

int rpl_for(int a_type) { return a_type == TOMOS_F32 ? 4 : 8; }

Plan make_plan(const tomos_lsqr_options& o, int nprob, const tomos_lsq_problem* pr, int g_req,
               int nt_req, int tiles_req) {
    Plan L;
    for (int i = 0; i < nprob; ++i) L.nmax = std::max(L.nmax, pr[i].n);
    int g = g_req, nt = nt_req, bpc = 0;
    if (g <= 0 && nt <= 0) lsqr_fused_config_get(o.a_type, L.nmax, &g, &bpc, &nt);
    if (nt <= 0) nt = 256;
    if (g <= 0 || fused_rounds(g, nt, L.nmax) == 0) {
        const FusedShape d = lsqr_fused_default_shape(L.nmax);
        g = d.g;
        nt = d.threads;
    }
    L.g = g;
    L.threads = nt;
    L.rounds = fused_rounds(g, nt, L.nmax);
    require(L.rounds > 0, "lsqr_batched: n too large for the single-pass kernel (max 2048)");
    const int64_t tr = fused_tile_rows(o.a_type, g);
    int64_t tiles_total = 0;
    for (int i = 0; i < nprob; ++i) tiles_total += ceil_div(std::max<int64_t>(pr[i].m, 1), tr);
    if (tiles_req > 0) {
        L.tiles_per_item = tiles_req;
    } else {
        // about four waves of resident blocks over the whole batch
        const int64_t resident = static_cast<int64_t>(grouped_fused_occupancy(o.a_type, g, nt, L.rounds)) * cu_count();
        L.tiles_per_item = std::max<int64_t>(1, ceil_div(tiles_total, 4 * resident));
    }
    L.prob_items.resize(nprob);
    for (int i = 0; i < nprob; ++i) {
        const int64_t nt = pr[i].n > 0 ? ceil_div(std::max<int64_t>(pr[i].m, 1), tr) : 0;
        int k = 0;
        for (int64_t t = 0; t < nt; t += L.tiles_per_item, ++k)
            L.items.push_back(GroupItem{i, k, t, std::min(nt, t + L.tiles_per_item)});
        L.prob_items[i] = k;
        const int64_t rows_per = static_cast<int64_t>(kT) * rpl_for(o.a_type);
        for (int64_t r = 0; pr[i].n > 0 && r < pr[i].m; r += rows_per)
            L.row_items.push_back(RowItem{i, r, std::min(pr[i].m, r + rows_per)});
    }
    L.nitems = static_cast<int>(L.items.size());
    L.nrow_items = static_cast<int>(L.row_items.size());
    int64_t pass_bytes = 0;
    for (int i = 0; i < nprob; ++i) pass_bytes += pr[i].m * pr[i].n * (o.a_type == TOMOS_F32 ? 4 : 2);
    L.stream = stream_loads(pass_bytes);
    const int64_t cap = static_cast<int64_t>(std::max(o.restarts, 0)) * (std::max(o.max_iters, 0) + 1) + 1;
    L.hist_cap = static_cast<int>(std::min<int64_t>(cap, 1 << 20));
    size_t off = 0;
    auto take = [&](size_t b) {
        const size_t at = off;
        off = a256(off + b);
        return at;
    };
    L.probs = take(sizeof(BProb) * nprob);
    L.gprobs_iter = take(sizeof(GroupProb) * nprob);
    L.gprobs_init = take(sizeof(GroupProb) * nprob);
    L.items_off = take(sizeof(GroupItem) * std::max(L.nitems, 1));
    L.row_items_off = take(sizeof(RowItem) * std::max(L.nrow_items, 1));
    L.zeros = take(static_cast<size_t>(std::max<int64_t>(L.nmax, 1)) * 8);
    L.states = take(sizeof(BState) * nprob);
    for (int i = 0; i < nprob; ++i) {
        L.hist_at.push_back(take(static_cast<size_t>(L.hist_cap) * 8));
        L.sumsq_at.push_back(take(static_cast<size_t>(std::max(L.prob_items[i], 1)) * 8));
        L.tpart_at.push_back(take(static_cast<size_t>(std::max(L.prob_items[i], 1)) *
                                  std::max<int64_t>(pr[i].n, 1) * 4));
        // u (FP32, m) then v, z, pv, temp, src (FP64, n)
        L.vec_at.push_back(take(a256(static_cast<size_t>(std::max<int64_t>(pr[i].m, 1)) * 4) +
                                5 * a256(static_cast<size_t>(std::max<int64_t>(pr[i].n, 1)) * 8)));
    }
    L.total = off;
    return L;
}

void validate(const tomos_lsqr_options& o, int nprob, const tomos_lsq_problem* pr) {
    require(nprob >= 0 && (nprob == 0 || pr), "lsqr_batched: bad problem list");
    require(o.a_type == TOMOS_F32 || o.a_type == TOMOS_FP32_BF16 || o.a_type == TOMOS_FP32_16,
            "lsqr_batched: a_type must be F32, FP32_16 or FP32_BF16");
    require(o.mv_tier == TOMOS_TIER_DEFAULT, "lsqr_batched: iterations run at the DEFAULT tier");
    require(o.value_type == TOMOS_F32 && o.refine_type == TOMOS_F64 && o.precond_type == TOMOS_F64,
            "lsqr_batched: value_type F32, refine_type F64 and precond_type F64 (R^-1) are required");
    require(o.max_iters >= 0 && o.restarts >= 0, "lsqr_batched: negative iteration limits");
    require(o.confirm == 0 || o.confirm == 1, "lsqr_batched: confirm must be 0 or 1");
    const int vec = o.a_type == TOMOS_F32 ? 4 : 8;
    const bool packed = o.a_type != TOMOS_F32;
    for (int i = 0; i < nprob; ++i) {
        const tomos_lsq_problem& p = pr[i];
        require(p.m >= 0 && p.n >= 0 && p.n <= 2048, "lsqr_batched: sizes (n <= 2048)");
        if (p.n == 0) continue;
        require(p.a && p.b && p.x && p.rinv, "lsqr_batched: null pointer in a problem");
        require(!packed || p.wr, "lsqr_batched: packed A needs the wr plane (exact restart residual)");
        require(p.lda >= p.m && p.lda % vec == 0, "lsqr_batched: lda must be >= m and a multiple of the 16-byte row vector");
        require(reinterpret_cast<uintptr_t>(p.a) % 16 == 0 && (!packed || reinterpret_cast<uintptr_t>(p.wr) % 16 == 0),
                "lsqr_batched: A planes must be 16-byte aligned");
        require(p.ldr >= p.n, "lsqr_batched: ldr < n");
    }
}

struct Batched {
    tomos_lsqr_options o{};
    int nprob;
    const tomos_lsq_problem* pr;
    Plan L;
    char* sc;
    hipStream_t s;
    BProb* dprobs() const { return reinterpret_cast<BProb*>(sc + L.probs); }
    GroupProb* giter() const { return reinterpret_cast<GroupProb*>(sc + L.gprobs_iter); }
    GroupProb* ginit() const { return reinterpret_cast<GroupProb*>(sc + L.gprobs_init); }
    GroupItem* ditems() const { return reinterpret_cast<GroupItem*>(sc + L.items_off); }
    RowItem* drows() const { return reinterpret_cast<RowItem*>(sc + L.row_items_off); }
    BState* states() const { return reinterpret_cast<BState*>(sc + L.states); }

    void upload() {
        std::vector<BProb> bp(nprob);
        std::vector<GroupProb> gi(nprob), gn(nprob);
        for (int i = 0; i < nprob; ++i) {
            const tomos_lsq_problem& p = pr[i];
            BProb& q = bp[i];
            q.m = p.m;
            q.n = p.n;
            q.lda = p.lda;
            q.ldr = p.ldr;
            q.a = p.a;
            q.wr = p.wr;
            q.b = p.b;
            q.x = p.x;
            q.rinv = p.rinv;
            char* vb = sc + L.vec_at[i];
            q.u = reinterpret_cast<float*>(vb);
            const size_t ub = a256(static_cast<size_t>(std::max<int64_t>(p.m, 1)) * 4);
            const size_t nb = a256(static_cast<size_t>(std::max<int64_t>(p.n, 1)) * 8);
            q.v = reinterpret_cast<double*>(vb + ub);
            q.z = reinterpret_cast<double*>(vb + ub + nb);
            q.pv = reinterpret_cast<double*>(vb + ub + 2 * nb);
            q.temp = reinterpret_cast<double*>(vb + ub + 3 * nb);
            q.src = reinterpret_cast<double*>(vb + ub + 4 * nb);
            q.st = states() + i;
            q.hist = reinterpret_cast<double*>(sc + L.hist_at[i]);
            q.hist_cap = L.hist_cap;
            q.nitems = L.prob_items[i];
            q.sumsq = reinterpret_cast<double*>(sc + L.sumsq_at[i]);
            q.tpart = reinterpret_cast<float*>(sc + L.tpart_at[i]);
            GroupProb g;
            g.m = p.m;
            g.n = p.n;
            g.lda = p.lda;
            g.a = p.a;
            g.p = q.pv;
            g.u = q.u;
            g.u_coef = &q.st->u_coef;
            g.skip = &q.st->done;
            g.sumsq = q.sumsq;
            g.tpart = q.tpart;
            g.rounds = fused_rounds(L.g, L.threads, p.n);
            gi[i] = g;
            g.p = reinterpret_cast<const double*>(sc + L.zeros);
            g.u_coef = &q.st->one;
            gn[i] = g;
        }
        TOMOS_HIP_CHECK(hipMemcpyAsync(dprobs(), bp.data(), sizeof(BProb) * nprob, hipMemcpyHostToDevice, s));
        TOMOS_HIP_CHECK(hipMemcpyAsync(giter(), gi.data(), sizeof(GroupProb) * nprob, hipMemcpyHostToDevice, s));
        TOMOS_HIP_CHECK(hipMemcpyAsync(ginit(), gn.data(), sizeof(GroupProb) * nprob, hipMemcpyHostToDevice, s));
        if (L.nitems)
            TOMOS_HIP_CHECK(hipMemcpyAsync(ditems(), L.items.data(), sizeof(GroupItem) * L.nitems,
                                           hipMemcpyHostToDevice, s));
        if (L.nrow_items)
            TOMOS_HIP_CHECK(hipMemcpyAsync(drows(), L.row_items.data(), sizeof(RowItem) * L.nrow_items,
                                           hipMemcpyHostToDevice, s));
        TOMOS_HIP_CHECK(hipMemsetAsync(sc + L.zeros, 0, static_cast<size_t>(std::max<int64_t>(L.nmax, 1)) * 8, s));
        // the host vectors must outlive the copies
        TOMOS_HIP_CHECK(hipStreamSynchronize(s));
    }

    int64_t tail_nmax() const { return L.nmax; }

    void tail(bool init) {
        const unsigned np = static_cast<unsigned>(nprob);
        k_batch_tail_a<<<np, kTail, 0, s>>>(dprobs());
        k_batch_tail_b<<<dim3(static_cast<unsigned>(ceil_div(L.nmax, static_cast<int64_t>(kTail / 64))), np), kTail, 0,
                         s>>>(dprobs());
        k_batch_tail_c<<<np, kTail, 0, s>>>(dprobs(), init ? 1 : 0, o.tol, o.max_relres, o.confirm);
        k_batch_tail_d<<<dim3(static_cast<unsigned>(ceil_div(L.nmax, static_cast<int64_t>(kTailRows))), np), kTail, 0,
                         s>>>(dprobs());
    }
    void residual(bool xzero) {
        if (!L.nrow_items) return;
        const unsigned nb = static_cast<unsigned>(L.nrow_items);
        if (xzero) {
            if (o.a_type == TOMOS_F32) k_batch_residual<0, true><<<nb, kT, 0, s>>>(dprobs(), drows());
            else if (o.a_type == TOMOS_FP32_BF16) k_batch_residual<1, true><<<nb, kT, 0, s>>>(dprobs(), drows());
            else k_batch_residual<2, true><<<nb, kT, 0, s>>>(dprobs(), drows());
            return;
        }
        if (o.a_type == TOMOS_F32) k_batch_residual<0, false><<<nb, kT, 0, s>>>(dprobs(), drows());
        else if (o.a_type == TOMOS_FP32_BF16) k_batch_residual<1, false><<<nb, kT, 0, s>>>(dprobs(), drows());
        else k_batch_residual<2, false><<<nb, kT, 0, s>>>(dprobs(), drows());
    }
    void fused(bool init) {
        grouped_fused_launch(o.a_type, L.g, L.threads, L.rounds, init ? ginit() : giter(), ditems(), L.nitems, s,
                             L.stream);
    }
    // One restart: check, exact residual, A^T u, tail(init), then max_iters
    // steps.  The first restart of a solve starts from x = 0.
    void restart_sequence(bool first) {
        k_batch_restart_check<<<static_cast<unsigned>(ceil_div(nprob, 64)), 64, 0, s>>>(dprobs(), nprob, o.tol,
                                                                                      o.max_relres, o.confirm);
        residual(first);
        fused(true);
        tail(true);
        for (int it = 0; it < o.max_iters; ++it) {
            fused(false);
            tail(false);
        }
    }

    // every problem stopped (done == 1; 2 waits for its next restart)
    bool all_done(std::vector<BState>& h) {
        TOMOS_HIP_CHECK(hipMemcpyAsync(h.data(), states(), sizeof(BState) * nprob, hipMemcpyDeviceToHost, s));
        TOMOS_HIP_CHECK(hipStreamSynchronize(s));
        for (const BState& b : h)
            if (b.done != 1) return false;
        return true;
    }

    // Everything a captured solve depends on: the launch grids, the tail's
    // LDS, the scalar arguments and the scratch address (the kernels read
    // their problem tables and work lists from the scratch, rewritten by
    // upload() before every launch).
    struct GraphKey {
        int device, nprob, nitems, nrow_items, g, threads, rounds, a_type, max_iters, restarts, confirm;
        size_t lds;
        double tol, maxrr;
        const void* scratch;
        bool stream;
        bool operator==(const GraphKey& k) const {
            return device == k.device && nprob == k.nprob && nitems == k.nitems && nrow_items == k.nrow_items &&
                   g == k.g && threads == k.threads && rounds == k.rounds && a_type == k.a_type &&
                   max_iters == k.max_iters && restarts == k.restarts && confirm == k.confirm && lds == k.lds &&
                   tol == k.tol && maxrr == k.maxrr && scratch == k.scratch && stream == k.stream;
        }
    };
    GraphKey key() const {
        GraphKey k;
        TOMOS_HIP_CHECK(hipGetDevice(&k.device));
        k.nprob = nprob;
        k.nitems = L.nitems;
        k.nrow_items = L.nrow_items;
        k.g = L.g;
        k.threads = L.threads;
        k.rounds = L.rounds;
        k.a_type = o.a_type;
        k.max_iters = o.max_iters;
        k.restarts = o.restarts;
        k.confirm = o.confirm;
        k.lds = static_cast<size_t>(tail_nmax());
        k.tol = o.tol;
        k.maxrr = o.max_relres;
        k.scratch = sc;
        k.stream = L.stream;
        return k;
    }

    // The whole solve: state set-up, then `restarts` restart sequences.
    // Problems that have stopped turn their blocks of the later launches
    // into immediate exits, so the sequence can run to its end blindly.
    void solve_sequence() {
        k_bstate_init<<<static_cast<unsigned>(ceil_div(nprob, 64)), 64, 0, s>>>(dprobs(), nprob);
        k_batch_setup<<<static_cast<unsigned>(nprob), kTail, 0, s>>>(dprobs());
        for (int r = 0; r < o.restarts; ++r) restart_sequence(r == 0);
    }

    void run(bool use_graph, int* iterations, double* final_relres) {
        upload();
        std::vector<BState> h(nprob);
        if (use_graph && o.restarts > 0) {
            // one graph launch for the whole solve, instantiated once per key
            // and kept (library-wide cache of the last few keys)
            hipGraphExec_t exec = graph_cache_find(key());
            if (!exec) {
                hipGraph_t graph = nullptr;
                TOMOS_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
                try {
                    solve_sequence();
                } catch (...) {
                    hipGraph_t g2 = nullptr;
                    (void)hipStreamEndCapture(s, &g2);
                    if (g2) (void)hipGraphDestroy(g2);
                    throw;
                }
                TOMOS_HIP_CHECK(hipStreamEndCapture(s, &graph));
                const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                TOMOS_HIP_CHECK(e);
                graph_cache_insert(key(), exec);
            }
            TOMOS_HIP_CHECK(hipGraphLaunch(exec, s));
            TOMOS_LAUNCH_CHECK("lsqr_batched graph");
            all_done(h);
        } else {
            k_bstate_init<<<static_cast<unsigned>(ceil_div(nprob, 64)), 64, 0, s>>>(dprobs(), nprob);
            k_batch_setup<<<static_cast<unsigned>(nprob), kTail, 0, s>>>(dprobs());
            TOMOS_LAUNCH_CHECK("lsqr_batched setup");
            for (int r = 0; r < o.restarts; ++r) {
                if (all_done(h)) break;
                restart_sequence(r == 0);
                TOMOS_LAUNCH_CHECK("lsqr_batched restart");
            }
            all_done(h);
        }
        for (int i = 0; i < nprob; ++i) {
            if (iterations) iterations[i] = h[i].completed;
            if (final_relres) final_relres[i] = h[i].rhsnorm > 0 ? std::fabs(h[i].phi_bar) / h[i].rhsnorm : 0.0;
        }
    }

    // Library-wide cache of instantiated solve graphs (a few keys, oldest
    // evicted).  Every solve synchronizes before it returns, so an evicted
    // graph is not running.
    static constexpr int kGraphCache = 8;
    struct CacheEntry {
        GraphKey key;
        hipGraphExec_t exec;
    };
    static std::vector<CacheEntry>& cache() {
        static std::vector<CacheEntry> c;
        return c;
    }
    static std::mutex& cache_mutex() {
        static std::mutex m;
        return m;
    }
    static hipGraphExec_t graph_cache_find(const GraphKey& k) {
        std::lock_guard<std::mutex> lock(cache_mutex());
        for (const CacheEntry& e : cache())
            if (e.key == k) return e.exec;
        return nullptr;
    }
    static void graph_cache_insert(const GraphKey& k, hipGraphExec_t exec) {
        std::lock_guard<std::mutex> lock(cache_mutex());
        auto& c = cache();
        if (static_cast<int>(c.size()) >= kGraphCache) {
            (void)hipGraphExecDestroy(c.front().exec);
            c.erase(c.begin());
        }
        c.push_back(CacheEntry{k, exec});
    }
};

}  // namespace

}  // namespace tomos_hip

using namespace tomos_hip;

extern "C" {

size_t tomos_lsqr_batched_scratch_size(const tomos_lsqr_options* o, int nprob,
                                       const tomos_lsq_problem* probs, int row_groups, int threads,
                                       int tiles_per_block) {
    size_t out = 0;
    const int rc = guarded([&] {
        require(o != nullptr, "tomos_lsqr_batched_scratch_size: null options");
        validate(*o, nprob, probs);
        out = make_plan(*o, nprob, probs, row_groups, threads, tiles_per_block).total;
    });
    return rc == 0 ? out : 0;
}

int tomos_lsqr_batched_solve(const tomos_lsqr_options* o, int nprob, const tomos_lsq_problem* probs,
                             int row_groups, int threads, int tiles_per_block, int use_graph, void* scratch,
                             size_t scratch_bytes, int* iterations, double* final_relres,
                             tomos_stream_t s) {
    return guarded([&] {
        require(o != nullptr, "tomos_lsqr_batched_solve: null options");
        validate(*o, nprob, probs);
        if (nprob == 0) return;
        Batched B;
        B.o = *o;
        B.nprob = nprob;
        B.pr = probs;
        B.L = make_plan(*o, nprob, probs, row_groups, threads, tiles_per_block);
        if (!scratch) scratch = fallback_scratch(B.L.total);
        else if (scratch_bytes < B.L.total)
            throw Error(kErrWorkspace, "lsqr_batched: scratch too small (need " + std::to_string(B.L.total) +
                                           " bytes, have " + std::to_string(scratch_bytes) + ")");
        B.sc = static_cast<char*>(scratch);
        // a graph cannot be captured from the legacy NULL stream: run on a
        // private stream (default priority), ordered after earlier NULL-stream work
        hipStream_t own = nullptr;
        if (use_graph && s == nullptr) TOMOS_HIP_CHECK(hipStreamCreate(&own));
        B.s = own ? own : s;
        try {
            B.run(use_graph != 0, iterations, final_relres);
        } catch (...) {
            if (own) {
                (void)hipStreamSynchronize(own);
                (void)hipStreamDestroy(own);
            }
            throw;
        }
        if (own) {
            TOMOS_HIP_CHECK(hipStreamSynchronize(own));
            TOMOS_HIP_CHECK(hipStreamDestroy(own));
        }
    });
}

}  // extern "C"
