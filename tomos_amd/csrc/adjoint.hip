// Implicit differentiation of x = argmin |A x - b| (full column rank A).
//
// With r = b - A x and an upstream gradient g = dL/dx:
//   w      = (A^T A)^-1 g
//   grad_b = A w
//   grad_A = r w^T - (A w) x^T
// (differentiate the normal equations A^T A x = A^T b).
//
// w is computed by restarted, R-preconditioned conjugate gradients on the
// normal equations, with the same kernels and the same precision split as
// the restarted LSQR solver:
//   restart: res = g - A^T (A w) at the restart tier (exact wh+wr
//            reconstruction, FP64 accumulation), stop if |res|/|g| < tol;
//            z = R^-1 R^-T res, d = z, rho = res.z
//   inner:   q = A d and t = A^T q from ONE read of A (the single-pass
//            kernel of the LSQR iteration, per-iteration tier), |q|^2
//            alpha = rho/|q|^2, w += alpha d, res -= alpha t,
//            z = R^-1 R^-T res, rho' = res.z, d = z + (rho'/rho) d;
//            the restart ends when |res|/|g| < tol.
// R^-1 is the sketch-QR preconditioner of the forward solve, so
// (A R^-1)^T (A R^-1) is well conditioned and CG converges in a few tens
// of iterations.  Scalars (rho, |q|^2, |res|, |g|) are FP64 in a device
// state block; every kernel checks the stop flags first, and the host
// polls them every check_every iterations (as the LSQR solver does).
//
// grad_A is written by one fused rank-2 kernel that reads r, A w (m each)
// and w, x (n each) and streams the m x n FP32 result: no copy of A is made.
// No reference counterpart: the reference has no adjoint.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "adjoint.hpp"
#include "common.hpp"
#include "gemv.hpp"
#include "lsqr_fused.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

constexpr int kT = 256;
constexpr int kBlk = 1024;  // one-block n-sized kernels
constexpr int kMaxAdjWorkers = 16;  // host threads (one HIP stream each) of the batched adjoint

struct AdjState {
    double rho;       // res . z
    double gnorm;     // |g|
    double zero;      // 0.0: u coefficient of the single-pass kernel (q = A d)
    double resnorm;   // |res| after the last update
    double res0;      // |res| at the start of the current restart (exact residual)
    int done;         // final: converged or restarts exhausted
    int inner_done;   // ends the inner iterations of the current restart
    int stop;         // done || inner_done (checked by every inner kernel)
    int completed;    // inner iterations run
    int hist_len;
};

__global__ void k_adj_init(AdjState* S) {
    if (threadIdx.x == 0) {
        S->rho = S->gnorm = S->zero = S->resnorm = S->res0 = 0.0;
        S->done = S->inner_done = S->stop = 0;
        S->completed = S->hist_len = 0;
    }
}

__device__ __forceinline__ double blk_sum(double v, double* sh) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < kBlk / 64; ++i) t += sh[i];  // same order in every thread
    return t;
}

__device__ __forceinline__ void push_hist(AdjState* S, double* hist, int cap, double v) {
    if (S->hist_len < cap) hist[S->hist_len] = v;
    S->hist_len += 1;
}

// res = g - rfull (rfull null: w = 0, res = g); |g| on the first call;
// exact-residual stop test.  Resets inner_done for the next restart.
__global__ __launch_bounds__(kBlk) void k_adj_residual(AdjState* S, const double* __restrict__ g,
                                                       const double* __restrict__ rfull,
                                                       double* __restrict__ res, int64_t n,
                                                       double* hist, int cap, double tol, int last) {
    if (S->done) return;
    __shared__ double sh[kBlk / 64];
    double gg = 0.0, rr = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += kBlk) {
        const double rj = g[j] - (rfull ? rfull[j] : 0.0);
        res[j] = rj;
        gg += g[j] * g[j];
        rr += rj * rj;
    }
    gg = blk_sum(gg, sh);
    rr = blk_sum(rr, sh);
    if (threadIdx.x != 0) return;
    if (!rfull) S->gnorm = sqrt(gg);
    S->resnorm = sqrt(rr);
    S->res0 = S->resnorm;
    S->inner_done = 0;
    S->stop = 0;
    if (S->gnorm == 0.0) { S->done = 1; S->stop = 1; return; }
    const double rel = S->resnorm / S->gnorm;
    push_hist(S, hist, cap, rel);
    if (rel < tol || last) { S->done = 1; S->stop = 1; }
}

// rho = res.z, d = z (start of a restart)
__global__ __launch_bounds__(kBlk) void k_adj_start(AdjState* S, const double* __restrict__ res,
                                                    const double* __restrict__ z,
                                                    double* __restrict__ d, int64_t n) {
    if (S->stop) return;
    __shared__ double sh[kBlk / 64];
    double a = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += kBlk) {
        a += res[j] * z[j];
        d[j] = z[j];
    }
    a = blk_sum(a, sh);
    if (threadIdx.x == 0) S->rho = a;
}

// |q|^2 from block partials, t = sum of the split partials (splits = 1:
// t given directly); alpha = rho/|q|^2; w += alpha d; res -= alpha t;
// relative residual estimate, inner stop test.
__global__ __launch_bounds__(kBlk) void k_adj_step(AdjState* S, const double* __restrict__ qsq,
                                                   int nq, const double* __restrict__ tsum,
                                                   int splits, const double* __restrict__ d,
                                                   double* __restrict__ w, double* __restrict__ res,
                                                   int64_t n, double* hist, int cap, double tol,
                                                   double drop) {
    if (S->stop) return;
    __shared__ double sh[kBlk / 64];
    double q2 = 0.0;
    for (int i = threadIdx.x; i < nq; i += kBlk) q2 += qsq[i];
    q2 = blk_sum(q2, sh);
    const double alpha = q2 > 0.0 ? S->rho / q2 : 0.0;
    double rr = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += kBlk) {
        double t = 0.0;
        for (int sp = 0; sp < splits; ++sp) t += tsum[static_cast<int64_t>(sp) * n + j];
        w[j] += alpha * d[j];
        const double rj = res[j] - alpha * t;
        res[j] = rj;
        rr += rj * rj;
    }
    rr = blk_sum(rr, sh);
    if (threadIdx.x != 0) return;
    S->completed += 1;
    S->resnorm = sqrt(rr);
    const double rel = S->resnorm / S->gnorm;
    push_hist(S, hist, cap, rel);
    if (rel < tol || q2 <= 0.0 || S->resnorm < drop * S->res0) { S->inner_done = 1; S->stop = 1; }
}

// rho' = res.z, d = z + (rho'/rho) d
__global__ __launch_bounds__(kBlk) void k_adj_dir(AdjState* S, const double* __restrict__ res,
                                                  const double* __restrict__ z,
                                                  double* __restrict__ d, int64_t n) {
    if (S->stop) return;
    __shared__ double sh[kBlk / 64];
    const double rho_old = S->rho;  // read before the barriers: thread 0 rewrites it
    double a = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += kBlk) a += res[j] * z[j];
    a = blk_sum(a, sh);
    const double beta = rho_old != 0.0 ? a / rho_old : 0.0;
    for (int64_t j = threadIdx.x; j < n; j += kBlk) d[j] = z[j] + beta * d[j];
    if (threadIdx.x == 0) S->rho = a;
}

__global__ __launch_bounds__(kT) void k_adj_zero(double* __restrict__ v, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (i < n) v[i] = 0.0;
}

// out[i + j*ldg] = r[i] w[j] - gb[i] x[j]: 4 rows per lane (one 16-byte
// store per column), a block covers 1024 rows x kCols columns; w[j], x[j]
// are block-uniform.
constexpr int kCols = 128;
__global__ __launch_bounds__(kT) void k_rank2_vec(int64_t m, int64_t n, const double* __restrict__ r,
                                                  const double* __restrict__ gb,
                                                  const double* __restrict__ w,
                                                  const double* __restrict__ x,
                                                  float* __restrict__ out, int64_t ldg) {
    const int64_t i0 = (static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x) * 4;
    if (i0 >= m) return;
    const int64_t j0 = static_cast<int64_t>(blockIdx.y) * kCols;
    const int64_t j1 = min(n, j0 + kCols);
    const int nv = static_cast<int>(min<int64_t>(4, m - i0));
    double rv[4], gv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        rv[k] = k < nv ? r[i0 + k] : 0.0;
        gv[k] = k < nv ? gb[i0 + k] : 0.0;
    }
    for (int64_t j = j0; j < j1; ++j) {
        const double wj = w[j], xj = x[j];
        float4 o;
        o.x = static_cast<float>(rv[0] * wj - gv[0] * xj);
        o.y = static_cast<float>(rv[1] * wj - gv[1] * xj);
        o.z = static_cast<float>(rv[2] * wj - gv[2] * xj);
        o.w = static_cast<float>(rv[3] * wj - gv[3] * xj);
        float* col = out + j * ldg + i0;
        if (nv == 4) {
            *reinterpret_cast<float4*>(col) = o;
        } else {
            const float ov[4] = {o.x, o.y, o.z, o.w};
            for (int k = 0; k < nv; ++k) col[k] = ov[k];
        }
    }
}

__global__ __launch_bounds__(kT) void k_rank2_scalar(int64_t m, int64_t n, const double* __restrict__ r,
                                                     const double* __restrict__ gb,
                                                     const double* __restrict__ w,
                                                     const double* __restrict__ x,
                                                     float* __restrict__ out, int64_t ldg) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (i >= m) return;
    const int64_t j0 = static_cast<int64_t>(blockIdx.y) * kCols;
    const int64_t j1 = min(n, j0 + kCols);
    const double ri = r[i], gi = gb[i];
    for (int64_t j = j0; j < j1; ++j) out[i + j * ldg] = static_cast<float>(ri * w[j] - gi * x[j]);
}

// ---- layout -------------------------------------------------------------
struct AdjLayout {
    size_t state, hist, res, z, y, d, t, rfull, qsq, tpart, tsum, q32, q64, gemv, total;
    size_t gemv_bytes;
    int hist_cap;
    bool fused_possible;
};

size_t a256(size_t v) { return (v + 255) / 256 * 256; }

AdjLayout adj_layout(const tomos_lsqr_options& o, int64_t m, int64_t n) {
    AdjLayout L;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = a256(off + bytes);
        return at;
    };
    const int64_t cap = static_cast<int64_t>(std::max(o.restarts, 0)) * (std::max(o.max_iters, 0) + 1) + 1;
    L.hist_cap = static_cast<int>(std::min<int64_t>(cap, 1 << 26));
    L.fused_possible = o.fused != 0 && o.mv_tier == TOMOS_TIER_DEFAULT && n >= 1 && n <= 2048 &&
                       (o.a_type == TOMOS_F32 || o.a_type == TOMOS_FP32_16 || o.a_type == TOMOS_FP32_BF16);
    const size_t nv = static_cast<size_t>(std::max<int64_t>(n, 1)) * 8;
    L.state = take(sizeof(AdjState));
    L.hist = take(static_cast<size_t>(L.hist_cap) * 8);
    L.res = take(nv);
    L.z = take(nv);
    L.y = take(nv);
    L.d = take(nv);
    L.t = take(nv);
    L.rfull = take(nv);
    const int64_t qblocks = std::max<int64_t>(gemv_sumsq_blocks_max(m),
                                              L.fused_possible ? lsqr_fused_max_blocks(m) : 0);
    L.qsq = take(static_cast<size_t>(qblocks) * 8);
    L.tpart = take(L.fused_possible ? lsqr_fused_max_tpart_bytes(m, n) : 0);
    L.tsum = take(L.fused_possible ? lsqr_fused_max_tsum_bytes(n) : 0);
    L.q32 = take(static_cast<size_t>(std::max<int64_t>(m, 1)) * 4);
    L.q64 = take(static_cast<size_t>(std::max<int64_t>(m, 1)) * 8);
    size_t sc = 0;
    for (int tr = 0; tr < 2; ++tr) {
        sc = std::max(sc, gemv_scratch_bytes(o.a_type, o.mv_tier, tr, m, n));
        sc = std::max(sc, gemv_scratch_bytes(o.a_type, o.restart_tier, tr, m, n));
        if (o.precond_type >= 0) sc = std::max(sc, gemv_scratch_bytes(o.precond_type, TOMOS_TIER_EXACT, tr, n, n));
    }
    L.gemv_bytes = sc;
    L.gemv = take(sc);
    L.total = off;
    return L;
}

void adj_validate(const tomos_lsqr_options& o, int64_t m, int64_t n) {
    require(o.a_type == TOMOS_F32 || o.a_type == TOMOS_F64 || o.a_type == TOMOS_FP32_16 ||
                o.a_type == TOMOS_FP32_BF16,
            "adjoint: unsupported a_type");
    require(o.precond_type == -1 || o.precond_type == TOMOS_F32 || o.precond_type == TOMOS_F64,
            "adjoint: precond_type must be F32, F64 or -1");
    require(o.max_iters >= 0 && o.restarts >= 1, "adjoint: need restarts >= 1");
    require(m >= 0 && n >= 0 && m >= n, "adjoint: need m >= n (full column rank least squares)");
}

struct Adj {
    tomos_lsqr_options o{};
    int64_t m, n, lda, ldr;
    const void* a;
    const uint16_t* wr;
    const double* g;
    const void* rinv;
    double* w;
    char* sc;
    AdjLayout L;
    FusedPlan plan;
    bool use_fused;
    hipStream_t s;

    AdjState* S() const { return reinterpret_cast<AdjState*>(sc + L.state); }
    double* buf(size_t off) const { return reinterpret_cast<double*>(sc + off); }
    void* gs() const { return sc + L.gemv; }
    int ptier() const { return o.precond_type == TOMOS_F64 ? TOMOS_TIER_DEFAULT : TOMOS_TIER_EXACT; }

    Coef cf(double alpha, double beta, const int* skip) const {
        Coef c;
        c.alpha = alpha;
        c.beta = beta;
        c.skip = skip;
        return c;
    }
    // z = R^-1 R^-T res (or z = res without a preconditioner: z aliases res)
    const double* precond(const int* skip) {
        if (o.precond_type < 0) return buf(L.res);
        gemv(o.precond_type, ptier(), TOMOS_TRANSPOSE, cf(1.0, 0.0, skip), n, n, rinv, nullptr, ldr,
             TOMOS_F64, buf(L.res), TOMOS_F64, buf(L.y), gs(), L.gemv_bytes, s);
        gemv(o.precond_type, ptier(), TOMOS_NO_TRANSPOSE, cf(1.0, 0.0, skip), n, n, rinv, nullptr, ldr,
             TOMOS_F64, buf(L.y), TOMOS_F64, buf(L.z), gs(), L.gemv_bytes, s);
        return buf(L.z);
    }

    void restart(bool first, bool last) {
        AdjState* st = S();
        const double* rfull = nullptr;
        if (!first) {
            // exact residual: res = g - A^T (A w) at the restart tier, FP64
            gemv(o.a_type, o.restart_tier, TOMOS_NO_TRANSPOSE, cf(1.0, 0.0, &st->done), m, n, a, wr, lda,
                 TOMOS_F64, w, TOMOS_F64, buf(L.q64), gs(), L.gemv_bytes, s);
            gemv(o.a_type, o.restart_tier, TOMOS_TRANSPOSE, cf(1.0, 0.0, &st->done), m, n, a, wr, lda,
                 TOMOS_F64, buf(L.q64), TOMOS_F64, buf(L.rfull), gs(), L.gemv_bytes, s);
            rfull = buf(L.rfull);
        }
        k_adj_residual<<<1, kBlk, 0, s>>>(st, g, rfull, buf(L.res), n, buf(L.hist), L.hist_cap, o.tol,
                                          last ? 1 : 0);
        const double* z = precond(&st->stop);
        k_adj_start<<<1, kBlk, 0, s>>>(st, buf(L.res), z, buf(L.d), n);
        TOMOS_LAUNCH_CHECK("adjoint restart");
    }

    void iterate() {
        AdjState* st = S();
        int nq = 0, splits = 1;
        const double* tsrc = buf(L.t);
        if (use_fused) {
            lsqr_fused_launch(o.a_type, m, n, a, lda, buf(L.d), true, sc + L.q32, false, &st->zero,
                              buf(L.qsq), reinterpret_cast<float*>(sc + L.tpart), buf(L.tsum), plan,
                              &st->stop, s);
            nq = plan.nblocks;
            splits = plan.splits;
            tsrc = buf(L.tsum);
        } else {
            Coef cn = cf(1.0, 0.0, &st->stop);
            cn.sumsq = buf(L.qsq);
            gemv(o.a_type, o.mv_tier, TOMOS_NO_TRANSPOSE, cn, m, n, a, wr, lda, TOMOS_F64, buf(L.d),
                 TOMOS_F32, sc + L.q32, gs(), L.gemv_bytes, s, &nq);
            gemv(o.a_type, o.mv_tier, TOMOS_TRANSPOSE, cf(1.0, 0.0, &st->stop), m, n, a, wr, lda,
                 TOMOS_F32, sc + L.q32, TOMOS_F64, buf(L.t), gs(), L.gemv_bytes, s);
        }
        k_adj_step<<<1, kBlk, 0, s>>>(st, buf(L.qsq), nq, tsrc, splits, buf(L.d), w, buf(L.res), n,
                                      buf(L.hist), L.hist_cap, o.tol, adjoint_inner_drop(o));
        const double* z = precond(&st->stop);
        k_adj_dir<<<1, kBlk, 0, s>>>(st, buf(L.res), z, buf(L.d), n);
    }

    void flags(int* done, int* stop) {
        AdjState h;
        TOMOS_HIP_CHECK(hipMemcpyAsync(&h, S(), sizeof(AdjState), hipMemcpyDeviceToHost, s));
        TOMOS_HIP_CHECK(hipStreamSynchronize(s));
        *done = h.done;
        *stop = h.stop;
    }

    void run(int* iterations, double* history, int cap, int* hist_len) {
        AdjState* st = S();
        k_adj_init<<<1, 64, 0, s>>>(st);
        k_adj_zero<<<static_cast<unsigned>(ceil_div(n, kT)), kT, 0, s>>>(w, n);
        const int every = std::max(1, o.check_every);
        for (int r = 0; r < o.restarts; ++r) {
            restart(r == 0, false);
            int done = 0, stop = 0;
            flags(&done, &stop);
            if (done) break;
            for (int it = 0; it < o.max_iters; ++it) {
                iterate();
                if ((it + 1) % every == 0 || it + 1 == o.max_iters) {
                    TOMOS_LAUNCH_CHECK("adjoint iteration");
                    flags(&done, &stop);
                    if (stop) break;
                }
            }
        }
        // final exact residual check (also records it in the history)
        int done = 0, stop = 0;
        flags(&done, &stop);
        if (!done) {
            restart(false, true);
            flags(&done, &stop);
        }
        AdjState h;
        TOMOS_HIP_CHECK(hipMemcpyAsync(&h, st, sizeof(AdjState), hipMemcpyDeviceToHost, s));
        TOMOS_HIP_CHECK(hipStreamSynchronize(s));
        if (iterations) *iterations = h.completed;
        if (hist_len) *hist_len = h.hist_len;
        if (history && cap > 0) {
            const int cnt = std::min(std::min(h.hist_len, L.hist_cap), cap);
            if (cnt > 0) {
                TOMOS_HIP_CHECK(hipMemcpyAsync(history, buf(L.hist), static_cast<size_t>(cnt) * 8,
                                               hipMemcpyDeviceToHost, s));
                TOMOS_HIP_CHECK(hipStreamSynchronize(s));
            }
        }
    }
};

}  // namespace

size_t grad_scratch_bytes(int a_type, int tier, int64_t m, int64_t n) {
    return a256(static_cast<size_t>(std::max<int64_t>(m, 1)) * 8) +
           gemv_scratch_bytes(a_type, tier, TOMOS_NO_TRANSPOSE, m, n);
}

namespace {

void adjoint_solve(const tomos_lsqr_options& o, int64_t m, int64_t n, const void* a, const uint16_t* wr,
                   int64_t lda, const double* g, const void* rinv, int64_t ldr, double* w, void* scratch,
                   size_t scratch_bytes, int* iterations, double* history, int history_cap, int* history_len,
                   hipStream_t s) {
    adj_validate(o, m, n);
    require(lda >= std::max<int64_t>(m, 1), "adjoint: lda < m");
    if (o.precond_type >= 0) {
        require(rinv != nullptr, "adjoint: precond_type set but no R^-1 given");
        require(ldr >= std::max<int64_t>(n, 1), "adjoint: ldr < n");
    }
    if (n == 0) {
        if (iterations) *iterations = 0;
        if (history_len) *history_len = 0;
        return;
    }
    require(a && g && w, "adjoint: null A, g or w");
    const bool packed = o.a_type == TOMOS_FP32_16 || o.a_type == TOMOS_FP32_BF16;
    if (packed && (o.mv_tier == TOMOS_TIER_EXACT || o.restart_tier == TOMOS_TIER_EXACT))
        require(wr != nullptr, "adjoint: EXACT tier needs the wr plane");
    Adj J;
    J.o = o;
    J.m = m;
    J.n = n;
    J.lda = lda;
    J.ldr = ldr;
    J.a = a;
    J.wr = wr;
    J.g = g;
    J.rinv = rinv;
    J.w = w;
    J.L = adj_layout(o, m, n);
    if (!scratch) scratch = fallback_scratch(J.L.total);
    else if (scratch_bytes < J.L.total)
        throw Error(kErrWorkspace, "adjoint: scratch too small (need " + std::to_string(J.L.total) +
                                       " bytes, have " + std::to_string(scratch_bytes) + ")");
    J.sc = static_cast<char*>(scratch);
    J.s = s;
    J.use_fused = J.L.fused_possible && lsqr_fused_supported(o.a_type, o.mv_tier, m, n, lda, a);
    if (J.use_fused) J.plan = lsqr_fused_plan(o.a_type, m, n);
    J.run(iterations, history, history_cap, history_len);
}

}  // namespace

void lsq_grad(int a_type, int tier, int64_t m, int64_t n, const void* a, const uint16_t* wr, int64_t lda,
              const double* b, const double* x, const double* w, double* grad_b, float* grad_a, int64_t ldg,
              void* scratch, size_t scratch_bytes, hipStream_t s) {
    require(m >= 0 && n >= 0 && lda >= std::max<int64_t>(m, 1), "tomos_lsq_grad: bad sizes");
    if (m == 0) return;
    if (n == 0) {
        // x is empty: grad_b = A w = 0, grad_A is m x 0
        require(grad_b != nullptr, "tomos_lsq_grad: null grad_b");
        k_adj_zero<<<static_cast<unsigned>(ceil_div(m, kT)), kT, 0, s>>>(grad_b, m);
        TOMOS_LAUNCH_CHECK("k_adj_zero");
        return;
    }
    require(a && b && x && w && grad_b, "tomos_lsq_grad: null pointer");
    require(!grad_a || ldg >= m, "tomos_lsq_grad: ldg < m");
    const size_t need = grad_scratch_bytes(a_type, tier, m, n);
    if (!scratch) scratch = fallback_scratch(need);
    else require(scratch_bytes >= need, "tomos_lsq_grad: scratch too small");
    char* base = static_cast<char*>(scratch);
    double* r = reinterpret_cast<double*>(base);
    void* gsc = base + a256(static_cast<size_t>(m) * 8);
    const size_t gbytes = gemv_scratch_bytes(a_type, tier, TOMOS_NO_TRANSPOSE, m, n);
    Coef c1;
    c1.alpha = 1.0;
    c1.beta = 0.0;
    gemv(a_type, tier, TOMOS_NO_TRANSPOSE, c1, m, n, a, wr, lda, TOMOS_F64, w, TOMOS_F64, grad_b, gsc, gbytes, s);
    if (!grad_a) return;
    TOMOS_HIP_CHECK(hipMemcpyAsync(r, b, static_cast<size_t>(m) * 8, hipMemcpyDeviceToDevice, s));
    Coef c2;
    c2.alpha = -1.0;
    c2.beta = 1.0;
    gemv(a_type, tier, TOMOS_NO_TRANSPOSE, c2, m, n, a, wr, lda, TOMOS_F64, x, TOMOS_F64, r, gsc, gbytes, s);
    const dim3 grid_cols(1, static_cast<unsigned>(ceil_div(n, kCols)));
    require(ceil_div(n, kCols) < 65536, "tomos_lsq_grad: n too large");
    if (reinterpret_cast<uintptr_t>(grad_a) % 16 == 0 && ldg % 4 == 0) {
        const dim3 grid(static_cast<unsigned>(ceil_div(ceil_div(m, 4), kT)), grid_cols.y);
        k_rank2_vec<<<grid, kT, 0, s>>>(m, n, r, grad_b, w, x, grad_a, ldg);
    } else {
        const dim3 grid(static_cast<unsigned>(ceil_div(m, kT)), grid_cols.y);
        k_rank2_scalar<<<grid, kT, 0, s>>>(m, n, r, grad_b, w, x, grad_a, ldg);
    }
    TOMOS_LAUNCH_CHECK("k_rank2");
}

namespace {

// One worker's scratch: the adjoint's and the gradient's (exact tier) for
// the largest problem of the batch, one after the other.
struct BatchedAdjLayout {
    size_t adj = 0, grad = 0, per_worker = 0;
};
BatchedAdjLayout batched_adj_layout(const tomos_lsqr_options& o, int nprob, const tomos_lsq_problem* probs) {
    BatchedAdjLayout L;
    for (int p = 0; p < nprob; ++p) {
        L.adj = std::max(L.adj, adj_layout(o, probs[p].m, probs[p].n).total);
        L.grad = std::max(L.grad, grad_scratch_bytes(o.a_type, o.restart_tier, probs[p].m, probs[p].n));
    }
    L.adj = a256(L.adj);
    L.grad = a256(L.grad);
    L.per_worker = L.adj + L.grad;
    return L;
}

}  // namespace
}  // namespace tomos_hip

using namespace tomos_hip;

extern "C" {

size_t tomos_lsq_adjoint_scratch_size(const tomos_lsqr_options* o, int64_t m, int64_t n) {
    size_t out = 0;
    const int rc = guarded([&] {
        require(o != nullptr, "tomos_lsq_adjoint_scratch_size: null options");
        adj_validate(*o, m, n);
        out = adj_layout(*o, m, n).total;
    });
    return rc == 0 ? out : 0;
}

int tomos_lsq_adjoint_solve(const tomos_lsqr_options* o, int64_t m, int64_t n, const void* a,
                            const uint16_t* wr, int64_t lda, const double* g, const void* rinv,
                            int64_t ldr, double* w, void* scratch, size_t scratch_bytes,
                            int* iterations, double* history, int history_cap, int* history_len,
                            tomos_stream_t s) {
    return guarded([&] {
        require(o != nullptr, "tomos_lsq_adjoint_solve: null options");
        adjoint_solve(*o, m, n, a, wr, lda, g, rinv, ldr, w, scratch, scratch_bytes, iterations, history,
                      history_cap, history_len, s);
    });
}

size_t tomos_lsq_grad_scratch_size(int a_type, int tier, int64_t m, int64_t n) {
    if (m < 0 || n < 0) return 0;
    return grad_scratch_bytes(a_type, tier, m, n);
}

int tomos_lsq_grad(int a_type, int tier, int64_t m, int64_t n, const void* a, const uint16_t* wr,
                   int64_t lda, const double* b, const double* x, const double* w, double* grad_b,
                   float* grad_a, int64_t ldg, void* scratch, size_t scratch_bytes,
                   tomos_stream_t s) {
    return guarded([&] {
        lsq_grad(a_type, tier, m, n, a, wr, lda, b, x, w, grad_b, grad_a, ldg, scratch, scratch_bytes, s);
    });
}

size_t tomos_lsq_adjoint_batched_scratch_size(const tomos_lsqr_options* o, int nprob,
                                              const tomos_lsq_problem* probs, int workers) {
    size_t out = 0;
    const int rc = guarded([&] {
        require(o != nullptr && (nprob == 0 || probs != nullptr), "tomos_lsq_adjoint_batched_scratch_size: null");
        require(nprob >= 0 && workers >= 0 && workers <= kMaxAdjWorkers,
                "tomos_lsq_adjoint_batched_scratch_size: bad nprob or workers");
        for (int p = 0; p < nprob; ++p) adj_validate(*o, probs[p].m, probs[p].n);
        if (workers == 0) out = adjoint_grouped_scratch_bytes(*o, nprob, probs);
        else out = batched_adj_layout(*o, nprob, probs).per_worker * static_cast<size_t>(std::min(workers, std::max(nprob, 1)));
    });
    return rc == 0 ? out : 0;
}

int tomos_lsq_adjoint_batched(const tomos_lsqr_options* o, int nprob, const tomos_lsq_problem* probs,
                              const double* const* g, double* const* w, double* const* grad_b,
                              float* const* grad_a, const int64_t* ldg, int workers, void* scratch,
                              size_t scratch_bytes, int* iterations, tomos_stream_t stream) {
    return guarded([&] {
        require(o != nullptr, "tomos_lsq_adjoint_batched: null options");
        require(nprob >= 0 && workers >= 0 && workers <= kMaxAdjWorkers,
                "tomos_lsq_adjoint_batched: bad nprob or workers (0..16)");
        if (nprob == 0) return;
        require(probs && g && w && grad_b, "tomos_lsq_adjoint_batched: null array");
        for (int p = 0; p < nprob; ++p) {
            const tomos_lsq_problem& q = probs[p];
            adj_validate(*o, q.m, q.n);
            require(q.a && q.b && q.x && g[p] && w[p] && grad_b[p], "tomos_lsq_adjoint_batched: null pointer");
            require(q.lda >= std::max<int64_t>(q.m, 1), "tomos_lsq_adjoint_batched: lda < m");
            require(!grad_a || !grad_a[p] || (ldg && ldg[p] >= q.m), "tomos_lsq_adjoint_batched: ldg < m");
        }
        if (workers == 0) {
            // every problem on the device in the same launches
            adjoint_grouped(*o, nprob, probs, g, w, grad_b, grad_a, ldg, scratch, scratch_bytes, iterations, stream);
            return;
        }
        const int nw = std::min(workers, nprob);
        const BatchedAdjLayout L = batched_adj_layout(*o, nprob, probs);
        require(scratch != nullptr && scratch_bytes >= L.per_worker * static_cast<size_t>(nw),
                "tomos_lsq_adjoint_batched: scratch missing or too small");
        // largest problems first, taken dynamically by the workers
        std::vector<int> order(static_cast<size_t>(nprob));
        for (int p = 0; p < nprob; ++p) order[static_cast<size_t>(p)] = p;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return probs[x].m * probs[x].n > probs[y].m * probs[y].n;
        });
        TOMOS_HIP_CHECK(hipStreamSynchronize(stream));  // inputs written on the caller's stream
        std::vector<hipStream_t> st(static_cast<size_t>(nw), nullptr);
        for (auto& x : st) TOMOS_HIP_CHECK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
        std::atomic<int> next{0};
        std::mutex err_mu;
        std::string err;
        int err_code = 0;
        auto work = [&](int wi) {
            char* base = static_cast<char*>(scratch) + L.per_worker * static_cast<size_t>(wi);
            hipStream_t s = st[static_cast<size_t>(wi)];
            try {
                for (int k = next++; k < nprob; k = next++) {
                    const int p = order[static_cast<size_t>(k)];
                    const tomos_lsq_problem& q = probs[p];
                    int its = 0;
                    adjoint_solve(*o, q.m, q.n, q.a, q.wr, q.lda, g[p], q.rinv, q.ldr, w[p], base, L.adj, &its,
                                  nullptr, 0, nullptr, s);
                    lsq_grad(o->a_type, o->restart_tier, q.m, q.n, q.a, q.wr, q.lda, q.b, q.x, w[p], grad_b[p],
                             grad_a ? grad_a[p] : nullptr, ldg ? ldg[p] : q.m, base + L.adj, L.grad, s);
                    if (iterations) iterations[p] = its;
                }
                TOMOS_HIP_CHECK(hipStreamSynchronize(s));
            } catch (const Error& e) {
                std::lock_guard<std::mutex> lk(err_mu);
                if (err.empty()) { err = e.what(); err_code = e.code; }
                next = nprob;
            } catch (const std::exception& e) {
                std::lock_guard<std::mutex> lk(err_mu);
                if (err.empty()) { err = e.what(); err_code = kErrInternal; }
                next = nprob;
            }
        };
        std::vector<std::thread> th;
        for (int wi = 1; wi < nw; ++wi) th.emplace_back(work, wi);
        work(0);
        for (auto& t : th) t.join();
        for (auto& x : st) (void)hipStreamSynchronize(x);
        for (auto& x : st) (void)hipStreamDestroy(x);
        if (!err.empty()) throw Error(err_code, err);
    });
}

}  // extern "C"
