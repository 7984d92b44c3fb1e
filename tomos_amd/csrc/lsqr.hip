// Restarted, right-preconditioned LSQR (Golub-Kahan bidiagonalization) on
// one GPU.
//
// Control flow follows fused_lsqr_impl (src/core/solver/lsqr.cpp:685-747):
//   rhsnorm = |b|; x = 0
//   for each restart:
//     stop check on |phi_bar|/rhsnorm (after the first iteration)
//     initialize (:431-513): u = b - A x (restart tier), beta = |u|, u /= beta,
//       v = P^T A^T u (or A^T u), alpha = |v|, v /= alpha, w = v,
//       rho_bar = alpha, phi_bar = beta; stop if beta/rhsnorm is small
//     iterations (:519-679): u = A P v - alpha u, beta = |u|, u /= beta,
//       v = P^T A^T u - beta v, alpha = |v|, v /= alpha; Givens update;
//       x += (phi/rho) P w; w = v - (theta/rho) w; stop check
// with P = R^-1 (precond_apply = a dense GEMV on R^-1, blas.cpp:88-96).
//
// MI355X design: every scalar of the recurrence (alpha, beta, rho_bar,
// phi_bar, rho, c, s, theta, phi) and every norm lives in FP64 in a device
// state block and is produced by device kernels (deterministic fixed-order
// reductions), so an iteration is a fixed sequence of kernel launches with
// no device->host traffic.  The stop decision is a device flag that every
// kernel of the sequence checks first (a converged solve turns the
// remaining launches into no-ops); the host polls the flag every
// `check_every` iterations.  Relative-residual history is recorded on the
// device, one entry per stop check, and copied back at the end.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "csr.hpp"
#include "gemv.hpp"
#include "lsqr_fused.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

constexpr int kT = 256;
constexpr int kRedBlocks = 1024;

struct LsqrState {
    double alpha, beta, rho_bar, phi_bar;
    double rhsnorm;
    double x_coef;    // phi / rho
    double w_coef;    // -theta / rho
    double inv_beta;  // 1 / beta: u = u_raw * inv_beta
    double u_coef;    // -alpha / beta: next u_raw = A p + u_coef * u_raw
    int done;
    int completed;
    int hist_len;
    int z_fresh;      // 1 right after initialize: z = P w restarts from z = P v
};

__device__ __forceinline__ double ld(const void* p, bool f64, int64_t i) {
    return f64 ? static_cast<const double*>(p)[i] : static_cast<double>(static_cast<const float*>(p)[i]);
}
__device__ __forceinline__ void st(void* p, bool f64, int64_t i, double v) {
    if (f64) static_cast<double*>(p)[i] = v;
    else static_cast<float*>(p)[i] = static_cast<float>(v);
}

__global__ void k_state_init(LsqrState* s) {
    if (threadIdx.x == 0) {
        s->alpha = s->beta = s->rho_bar = s->phi_bar = 0.0;
        s->rhsnorm = 0.0;
        s->x_coef = s->w_coef = 0.0;
        s->inv_beta = s->u_coef = 0.0;
        s->z_fresh = 1;
        s->done = 0;
        s->completed = 0;
        s->hist_len = 0;
    }
}

// A restart that follows an unconfirmed estimate stop (done == 2) runs
// again: the stop flag is cleared before its true-residual test.
__global__ void k_resume(LsqrState* s) {
    if (threadIdx.x == 0 && s->done == 2) s->done = 0;
}

__global__ __launch_bounds__(kT) void k_zero(void* x, bool f64, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (i < n) st(x, f64, i, 0.0);
}

__global__ __launch_bounds__(kT) void k_copy(const LsqrState* __restrict__ s, const void* src,
                                             bool s64, void* dst, bool d64, int64_t n) {
    if (s->done) return;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kT)
        st(dst, d64, i, ld(src, s64, i));
}

// sum of squares, per-block partials
__global__ __launch_bounds__(kT) void k_sumsq(const LsqrState* __restrict__ s, const void* x,
                                              bool f64, int64_t n, double* __restrict__ red) {
    if (s && s->done) return;
    double acc = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kT) {
        const double v = ld(x, f64, i);
        acc += v * v;
    }
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
    __shared__ double w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) red[blockIdx.x] = (w[0] + w[1]) + (w[2] + w[3]);
}

// out = sqrt(sum of red[0..nb)) in a fixed order
__global__ __launch_bounds__(kT) void k_norm_finish(const LsqrState* s, const double* __restrict__ red,
                                                    int nb, double* __restrict__ out) {
    if (s && s->done) return;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += kT) acc += red[i];
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
    __shared__ double w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) *out = sqrt((w[0] + w[1]) + (w[2] + w[3]));
}

__device__ __forceinline__ void push_history(LsqrState* s, double* hist, int cap, double rr) {
    if (s->hist_len < cap) hist[s->hist_len] = rr;
    s->hist_len += 1;
}

// check_stopping_criteria at the top of a restart (lsqr.cpp:417-425, :708).
// confirm: the estimate alone does not stop the solve; the restart's true
// residual decides.
__global__ void k_restart_check(LsqrState* s, double* hist, int cap, double tol, double maxrr, int confirm) {
    if (threadIdx.x != 0 || s->done) return;
    if (s->completed == 0) return;
    if (s->rhsnorm == 0.0) { s->done = 1; return; }
    const double rr = fabs(s->phi_bar) / s->rhsnorm;
    push_history(s, hist, cap, rr);
    if (!confirm && (rr < tol || rr < maxrr)) s->done = 1;
}

// beta = sqrt(sum red[0..nb)) in a fixed order; inv_beta = 1/beta
__global__ __launch_bounds__(kT) void k_beta_finish(LsqrState* s, const double* __restrict__ red, int nb) {
    if (s->done) return;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += kT) acc += red[i];
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
    __shared__ double w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b = sqrt((w[0] + w[1]) + (w[2] + w[3]));
        s->beta = b;
        s->inv_beta = 1.0 / b;
    }
}

// beta = |u_raw| (as k_beta_finish) and temp[j] = sum_s tsum[s][j] / beta
// (A^T u from the single-pass kernel's level-1 sums, fixed order) in one
// launch: every block forms beta from the same partials in the same order,
// block 0 stores it, and each block scales its columns by its own copy
__global__ __launch_bounds__(kT) void k_beta_temp_finish(LsqrState* s, const double* __restrict__ red, int nb,
                                                         int64_t n, int splits, const double* __restrict__ tsum,
                                                         void* temp, bool t64) {
    if (s->done) return;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += kT) acc += red[i];
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
    __shared__ double w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = acc;
    __syncthreads();
    const double b = sqrt((w[0] + w[1]) + (w[2] + w[3]));
    const double inv_b = 1.0 / b;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        s->beta = b;
        s->inv_beta = inv_b;
    }
    const int64_t j = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (j >= n) return;
    double a = 0.0;
    for (int sp = 0; sp < splits; ++sp) a += tsum[static_cast<int64_t>(sp) * n + j];
    st(temp, t64, j, a * inv_b);
}

constexpr int kTail = 1024;

__device__ __forceinline__ double tail_block_sum(double v, double* sh) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < kTail / 64; ++i) t += sh[i];  // same order in every thread
    return t;
}

// The n-sized part of an LSQR step in one block (lsqr.cpp:557-590 after
// the A^T u matvec, and :632-679):
//   src = P^T (A^T u)  (pre: sum of the R^-T chunk partials; else temp)
//   q = src - beta*v (iteration) or src (initialize)
//   alpha = |q|, v = q/alpha
//   iteration: Givens update; x += (phi/rho) * (pre ? z : w) with
//     z = P w carried by z <- P v_new + (-theta/rho) z (no extra GEMV);
//     w <- v_new + (-theta/rho) w
//   initialize: w = v (z restarts), rho_bar = alpha, phi_bar = beta,
//     true-residual stop test
//   next u_raw coefficient: -alpha/beta; stop check + history.
__global__ __launch_bounds__(kTail) void k_tail(LsqrState* S, int init, int pre,
                                                const double* __restrict__ rpart, int rchunks,
                                                const void* temp, void* v, void* w, void* z,
                                                const void* p, bool v64, bool p64, void* x,
                                                bool x64, int64_t n, double* hist, int cap,
                                                double tol, double maxrr, int confirm) {
    if (S->done) return;
    __shared__ double sh[kTail / 64];
    __shared__ double bc[4];
    const double beta = S->beta;
    auto src = [&](int64_t j) -> double {
        if (pre) {
            double a = 0.0;
            for (int c = 0; c < rchunks; ++c) a += rpart[static_cast<int64_t>(c) * n + j];
            return a;
        }
        return ld(temp, v64, j);
    };
    double ss = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += kTail) {
        const double q = init ? src(j) : src(j) - beta * ld(v, v64, j);
        ss += q * q;
    }
    const double alpha = sqrt(tail_block_sum(ss, sh));
    if (threadIdx.x == 0) {
        if (init) {
            bc[0] = bc[1] = bc[2] = 0.0;
        } else {
            const double rho = hypot(S->rho_bar, beta);
            const double c = S->rho_bar / rho, sn = beta / rho;
            const double theta = sn * alpha;
            const double phi = c * S->phi_bar;
            S->rho_bar = -c * alpha;
            S->phi_bar = sn * S->phi_bar;
            bc[0] = phi / rho;       // x coefficient
            bc[1] = -theta / rho;    // new w coefficient
            bc[2] = S->w_coef;       // previous w coefficient (z recurrence)
        }
    }
    __syncthreads();
    const double xc = bc[0], wc = bc[1], wc_prev = bc[2];
    const bool fresh = S->z_fresh != 0;
    const double inv_alpha = 1.0 / alpha;
    for (int64_t j = threadIdx.x; j < n; j += kTail) {
        const double q = init ? src(j) : src(j) - beta * ld(v, v64, j);
        const double vn = q * inv_alpha;
        if (init) {
            if (!pre) st(w, v64, j, vn);
        } else if (pre) {
            const double pj = ld(p, p64, j);
            const double zj = fresh ? pj : pj + wc_prev * ld(z, p64, j);
            st(z, p64, j, zj);
            st(x, x64, j, ld(x, x64, j) + xc * zj);
        } else {
            const double wj = ld(w, v64, j);
            st(x, x64, j, ld(x, x64, j) + xc * wj);
            st(w, v64, j, vn + wc * wj);
        }
        st(v, v64, j, vn);
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    S->alpha = alpha;
    S->u_coef = -alpha / beta;
    if (init) {
        S->rho_bar = alpha;
        S->phi_bar = beta;
        S->z_fresh = 1;
        S->w_coef = 0.0;
        if (S->rhsnorm == 0.0) { S->done = 1; return; }
        const double rr = beta / S->rhsnorm;
        if (rr < tol || rr < maxrr) S->done = 1;
        return;
    }
    S->z_fresh = 0;
    S->w_coef = wc;
    S->x_coef = xc;
    S->completed += 1;
    if (S->rhsnorm == 0.0) { S->done = 1; return; }
    const double rr = fabs(S->phi_bar) / S->rhsnorm;
    push_history(S, hist, cap, rr);
    // confirm: end this restart early (2); the next restart's exact
    // residual decides whether the solve stops
    if (rr < tol || rr < maxrr) S->done = confirm ? 2 : 1;
}

// Two buffers, as the reference's Workspace: the arena holds the Krylov
// vectors carved in the reference's order (u, v, w, temp1 in value_type,
// then precond_v in precond_type, each offset rounded up to its element
// size; lsqr_registries.cpp:23-77, sized as create_lsqr_workspace,
// lsqr.cpp:389-406), the scratch holds this implementation's device state,
// history, reduction partials and GEMV partial sums.
struct Layout {
    size_t u, v, w, t, pv, arena_total;
    size_t state, hist, red, z, tpart, tsum, scratch, scratch_total;
    bool fused_possible;
    FusedPlan fplan;
    size_t gemv_bytes;
    int hist_cap;
    // CSR A: A^T as CSR (values, row pointers, column indices) and the
    // transpose's working space
    size_t csr_vals = 0, csr_ptr = 0, csr_idx = 0, csr_tmp = 0, csr_tmp_bytes = 0, csr_part = 0,
           csr_part_bytes = 0;
    int64_t csr_panel_rows = 0, csr_panels = 1;
};

size_t align256(size_t x) { return (x + 255) / 256 * 256; }
size_t tsize(int t) { return t == TOMOS_F64 ? 8 : 4; }
size_t carve(size_t& off, size_t elem, size_t count) {
    off = (off + elem - 1) / elem * elem;
    const size_t at = off;
    off += count * elem;
    return at;
}

size_t arena_formula(const tomos_lsqr_options& o, int64_t m, int64_t n) {
    const size_t vs = tsize(o.value_type);
    const size_t ps = o.precond_type >= 0 ? tsize(o.precond_type) : vs;
    return static_cast<size_t>(m) * vs + 3 * static_cast<size_t>(n) * vs + (ps - 1) +
           static_cast<size_t>(n) * ps;
}

Layout layout_for(const tomos_lsqr_options& o, int64_t m, int64_t n) {
    Layout L;
    const size_t vs = tsize(o.value_type);
    const size_t ps = o.precond_type >= 0 ? tsize(o.precond_type) : vs;
    size_t off = 0;
    L.u = carve(off, vs, static_cast<size_t>(m));
    L.v = carve(off, vs, static_cast<size_t>(n));
    L.w = carve(off, vs, static_cast<size_t>(n));
    L.t = carve(off, vs, static_cast<size_t>(n));
    L.pv = carve(off, ps, static_cast<size_t>(n));
    L.arena_total = off;
    off = 0;
    L.state = off; off = align256(off + sizeof(LsqrState));
    const int64_t cap = static_cast<int64_t>(std::max(o.restarts, 0)) * (std::max(o.max_iters, 0) + 1) + 1;
    L.hist_cap = static_cast<int>(std::min<int64_t>(cap, 1 << 26));
    L.hist = off; off = align256(off + static_cast<size_t>(L.hist_cap) * 8);
    L.fused_possible = o.fused != 0 && o.mv_tier == TOMOS_TIER_DEFAULT && n >= 1 && n <= 2048 &&
                       (o.a_type == TOMOS_F32 || o.a_type == TOMOS_FP32_16 || o.a_type == TOMOS_FP32_BF16 ||
                        o.a_type == TOMOS_BF16);
    if (L.fused_possible) L.fplan = lsqr_fused_plan(o.a_type, m, n);
    // sized for every fused plan of (m, n), so tuning after sizing is safe
    const int64_t red_n = std::max<int64_t>(std::max<int64_t>(gemv_sumsq_blocks_max(m), kRedBlocks),
                                            L.fused_possible ? lsqr_fused_max_blocks(m) : 0);
    L.red = off; off = align256(off + static_cast<size_t>(red_n) * 8);
    L.z = off; off = align256(off + static_cast<size_t>(n) * 8);  // z = P w (precond type)
    L.tpart = L.tsum = off;
    if (L.fused_possible) {
        L.tpart = off; off = align256(off + lsqr_fused_max_tpart_bytes(m, n));
        L.tsum = off; off = align256(off + lsqr_fused_max_tsum_bytes(n));
    }
    size_t sc = 0;
    sc = std::max(sc, gemv_scratch_bytes(o.a_type, o.mv_tier, TOMOS_NO_TRANSPOSE, m, n));
    sc = std::max(sc, gemv_scratch_bytes(o.a_type, o.mv_tier, TOMOS_TRANSPOSE, m, n));
    sc = std::max(sc, gemv_scratch_bytes(o.a_type, o.restart_tier, TOMOS_NO_TRANSPOSE, m, n));
    // an explicit F32 / F64 duplicate of A may serve the restarts
    sc = std::max(sc, gemv_scratch_bytes(TOMOS_F32, TOMOS_TIER_EXACT, TOMOS_NO_TRANSPOSE, m, n));
    sc = std::max(sc, gemv_scratch_bytes(TOMOS_F64, TOMOS_TIER_EXACT, TOMOS_NO_TRANSPOSE, m, n));
    if (o.precond_type >= 0) {
        const int pt = o.precond_type == TOMOS_F64 ? TOMOS_TIER_DEFAULT : TOMOS_TIER_EXACT;
        sc = std::max(sc, gemv_scratch_bytes(o.precond_type, pt, TOMOS_NO_TRANSPOSE, n, n));
        sc = std::max(sc, gemv_scratch_bytes(o.precond_type, pt, TOMOS_TRANSPOSE, n, n));
    }
    L.scratch = off;
    L.gemv_bytes = sc;
    off = align256(off + sc);
    L.scratch_total = off;
    return L;
}

// Dense layout plus A^T for a CSR matrix (the single-pass kernels are
// dense-only).
Layout csr_layout_for(const tomos_lsqr_options& o, const CsrMat& A) {
    tomos_lsqr_options od = o;
    od.fused = 0;
    Layout L = layout_for(od, A.m, A.n);
    const size_t e = static_cast<size_t>(std::max<int64_t>(A.nnz, 1));
    size_t off = L.scratch_total;
    // the row-group kernel may use more blocks (sum-of-squares partials)
    // than the dense GEMV
    L.red = off; off = align256(off + static_cast<size_t>(std::max<int64_t>(csr_sumsq_blocks_max(A.m), kRedBlocks)) * 8);
    L.csr_panel_rows = csr_panel_rows(A, o.value_type);
    L.csr_panels = csr_panels(A, L.csr_panel_rows);
    L.csr_vals = off; off = align256(off + e * csr_value_bytes(A.val_type));
    L.csr_ptr = off;
    off = align256(off + static_cast<size_t>(L.csr_panels * A.n + 1) * csr_index_bytes(A.idx_type));
    L.csr_idx = off; off = align256(off + e * csr_index_bytes(A.idx_type));
    L.csr_tmp_bytes = csr_transpose_scratch_bytes(A, L.csr_panel_rows);
    L.csr_tmp = off; off = align256(off + L.csr_tmp_bytes);
    CsrMat T = A;
    T.m = L.csr_panels * A.n;
    T.n = A.m;
    L.csr_part_bytes =
        std::max<size_t>(std::max(csr_mv_scratch_bytes(A), csr_tmv_scratch_bytes(T, L.csr_panels, A.n)), 256);
    L.csr_part = off; off = align256(off + L.csr_part_bytes);
    L.scratch_total = off;
    return L;
}

void validate(const tomos_lsqr_options& o, int64_t m, int64_t n) {
    require(o.a_type == TOMOS_F32 || o.a_type == TOMOS_F64 || o.a_type == TOMOS_FP32_16 ||
                o.a_type == TOMOS_FP32_BF16 || o.a_type == TOMOS_F16 || o.a_type == TOMOS_BF16,
            "lsqr: unsupported a_type");
    require(o.value_type == TOMOS_F32 || o.value_type == TOMOS_F64, "lsqr: value_type must be F32/F64");
    require(o.refine_type == TOMOS_F32 || o.refine_type == TOMOS_F64, "lsqr: refine_type must be F32/F64");
    require(o.precond_type == -1 || o.precond_type == TOMOS_F32 || o.precond_type == TOMOS_F64,
            "lsqr: precond_type must be F32, F64 or -1");
    require(o.max_iters >= 0 && o.restarts >= 0, "lsqr: negative iteration limits");
    require(o.confirm == 0 || o.confirm == 1, "lsqr: confirm must be 0 or 1");
    require(m >= 0 && n >= 0, "lsqr: negative size");
}

unsigned grid_for(int64_t n) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kT), kRedBlocks)));
}

struct Solver {
    tomos_lsqr_options o;
    int64_t m, n;
    const void* a;
    const uint16_t* wr;
    int64_t lda;
    // explicit refine-precision duplicate of A (F32/F64) for the restart
    // residuals, or none (lsqr.cpp:463-469 with mtx_refine_precision)
    const void* aref = nullptr;
    int ref_type = -1;
    int64_t lda_ref = 0;
    const void* b;
    void* x;
    const void* rinv;
    int64_t ldr;
    char* arena;
    char* sc;
    Layout L;
    hipStream_t s;
    bool v64, r64, p64;

    LsqrState* state() const { return reinterpret_cast<LsqrState*>(sc + L.state); }
    void* U() const { return arena + L.u; }
    void* V() const { return arena + L.v; }
    void* W() const { return arena + L.w; }
    void* Tm() const { return arena + L.t; }
    void* PV() const { return arena + L.pv; }
    double* hist() const { return reinterpret_cast<double*>(sc + L.hist); }
    double* red() const { return reinterpret_cast<double*>(sc + L.red); }
    void* Z() const { return sc + L.z; }
    // CSR A (instead of the dense a/lda): its transpose, and the refine
    // duplicate's values over the same pattern
    const CsrMat* csr = nullptr;
    CsrMat csr_t, csr_ref;
    bool use_fused = false;
    bool use_fused_restart = false;  // initialize() in one pass (exact residual + A_hi^T u)
    bool x_zero = false;             // x is still the zero vector of run()'s start
    FusedPlan rplan;
    void* scratch() const { return sc + L.scratch; }

    void norm(const void* vec, bool f64, int64_t len, double* out, bool guarded_by_state = true) {
        const unsigned g = grid_for(len);
        const LsqrState* sp = guarded_by_state ? state() : nullptr;
        k_sumsq<<<g, kT, 0, s>>>(sp, vec, f64, len, red());
        k_norm_finish<<<1, kT, 0, s>>>(sp, red(), static_cast<int>(g), out);
    }
    Coef coef(double alpha, double beta) const {
        Coef c;
        c.alpha = alpha;
        c.beta = beta;
        c.skip = &state()->done;
        return c;
    }
    // A-matrix GEMV at a tier
    void gemv_a(int tier, int trans, const Coef& c, const void* xin, bool x64, void* yout, bool y64,
                int* nb = nullptr) {
        if (csr) {  // FP64 accumulation at every tier
            if (trans == TOMOS_TRANSPOSE)
                csr_tmv(csr_t, L.csr_panels, L.csr_panel_rows, n, c, x64 ? TOMOS_F64 : TOMOS_F32, xin,
                        y64 ? TOMOS_F64 : TOMOS_F32, yout, sc + L.csr_part, L.csr_part_bytes, s, nb);
            else
                csr_mv(*csr, c, x64 ? TOMOS_F64 : TOMOS_F32, xin, y64 ? TOMOS_F64 : TOMOS_F32, yout,
                       sc + L.csr_part, L.csr_part_bytes, s, nb);
            return;
        }
        gemv(o.a_type, tier, trans, c, m, n, a, wr, lda, x64 ? TOMOS_F64 : TOMOS_F32, xin,
             y64 ? TOMOS_F64 : TOMOS_F32, yout, scratch(), L.gemv_bytes, s, nb);
    }
    // P GEMVs accumulate in FP64 whatever the storage type of R^-1
    int ptier() const { return o.precond_type == TOMOS_F64 ? TOMOS_TIER_DEFAULT : TOMOS_TIER_EXACT; }
    // P = R^-1 GEMV
    void gemv_p(int trans, const void* xin, bool x64, void* yout, bool y64) {
        gemv(o.precond_type, ptier(), trans, coef(1.0, 0.0), n, n, rinv, nullptr, ldr,
             x64 ? TOMOS_F64 : TOMOS_F32, xin, y64 ? TOMOS_F64 : TOMOS_F32, yout, scratch(),
             L.gemv_bytes, s);
    }

    void initialize() {
        LsqrState* S = state();
        // u_raw = b - A x at the restart tier; beta = |u_raw| (fused partials)
        k_copy<<<grid_for(m), kT, 0, s>>>(S, b, r64, U(), v64, m);
        if (aref) {
            // the duplicate at its own precision, FP64 accumulation
            Coef cr = coef(-1.0, 1.0);
            cr.sumsq = red();
            int nb = 0;
            if (csr)
                csr_mv(csr_ref, cr, r64 ? TOMOS_F64 : TOMOS_F32, x, v64 ? TOMOS_F64 : TOMOS_F32, U(),
                       sc + L.csr_part, L.csr_part_bytes, s, &nb);
            else
                gemv(ref_type, TOMOS_TIER_EXACT, TOMOS_NO_TRANSPOSE, cr, m, n, aref, nullptr, lda_ref,
                     r64 ? TOMOS_F64 : TOMOS_F32, x, v64 ? TOMOS_F64 : TOMOS_F32, U(), scratch(), L.gemv_bytes, s,
                     &nb);
            k_beta_finish<<<1, kT, 0, s>>>(S, red(), nb);
            tail(true);
            return;
        }
        if (use_fused_restart) {
            float* tpart = reinterpret_cast<float*>(sc + L.tpart);
            double* tsum = reinterpret_cast<double*>(sc + L.tsum);
            lsqr_restart_launch(o.a_type, m, n, a, wr, lda, static_cast<const double*>(x), x_zero, U(), v64,
                                red(), tpart, tsum, rplan, &S->done, s);
            k_beta_temp_finish<<<grid_for(n), kT, 0, s>>>(S, red(), rplan.nblocks, n, rplan.splits, tsum, Tm(), v64);
            tail(true, true);
            return;
        }
        Coef cr = coef(-1.0, 1.0);
        cr.sumsq = red();
        int nb = 0;
        gemv_a(o.restart_tier, TOMOS_NO_TRANSPOSE, cr, x, r64, U(), v64, &nb);
        k_beta_finish<<<1, kT, 0, s>>>(S, red(), nb);
        tail(true);
    }

    // temp = A^T u = (A^T u_raw) / beta, then (pre) the R^-T chunk partials,
    // then the one-block tail.
    void tail(bool init, bool temp_ready = false) {
        LsqrState* S = state();
        const bool pre = o.precond_type >= 0;
        if (!temp_ready) {
            Coef ct = coef(0.0, 0.0);
            ct.dalpha = &S->inv_beta;
            gemv_a(o.mv_tier, TOMOS_TRANSPOSE, ct, U(), v64, Tm(), v64);
        }
        int rch = 0;
        if (pre)
            rch = gemv_t_partials(o.precond_type, ptier(), n, n, rinv, nullptr, ldr,
                                  v64 ? TOMOS_F64 : TOMOS_F32, Tm(), scratch(), L.gemv_bytes,
                                  &S->done, s);
        k_tail<<<1, kTail, 0, s>>>(S, init ? 1 : 0, pre ? 1 : 0, static_cast<const double*>(scratch()),
                                   rch, Tm(), V(), W(), Z(), PV(), v64, p64, x, r64, n, hist(),
                                   L.hist_cap, o.tol, o.max_relres, o.confirm);
    }

    void iterate() {
        LsqrState* S = state();
        const bool pre = o.precond_type >= 0;
        // u_raw = A (P v) + (-alpha/beta) u_raw, |u_raw|^2 fused
        Coef cn = coef(1.0, 0.0);
        cn.dbeta = &S->u_coef;
        cn.sumsq = red();
        int nb = 0;
        if (pre) gemv_p(TOMOS_NO_TRANSPOSE, V(), v64, PV(), p64);
        const void* xin = pre ? PV() : V();
        const bool x64 = pre ? p64 : v64;
        if (use_fused) {
            // one pass over A: u_raw and the A^T u_raw partials together
            lsqr_fused_launch(o.a_type, m, n, a, lda, xin, x64, U(), v64, &S->u_coef, red(),
                              reinterpret_cast<float*>(sc + L.tpart),
                              reinterpret_cast<double*>(sc + L.tsum), L.fplan, &S->done, s);
            k_beta_temp_finish<<<grid_for(n), kT, 0, s>>>(S, red(), L.fplan.nblocks, n, L.fplan.splits,
                                                         reinterpret_cast<const double*>(sc + L.tsum), Tm(), v64);
            tail(false, true);
            return;
        }
        gemv_a(o.mv_tier, TOMOS_NO_TRANSPOSE, cn, xin, x64, U(), v64, &nb);
        k_beta_finish<<<1, kT, 0, s>>>(S, red(), nb);
        tail(false);
    }

    // 0 running, 1 stopped, 2 estimate below tol awaiting the next
    // restart's true residual (confirm mode)
    int done_flag() {
        int d = 0;
        TOMOS_HIP_CHECK(hipMemcpyAsync(&d, &state()->done, sizeof(int), hipMemcpyDeviceToHost, s));
        TOMOS_HIP_CHECK(hipStreamSynchronize(s));
        return d;
    }

    void run(int* iterations, double* history, int history_cap, int* history_len) {
        LsqrState* S = state();
        k_state_init<<<1, 64, 0, s>>>(S);
        k_zero<<<static_cast<unsigned>(ceil_div(std::max<int64_t>(n, 1), kT)), kT, 0, s>>>(x, r64, n);
        x_zero = true;
        norm(b, r64, m, &S->rhsnorm, false);
        TOMOS_LAUNCH_CHECK("lsqr setup");
        const int every = std::max(1, o.check_every);
        bool stop = false;
        for (int r = 0; r < o.restarts && !stop; ++r) {
            if (o.confirm) k_resume<<<1, 64, 0, s>>>(S);
            k_restart_check<<<1, 64, 0, s>>>(S, hist(), L.hist_cap, o.tol, o.max_relres, o.confirm);
            initialize();
            x_zero = false;
            TOMOS_LAUNCH_CHECK("lsqr initialize");
            if (done_flag() == 1) break;
            for (int it = 0; it < o.max_iters; ++it) {
                iterate();
                if ((it + 1) % every == 0 || it + 1 == o.max_iters) {
                    TOMOS_LAUNCH_CHECK("lsqr iteration");
                    const int d = done_flag();
                    if (d == 1) stop = true;
                    if (d != 0) break;
                }
            }
        }
        LsqrState hs;
        TOMOS_HIP_CHECK(hipMemcpyAsync(&hs, S, sizeof(LsqrState), hipMemcpyDeviceToHost, s));
        TOMOS_HIP_CHECK(hipStreamSynchronize(s));
        if (iterations) *iterations = hs.completed;
        if (history_len) *history_len = hs.hist_len;
        if (history && history_cap > 0) {
            const int cnt = std::min(std::min(hs.hist_len, L.hist_cap), history_cap);
            if (cnt > 0) {
                TOMOS_HIP_CHECK(hipMemcpyAsync(history, hist(), static_cast<size_t>(cnt) * 8,
                                               hipMemcpyDeviceToHost, s));
                TOMOS_HIP_CHECK(hipStreamSynchronize(s));
            }
        }
    }
};

}  // namespace

size_t lsqr_arena_bytes(const tomos_lsqr_options& o, int64_t m, int64_t n) {
    validate(o, m, n);
    return std::max(arena_formula(o, m, n), layout_for(o, m, n).arena_total);
}

size_t lsqr_scratch_bytes(const tomos_lsqr_options& o, int64_t m, int64_t n) {
    validate(o, m, n);
    return layout_for(o, m, n).scratch_total;
}

void lsqr_solve(const tomos_lsqr_options& o, int64_t m, int64_t n, const void* a, const uint16_t* wr,
                int64_t lda, const void* b, void* x, const void* rinv, int64_t ldr, void* arena,
                size_t arena_bytes, void* scratch, size_t scratch_bytes, int* iterations,
                double* history, int history_cap, int* history_len, hipStream_t s, int ref_type = -1,
                const void* aref = nullptr, int64_t lda_ref = 0) {
    validate(o, m, n);
    require(lda >= std::max<int64_t>(m, 1), "lsqr: lda < m");
    if (aref) {
        require(ref_type == TOMOS_F32 || ref_type == TOMOS_F64, "lsqr: the refine duplicate must be F32 or F64");
        require(lda_ref >= std::max<int64_t>(m, 1), "lsqr: lda of the refine duplicate < m");
    }
    if (o.precond_type >= 0) {
        require(rinv != nullptr, "lsqr: precond_type set but no R^-1 given");
        require(ldr >= std::max<int64_t>(n, 1), "lsqr: ldr < n");
    }
    if (n == 0) {
        if (iterations) *iterations = 0;
        if (history_len) *history_len = 0;
        return;
    }
    require(a && b && x, "lsqr: null A, b or x");
    const bool packed = o.a_type == TOMOS_FP32_16 || o.a_type == TOMOS_FP32_BF16;
    if (packed && (o.mv_tier == TOMOS_TIER_EXACT || o.restart_tier == TOMOS_TIER_EXACT))
        require(wr != nullptr, "lsqr: EXACT tier needs the wr plane");
    Solver sv;
    sv.o = o;
    sv.m = m;
    sv.n = n;
    sv.a = a;
    sv.wr = wr;
    sv.lda = lda;
    sv.b = b;
    sv.x = x;
    sv.rinv = rinv;
    sv.ldr = ldr;
    sv.aref = aref;
    sv.ref_type = ref_type;
    sv.lda_ref = lda_ref;
    sv.L = layout_for(o, m, n);
    // lsqr_registries.cpp:59-63: the arena must hold the carved vectors
    if (arena == nullptr || arena_bytes < sv.L.arena_total)
        throw Error(kErrWorkspace, "lsqr: workspace is too small for LSQR's working vectors (need " +
                                       std::to_string(sv.L.arena_total) + " bytes, have " +
                                       std::to_string(arena ? arena_bytes : 0) + ")");
    if (scratch == nullptr) scratch = fallback_scratch(sv.L.scratch_total);
    else if (scratch_bytes < sv.L.scratch_total)
        throw Error(kErrWorkspace, "lsqr: scratch too small (need " + std::to_string(sv.L.scratch_total) +
                                       " bytes, have " + std::to_string(scratch_bytes) + ")");
    sv.arena = static_cast<char*>(arena);
    sv.sc = static_cast<char*>(scratch);
    sv.s = s;
    sv.v64 = o.value_type == TOMOS_F64;
    sv.r64 = o.refine_type == TOMOS_F64;
    sv.p64 = o.precond_type >= 0 ? o.precond_type == TOMOS_F64 : sv.v64;
    sv.use_fused = sv.L.fused_possible && lsqr_fused_supported(o.a_type, o.mv_tier, m, n, lda, a);
    sv.use_fused_restart = sv.use_fused && sv.r64 && aref == nullptr &&
                           lsqr_restart_supported(o.a_type, o.restart_tier, m, n, lda, a, wr);
    if (sv.use_fused_restart) sv.rplan = lsqr_restart_plan(o.a_type, m, n);
    sv.run(iterations, history, history_cap, history_len);
}

// LSQR over a CSR matrix (lsqr_sparse_impl, lsqr_registries.cpp:164-196):
// the same iteration with A p from the CSR rows and A^T u from a CSR
// transpose built once per solve.  ref_vals (ref_type F32/F64, or null):
// the refine duplicate's values over A's pattern, read by the restarts.
void lsqr_solve_csr(const tomos_lsqr_options& o, const CsrMat& A, int ref_type, const void* ref_vals,
                    const void* b, void* x, const void* rinv, int64_t ldr, void* arena, size_t arena_bytes,
                    void* scratch, size_t scratch_bytes, int* iterations, double* history, int history_cap,
                    int* history_len, hipStream_t s) {
    validate(o, A.m, A.n);
    csr_require_valid_types(A, "lsqr (CSR)");
    require(o.a_type == A.val_type, "lsqr (CSR): options.a_type must be the CSR value type");
    if (ref_vals) require(ref_type == TOMOS_F32 || ref_type == TOMOS_F64, "lsqr: the refine duplicate must be F32 or F64");
    if (o.precond_type >= 0) {
        require(rinv != nullptr, "lsqr: precond_type set but no R^-1 given");
        require(ldr >= std::max<int64_t>(A.n, 1), "lsqr: ldr < n");
    }
    if (A.n == 0) {
        if (iterations) *iterations = 0;
        if (history_len) *history_len = 0;
        return;
    }
    require(b && x, "lsqr: null b or x");
    Solver sv;
    sv.o = o;
    sv.o.fused = 0;
    sv.m = A.m;
    sv.n = A.n;
    sv.b = b;
    sv.x = x;
    sv.rinv = rinv;
    sv.ldr = ldr;
    sv.L = csr_layout_for(o, A);
    if (arena == nullptr || arena_bytes < sv.L.arena_total)
        throw Error(kErrWorkspace, "lsqr: workspace is too small for LSQR's working vectors (need " +
                                       std::to_string(sv.L.arena_total) + " bytes, have " +
                                       std::to_string(arena ? arena_bytes : 0) + ")");
    if (scratch == nullptr) scratch = fallback_scratch(sv.L.scratch_total);
    else if (scratch_bytes < sv.L.scratch_total)
        throw Error(kErrWorkspace, "lsqr: scratch too small (need " + std::to_string(sv.L.scratch_total) +
                                       " bytes, have " + std::to_string(scratch_bytes) + ")");
    sv.arena = static_cast<char*>(arena);
    sv.sc = static_cast<char*>(scratch);
    sv.s = s;
    sv.v64 = o.value_type == TOMOS_F64;
    sv.r64 = o.refine_type == TOMOS_F64;
    sv.p64 = o.precond_type >= 0 ? o.precond_type == TOMOS_F64 : sv.v64;
    // indices validated once (synchronous), then A^T built in the scratch
    csr_check(A, sv.sc + sv.L.csr_tmp, sv.L.csr_tmp_bytes, s);
    csr_transpose(A, sv.L.csr_panel_rows, sv.sc + sv.L.csr_vals, sv.sc + sv.L.csr_ptr, sv.sc + sv.L.csr_idx,
                  sv.sc + sv.L.csr_tmp, sv.L.csr_tmp_bytes, s);
    sv.csr = &A;
    sv.csr_t = A;
    sv.csr_t.m = sv.L.csr_panels * A.n;
    sv.csr_t.n = A.m;
    sv.csr_t.vals = sv.sc + sv.L.csr_vals;
    sv.csr_t.row_ptrs = sv.sc + sv.L.csr_ptr;
    sv.csr_t.col_idxs = sv.sc + sv.L.csr_idx;
    if (ref_vals) {
        sv.csr_ref = A;
        sv.csr_ref.val_type = ref_type;
        sv.csr_ref.vals = ref_vals;
        sv.aref = ref_vals;
        sv.ref_type = ref_type;
    }
    sv.run(iterations, history, history_cap, history_len);
}

}  // namespace tomos_hip

using namespace tomos_hip;

extern "C" {

size_t tomos_lsqr_csr_scratch_size(const tomos_lsqr_options* o, int idx_type, int64_t m, int64_t n,
                                   int64_t nnz) {
    size_t out = 0;
    int rc = guarded([&] {
        require(o != nullptr, "tomos_lsqr_csr_scratch_size: null options");
        validate(*o, m, n);
        CsrMat A;
        A.val_type = o->a_type;
        A.idx_type = idx_type;
        A.m = m;
        A.n = n;
        A.nnz = nnz;
        require(nnz >= 0 && (idx_type == TOMOS_INDEX_INT32 || idx_type == TOMOS_INDEX_INT64),
                "tomos_lsqr_csr_scratch_size: bad nnz or index type");
        out = csr_layout_for(*o, A).scratch_total;
    });
    return rc == 0 ? out : 0;
}

int tomos_lsqr_solve_csr(const tomos_lsqr_options* o, int idx_type, int64_t m, int64_t n, int64_t nnz,
                         const void* vals, const void* row_ptrs, const void* col_idxs, int ref_type,
                         const void* ref_vals, const void* b, void* x, const void* rinv, int64_t ldr, void* arena,
                         size_t arena_bytes, void* scratch, size_t scratch_bytes, int* iterations, double* history,
                         int history_cap, int* history_len, tomos_stream_t s) {
    return guarded([&] {
        require(o != nullptr, "tomos_lsqr_solve_csr: null options");
        CsrMat A;
        A.val_type = o->a_type;
        A.idx_type = idx_type;
        A.m = m;
        A.n = n;
        A.nnz = nnz;
        A.vals = vals;
        A.row_ptrs = row_ptrs;
        A.col_idxs = col_idxs;
        lsqr_solve_csr(*o, A, ref_type, ref_vals, b, x, rinv, ldr, arena, arena_bytes, scratch, scratch_bytes,
                       iterations, history, history_cap, history_len, s);
    });
}

size_t tomos_lsqr_arena_size(const tomos_lsqr_options* o, int64_t m, int64_t n) {
    size_t out = 0;
    int rc = guarded([&] {
        require(o != nullptr, "tomos_lsqr_arena_size: null options");
        out = lsqr_arena_bytes(*o, m, n);
    });
    return rc == 0 ? out : 0;
}

size_t tomos_lsqr_scratch_size(const tomos_lsqr_options* o, int64_t m, int64_t n) {
    size_t out = 0;
    int rc = guarded([&] {
        require(o != nullptr, "tomos_lsqr_scratch_size: null options");
        out = lsqr_scratch_bytes(*o, m, n);
    });
    return rc == 0 ? out : 0;
}

int tomos_lsqr_solve(const tomos_lsqr_options* o, int64_t m, int64_t n, const void* a,
                     const uint16_t* wr, int64_t lda, const void* b, void* x, const void* rinv,
                     int64_t ldr, void* arena, size_t arena_bytes, void* scratch,
                     size_t scratch_bytes, int* iterations, double* history, int history_cap,
                     int* history_len, tomos_stream_t s) {
    return guarded([&] {
        require(o != nullptr, "tomos_lsqr_solve: null options");
        lsqr_solve(*o, m, n, a, wr, lda, b, x, rinv, ldr, arena, arena_bytes, scratch,
                   scratch_bytes, iterations, history, history_cap, history_len, s);
    });
}

int tomos_lsqr_solve_refined(const tomos_lsqr_options* o, int64_t m, int64_t n, const void* a,
                             const uint16_t* wr, int64_t lda, int ref_type, const void* a_ref, int64_t lda_ref,
                             const void* b, void* x, const void* rinv, int64_t ldr, void* arena,
                             size_t arena_bytes, void* scratch, size_t scratch_bytes, int* iterations,
                             double* history, int history_cap, int* history_len, tomos_stream_t s) {
    return guarded([&] {
        require(o != nullptr, "tomos_lsqr_solve_refined: null options");
        require(a_ref != nullptr, "tomos_lsqr_solve_refined: null refine duplicate");
        lsqr_solve(*o, m, n, a, wr, lda, b, x, rinv, ldr, arena, arena_bytes, scratch, scratch_bytes,
                   iterations, history, history_cap, history_len, s, ref_type, a_ref, lda_ref);
    });
}

}  // extern "C"
