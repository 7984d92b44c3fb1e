// FP64 Householder QR of tall k x n matrices (CountSketch outputs) and the
// inverse of the triangular factor, for the sketch-QR preconditioner
// (src/core/precond/sketch_qr.cpp:38-64: geqrf, zero the reflectors,
// trtri('U','N')).  One problem or many in the same launches: every kernel
// takes a device table of problem descriptors and picks its problem with
// blockIdx.z, so a batch of uneven problems costs the launches of the
// largest one.
//
// QR, blocked by panels of NB = 32 columns (LAPACK dgeqrf's structure).
// Three panel factorizations, all giving R up to the signs of its rows:
//  * CholeskyQR2 panels (the default, tomos_qr_panels_set(0)): two Gram
//    passes over row chunks, the 32 x 32 Cholesky factors, and the
//    panel's Householder vectors reconstructed from Q by an LU without
//    pivoting, then the blocked update A -= Y T^T (Y^T A) with a look-ahead
//    (see qr_cq_gram_kernel).  A sketch where a panel is too
//    ill-conditioned for two Gram passes (or rank deficient) is factored
//    again from a saved copy with the column-by-column panels below.
//  * TSQR panels (k <= 8192): leaves of 256 or 512 rows factored column by
//    column in registers, their stacked R factors likewise, and the trailing
//    columns updated with each level's block reflector (qr_tsqr_*).
//  * wavefront panels (k > 8192): one launch per panel, one block per
//    panel column.  The block owning column j applies the reflectors c < j
//    as each column c becomes final (a per-column flag, published with
//    release/acquire at device scope), then forms column j's reflector
//    (LAPACK dlarfg convention: beta = -sign(x0)|x|, v = [1,
//    x(1:)/(x0-beta)], tau = (beta-x0)/beta; a zero column gives no
//    reflection) and publishes it; V, T (dlarft's recurrence from V^T V)
//    and the trailing update A <- (I - V T^T V^T) A as two tiled FP64
//    products: W = V^T A (split over rows, partials summed in a fixed
//    order, then W2 = T^T W once) and A -= V W2.
// All reductions run in a fixed order (deterministic).
//
// Triangular inverse: one block per (column j, problem) solves R x = e_j by
// column-oriented back substitution with x kept in LDS.
#include <algorithm>
#include <atomic>
#include <vector>

#include "common.hpp"
#include "qr.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

constexpr int kT = 256;
constexpr int kNB = 32;   // panel width
constexpr int kGR = 128;  // rows per Gram chunk

// Per-problem device descriptor.
struct QrDesc {
    int64_t k, n, lda, ldr;
    double* a;       // k x n, overwritten by the factorization
    double* r;       // R output (n x n, ldr)
    double* diag;    // n
    double* tau;     // n
    double* inv_v0;  // n
    double* v;       // k x kNB (ld k); the CholeskyQR2 path keeps a second set at v + k kNB
    double* gp;      // Gram partials: ceil(k/kGR) x kNB^2
    double* t;       // kNB^2; the CholeskyQR2 path keeps a second set at t + kNB^2
    double* wp;      // W partials: nsplit x n x kNB; the inverse's copy of R (n x n)
    double* w2;      // T^T W: n x kNB (column jt at w2 + jt*kNB)
    int* bad;        // trtri: first zero diagonal (INT_MAX if none)
    int* sync;       // panel kernel: [0] ticket counter, [kSyncHead + c] column c final
    double* tl;      // TSQR: leaf T factors, kTsqrMaxLeaves x kNB^2
    double* vc;      // TSQR: combine reflectors, kTsqrComb x kNB (row-major)
    double* tc;      // TSQR: combine T, kNB^2
    double* cq;      // CholeskyQR2 panels: R1^-1, R1, N (kNB^2 each, row-major), then cq_ints
    double* bk;      // copy of A (k x n, ld k) taken before the CholeskyQR2 panels
};
constexpr int kSyncHead = 64;  // ints before the column flags (own cache lines)
constexpr int kTsqrMaxLeaves = 16;           // leaves per panel (tile rows 256 or 512)
constexpr int kTsqrComb = kTsqrMaxLeaves * kNB;  // stacked rows of the combine step
// doubles of one set of TSQR block-reflector factors (leaf T's, combine V
// and T); two sets alternate by panel parity, so a panel's trailing update
// can still read its set while the next panel is factored
constexpr int64_t kTsqrSet = kTsqrMaxLeaves * kNB * kNB + kTsqrComb * kNB + kNB * kNB;
constexpr int64_t kTsqrMaxK = 16 * 512;      // k beyond this: the wavefront panel path

// CholeskyQR2 panel state per problem: three kNB x kNB matrices, then ints
// [0], [1] (unused), [2] the failure flag (nonzero: a panel could not be
// factored this way) and [3] "one Gram pass suffices" for the current
// panel.
constexpr int kCqMat = kNB * kNB;
constexpr int kCqDoubles = 3 * kCqMat;
static_assert(kCqMat == 4 * 256, "the CholeskyQR2 kernels give each of 256 threads four kNB x kNB entries");
__device__ __forceinline__ int* cq_ints(const QrDesc& d) { return reinterpret_cast<int*>(d.cq + kCqDoubles); }
__device__ __forceinline__ bool cq_failed(const QrDesc& d) { return cq_ints(d)[2] != 0; }

// Row splits of W = V^T A: 128-row splits, at most 32, so a panel's W has
// ~1000 blocks in flight at k = 4000.
__host__ __device__ inline int64_t qr_nsplit(int64_t k) {
    const int64_t s = (k + 127) / 128;
    return s < 1 ? 1 : (s > 32 ? 32 : s);
}

// Fixed-order sum of cnt doubles at p[0], p[stride], ...: four running sums
// so that several loads are in flight, combined as (s0 + s1) + (s2 + s3).
__device__ __forceinline__ double strided_sum(const double* __restrict__ p, int cnt, int64_t stride) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = 0;
    for (; i + 4 <= cnt; i += 4) {
        s0 += p[(i + 0) * stride];
        s1 += p[(i + 1) * stride];
        s2 += p[(i + 2) * stride];
        s3 += p[(i + 3) * stride];
    }
    for (; i < cnt; ++i) s0 += p[i * stride];
    return (s0 + s1) + (s2 + s3);
}

// Descriptors travel as kernel arguments, kQrDescChunk a launch (under 4 KB)
constexpr int kQrDescChunk = 24;
struct QrDescChunk {
    QrDesc d[kQrDescChunk];
};
static_assert(sizeof(QrDescChunk) + 16 <= 4096, "descriptor chunk exceeds the kernel-argument limit");

__global__ __launch_bounds__(64) void qr_set_desc_kernel(QrDesc* __restrict__ dst, QrDescChunk c, int count) {
    if (static_cast<int>(threadIdx.x) < count) dst[threadIdx.x] = c.d[threadIdx.x];
}
__global__ void qr_status_reset_kernel(int* status) { *status = 0; }

constexpr int kPT = 1024;  // panel-step block

// Block-wide FP64 sum over kPT threads, same order in every thread; red
// must not be reused before a barrier that follows every thread's read.
__device__ __forceinline__ double panel_sum(double v, double* red) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kPT / 64; ++w) t += red[w];
    return t;
}

// Coherent read of data another block of this launch published (it may run
// on another XCD, whose L2 this one does not see).
__device__ __forceinline__ double load_published(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr int kPanelSpinLimit = 1 << 18;  // ~0.3 s of polling; a legitimate wait is microseconds

// Whole panel in one launch.  The block with logical id b (from the
// problem's ticket counter: a block only ever waits on blocks that started
// before it, so every wait ends) owns panel column j = c0 + b.  For each
// reflector c = c0 .. j-1 it waits until column c is final, reads column c
// and its tau / inv_v0, and applies H_c to column j (one pass for
// v^T A(:,j), one for the update; same arithmetic as LAPACK dlarf).  Then it
// forms column j's own reflector (dlarfg convention: beta = -sign(x0)|x|,
// tau = (beta - x0)/beta, v = [1, x(j+1:)/(x0 - beta)]; a zero column gives
// none), records diag/tau/inv_v0, and publishes column j.  Column j keeps x
// (the reflector is stored implicitly).  A wait that exceeds the spin limit
// sets *status and ends the block (the host turns it into an error).
__global__ __launch_bounds__(kPT) void qr_panel_kernel(const QrDesc* __restrict__ D, int64_t c0, int64_t panel,
                                                       int* __restrict__ status) {
    const QrDesc& d = D[blockIdx.z];
    if (c0 >= d.n) return;
    int* ticket = d.sync;
    int* flag = d.sync + kSyncHead;
    __shared__ double red[kPT / 64];
    __shared__ int sb, sfail;
    if (threadIdx.x == 0) {
        sb = atomicAdd(ticket, 1) - static_cast<int>(panel * kNB);
        sfail = 0;
    }
    __syncthreads();
    const int64_t j = c0 + sb;
    if (sb < 0 || j >= min(c0 + kNB, d.n)) return;
    double* cj = d.a + j * d.lda;
    for (int64_t c = c0; c < j; ++c) {
        if (threadIdx.x == 0) {
            int spins = 0;
            while (__hip_atomic_load(flag + c, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
                if (++spins > kPanelSpinLimit) {
                    sfail = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __syncthreads();
        if (sfail) {
            if (threadIdx.x == 0) atomicOr(status, 1);
            return;
        }
        const double* col = d.a + c * d.lda;
        const double tau = load_published(d.tau + c), inv_v0 = load_published(d.inv_v0 + c);
        const double xc = cj[c];  // read before the reduction's barrier: thread 0 rewrites it
        double dot = 0.0;
        for (int64_t i = c + 1 + threadIdx.x; i < d.k; i += kPT) dot += load_published(col + i) * cj[i];
        dot = panel_sum(dot, red);
        if (tau != 0.0) {
            const double w = tau * (xc + inv_v0 * dot);
            const double wv = w * inv_v0;
            for (int64_t i = c + 1 + threadIdx.x; i < d.k; i += kPT) cj[i] -= wv * load_published(col + i);
            if (threadIdx.x == 0) cj[c] = xc - w;
        }
        __syncthreads();
    }
    const double x0 = cj[j];
    double ss = 0.0;
    for (int64_t i = j + 1 + threadIdx.x; i < d.k; i += kPT) ss += cj[i] * cj[i];
    const double sigma = panel_sum(ss, red);
    const double nrm = sqrt(x0 * x0 + sigma);
    const bool active = nrm != 0.0;
    const double beta = x0 >= 0.0 ? -nrm : nrm;
    if (threadIdx.x == 0) {
        d.diag[j] = active ? beta : x0;
        d.tau[j] = active ? (beta - x0) / beta : 0.0;
        d.inv_v0[j] = active ? 1.0 / (x0 - beta) : 0.0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        __hip_atomic_store(flag + j, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------- TSQR panel
// Panel of kNB columns over rows c0..k-1 as a two-level tree (tall-skinny
// QR): the rows are cut into leaves of LR = 256 or 512 rows; each leaf is
// factored in registers by one block (one row per thread), the leaves'
// R factors (stacked, up to 16 x 32 rows) by one more block, and the
// trailing columns are updated with each level's block reflector
// I - V T V^T.  The sequential part per column is a block-local reduction,
// never a device-wide one: four launches per panel.  R is unique up to the
// signs of its rows, so the result matches column-by-column Householder
// up to those signs and rounding.

// Wave transpose-reduce of 32 values per lane: afterwards v[0] of lane L
// holds the wave sum of value (L >> 1) & 31 (lanes 2i and 2i+1 agree).
// Stage b (b = 5 .. 1) exchanges half of each lane's values with a partner
// across bit b: the lane keeps the half its own bit selects and adds the
// partner's copy of that half.
// The two widest stages with gfx950's lane-swap instructions: for
// HALF = 16 (partner lane ^ 32) v_permlane32_swap(v[t], v[t+16]) leaves
// each lane its own half and its partner's matching half, for HALF = 8
// (partner ^ 16) v_permlane16_swap does the same between 16-lane rows; the
// sum of the two swapped registers is the stage's result, with no selects.
__device__ __forceinline__ void swap_add(double& lo, double hi, bool rows16) {
    const uint64_t a = __builtin_bit_cast(uint64_t, lo), b = __builtin_bit_cast(uint64_t, hi);
    uint32_t a0, a1, b0, b1;
    if (rows16) {
        const auto l = __builtin_amdgcn_permlane16_swap(static_cast<uint32_t>(a), static_cast<uint32_t>(b), false, false);
        const auto h = __builtin_amdgcn_permlane16_swap(static_cast<uint32_t>(a >> 32), static_cast<uint32_t>(b >> 32),
                                                        false, false);
        a0 = l[0]; b0 = l[1]; a1 = h[0]; b1 = h[1];
    } else {
        const auto l = __builtin_amdgcn_permlane32_swap(static_cast<uint32_t>(a), static_cast<uint32_t>(b), false, false);
        const auto h = __builtin_amdgcn_permlane32_swap(static_cast<uint32_t>(a >> 32), static_cast<uint32_t>(b >> 32),
                                                        false, false);
        a0 = l[0]; b0 = l[1]; a1 = h[0]; b1 = h[1];
    }
    lo = __builtin_bit_cast(double, (static_cast<uint64_t>(a1) << 32) | a0) +
         __builtin_bit_cast(double, (static_cast<uint64_t>(b1) << 32) | b0);
}

// The narrow stages with DPP lane reads inside 16-lane rows: the partner
// only has to differ in the stage's bit and agree in the higher ones, so
// row_mirror (lane ^ 15) serves for bit 3, row_half_mirror (lane ^ 7) for
// bit 2, and quad permutations for bits 1 and 0.
template <int CTRL>
__device__ __forceinline__ double dpp_read(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(u)), CTRL, 0xF, 0xF, true));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(u >> 32)), CTRL, 0xF, 0xF, true));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}
template <int HALF, int CTRL>
__device__ __forceinline__ void tr_stage_dpp(double* v, int lane) {
    const bool up = (lane & (2 * HALF)) != 0;
#pragma unroll
    for (int t = 0; t < HALF; ++t) {
        const double send = up ? v[t] : v[t + HALF];
        const double keep = up ? v[t + HALF] : v[t];
        v[t] = keep + dpp_read<CTRL>(send);
    }
}
constexpr int kDppRowMirror = 0x140, kDppRowHalfMirror = 0x141, kDppQuadXor2 = 0x4E, kDppQuadXor1 = 0xB1;

__device__ __forceinline__ void wave_tr32(double* v, int lane) {
#pragma unroll
    for (int t = 0; t < 16; ++t) swap_add(v[t], v[t + 16], false);
#pragma unroll
    for (int t = 0; t < 8; ++t) swap_add(v[t], v[t + 8], true);
    tr_stage_dpp<4, kDppRowMirror>(v, lane);
    tr_stage_dpp<2, kDppRowHalfMirror>(v, lane);
    tr_stage_dpp<1, kDppQuadXor2>(v, lane);
    v[0] += dpp_read<kDppQuadXor1>(v[0]);
}

__device__ __forceinline__ double read_lane(double v, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(u)), l));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(u >> 32)), l));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// One column step of block_qr32 with the column index J a constant, so
// every register index is static; recurses to the next column.  One
// barrier per column: after it every wavefront reads all waves' column sums
// and forms the reflector and u itself (the same arithmetic in every wave,
// so every wave gets the same bits); lane q holds u_q.
template <int NT, int J>
__device__ __forceinline__ void block_qr32_step(double (&a)[kNB], int nbp, double* Ts, double* sh) {
    if constexpr (J < kNB) {
        if (J >= nbp) return;
        constexpr int NW = NT / 64;
        const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
        // sh: red[2][NW][kNB] | rowj[2][kNB]; the buffers alternate by step
        // parity, so a buffer is rewritten only after every wave has passed
        // the next step's barrier, i.e. finished reading it
        constexpr int b = J & 1;
        double* rd = sh + b * NW * kNB;
        double* rj = sh + 2 * NW * kNB + b * kNB;
        if (t == J) {
#pragma unroll
            for (int q = 0; q < kNB; ++q) rj[q] = a[q];
        }
        const double x = t > J ? a[J] : 0.0;
        double vals[kNB];
#pragma unroll
        for (int q = 0; q < kNB; ++q) vals[q] = x * a[q];  // vals[J] = x^2
        wave_tr32(vals, lane);
        if ((lane & 1) == 0) rd[wave * kNB + ((lane >> 1) & 31)] = vals[0];
        __syncthreads();
        // tot_q = sum over rows below J of x * A(:, q) (fixed order), the
        // reflector, and u_q = V(J, q) + inv * tot_q: s_q = v_J^T A(:, q) for
        // q > J and z_q = v_q^T v_J for q < J
        const int q = lane & 31;
        double tq = rd[q], tj = rd[J];
#pragma unroll
        for (int w = 1; w < NW; ++w) {
            tq += rd[w * kNB + q];
            tj += rd[w * kNB + J];
        }
        const double x0 = rj[J];
        const double nrm = sqrt(x0 * x0 + tj);
        const bool active = nrm != 0.0;
        const double beta = x0 >= 0.0 ? -nrm : nrm;
        const double inv = active ? 1.0 / (x0 - beta) : 0.0;
        const double tau = active ? (beta - x0) / beta : 0.0;
        const double uq = rj[q] + inv * tq;
        const bool upd = active && t >= J;
        const double v = t > J ? x * inv : 1.0;
        if (upd) {
            const double tv = tau * v;
#pragma unroll
            for (int l = J + 1; l < kNB; ++l) a[l] -= tv * read_lane(uq, l);
            a[J] = t > J ? v : beta;
        }
        // T(0:J, J) = -tau T(0:J, 0:J) z (wave 0 holds rows t < kNB of T;
        // its earlier columns were written by this same wave)
        if (t < J) {
            double acc = 0.0;
#pragma unroll
            for (int pp = 0; pp < J; ++pp) {
                const double term = Ts[pp * kNB + t] * read_lane(uq, pp);
                acc += pp >= t ? term : 0.0;
            }
            Ts[J * kNB + t] = -tau * acc;
        }
        if (t == J) Ts[J * kNB + J] = tau;
        block_qr32_step<NT, J + 1>(a, nbp, Ts, sh);
    }
}

// LDS scratch of block_qr32, in doubles.
__host__ __device__ constexpr int block_qr32_lds_doubles(int nt) { return 2 * (nt / 64) * kNB + 4 * kNB + 8; }

// a[] = this thread's row (row index t of the block's rows; rows past the
// tile are zero).  Householder QR of the NT x nbp block in place: rows
// t <= q of column q end as R, rows t > q as the reflector entries
// (unit diagonal implicit).  Ts (LDS, column-major 32 x 32) gets the upper
// triangular T with H_0 ... H_{nbp-1} = I - V T V^T (dlarft forward
// columnwise, formed column by column).  dlarfg convention per column:
// beta = -sign(x0)|x|, tau = (beta - x0)/beta, v = [1, x(1:)/(x0 - beta)];
// a zero column gives tau = 0.  sh: block_qr32_lds_doubles(NT) doubles.
template <int NT>
__device__ __forceinline__ void block_qr32(double (&a)[kNB], int nbp, double* Ts, double* sh) {
    for (int i = threadIdx.x; i < kNB * kNB; i += NT) Ts[i] = 0.0;
    __syncthreads();
    block_qr32_step<NT, 0>(a, nbp, Ts, sh);
    __syncthreads();
}

// Leaf factorization: block (leaf l, problem z) factors rows
// r0 = c0 + l*NT .. of the panel; R and V back into A, T into d.tl[l].
template <int NT>
__global__ __launch_bounds__(NT) void qr_tsqr_leaf_kernel(const QrDesc* __restrict__ D, int64_t c0, int par) {
    const QrDesc& d = D[blockIdx.z];
    if (c0 >= d.n) return;
    const int nbp = static_cast<int>(min<int64_t>(kNB, d.n - c0));
    const int64_t r0 = c0 + static_cast<int64_t>(blockIdx.x) * NT;
    if (r0 >= d.k) return;
    const int rows = static_cast<int>(min<int64_t>(NT, d.k - r0));
    __shared__ double Ts[kNB * kNB];
    __shared__ double sh[block_qr32_lds_doubles(NT)];
    const int t = threadIdx.x;
    double a[kNB];
#pragma unroll
    for (int q = 0; q < kNB; ++q) a[q] = (t < rows && q < nbp) ? d.a[r0 + t + (c0 + q) * d.lda] : 0.0;
    block_qr32<NT>(a, nbp, Ts, sh);
    if (t < rows) {
#pragma unroll
        for (int q = 0; q < kNB; ++q)
            if (q < nbp) d.a[r0 + t + (c0 + q) * d.lda] = a[q];
    }
    double* tl = d.tl + par * kTsqrSet + static_cast<int64_t>(blockIdx.x) * kNB * kNB;
    for (int i = t; i < kNB * kNB; i += NT) tl[i] = Ts[i];
}

// Combine: the leaves' R factors stacked (leaf l's row r at stacked row
// 32 l + r) factored by one block of kTsqrComb threads.  The final R goes
// to the first leaf's top rows (upper triangle) and its diagonal to d.diag;
// the stacked reflectors to d.vc (explicit, unit diagonal), T to d.tc.
template <int NC>
__global__ __launch_bounds__(NC) void qr_tsqr_combine_kernel(const QrDesc* __restrict__ D, int64_t c0, int lr,
                                                             int par) {
    const QrDesc& d = D[blockIdx.z];
    if (c0 >= d.n) return;
    const int nbp = static_cast<int>(min<int64_t>(kNB, d.n - c0));
    const int64_t kk = d.k - c0;
    const int nleaves = static_cast<int>((kk + lr - 1) / lr);
    __shared__ double Ts[kNB * kNB];
    __shared__ double sh[block_qr32_lds_doubles(NC)];
    const int t = threadIdx.x, l = t >> 5, r = t & 31;
    const int64_t row = c0 + static_cast<int64_t>(l) * lr + r;
    const bool valid = l < nleaves && r < min<int64_t>(lr, kk - static_cast<int64_t>(l) * lr);
    double a[kNB];
#pragma unroll
    for (int q = 0; q < kNB; ++q) a[q] = (valid && q >= r && q < nbp) ? d.a[row + (c0 + q) * d.lda] : 0.0;
    block_qr32<NC>(a, nbp, Ts, sh);
    if (t < nbp) {
#pragma unroll
        for (int q = 0; q < kNB; ++q)
            if (q >= t && q < nbp) d.a[c0 + t + (c0 + q) * d.lda] = a[q];
#pragma unroll
        for (int q = 0; q < kNB; ++q)
            if (q == t) d.diag[c0 + t] = a[q];
    }
    double* vc = d.vc + par * kTsqrSet;
    double* tc = d.tc + par * kTsqrSet;
#pragma unroll
    for (int q = 0; q < kNB; ++q) vc[t * kNB + q] = t > q ? a[q] : (t == q ? 1.0 : 0.0);
    for (int i = t; i < kNB * kNB; i += NC) tc[i] = Ts[i];
}

// Trailing update with one level's block reflector: rows S of the trailing
// columns j0 + jt0 .. j0 + jt0 + kCT - 1 become S - V T^T V^T S.
// COMBINE = false: S = leaf blockIdx.x's rows, V from the panel columns
// below the leaf's diagonal, T = d.tl[leaf].  COMBINE = true: S = the
// stacked top kNB rows of every leaf, V = d.vc, T = d.tc.  One row per
// thread (V and S rows kept in registers); W = V^T S accumulates over
// 256-row chunks staged in LDS, then T^T W, all in a fixed order.
constexpr int kCT = 16;
constexpr int kChunk = 128;

template <int NT, bool COMBINE>
__global__ __launch_bounds__(NT) void qr_tsqr_update_kernel(const QrDesc* __restrict__ D, int64_t c0, int lr,
                                                            int par, int tile0) {
    const QrDesc& d = D[blockIdx.z];
    if (c0 >= d.n) return;
    const int nbp = static_cast<int>(min<int64_t>(kNB, d.n - c0));
    const int64_t j0 = c0 + nbp, nt = d.n - j0;
    const int64_t jt0 = static_cast<int64_t>(tile0 + static_cast<int>(blockIdx.y)) * kCT;
    if (nt <= 0 || jt0 >= nt) return;
    const int64_t kk = d.k - c0;
    const int t = threadIdx.x;
    int64_t row;
    bool valid;
    const double* tsrc;
    if constexpr (COMBINE) {
        const int l = t >> 5, r = t & 31;
        const int nleaves = static_cast<int>((kk + lr - 1) / lr);
        row = c0 + static_cast<int64_t>(l) * lr + r;
        valid = l < nleaves && r < min<int64_t>(lr, kk - static_cast<int64_t>(l) * lr);
        tsrc = d.tc + par * kTsqrSet;
    } else {
        const int64_t r0 = c0 + static_cast<int64_t>(blockIdx.x) * NT;
        if (r0 >= d.k) return;
        row = r0 + t;
        valid = row < d.k;
        tsrc = d.tl + par * kTsqrSet + static_cast<int64_t>(blockIdx.x) * kNB * kNB;
    }
    __shared__ double sV[kChunk][kNB + 1];
    __shared__ double sS[kChunk][kCT + 1];
    __shared__ double sT[kNB * kNB];
    __shared__ double sW[kNB][kCT + 1];
    __shared__ double sW2[kNB][kCT + 1];
    double v[kNB], srow[kCT];
#pragma unroll
    for (int q = 0; q < kNB; ++q) {
        double x = 0.0;
        if (valid && q < nbp) {
            if constexpr (COMBINE) x = d.vc[par * kTsqrSet + t * kNB + q];
            else x = t > q ? d.a[row + (c0 + q) * d.lda] : (t == q ? 1.0 : 0.0);
        }
        v[q] = x;
    }
#pragma unroll
    for (int c = 0; c < kCT; ++c) srow[c] = (valid && jt0 + c < nt) ? d.a[row + (j0 + jt0 + c) * d.lda] : 0.0;
    for (int i = t; i < kNB * kNB; i += NT) sT[i] = tsrc[i];
    constexpr int kPer = (kNB * kCT + NT - 1) / NT;  // W entries per thread
    double wacc[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) wacc[u] = 0.0;
    for (int ch = 0; ch < NT / kChunk; ++ch) {  // W = V^T S, chunk by chunk
        if (t / kChunk == ch) {
            const int lt = t - ch * kChunk;
#pragma unroll
            for (int q = 0; q < kNB; ++q) sV[lt][q] = v[q];
#pragma unroll
            for (int c = 0; c < kCT; ++c) sS[lt][c] = srow[c];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = t + u * NT;
            if (e < kNB * kCT) {
                const int q = e % kNB, c = e / kNB;
                double acc = wacc[u];
#pragma unroll 16
                for (int i = 0; i < kChunk; ++i) acc += sV[i][q] * sS[i][c];
                wacc[u] = acc;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int e = t + u * NT;
        if (e < kNB * kCT) sW[e % kNB][e / kNB] = wacc[u];
    }
    __syncthreads();
    for (int e = t; e < kNB * kCT; e += NT) {  // W2 = T^T W (T upper triangular)
        const int q = e % kNB, c = e / kNB;
        double acc = 0.0;
#pragma unroll
        for (int pp = 0; pp < kNB; ++pp) {
            const double term = sT[q * kNB + pp] * sW[pp][c];
            acc += pp <= q ? term : 0.0;
        }
        sW2[q][c] = acc;
    }
    __syncthreads();
    if (!valid) return;
#pragma unroll
    for (int c = 0; c < kCT; ++c) {
        if (jt0 + c >= nt) continue;
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < kNB; ++q) acc += v[q] * sW2[q][c];
        d.a[row + (j0 + jt0 + c) * d.lda] = srow[c] - acc;
    }
}

// Zero the panel kernel's ticket counters and column flags; status = 0.
__global__ __launch_bounds__(kT) void qr_sync_reset_kernel(const QrDesc* __restrict__ D, int* __restrict__ status) {
    const QrDesc& d = D[blockIdx.z];
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (t < kSyncHead + d.n) d.sync[t] = 0;
    if (blockIdx.z == 0 && t == 0) *status = 0;
}

// V (rows c0..k-1 used, ld = k): unit lower trapezoidal reflector columns.
__global__ __launch_bounds__(kT) void qr_form_v_kernel(const QrDesc* __restrict__ D, int64_t c0) {
    const QrDesc& d = D[blockIdx.z];
    const int64_t nbp = min<int64_t>(kNB, d.n - c0);
    if (nbp <= 0 || c0 + nbp >= d.n) return;  // no trailing columns: V unused
    const int64_t rows = d.k - c0;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (t >= rows * nbp) return;
    const int64_t jj = t / rows, i = c0 + (t - jj * rows);
    const int64_t c = c0 + jj;
    d.v[i + jj * d.k] = i < c ? 0.0 : (i == c ? 1.0 : d.a[i + c * d.lda] * d.inv_v0[c]);
}

// Partial Gram V^T V over a chunk of kGR rows: gp[chunk][p + q*kNB].
__global__ __launch_bounds__(kT) void qr_gram_kernel(const QrDesc* __restrict__ D, int64_t c0) {
    const QrDesc& d = D[blockIdx.z];
    const int64_t nbp = min<int64_t>(kNB, d.n - c0);
    if (nbp <= 0 || c0 + nbp >= d.n) return;
    const int64_t r0 = c0 + static_cast<int64_t>(blockIdx.x) * kGR;
    if (r0 >= d.k) return;
    __shared__ double sv[kGR][kNB + 1];
    for (int e = threadIdx.x; e < kGR * kNB; e += kT) {
        const int rr = e % kGR, q = e / kGR;
        const int64_t i = r0 + rr;
        sv[rr][q] = (i < d.k && q < nbp) ? d.v[i + q * d.k] : 0.0;
    }
    __syncthreads();
    const int rows = static_cast<int>(min<int64_t>(kGR, d.k - r0));
    for (int e = threadIdx.x; e < kNB * kNB; e += kT) {
        const int p = e % kNB, q = e / kNB;
        double acc = 0.0;
        if (p < nbp && q < nbp)
            for (int r = 0; r < rows; ++r) acc += sv[r][p] * sv[r][q];
        d.gp[static_cast<int64_t>(blockIdx.x) * kNB * kNB + e] = acc;
    }
}

// T (kNB x kNB, column-major) from the Gram partials and tau (dlarft,
// forward columnwise): T(0:i,i) = -tau_i T(0:i,0:i) G(0:i,i), T(i,i) = tau_i.
constexpr int kTT = kNB * kNB;  // form-T block: one thread per entry

__global__ __launch_bounds__(kTT) void qr_form_t_kernel(const QrDesc* __restrict__ D, int64_t c0) {
    const QrDesc& d = D[blockIdx.z];
    const int64_t nbp64 = min<int64_t>(kNB, d.n - c0);
    if (nbp64 <= 0 || c0 + nbp64 >= d.n) return;
    const int nbp = static_cast<int>(nbp64);
    const int nchunks = static_cast<int>((d.k - c0 + kGR - 1) / kGR);
    const double* tau = d.tau + c0;
    __shared__ double g[kNB * kNB];
    __shared__ double t[kNB * kNB];
    __shared__ double z[kNB];
    {
        const int e = threadIdx.x;
        g[e] = strided_sum(d.gp + e, nchunks, kNB * kNB);
        t[e] = 0.0;
    }
    __syncthreads();
    for (int i = 0; i < nbp; ++i) {
        const double ti = tau[i];
        if (threadIdx.x < i) z[threadIdx.x] = -ti * g[threadIdx.x + i * kNB];
        __syncthreads();
        if (threadIdx.x < i) {
            double acc = 0.0;
            for (int q = threadIdx.x; q < i; ++q) acc += t[threadIdx.x + q * kNB] * z[q];
            t[threadIdx.x + i * kNB] = acc;
        }
        if (threadIdx.x == 0) t[i + i * kNB] = ti;
        __syncthreads();
    }
    d.t[threadIdx.x] = t[threadIdx.x];
}

// W partials: wp[split][jt][p] = sum over the split's rows of V(i,p) A(i, c_end+jt).
// grid (ceil(nt/32), nsplit, problems); 256 threads = 32 p x 8 groups of 4 columns.
__global__ __launch_bounds__(kT) void qr_w_kernel(const QrDesc* __restrict__ D, int64_t c0) {
    const QrDesc& d = D[blockIdx.z];
    if (cq_failed(d)) return;  // a CholeskyQR2 panel failed: the host redoes this problem
    const int64_t nbp64 = min<int64_t>(kNB, d.n - c0);
    if (nbp64 <= 0) return;
    const int nbp = static_cast<int>(nbp64);
    const int64_t j0 = c0 + nbp, nt = d.n - j0;
    const int64_t rows = d.k - c0;
    const int64_t rps = (rows + qr_nsplit(d.k) - 1) / qr_nsplit(d.k);
    const int64_t jt0 = static_cast<int64_t>(blockIdx.x) * kNB;
    const int64_t ra = c0 + static_cast<int64_t>(blockIdx.y) * rps;
    if (nt <= 0 || jt0 >= nt || ra >= d.k) return;
    const int64_t rb = min(d.k, ra + rps);
    __shared__ double sv[64][kNB + 1];
    __shared__ double sa[64][kNB + 1];
    const int p = threadIdx.x % kNB, qg = threadIdx.x / kNB;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t r = ra; r < rb; r += 64) {
        for (int e = threadIdx.x; e < 64 * kNB; e += kT) {
            const int rr = e % 64, q = e / 64;
            const int64_t i = r + rr;
            sv[rr][q] = (i < rb && q < nbp) ? d.v[i + q * d.k] : 0.0;
            sa[rr][q] = (i < rb && jt0 + q < nt) ? d.a[i + (j0 + jt0 + q) * d.lda] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int rr = 0; rr < 64; ++rr) {
            const double vv = sv[rr][p];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[u] += vv * sa[rr][qg * 4 + u];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int64_t jt = jt0 + qg * 4 + u;
        if (jt < nt && p < nbp) d.wp[(static_cast<int64_t>(blockIdx.y) * nt + jt) * kNB + p] = acc[u];
    }
}

// W2 = T^T W for 64 trailing columns per block: the split partials of W
// summed in a fixed order, then the triangular product.
__global__ __launch_bounds__(kT) void qr_w2_kernel(const QrDesc* __restrict__ D, int64_t c0) {
    const QrDesc& d = D[blockIdx.z];
    if (cq_failed(d)) return;  // a CholeskyQR2 panel failed: the host redoes this problem
    const int64_t nbp64 = min<int64_t>(kNB, d.n - c0);
    if (nbp64 <= 0) return;
    const int nbp = static_cast<int>(nbp64);
    const int64_t j0 = c0 + nbp, nt = d.n - j0;
    const int64_t rows = d.k - c0;
    const int64_t jt0 = static_cast<int64_t>(blockIdx.x) * 64;
    if (nt <= 0 || jt0 >= nt) return;
    const int64_t rps = (rows + qr_nsplit(d.k) - 1) / qr_nsplit(d.k);
    const int nsplit = static_cast<int>((rows + rps - 1) / rps);
    __shared__ double sw[kNB][64 + 1];
    __shared__ double st[kNB * kNB];
    for (int e = threadIdx.x; e < kNB * kNB; e += kT) st[e] = d.t[e];
    for (int e = threadIdx.x; e < kNB * 64; e += kT) {
        const int p = e % kNB, jj = e / kNB;
        const int64_t jt = jt0 + jj;
        sw[p][jj] = (jt < nt && p < nbp) ? strided_sum(d.wp + jt * kNB + p, nsplit, nt * kNB) : 0.0;
    }
    __syncthreads();
    // W2(p, j) = sum_{q <= p} T(q, p) W(q, j)
    for (int e = threadIdx.x; e < kNB * 64; e += kT) {
        const int p = e % kNB, jj = e / kNB;
        const int64_t jt = jt0 + jj;
        if (jt >= nt) continue;
        double acc = 0.0;
        for (int q = 0; q <= p; ++q) acc += st[q + p * kNB] * sw[q][jj];
        d.w2[jt * kNB + p] = p < nbp ? acc : 0.0;
    }
}

// A(c0:k, c_end+jt) -= V W2 on 64 x 64 tiles; V is set `par` of d.v and the
// column tiles start at tile y0.
__global__ __launch_bounds__(kT) void qr_update_kernel(const QrDesc* __restrict__ D, int64_t c0, int par, int y0) {
    const QrDesc& d = D[blockIdx.z];
    if (cq_failed(d)) return;  // a CholeskyQR2 panel failed: the host redoes this problem
    const int64_t nbp64 = min<int64_t>(kNB, d.n - c0);
    if (nbp64 <= 0) return;
    const int nbp = static_cast<int>(nbp64);
    const int64_t j0 = c0 + nbp, nt = d.n - j0;
    const int64_t r0 = c0 + static_cast<int64_t>(blockIdx.x) * 64;
    const int64_t jt0 = static_cast<int64_t>(y0 + static_cast<int>(blockIdx.y)) * 64;
    if (nt <= 0 || r0 >= d.k || jt0 >= nt) return;
    const double* vv = d.v + par * d.k * kNB;
    // 256 threads: 16 row groups x 16 column groups, 4 x 4 outputs each; this
    // thread's A entries are loaded first, overlapping the staging of V and W2
    const int rg = threadIdx.x % 16, cgp = threadIdx.x / 16;
    double aold[4][4];
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        const int64_t jt = jt0 + cgp + 16 * y;
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            const int64_t i = r0 + rg + 16 * x;
            aold[x][y] = (jt < nt && i < d.k) ? d.a[i + (j0 + jt) * d.lda] : 0.0;
        }
    }
    __shared__ double sw2[kNB][64 + 1];
    __shared__ double sv[64][kNB + 1];
    for (int e = threadIdx.x; e < kNB * 64; e += kT) {
        const int p = e % kNB, jj = e / kNB;
        const int64_t jt = jt0 + jj;
        sw2[p][jj] = (jt < nt && p < nbp) ? d.w2[jt * kNB + p] : 0.0;
    }
    for (int e = threadIdx.x; e < 64 * kNB; e += kT) {
        const int rr = e % 64, q = e / 64;
        const int64_t i = r0 + rr;
        sv[rr][q] = (i < d.k && q < nbp) ? vv[i + q * d.k] : 0.0;
    }
    __syncthreads();
    double acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = 0.0;
    for (int p = 0; p < nbp; ++p) {
        double vv[4], ww[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) vv[x] = sv[rg + 16 * x][p];
#pragma unroll
        for (int y = 0; y < 4; ++y) ww[y] = sw2[p][cgp + 16 * y];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) acc[x][y] += vv[x] * ww[y];
    }
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        const int64_t jt = jt0 + cgp + 16 * y;
        if (jt >= nt) continue;
        double* colp = d.a + (j0 + jt) * d.lda;
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            const int64_t i = r0 + rg + 16 * x;
            if (i < d.k) colp[i] = aold[x][y] - acc[x][y];
        }
    }
}

// ---------------------------------------------------------------- CholeskyQR2 panels
// A panel of nbp <= kNB columns over rows c0..k-1 (P) factored without a
// column-by-column pass over its rows:
//   G1 = P^T P, R1 = chol(G1);  G2 = (P R1^-1)^T (P R1^-1), R2 = chol(G2);
//   Rc = R2 R1, Q = P Rc^-1 (CholeskyQR2: Q has orthonormal columns to
//   working precision while cond(P) stays below ~1e7).
// The Householder form of Q is then reconstructed (Ballard, Demmel,
// Grigori, Jacquelin, Knight, Nguyen, "Reconstructing Householder vectors
// from tall-skinny QR", 2015): with S = diag(-sign(Q_ii)) and Q S - I =
// Y U (LU without pivoting, Y unit lower trapezoidal), T = -U Y1^-T makes
// H = I - Y T Y^T orthogonal with Q S as its first nbp columns, so the
// panel's R is S Rc and the trailing columns get H^T A exactly as after a
// Householder panel (qr_cq_w / qr_update kernels).  Y's rows below the
// panel's top block are P2 N with N = Rc^-1 S U^-1.  Every row-sized step
// is one launch over all row chunks; the kNB x kNB algebra after each Gram
// pass is one block per problem (qr_cq_small_kernel), its factorizations
// on one wavefront with a matrix column per lane.
// A Gram matrix that is not numerically positive definite (a pivot below
// kCqPivot times its diagonal, or not finite), or a second Gram matrix
// further than kCqOrth from I, marks the problem failed: its later
// launches return at once and the host factors it again from the copy
// taken first, with the Householder panels.
constexpr int kCqT = 256;            // threads of the Gram and form-Y kernels
constexpr int kCqTile = 64;          // rows per LDS tile
constexpr int kCqLD = kNB + 2;       // LDS row stride of a tile (16-byte aligned rows)
constexpr double kCqPivot = 1e-15;   // relative Cholesky pivot floor
constexpr double kCqOrth = 0.1;      // max |G2 - I| for which two passes give an orthonormal Q
// One Gram pass suffices when the column-normalized Gram matrix Gn = D^-1 G1
// D^-1 (D = sqrt(diag G1)) has every off-diagonal absolute row sum below
// kCqOnePass: by Gershgorin its eigenvalues then lie in [1 - 0.8, 1 + 0.8],
// cond(P D^-1) < 3, and Q = P R1^-1 is orthonormal to a few ulps (Cholesky
// is invariant under column scaling, so the graded columns of a sketch do
// not count against it).
constexpr double kCqOnePass = 0.8;

// Rows per Gram chunk: a multiple of the tile, at least two tiles, at most
// 64 chunks.  Chunks never exceed ceil(rows / 128).
__host__ __device__ inline int64_t cq_chunk_rows(int64_t rows) {
    int64_t c = (rows + 63) / 64;
    c = (c + kCqTile - 1) / kCqTile * kCqTile;
    return c < 2 * kCqTile ? 2 * kCqTile : c;
}

// Row splits of the CholeskyQR2 path's W = Y^T A: about 256 rows each, at
// most 16.
__host__ __device__ inline int cq_wsplits(int64_t rows) {
    const int64_t s = (rows + 255) / 256;
    return static_cast<int>(s < 1 ? 1 : (s > 16 ? 16 : s));
}

// The kNB x kNB algebra of the last Gram block runs on one wavefront, lane
// j (mod 32) owning matrix column j; what a step needs from other lanes
// goes through a kNB-double LDS buffer (bc) and is read back as broadcasts.
// LDS accesses of one wavefront complete in order, so a write followed by
// reads of the same buffer needs no barrier; a compiler memory barrier
// (cq_order) around each exchange keeps the accesses in program order and
// stops the compiler from replacing a read by the value this lane wrote
// (the other lanes' writes are what is read).  All matrices are row-major
// in LDS.
__device__ __forceinline__ void cq_order() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// 1/sqrt(x) and 1/x to full FP64 precision: the hardware estimates
// refined by Newton steps (two for the square root, two for the
// reciprocal; each doubles the correct bits of the ~2^-23 estimate).
__device__ __forceinline__ double cq_rsqrt(double x) {
    double y = __builtin_amdgcn_rsq(x);
    y = y * (1.5 - 0.5 * x * y * y);
    return y * (1.5 - 0.5 * x * y * y);
}
__device__ __forceinline__ double cq_rcp(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = y * (2.0 - x * y);
    return y * (2.0 - x * y);
}

// Upper Cholesky factor R (G = R^T R) of the nb x nb leading block of G,
// padded with the identity to kNB x kNB, into r.  Step i: the pivot comes
// from lane i (readlane), every lane scales its own entry of row i and
// writes it to bc, and every lane updates its column with the row read
// back.  Returns false (on every lane) if a pivot is not finite or not
// above kCqPivot times its original diagonal.
__device__ __forceinline__ bool cq_wave_chol(const double* g, double* r, double* bc, int nb) {
    const int l = threadIdx.x & 63, j = l & 31;
    double c[kNB];
#pragma unroll
    for (int i = 0; i < kNB; ++i) c[i] = (i < nb && j < nb) ? g[i * kNB + j] : (i == j ? 1.0 : 0.0);
    int ok = 1;
#pragma unroll
    for (int i = 0; i < kNB; ++i) {
        const double piv = read_lane(c[i], i);
        const double gii = i < nb ? g[i * kNB + i] : 1.0;  // broadcast read
        ok &= static_cast<int>(piv > kCqPivot * gii) & static_cast<int>(piv < __builtin_huge_val());
        const double rs = cq_rsqrt(piv);
        const double rij = j >= i ? c[i] * rs : 0.0;
        c[i] = rij;
        cq_order();
        bc[j] = rij;
        cq_order();
#pragma unroll
        for (int p = i + 1; p < kNB; ++p) c[p] -= bc[p] * rij;
    }
    if (l < kNB) {
#pragma unroll
        for (int i = 0; i < kNB; ++i) r[i * kNB + j] = c[i];
    }
    return ok != 0;
}

// x = r^-1 for r upper triangular (kNB x kNB, nonzero diagonal): lane j
// solves r x = e_j by back substitution, x_i = (e_ij - r(i, >i) x(>i)) /
// r_ii, the dot product over row i (contiguous, broadcast reads) in four
// running sums.
__device__ __forceinline__ void cq_wave_inv_upper(const double* r, double* x_out) {
    const int l = threadIdx.x & 63, j = l & 31;
    double x[kNB];
#pragma unroll
    for (int i = kNB - 1; i >= 0; --i) {
        double s0 = i == j ? 1.0 : 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
        for (int k = i + 1; k < kNB; ++k) {
            const double t = r[i * kNB + k] * x[k];
            if ((k - i - 1) % 4 == 0) s0 -= t;
            else if ((k - i - 1) % 4 == 1) s1 -= t;
            else if ((k - i - 1) % 4 == 2) s2 -= t;
            else s3 -= t;
        }
        x[i] = ((s0 + s1) + (s2 + s3)) * cq_rcp(r[i * kNB + i]);
    }
    if (l < kNB) {
#pragma unroll
        for (int i = 0; i < kNB; ++i) x_out[i * kNB + j] = x[i];
    }
}

// x = L^-1 for L unit lower triangular (diagonal not read), forward
// substitution over row i the same way.
__device__ __forceinline__ void cq_wave_inv_unit_lower(const double* lo, double* x_out) {
    const int l = threadIdx.x & 63, j = l & 31;
    double x[kNB];
#pragma unroll
    for (int i = 0; i < kNB; ++i) {
        double s0 = i == j ? 1.0 : 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
        for (int k = 0; k < i; ++k) {
            const double t = lo[i * kNB + k] * x[k];
            if (k % 4 == 0) s0 -= t;
            else if (k % 4 == 1) s1 -= t;
            else if (k % 4 == 2) s2 -= t;
            else s3 -= t;
        }
        x[i] = (s0 + s1) + (s2 + s3);
    }
    if (l < kNB) {
#pragma unroll
        for (int i = 0; i < kNB; ++i) x_out[i * kNB + j] = x[i];
    }
}

// LU without pivoting of v (kNB x kNB): lo gets the unit lower factor, up
// the upper one.  Step i: lane i publishes its column below the pivot (the
// multipliers before scaling) in bc and writes the scaled multipliers to
// lo; every lane j > i then subtracts multiplier x U(i, j) from its column.
__device__ __forceinline__ void cq_wave_lu(const double* v, double* lo, double* up, double* bc) {
    const int l = threadIdx.x & 63, j = l & 31;
    double c[kNB];
#pragma unroll
    for (int i = 0; i < kNB; ++i) c[i] = v[i * kNB + j];
#pragma unroll
    for (int i = 0; i < kNB; ++i) {
        const double inv = cq_rcp(read_lane(c[i], i));
        cq_order();
        if (j == i) {
#pragma unroll
            for (int p = i + 1; p < kNB; ++p) bc[p] = c[p];
        }
        cq_order();
        const double uij = j > i ? c[i] : 0.0;
#pragma unroll
        for (int p = i + 1; p < kNB; ++p) c[p] -= (bc[p] * inv) * uij;
        if (l == i) {
#pragma unroll
            for (int p = i + 1; p < kNB; ++p) lo[p * kNB + i] = bc[p] * inv;
        }
    }
    if (l < kNB) {
#pragma unroll
        for (int i = 0; i < kNB; ++i) {
            up[i * kNB + j] = i <= j ? c[i] : 0.0;
            if (i <= j) lo[i * kNB + j] = i == j ? 1.0 : 0.0;
        }
    }
}

// c = a b for kNB x kNB row-major LDS matrices, all kCqT threads (a 2 x 2
// block of c each); the caller separates producers and consumers with
// barriers.
__device__ __forceinline__ void cq_block_mm(const double* a, const double* b, double* c) {
    const int i0 = 2 * (threadIdx.x / 16), j0 = 2 * (threadIdx.x % 16);
    double c00 = 0.0, c01 = 0.0, c10 = 0.0, c11 = 0.0;
#pragma unroll 8
    for (int k = 0; k < kNB; ++k) {
        const double a0 = a[i0 * kNB + k], a1 = a[(i0 + 1) * kNB + k];
        const double b0 = b[k * kNB + j0], b1 = b[k * kNB + j0 + 1];
        c00 += a0 * b0;
        c01 += a0 * b1;
        c10 += a1 * b0;
        c11 += a1 * b1;
    }
    c[i0 * kNB + j0] = c00;
    c[i0 * kNB + j0 + 1] = c01;
    c[(i0 + 1) * kNB + j0] = c10;
    c[(i0 + 1) * kNB + j0 + 1] = c11;
}

// Reset tickets and failure flags (and the panel kernel's sync words).
__global__ void qr_cq_reset_kernel(const QrDesc* __restrict__ D, int nprob) {
    const int p = blockIdx.x * 64 + threadIdx.x;
    if (p < nprob) {
        int* ci = cq_ints(D[p]);
        ci[0] = 0;
        ci[1] = 0;
        ci[2] = 0;
        ci[3] = 0;
    }
}

// A (k x n, lda) <-> the copy bk (ld k): TO_COPY = true saves, false restores.
template <bool TO_COPY>
__global__ __launch_bounds__(kT) void qr_cq_copy_kernel(const QrDesc* __restrict__ D) {
    const QrDesc& d = D[blockIdx.z];
    const int64_t total = d.k * d.n;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x; t < total;
         t += static_cast<int64_t>(gridDim.x) * kT) {
        const int64_t j = t / d.k, i = t - j * d.k;
        if (TO_COPY) d.bk[t] = d.a[i + j * d.lda];
        else d.a[i + j * d.lda] = d.bk[t];
    }
}

// Partial Gram matrix of one row chunk of the panel (grid (chunks, 1,
// problems)).  PASS 1: P^T P; PASS 2: (P R1^-1)^T (P R1^-1).  Tiles of 64
// rows go through LDS; the next tile is loaded into registers while the
// current one is multiplied.  Each lane keeps a 4 x 4 block of the Gram
// matrix; the four waves' blocks are added in LDS and the chunk's partial
// stored to d.gp (the small-algebra kernel adds the chunks in order).
template <int PASS>
__global__ __launch_bounds__(kCqT) void qr_cq_gram_kernel(const QrDesc* __restrict__ D, int64_t c0) {
    const QrDesc& d = D[blockIdx.z];
    if (c0 >= d.n || cq_failed(d)) return;
    if (PASS == 2 && cq_ints(d)[3]) return;  // one pass suffices for this panel
    const int nbp = static_cast<int>(min<int64_t>(kNB, d.n - c0));
    const int64_t rows = d.k - c0;
    const int64_t cr = cq_chunk_rows(rows);
    const int nch = static_cast<int>((rows + cr - 1) / cr);
    if (static_cast<int>(blockIdx.x) >= nch) return;
    const int64_t ra = c0 + static_cast<int64_t>(blockIdx.x) * cr, rb = min(d.k, ra + cr);
    __shared__ __align__(16) double sred[4 * kCqMat];
    __shared__ __align__(16) double sp[kCqTile * kCqLD];
    __shared__ __align__(16) double sx[kCqMat];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    if (PASS == 2)
        for (int e = t; e < kCqMat; e += kCqT) sx[e] = d.cq[e];  // R1^-1, row-major
    constexpr int kPer = kCqTile * kNB / kCqT;  // tile elements per thread
    double nxt[kPer];
    auto load_tile = [&](int64_t r) {
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = t + u * kCqT, rr = e & (kCqTile - 1), q = e / kCqTile;
            const int64_t i = r + rr;
            nxt[u] = (i < rb && q < nbp) ? d.a[i + (c0 + q) * d.lda] : 0.0;
        }
    };
    load_tile(ra);
    const int p0 = 4 * (l & 7), q0 = 4 * (l >> 3);  // this lane's 4 x 4 block of G
    double acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = 0.0;
    for (int64_t r = ra; r < rb; r += kCqTile) {
        __syncthreads();  // the previous tile has been read
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = t + u * kCqT;
            sp[(e & (kCqTile - 1)) * kCqLD + e / kCqTile] = nxt[u];
        }
        __syncthreads();
        if (r + kCqTile < rb) load_tile(r + kCqTile);
        if (PASS == 2) {  // rows of P R1^-1 (R1^-1 upper: column j uses rows i <= j)
            const int rr = t >> 2, j0 = (t & 3) * 8;
            double o[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) o[u] = 0.0;
            for (int i = 0; i < j0 + 8; ++i) {
                const double pv = sp[rr * kCqLD + i];
#pragma unroll
                for (int u = 0; u < 8; ++u) o[u] += pv * sx[i * kNB + j0 + u];
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 8; ++u) sp[rr * kCqLD + j0 + u] = o[u];
            __syncthreads();
        }
        for (int rr = 16 * w; rr < 16 * w + 16; ++rr) {
            double a4[4], b4[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                a4[x] = sp[rr * kCqLD + p0 + x];
                b4[x] = sp[rr * kCqLD + q0 + x];
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) acc[x][y] += a4[x] * b4[y];
        }
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) sred[w * kCqMat + (p0 + x) * kNB + q0 + y] = acc[x][y];
    __syncthreads();
    double* gpart = d.gp + static_cast<int64_t>(blockIdx.x) * kCqMat;
    for (int e = t; e < kCqMat; e += kCqT)
        gpart[e] = (sred[e] + sred[kCqMat + e]) + (sred[2 * kCqMat + e] + sred[3 * kCqMat + e]);
}

// The Gram matrix from the chunk partials (added in chunk order, all loads
// of a batch of 16 chunks in flight), into g (row-major, LDS).
__device__ __forceinline__ void cq_sum_partials(const QrDesc& d, int nch, double* g) {
    const int t = threadIdx.x;
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    for (int cb = 0; cb < nch; cb += 16) {
        double v[16][4];
#pragma unroll
        for (int u = 0; u < 16; ++u)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                v[u][q] = cb + u < nch ? d.gp[static_cast<int64_t>(cb + u) * kCqMat + t + q * kCqT] : 0.0;
#pragma unroll
        for (int u = 0; u < 16; ++u)
#pragma unroll
            for (int q = 0; q < 4; ++q) sum[q] += v[u][q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) g[t + q * kCqT] = sum[q];
}

// The kNB x kNB algebra of a panel, one block per problem (grid (1, 1,
// problems)).  PASS 1: R1 = chol(G1) and R1^-1.  PASS 2: R2 = chol(G2),
// Rc = R2 R1, the reconstruction (S, L1, U, T, N), the top rows of Y and
// the panel's rows of R.
template <int PASS>
__global__ __launch_bounds__(kCqT) void qr_cq_small_kernel(const QrDesc* __restrict__ D, int64_t c0, int par) {
    const QrDesc& d = D[blockIdx.z];
    if (c0 >= d.n || cq_failed(d)) return;
    const int nbp = static_cast<int>(min<int64_t>(kNB, d.n - c0));
    const int64_t rows = d.k - c0;
    const int nch = static_cast<int>((rows + cq_chunk_rows(rows) - 1) / cq_chunk_rows(rows));
    __shared__ __align__(16) double sm[7 * kCqMat];
    __shared__ __align__(16) double bc[kNB];
    __shared__ __align__(16) double sg[kNB];
    __shared__ int s_ok;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    double* B0 = sm;
    double* B1 = sm + kCqMat;
    double* B2 = sm + 2 * kCqMat;
    double* B3 = sm + 3 * kCqMat;
    double* B4 = sm + 4 * kCqMat;
    double* B5 = sm + 5 * kCqMat;
    double* B6 = sm + 6 * kCqMat;
    int* ci = cq_ints(d);
    const bool one_pass = PASS == 2 && ci[3] != 0;  // read before thread 0 can rewrite it
    if (!one_pass) cq_sum_partials(d, nch, B0);
    if (PASS == 2) {
        for (int e = t; e < kCqMat; e += kCqT) {
            B3[e] = d.cq[e];           // R1^-1
            B5[e] = d.cq[kCqMat + e];  // R1
            const int i = e / kNB, j = e % kNB;
            B6[e] = (i < nbp && j < nbp) ? d.a[(c0 + i) + (c0 + j) * d.lda] : 0.0;  // top rows of P
        }
    }
    __syncthreads();
    if (one_pass) {  // R2 = R2^-1 = I
        for (int e = t; e < kCqMat; e += kCqT) B1[e] = B2[e] = (e / kNB == e % kNB) ? 1.0 : 0.0;
        if (t == 0) s_ok = 1;
    } else if (w == 0) {
        bool far = false;
        if (PASS == 2) {
            const int j = l & 31;
            for (int i = 0; i < nbp; ++i)
                if (j < nbp) far = far || !(fabs(B0[i * kNB + j] - (i == j ? 1.0 : 0.0)) < kCqOrth);
        }
        const bool ok = cq_wave_chol(B0, B1, bc, nbp) && !__any(far);  // R1 or R2
        cq_wave_inv_upper(B1, B2);                                     // its inverse
        int one = 0;
        if (PASS == 1) {  // Gershgorin row sums of the column-normalized G1 (lane j: column j)
            const int j = l & 31;
            const double dj = j < nbp ? sqrt(B0[j * kNB + j]) : 1.0;
            cq_order();
            bc[j] = dj;
            cq_order();
            double rs = 0.0;
            for (int k = 0; k < nbp; ++k)
                if (k != j && j < nbp) rs += fabs(B0[k * kNB + j]) / (bc[k] * dj);
            for (int o = 16; o >= 1; o >>= 1) rs = fmax(rs, __shfl_xor(rs, o));
            one = rs < kCqOnePass ? 1 : 0;  // NaN (a failed Cholesky) gives 0
        }
        if (l == 0) {
            s_ok = ok;
            if (PASS == 1) ci[3] = one;
        }
    }
    __syncthreads();
    if (!s_ok) {
        if (t == 0) ci[2] = 1;
        return;
    }
    if (PASS == 1) {
        for (int e = t; e < kCqMat; e += kCqT) {
            d.cq[e] = B2[e];
            d.cq[kCqMat + e] = B1[e];
        }
        return;
    }
    cq_block_mm(B3, B2, B4);  // M = R1^-1 R2^-1 = Rc^-1
    cq_block_mm(B1, B5, B0);  // Rc = R2 R1
    __syncthreads();
    cq_block_mm(B6, B4, B2);  // Q top rows = P_top M
    __syncthreads();
    // S (sg) and V1 = Q_top S - I (identity-padded beyond nbp)
    if (t < kNB) sg[t] = B2[t * kNB + t] >= 0.0 ? -1.0 : 1.0;
    for (int e = t; e < kCqMat; e += kCqT) {
        const int i = e / kNB, j = e % kNB;
        const double sj = B2[j * kNB + j] >= 0.0 ? -1.0 : 1.0;
        B1[e] = (i < nbp && j < nbp) ? B2[e] * sj - (i == j ? 1.0 : 0.0) : (i == j ? 1.0 : 0.0);
    }
    __syncthreads();
    if (w == 0) cq_wave_lu(B1, B5, B3, bc);  // L1 -> B5, U -> B3
    __syncthreads();
    if (w == 0) cq_wave_inv_upper(B3, B1);       // U^-1
    if (w == 1) cq_wave_inv_unit_lower(B5, B2);  // L1^-1
    __syncthreads();
    // T = -U L1^-T into B6, S U^-1 into B3, then N = M (S U^-1) into B2
    // (B5 keeps L1 for the outputs)
    {
        const int i0 = 2 * (t / 16), j0 = 2 * (t % 16);
        double c00 = 0.0, c01 = 0.0, c10 = 0.0, c11 = 0.0;
#pragma unroll 8
        for (int k = 0; k < kNB; ++k) {
            const double a0 = B3[i0 * kNB + k], a1 = B3[(i0 + 1) * kNB + k];
            const double b0 = B2[j0 * kNB + k], b1 = B2[(j0 + 1) * kNB + k];
            c00 -= a0 * b0;
            c01 -= a0 * b1;
            c10 -= a1 * b0;
            c11 -= a1 * b1;
        }
        B6[i0 * kNB + j0] = c00;
        B6[i0 * kNB + j0 + 1] = c01;
        B6[(i0 + 1) * kNB + j0] = c10;
        B6[(i0 + 1) * kNB + j0 + 1] = c11;
    }
    __syncthreads();
    for (int e = t; e < kCqMat; e += kCqT) B3[e] = sg[e / kNB] * B1[e];  // S U^-1 (U no longer needed)
    __syncthreads();
    cq_block_mm(B4, B3, B2);  // N = M S U^-1 (L1^-1 no longer needed)
    __syncthreads();
    // outputs: T (column-major, zero beyond nbp), N (row-major), Y's top rows, the panel's R rows
    for (int e = t; e < kCqMat; e += kCqT) {
        const int i = e / kNB, j = e % kNB;
        const bool in = i < nbp && j < nbp;
        d.t[par * kCqMat + i + j * kNB] = in ? B6[e] : 0.0;
        d.cq[2 * kCqMat + e] = in ? B2[e] : 0.0;
        if (i < nbp)
            d.v[par * d.k * kNB + (c0 + i) + static_cast<int64_t>(j) * d.k] =
                j < nbp ? (i > j ? B5[e] : (i == j ? 1.0 : 0.0)) : 0.0;
        if (i < nbp && j < nbp && i <= j) {
            const double rv = sg[i] * B0[e];  // S Rc
            if (i == j) d.diag[c0 + i] = rv;
            else d.a[(c0 + i) + (c0 + j) * d.lda] = rv;
        }
    }
}

// Y's rows below the panel's top block: row i of P times N.  Grid
// (row blocks of kCqYT, 1, problems); one row per thread.
constexpr int kCqYT = 128;
__global__ __launch_bounds__(kCqYT) void qr_cq_form_y_kernel(const QrDesc* __restrict__ D, int64_t c0, int par) {
    const QrDesc& d = D[blockIdx.z];
    if (c0 >= d.n || cq_failed(d)) return;
    const int nbp = static_cast<int>(min<int64_t>(kNB, d.n - c0));
    if (c0 + nbp >= d.n) return;  // no trailing columns: Y is not used
    __shared__ __align__(16) double sn[kCqMat];
    const int64_t i = c0 + nbp + static_cast<int64_t>(blockIdx.x) * kCqYT + threadIdx.x;
    double p[kNB];  // the row is loaded first so that its loads and N's overlap
#pragma unroll
    for (int q = 0; q < kNB; ++q) p[q] = (i < d.k && q < nbp) ? d.a[i + (c0 + q) * d.lda] : 0.0;
    for (int e = threadIdx.x; e < kCqMat; e += kCqYT) sn[e] = d.cq[2 * kCqMat + e];
    __syncthreads();
    if (i >= d.k) return;
#pragma unroll
    for (int j = 0; j < kNB; ++j) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q <= j; ++q) acc += p[q] * sn[q * kNB + j];  // N upper triangular
        d.v[par * d.k * kNB + i + static_cast<int64_t>(j) * d.k] = acc;
    }
}

// Split partials of W = Y^T A(c0:k, trailing) for the CholeskyQR2 panels.
// Grid (32-column tiles from tile0, 16, problems): block (tile, split)
// forms the split's partial Y^T A over ~256 rows (64-row tiles of Y and A
// staged in LDS with two tiles in flight, a 4 x 4 block of W per lane, the
// four waves' sums added in LDS) and stores it to d.wp; qr_cq_w2_kernel
// adds them.
__global__ __launch_bounds__(kCqT) void qr_cq_w_kernel(const QrDesc* __restrict__ D, int64_t c0, int par, int tile0) {
    const QrDesc& d = D[blockIdx.z];
    if (c0 >= d.n || cq_failed(d)) return;
    const int nbp = static_cast<int>(min<int64_t>(kNB, d.n - c0));
    const int64_t j0 = c0 + nbp, nt = d.n - j0;
    const int tile = tile0 + static_cast<int>(blockIdx.x);
    const int64_t jt0 = static_cast<int64_t>(tile) * kNB;
    const double* yv = d.v + par * d.k * kNB;
    if (nt <= 0 || jt0 >= nt) return;
    const int64_t rows = d.k - c0;
    const int ns = cq_wsplits(rows);
    if (static_cast<int>(blockIdx.y) >= ns) return;
    const int64_t rps = (rows + ns - 1) / ns;
    const int64_t ra = c0 + static_cast<int64_t>(blockIdx.y) * rps, rb = min(d.k, ra + rps);
    const int ncols = static_cast<int>(min<int64_t>(kNB, nt - jt0));
    __shared__ __align__(16) double sya[2 * kCqTile * kCqLD];  // Y tile, A tile; then the per-wave W blocks
    static_assert(2 * kCqTile * kCqLD >= 4 * kCqMat, "the W blocks reuse the tile buffers");
    double* sy = sya;
    double* sa = sya + kCqTile * kCqLD;
    double* sred = sya;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int p0 = 4 * (l & 7), q0 = 4 * (l >> 3);  // W rows (Y columns) p0.., W columns q0..
    double acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = 0.0;
    constexpr int kPer = kCqTile * kNB / kCqT;  // tile elements per thread
    // two tiles in flight: while tile r is multiplied, tiles r+1 (buffer b ^ 1)
    // and r+2 are being loaded
    double ny[2][kPer], na[2][kPer];
    auto load_tile = [&](int64_t r, double (&by)[kPer], double (&ba)[kPer]) {
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = t + u * kCqT, rr = e & (kCqTile - 1), q = e / kCqTile;
            const int64_t i = r + rr;
            const bool in = i < rb;
            by[u] = (in && q < nbp) ? yv[i + static_cast<int64_t>(q) * d.k] : 0.0;
            ba[u] = (in && q < ncols) ? d.a[i + (j0 + jt0 + q) * d.lda] : 0.0;
        }
    };
    auto mult_tile = [&](const double (&by)[kPer], const double (&ba)[kPer]) {
        __syncthreads();  // the previous tile has been read
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = t + u * kCqT, o = (e & (kCqTile - 1)) * kCqLD + e / kCqTile;
            sy[o] = by[u];
            sa[o] = ba[u];
        }
        __syncthreads();
    };
    auto fma_tile = [&]() {
        for (int rr = 16 * w; rr < 16 * w + 16; ++rr) {
            double a4[4], b4[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                a4[x] = sy[rr * kCqLD + p0 + x];
                b4[x] = sa[rr * kCqLD + q0 + x];
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) acc[x][y] += a4[x] * b4[y];
        }
    };
    load_tile(ra, ny[0], na[0]);
    if (ra + kCqTile < rb) load_tile(ra + kCqTile, ny[1], na[1]);
    for (int64_t r = ra; r < rb; r += 2 * kCqTile) {
        mult_tile(ny[0], na[0]);
        if (r + 2 * kCqTile < rb) load_tile(r + 2 * kCqTile, ny[0], na[0]);
        fma_tile();
        if (r + kCqTile >= rb) break;
        mult_tile(ny[1], na[1]);
        if (r + 3 * kCqTile < rb) load_tile(r + 3 * kCqTile, ny[1], na[1]);
        fma_tile();
    }
    __syncthreads();  // the last tile has been read: the buffers take the W blocks
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) sred[w * kCqMat + (p0 + x) * kNB + q0 + y] = acc[x][y];
    __syncthreads();
    const int64_t ntiles = (nt + kNB - 1) / kNB;
    double* part = d.wp + (static_cast<int64_t>(blockIdx.y) * ntiles + tile) * kCqMat;
    for (int e = t; e < kCqMat; e += kCqT)
        part[e] = (sred[e] + sred[kCqMat + e]) + (sred[2 * kCqMat + e] + sred[3 * kCqMat + e]);
}

// W2 = T^T W for 32-column tiles tile0.. (grid (tiles, 1, problems)): W is
// the split partials of qr_cq_w_kernel added in split order (every load of
// a thread in flight at once), stored at w2[(jt0 + q) * kNB + p].  A launch
// of its own: a last-block ticket would need a device-scope release fence
// in every W block, which writes back each XCD's L2.
__global__ __launch_bounds__(kCqT) void qr_cq_w2_kernel(const QrDesc* __restrict__ D, int64_t c0, int par, int tile0) {
    const QrDesc& d = D[blockIdx.z];
    if (c0 >= d.n || cq_failed(d)) return;
    const int nbp = static_cast<int>(min<int64_t>(kNB, d.n - c0));
    const int64_t j0 = c0 + nbp, nt = d.n - j0;
    const int tile = tile0 + static_cast<int>(blockIdx.x);
    const int64_t jt0 = static_cast<int64_t>(tile) * kNB;
    if (nt <= 0 || jt0 >= nt) return;
    const int ns = cq_wsplits(d.k - c0);
    const int ncols = static_cast<int>(min<int64_t>(kNB, nt - jt0));
    const int64_t ntiles = (nt + kNB - 1) / kNB;
    __shared__ __align__(16) double sw[kCqMat];
    __shared__ __align__(16) double st[kCqMat];
    const int t = threadIdx.x;
    {
        double v[16][4];
#pragma unroll
        for (int sp = 0; sp < 16; ++sp)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                v[sp][q] = sp < ns ? d.wp[(static_cast<int64_t>(sp) * ntiles + tile) * kCqMat + t + q * kCqT] : 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double sum = 0.0;
#pragma unroll
            for (int sp = 0; sp < 16; ++sp) sum += v[sp][q];
            sw[t + q * kCqT] = sum;
            st[t + q * kCqT] = d.t[par * kCqMat + t + q * kCqT];  // T, column-major
        }
    }
    __syncthreads();
    // W2(p, q) = sum_{r <= p} T(r, p) W(r, q); a 2 x 2 block per thread (T is
    // zero below its diagonal)
    const int p0 = 2 * (t % 16), q0 = 2 * (t / 16);
    double a00 = 0.0, a01 = 0.0, a10 = 0.0, a11 = 0.0;
#pragma unroll 8
    for (int r = 0; r < kNB; ++r) {
        const double t0 = st[r + p0 * kNB], t1 = st[r + (p0 + 1) * kNB];
        const double w0 = sw[r * kNB + q0], w1 = sw[r * kNB + q0 + 1];
        a00 += t0 * w0;
        a01 += t0 * w1;
        a10 += t1 * w0;
        a11 += t1 * w1;
    }
    const bool p0in = p0 < nbp, p1in = p0 + 1 < nbp;
    if (q0 < ncols) {
        d.w2[(jt0 + q0) * kNB + p0] = p0in ? a00 : 0.0;
        d.w2[(jt0 + q0) * kNB + p0 + 1] = p1in ? a10 : 0.0;
    }
    if (q0 + 1 < ncols) {
        d.w2[(jt0 + q0 + 1) * kNB + p0] = p0in ? a01 : 0.0;
        d.w2[(jt0 + q0 + 1) * kNB + p0 + 1] = p1in ? a11 : 0.0;
    }
}

__global__ __launch_bounds__(kT) void extract_r_kernel(const QrDesc* __restrict__ D) {
    const QrDesc& d = D[blockIdx.z];
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (t >= d.n * d.n) return;
    const int64_t j = t / d.n, i = t - j * d.n;
    d.r[i + j * d.ldr] = (i < j) ? d.a[i + j * d.lda] : (i == j ? d.diag[j] : 0.0);
}

__global__ void reset_bad_kernel(const QrDesc* __restrict__ D, int nprob) {
    const int p = blockIdx.x * 64 + threadIdx.x;
    if (p < nprob) *D[p].bad = 0x7FFFFFFF;
}

// first zero diagonal of R per problem
__global__ __launch_bounds__(kT) void diag_zero_kernel(const QrDesc* __restrict__ D) {
    const QrDesc& d = D[blockIdx.z];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (i < d.n && d.r[i + i * d.ldr] == 0.0) atomicMin(d.bad, static_cast<int>(i));
}

// the inverse's input copy: wp (ld n) <- R (ld ldr)
__global__ __launch_bounds__(kT) void copy_r_kernel(const QrDesc* __restrict__ D) {
    const QrDesc& d = D[blockIdx.z];
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (t >= d.n * d.n) return;
    const int64_t j = t / d.n, i = t - j * d.n;
    d.wp[i + j * d.n] = d.r[i + j * d.ldr];
}

// R^-1 (d.r, ld ldr) over the leading n x n block of the problem's A (ld lda)
__global__ __launch_bounds__(kT) void copy_back_kernel(const QrDesc* __restrict__ D) {
    const QrDesc& d = D[blockIdx.z];
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (t >= d.n * d.n) return;
    const int64_t j = t / d.n, i = t - j * d.n;
    d.a[i + j * d.lda] = d.r[i + j * d.ldr];
}

// Column j of R^-1 (problem blockIdx.z) from the copy in wp into d.r:
// R x = e_j by back substitution in 32-row blocks.  For block I (bottom
// up) the block's right-hand side e_j - R(I, >I) x(>I) is a GEMV over the
// rows already solved (8 column strides x 32 rows, partial sums added in a
// fixed order), then wavefront 0 solves the 32 x 32 diagonal block, one
// lane per row, the solved entry broadcast by readlane.
constexpr int kTri = 256;  // threads of trtri_col_kernel: 8 column slots x 32 rows

__global__ __launch_bounds__(kTri) void trtri_col_kernel(const QrDesc* __restrict__ D) {
    extern __shared__ double b[];
    __shared__ double part[kTri / 32][32];
    const QrDesc& d = D[blockIdx.z];
    const int64_t n = d.n;
    const int64_t j = blockIdx.x;
    if (j >= n) return;
    const double* rc = d.wp;
    const int t = threadIdx.x, r = t & 31, cs = t >> 5;  // row in block, column stride slot
    for (int64_t I = j / 32; I >= 0; --I) {
        const int64_t r0 = I * 32, r1 = min<int64_t>(r0 + 32, j + 1);
        // partial sums of R(r0 + r, c) x(c) over solved columns c in [r1, j]
        double acc = 0.0;
        if (r0 + r < r1) {
            constexpr int S = kTri / 32;
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;  // four loads in flight
            int64_t c = r1 + cs;
            for (; c + 3 * S <= j; c += 4 * S) {
                a0 += rc[r0 + r + c * n] * b[c];
                a1 += rc[r0 + r + (c + S) * n] * b[c + S];
                a2 += rc[r0 + r + (c + 2 * S) * n] * b[c + 2 * S];
                a3 += rc[r0 + r + (c + 3 * S) * n] * b[c + 3 * S];
            }
            for (; c <= j; c += S) a0 += rc[r0 + r + c * n] * b[c];
            acc = (a0 + a1) + (a2 + a3);
        }
        part[cs][r] = acc;
        __syncthreads();
        if (t < 64) {
            // lane t < 32 holds row r0 + t of the diagonal block
            double rowv[32];
#pragma unroll
            for (int q = 0; q < 32; ++q)
                rowv[q] = (t < 32 && r0 + t < r1 && r0 + q < r1) ? rc[r0 + t + (r0 + q) * n] : 0.0;
            double rhs = 0.0;
            if (r0 + t < r1) {
                rhs = (r0 + t == j) ? 1.0 : 0.0;
#pragma unroll
                for (int q = 0; q < kTri / 32; ++q) rhs -= part[q][t];
            }
            // diagonal block: rows r0 .. r1-1, bottom up
#pragma unroll
            for (int lc = 31; lc >= 0; --lc) {
                if (r0 + lc < r1) {
                    const double mine = t == lc ? rhs / rowv[lc] : 0.0;
                    const double xc = read_lane(mine, lc);
                    if (t < lc) rhs -= rowv[lc] * xc;
                    if (t == lc) rhs = xc;
                }
            }
            if (r0 + t < r1 && t < 32) b[r0 + t] = rhs;
        }
        __syncthreads();
    }
    for (int64_t i = threadIdx.x; i < n; i += kTri) d.r[i + j * d.ldr] = (i <= j) ? b[i] : 0.0;
}

size_t a256(size_t v) { return (v + 255) / 256 * 256; }

// per-problem scratch: diag, tau, inv_v0 | V | Gram partials | T |
// W partials (also the inverse's n x n copy of R) | bad flag
struct Sizes {
    size_t vec3, v, gp, t, wp, w2, sync, tsqr, cq, bk, total;
};

Sizes sizes(int64_t k, int64_t n) {
    const int64_t nn = std::max<int64_t>(n, 1), kk = std::max<int64_t>(k, 1);
    const int64_t nch = (kk + kGR - 1) / kGR;
    Sizes z;
    z.vec3 = a256(static_cast<size_t>(3 * nn) * 8);
    z.v = a256(static_cast<size_t>(2 * kk * kNB) * 8);  // two sets (panel parity) for the look-ahead
    z.gp = a256(static_cast<size_t>(nch * kNB * kNB) * 8);
    z.t = a256(static_cast<size_t>(2 * kNB * kNB) * 8);
    z.wp = a256(static_cast<size_t>(std::max({qr_nsplit(kk) * nn * kNB, nn * nn,
                                              int64_t(16) * ((nn + kNB - 1) / kNB) * kCqMat})) * 8);
    z.w2 = a256(static_cast<size_t>(nn * kNB) * 8);
    z.sync = a256(static_cast<size_t>(kSyncHead + nn) * 4);
    z.tsqr = a256(static_cast<size_t>(2 * kTsqrSet) * 8);
    z.cq = a256(static_cast<size_t>(kCqDoubles) * 8 + 4 * sizeof(int));
    z.bk = a256(static_cast<size_t>(kk * nn) * 8);
    z.total = z.vec3 + z.v + z.gp + z.t + z.wp + z.w2 + z.sync + z.tsqr + z.cq + z.bk + 256;
    return z;
}

QrDesc carve(char*& p, const QrProblem& q) {
    const Sizes z = sizes(q.k, q.n);
    const int64_t nn = std::max<int64_t>(q.n, 1);
    QrDesc d;
    d.k = q.k;
    d.n = q.n;
    d.lda = q.lda;
    d.ldr = q.ldr;
    d.a = q.a;
    d.r = q.r;
    d.diag = reinterpret_cast<double*>(p);
    d.tau = d.diag + nn;
    d.inv_v0 = d.tau + nn;
    p += z.vec3;
    d.v = reinterpret_cast<double*>(p);
    p += z.v;
    d.gp = reinterpret_cast<double*>(p);
    p += z.gp;
    d.t = reinterpret_cast<double*>(p);
    p += z.t;
    d.wp = reinterpret_cast<double*>(p);
    p += z.wp;
    d.w2 = reinterpret_cast<double*>(p);
    p += z.w2;
    d.sync = reinterpret_cast<int*>(p);
    p += z.sync;
    d.tl = reinterpret_cast<double*>(p);
    d.vc = d.tl + kTsqrMaxLeaves * kNB * kNB;
    d.tc = d.vc + kTsqrComb * kNB;
    p += z.tsqr;
    d.cq = reinterpret_cast<double*>(p);
    p += z.cq;
    d.bk = reinterpret_cast<double*>(p);
    p += z.bk;
    d.bad = reinterpret_cast<int*>(p);
    p += 256;
    return d;
}

size_t status_bytes(int nprob) { return a256(static_cast<size_t>(2 * std::max(nprob, 1) + 1) * 4); }

size_t batch_bytes(int nprob, const QrProblem* probs) {
    size_t b = a256(static_cast<size_t>(std::max(nprob, 1)) * sizeof(QrDesc)) + status_bytes(nprob);
    for (int i = 0; i < nprob; ++i) b += sizes(probs[i].k, probs[i].n).total;
    return b;
}

// Descriptor table at the start of scratch, then the status words
// ([0] panel wait failure, [1 + i] problem i's first zero diagonal,
// [1 + nprob + i] problem i's CholeskyQR2 failure flag), then the
// problems' regions.
struct Batch {
    QrDesc* dev = nullptr;
    int* status = nullptr;
    int nprob = 0;
    int64_t kmax = 0, nmax = 0;
};

Batch setup(int nprob, const QrProblem* probs, void* scratch, size_t scratch_bytes, hipStream_t s) {
    const size_t need = batch_bytes(nprob, probs);
    if (!scratch) scratch = fallback_scratch(need);
    else require(scratch_bytes >= need, "geqrf: scratch too small");
    Batch B;
    B.nprob = nprob;
    char* base = static_cast<char*>(scratch);
    B.dev = reinterpret_cast<QrDesc*>(base);
    B.status = reinterpret_cast<int*>(base + a256(static_cast<size_t>(std::max(nprob, 1)) * sizeof(QrDesc)));
    char* p = base + a256(static_cast<size_t>(std::max(nprob, 1)) * sizeof(QrDesc)) + status_bytes(nprob);
    for (int i0 = 0; i0 < nprob; i0 += kQrDescChunk) {
        QrDescChunk c{};  // by value: no host buffer to keep alive
        const int cnt = std::min(kQrDescChunk, nprob - i0);
        for (int k = 0; k < cnt; ++k) {
            const QrProblem& q = probs[i0 + k];
            c.d[k] = carve(p, q);
            B.kmax = std::max(B.kmax, q.k);
            B.nmax = std::max(B.nmax, q.n);
        }
        qr_set_desc_kernel<<<1, 64, 0, s>>>(B.dev + i0, c, cnt);
    }
    TOMOS_LAUNCH_CHECK("qr_set_desc_kernel");
    qr_status_reset_kernel<<<1, 1, 0, s>>>(B.status);
    qr_cq_reset_kernel<<<static_cast<unsigned>(ceil_div(nprob, 64)), 64, 0, s>>>(B.dev, nprob);
    TOMOS_LAUNCH_CHECK("qr_status_reset_kernel");
    return B;
}

// Which panels factor(): 0 = CholeskyQR2 panels, the Householder panels for
// problems where they fail; 1 = the Householder panels only.
std::atomic<int> g_qr_panels{0};

// A (k x n, lda) <-> its saved copy, every problem of the batch.
template <bool TO_COPY>
void cq_copy(const Batch& B, hipStream_t s) {
    const int64_t blocks = std::min<int64_t>(ceil_div(B.kmax * B.nmax, kT), 2048);
    qr_cq_copy_kernel<TO_COPY><<<dim3(static_cast<unsigned>(blocks), 1, static_cast<unsigned>(B.nprob)), kT, 0, s>>>(
        B.dev);
    TOMOS_LAUNCH_CHECK("qr_cq_copy_kernel");
}

// CholeskyQR2 panels (see qr_cq_gram_kernel) with the blocked Householder
// trailing update, all problems in the same launches; A is saved first.
// Per panel p: two Gram passes (the second skipped when one suffices) and
// their kNB x kNB algebra, Y, then W = Y^T A, T^T W and A -= Y T^T W, in
// order on one stream.  (An overlap of each panel's far trailing update
// with the next panel's factorization on a second stream was measured
// slower at 4000 x 1000: the cross-stream events cost ~20 us per panel and
// the update's blocks slowed the latency-bound panel kernels.)  Y and T
// alternate between two sets by panel parity; the kernels take the set.
void factor_cq(const Batch& B, hipStream_t s) {
    const unsigned P = static_cast<unsigned>(B.nprob);
    cq_copy<true>(B, s);
    for (int64_t c0 = 0; c0 < B.nmax; c0 += kNB) {
        const int par = static_cast<int>(c0 / kNB) & 1;
        const int64_t rows = B.kmax - c0;
        const unsigned gx = static_cast<unsigned>(std::min<int64_t>(64, ceil_div(rows, 2 * kCqTile)));
        qr_cq_gram_kernel<1><<<dim3(gx, 1, P), kCqT, 0, s>>>(B.dev, c0);
        qr_cq_small_kernel<1><<<dim3(1, 1, P), kCqT, 0, s>>>(B.dev, c0, par);
        qr_cq_gram_kernel<2><<<dim3(gx, 1, P), kCqT, 0, s>>>(B.dev, c0);
        qr_cq_small_kernel<2><<<dim3(1, 1, P), kCqT, 0, s>>>(B.dev, c0, par);
        TOMOS_LAUNCH_CHECK("qr_cq_gram_kernel");
        if (c0 + kNB >= B.nmax) break;  // no problem has columns right of this panel
        const int64_t nt = B.nmax - c0 - kNB;
        const unsigned wtiles = static_cast<unsigned>(ceil_div(nt, kNB));
        qr_cq_form_y_kernel<<<dim3(static_cast<unsigned>(ceil_div(rows, kCqYT)), 1, P), kCqYT, 0, s>>>(B.dev, c0, par);
        qr_cq_w_kernel<<<dim3(wtiles, 16, P), kCqT, 0, s>>>(B.dev, c0, par, 0);
        qr_cq_w2_kernel<<<dim3(wtiles, 1, P), kCqT, 0, s>>>(B.dev, c0, par, 0);
        qr_update_kernel<<<dim3(static_cast<unsigned>(ceil_div(rows, 64)), static_cast<unsigned>(ceil_div(nt, 64)), P),
                           kT, 0, s>>>(B.dev, c0, par, 0);
        TOMOS_LAUNCH_CHECK("qr_cq trailing update");
    }
    extract_r_kernel<<<dim3(static_cast<unsigned>(ceil_div(B.nmax * B.nmax, kT)), 1, P), kT, 0, s>>>(B.dev);
    TOMOS_LAUNCH_CHECK("extract_r_kernel");
}


void factor(const Batch& B, hipStream_t s) {
    const unsigned P = static_cast<unsigned>(B.nprob);
    qr_sync_reset_kernel<<<dim3(static_cast<unsigned>(ceil_div(kSyncHead + B.nmax, kT)), 1, P), kT, 0, s>>>(
        B.dev, B.status);
    TOMOS_LAUNCH_CHECK("qr_sync_reset_kernel");
    // TSQR panels when every problem has at most 16 leaves of 256 or 512 rows;
    // at most 8 leaves (a 256-row combine) whenever 512-row leaves allow it
    const bool tsqr = B.kmax <= kTsqrMaxK;
    const int lr = B.kmax <= 8 * 256 ? 256 : 512;
    const bool comb256 = ceil_div(B.kmax, lr) <= 8;
    const bool lookahead = tsqr && P > 1;
    const SideStream& side = lookahead ? side_stream() : SideStream{};
    for (int64_t c0 = 0; c0 < B.nmax; c0 += kNB) {
        if (tsqr) {
            // Look-ahead (batches of several problems, where the updates are
            // large enough to repay the cross-stream waits): after panel p
            // is factored, only the next panel's columns get its update on
            // `s`; the rest of its trailing update runs on the side stream
            // while panel p+1 is factored.  Panel p waits for the side update
            // of p-2 (its own columns) and the next-panel update for the side
            // update of p-1.
            const int p = static_cast<int>(c0 / kNB), par = p & 1;
            const unsigned nleaves = static_cast<unsigned>(ceil_div(B.kmax - c0, lr));
            if (lookahead && p >= 2) TOMOS_HIP_CHECK(hipStreamWaitEvent(s, side.join[par], 0));
            if (lr == 256) qr_tsqr_leaf_kernel<256><<<dim3(nleaves, 1, P), 256, 0, s>>>(B.dev, c0, par);
            else qr_tsqr_leaf_kernel<512><<<dim3(nleaves, 1, P), 512, 0, s>>>(B.dev, c0, par);
            if (comb256) qr_tsqr_combine_kernel<256><<<dim3(1, 1, P), 256, 0, s>>>(B.dev, c0, lr, par);
            else qr_tsqr_combine_kernel<kTsqrComb><<<dim3(1, 1, P), kTsqrComb, 0, s>>>(B.dev, c0, lr, par);
            TOMOS_LAUNCH_CHECK("qr_tsqr panel");
            if (c0 + kNB >= B.nmax) break;
            const int ctiles = static_cast<int>(ceil_div(B.nmax - c0 - kNB, kCT));
            const int near = std::min(ctiles, kNB / kCT);  // the next panel's columns
            auto update = [&](hipStream_t st, int t0, int count) {
                const dim3 gl(nleaves, static_cast<unsigned>(count), P), gc(1, static_cast<unsigned>(count), P);
                if (lr == 256) qr_tsqr_update_kernel<256, false><<<gl, 256, 0, st>>>(B.dev, c0, lr, par, t0);
                else qr_tsqr_update_kernel<512, false><<<gl, 512, 0, st>>>(B.dev, c0, lr, par, t0);
                if (comb256) qr_tsqr_update_kernel<256, true><<<gc, 256, 0, st>>>(B.dev, c0, lr, par, t0);
                else qr_tsqr_update_kernel<kTsqrComb, true><<<gc, kTsqrComb, 0, st>>>(B.dev, c0, lr, par, t0);
                TOMOS_LAUNCH_CHECK("qr_tsqr update");
            };
            if (!lookahead) {
                update(s, 0, ctiles);
                continue;
            }
            if (ctiles > near) {
                TOMOS_HIP_CHECK(hipEventRecord(side.fork, s));
                TOMOS_HIP_CHECK(hipStreamWaitEvent(side.stream, side.fork, 0));
                update(side.stream, near, ctiles - near);
                TOMOS_HIP_CHECK(hipEventRecord(side.join[par], side.stream));
            }
            if (p >= 1) TOMOS_HIP_CHECK(hipStreamWaitEvent(s, side.join[par ^ 1], 0));
            update(s, 0, near);
            continue;
        }
        qr_panel_kernel<<<dim3(kNB, 1, P), kPT, 0, s>>>(B.dev, c0, c0 / kNB, B.status);
        TOMOS_LAUNCH_CHECK("qr_panel_kernel");
        if (c0 + kNB >= B.nmax) break;  // no problem has columns right of this panel
        const int64_t rows = B.kmax - c0;
        const int64_t nt = B.nmax - c0 - kNB;
        qr_form_v_kernel<<<dim3(static_cast<unsigned>(ceil_div(rows * kNB, kT)), 1, P), kT, 0, s>>>(B.dev, c0);
        qr_gram_kernel<<<dim3(static_cast<unsigned>(ceil_div(rows, kGR)), 1, P), kT, 0, s>>>(B.dev, c0);
        qr_form_t_kernel<<<dim3(1, 1, P), kTT, 0, s>>>(B.dev, c0);
        qr_w_kernel<<<dim3(static_cast<unsigned>(ceil_div(nt, kNB)), static_cast<unsigned>(qr_nsplit(B.kmax)), P),
                      kT, 0, s>>>(B.dev, c0);
        qr_w2_kernel<<<dim3(static_cast<unsigned>(ceil_div(nt, 64)), 1, P), kT, 0, s>>>(B.dev, c0);
        qr_update_kernel<<<dim3(static_cast<unsigned>(ceil_div(rows, 64)), static_cast<unsigned>(ceil_div(nt, 64)), P),
                           kT, 0, s>>>(B.dev, c0, 0, 0);
        TOMOS_LAUNCH_CHECK("qr trailing update");
    }
    if (lookahead) {  // every side-stream update is done before R is read
        TOMOS_HIP_CHECK(hipStreamWaitEvent(s, side.join[0], 0));
        TOMOS_HIP_CHECK(hipStreamWaitEvent(s, side.join[1], 0));
    }
    extract_r_kernel<<<dim3(static_cast<unsigned>(ceil_div(B.nmax * B.nmax, kT)), 1, P), kT, 0, s>>>(B.dev);
    TOMOS_LAUNCH_CHECK("extract_r_kernel");
}

// status[1 + i] = problem i's first zero diagonal (when with_bad), else
// INT_MAX; status[1 + nprob + i] its CholeskyQR2 failure flag; status[0]
// is the panel kernel's wait-failure flag.
__global__ void gather_status_kernel(const QrDesc* __restrict__ D, int nprob, bool with_bad, int* __restrict__ status) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < nprob) {
        status[1 + i] = (with_bad && D[i].n > 0) ? *D[i].bad : 0x7FFFFFFF;
        status[1 + nprob + i] = cq_ints(D[i])[2];
    }
}

// The status words on the host (one copy, one synchronisation); throws if
// a panel wait failed.
std::vector<int> read_status(const Batch& B, bool with_bad, hipStream_t s) {
    gather_status_kernel<<<static_cast<unsigned>(ceil_div(B.nprob, 64)), 64, 0, s>>>(B.dev, B.nprob, with_bad,
                                                                                    B.status);
    TOMOS_LAUNCH_CHECK("gather_status_kernel");
    std::vector<int> h(2 * static_cast<size_t>(B.nprob) + 1, 0);
    TOMOS_HIP_CHECK(hipMemcpyAsync(h.data(), B.status, h.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    TOMOS_HIP_CHECK(hipStreamSynchronize(s));
    if (h[0] != 0) throw Error(kErrInternal, "geqrf: panel kernel wait limit reached");
    return h;
}

// The status words after marking each R's first zero diagonal (with_bad).
std::vector<int> check_status(const Batch& B, bool with_bad, hipStream_t s) {
    if (with_bad) {
        reset_bad_kernel<<<static_cast<unsigned>(ceil_div(B.nprob, 64)), 64, 0, s>>>(B.dev, B.nprob);
        diag_zero_kernel<<<dim3(static_cast<unsigned>(ceil_div(B.nmax, kT)), 1, static_cast<unsigned>(B.nprob)), kT,
                           0, s>>>(B.dev);
        TOMOS_LAUNCH_CHECK("diag_zero_kernel");
    }
    return read_status(B, with_bad, s);
}

// Factor every problem of the batch and return the status words (one
// synchronisation when the CholeskyQR2 panels succeed).  If any problem's
// CholeskyQR2 panel failed, every A is restored from its copy and the
// batch is factored again with the Householder panels.
std::vector<int> factor_checked(const Batch& B, bool with_bad, hipStream_t s) {
    if (g_qr_panels.load(std::memory_order_relaxed) == 0) {
        factor_cq(B, s);
        std::vector<int> h = check_status(B, with_bad, s);
        bool failed = false;
        for (int i = 0; i < B.nprob; ++i) failed = failed || h[1 + B.nprob + i] != 0;
        if (!failed) return h;
        cq_copy<false>(B, s);
        qr_cq_reset_kernel<<<static_cast<unsigned>(ceil_div(B.nprob, 64)), 64, 0, s>>>(B.dev, B.nprob);
        TOMOS_LAUNCH_CHECK("qr_cq_reset_kernel");
    }
    factor(B, s);
    return check_status(B, with_bad, s);
}

// R -> R^-1 in place for every problem (st from check_status with
// with_bad); throws on a zero diagonal.
void invert(const Batch& B, const std::vector<int>& st, hipStream_t s) {
    const unsigned P = static_cast<unsigned>(B.nprob);
    require(B.nmax * 8 <= 128 * 1024, "trtri: n too large (max 16384)");
    for (int i = 0; i < B.nprob; ++i) {
        const int bad = st[1 + i];
        if (bad != 0x7FFFFFFF)
            throw Error(kErrSingular, "trtri: R is singular (zero diagonal at " + std::to_string(bad) +
                                          (B.nprob > 1 ? ", problem " + std::to_string(i) : std::string()) + ")");
    }
    copy_r_kernel<<<dim3(static_cast<unsigned>(ceil_div(B.nmax * B.nmax, kT)), 1, P), kT, 0, s>>>(B.dev);
    const size_t lds = static_cast<size_t>(std::max<int64_t>(B.nmax, 1)) * 8;
    TOMOS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(trtri_col_kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
    trtri_col_kernel<<<dim3(static_cast<unsigned>(B.nmax), 1, P), kTri, lds, s>>>(B.dev);
    TOMOS_LAUNCH_CHECK("trtri_col_kernel");
}

void check_problem(const QrProblem& q) {
    require(q.n >= 0 && q.k >= q.n, "geqrf: need k >= n (generate_sketch_qr_precond contract)");
    require(q.lda >= std::max<int64_t>(q.k, 1) && q.ldr >= std::max<int64_t>(q.n, 1), "geqrf: bad ld");
    require(q.n == 0 || (q.a && q.r), "geqrf: null pointer");
}

}  // namespace

size_t geqrf_scratch_bytes(int64_t k, int64_t n) {
    const QrProblem q{k, n, std::max<int64_t>(k, 1), std::max<int64_t>(n, 1), nullptr, nullptr};
    return batch_bytes(1, &q);
}

size_t geqrf_batched_scratch_bytes(int nprob, const QrProblem* probs) { return batch_bytes(nprob, probs); }

void geqrf_r_batched(int nprob, const QrProblem* probs, bool invert_r, void* scratch,
                     size_t scratch_bytes, hipStream_t s, bool copy_back) {
    require(nprob >= 0 && (nprob == 0 || probs), "geqrf: bad problem list");
    int64_t nmax = 0;
    for (int i = 0; i < nprob; ++i) {
        check_problem(probs[i]);
        nmax = std::max(nmax, probs[i].n);
    }
    if (nprob == 0 || nmax == 0) return;
    const Batch B = setup(nprob, probs, scratch, scratch_bytes, s);
    const std::vector<int> st = factor_checked(B, invert_r, s);
    if (invert_r) invert(B, st, s);
    if (copy_back) {
        copy_back_kernel<<<dim3(static_cast<unsigned>(ceil_div(nmax * nmax, kT)), 1, static_cast<unsigned>(nprob)),
                           kT, 0, s>>>(B.dev);
        TOMOS_LAUNCH_CHECK("copy_back_kernel");
    }
    TOMOS_HIP_CHECK(hipStreamSynchronize(s));
}

void geqrf_r(int64_t k, int64_t n, double* a, int64_t lda, double* r, int64_t ldr, void* scratch,
             size_t scratch_bytes, hipStream_t s) {
    const QrProblem q{k, n, lda, ldr, a, r};
    geqrf_r_batched(1, &q, false, scratch, scratch_bytes, s);
}

void trtri_upper(int64_t n, double* r, int64_t ldr, hipStream_t s) {
    require(n >= 0 && ldr >= std::max<int64_t>(n, 1), "trtri: bad sizes");
    if (n == 0) return;
    require(r != nullptr, "trtri: null R");
    require(n * 8 <= 128 * 1024, "trtri: n too large (max 16384)");
    const QrProblem q{n, n, n, ldr, nullptr, r};
    const Batch B = setup(1, &q, nullptr, 0, s);
    invert(B, check_status(B, true, s), s);
    TOMOS_HIP_CHECK(hipStreamSynchronize(s));
}

size_t sketch_precond_batched_bytes(int nprob, const QrProblem* probs) {
    size_t b = batch_bytes(nprob, probs);
    for (int i = 0; i < nprob; ++i)
        b += a256(static_cast<size_t>(std::max<int64_t>(probs[i].n, 1)) * probs[i].n * 8);
    return b;
}

void sketch_precond_batched(int nprob, QrProblem* probs, void* scratch, size_t scratch_bytes,
                            hipStream_t s) {
    // R^-1 of each sketch goes to its own n x n buffer, then overwrites the
    // sketch's leading block (as generate_sketch_qr_precond writes it back)
    require(scratch_bytes >= sketch_precond_batched_bytes(nprob, probs), "sketch_qr_precond: scratch too small");
    char* p = static_cast<char*>(scratch);
    for (int i = 0; i < nprob; ++i) {
        probs[i].ldr = std::max<int64_t>(probs[i].n, 1);
        probs[i].r = reinterpret_cast<double*>(p);
        p += a256(static_cast<size_t>(probs[i].ldr) * probs[i].n * 8);
    }
    const size_t rest = scratch_bytes - static_cast<size_t>(p - static_cast<char*>(scratch));
    geqrf_r_batched(nprob, probs, true, p, rest, s, true);
}

}  // namespace tomos_hip

using namespace tomos_hip;

extern "C" {

size_t tomos_geqrf_scratch_size(int64_t k, int64_t n) {
    if (k < 0 || n < 0) return 0;
    return geqrf_scratch_bytes(k, n);
}

int tomos_geqrf_r_f64(int64_t k, int64_t n, double* a, int64_t lda, double* r, int64_t ldr,
                      void* scratch, size_t scratch_bytes, tomos_stream_t s) {
    return guarded([&] { geqrf_r(k, n, a, lda, r, ldr, scratch, scratch_bytes, s); });
}

int tomos_trtri_upper_f64(int64_t n, double* r, int64_t ldr, tomos_stream_t s) {
    return guarded([&] { trtri_upper(n, r, ldr, s); });
}

int tomos_qr_panels_set(int mode) {
    return guarded([&] {
        require(mode == 0 || mode == 1, "tomos_qr_panels_set: mode must be 0 or 1");
        g_qr_panels.store(mode, std::memory_order_relaxed);
    });
}

int tomos_sketch_qr_precond_f64(int64_t k, int64_t n, double* a, int64_t lda, tomos_stream_t s) {
    const tomos_qr_problem p{k, n, lda, a};
    return tomos_sketch_qr_precond_batched_f64(1, &p, nullptr, 0, s);
}

size_t tomos_sketch_qr_precond_batched_scratch_size(int nprob, const tomos_qr_problem* probs) {
    size_t out = 0;
    const int rc = guarded([&] {
        require(nprob >= 0 && (nprob == 0 || probs), "sketch_qr_precond_batched: bad problem list");
        std::vector<QrProblem> q(nprob);
        for (int i = 0; i < nprob; ++i)
            q[i] = QrProblem{probs[i].k, probs[i].n, probs[i].lda, std::max<int64_t>(probs[i].n, 1),
                             probs[i].a, nullptr};
        out = sketch_precond_batched_bytes(nprob, q.data());
    });
    return rc == 0 ? out : 0;
}

int tomos_sketch_qr_precond_batched_f64(int nprob, const tomos_qr_problem* probs, void* scratch,
                                        size_t scratch_bytes, tomos_stream_t s) {
    return guarded([&] {
        require(nprob >= 0 && (nprob == 0 || probs), "sketch_qr_precond_batched: bad problem list");
        std::vector<QrProblem> q(nprob);
        for (int i = 0; i < nprob; ++i) {
            require(probs[i].n >= 0 && probs[i].k >= probs[i].n,
                    "sketch_qr_precond: the sketched matrix must have k >= n");
            q[i] = QrProblem{probs[i].k, probs[i].n, probs[i].lda, std::max<int64_t>(probs[i].n, 1),
                             probs[i].a, nullptr};
        }
        if (nprob == 0) return;
        const size_t need = sketch_precond_batched_bytes(nprob, q.data());
        if (scratch) {
            require(scratch_bytes >= need, "sketch_qr_precond: scratch too small");
            sketch_precond_batched(nprob, q.data(), scratch, scratch_bytes, s);
            return;
        }
        void* buf = nullptr;
        TOMOS_HIP_CHECK(hipMalloc(&buf, need));
        try {
            sketch_precond_batched(nprob, q.data(), buf, need, s);
        } catch (...) {
            (void)hipFree(buf);
            throw;
        }
        TOMOS_HIP_CHECK(hipFree(buf));
    });
}

}  // extern "C"
