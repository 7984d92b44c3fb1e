// FP64 Householder QR of tall k x n matrices (CountSketch outputs) and the
// inverse of the triangular factor, for the sketch-QR preconditioner
// (src/core/precond/sketch_qr.cpp:38-64: geqrf, zero the reflectors,
// trtri('U','N')).  One problem or many in the same launches: every kernel
// takes a device table of problem descriptors and picks its problem with
// blockIdx.z, so a batch of uneven problems costs the launches of the
// largest one.
//
// QR, blocked by panels of NB = 32 columns (LAPACK dgeqrf's structure):
//  * panel: one launch per column c.  Every block recomputes the reflector
//    of column c (LAPACK dlarfg convention: beta = -sign(x0)|x|,
//    v = [1, x(1:)/(x0-beta)], tau = (beta-x0)/beta; a zero column gives no
//    reflection) and applies it to the remaining panel column it owns: one
//    pass computes |x|^2 and x^T A(:,j) together, a second does
//    A(:,j) -= tau*(v^T A(:,j))*v.  Column c itself is never written during
//    its own step (other blocks are reading it); R's diagonal, tau and the
//    reflector scale go to separate arrays.
//  * V and T: the panel's reflectors are written as an explicit unit lower
//    trapezoidal V, and T (upper triangular, H_1...H_nb = I - V T V^T) is
//    formed from the Gram matrix V^T V (dlarft's recurrence).
//  * trailing update A <- (I - V T^T V^T) A over the columns right of the
//    panel, as two tiled FP64 products: W = V^T A (split over rows, partials
//    summed in a fixed order, then W2 = T^T W once) and A -= V W2.
// All reductions run in a fixed order (deterministic).
//
// Triangular inverse: one block per (column j, problem) solves R x = e_j by
// column-oriented back substitution with x kept in LDS.
#include <algorithm>
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
    double* v;       // k x kNB (ld k)
    double* gp;      // Gram partials: ceil(k/kGR) x kNB^2
    double* t;       // kNB^2
    double* wp;      // W partials: nsplit x n x kNB; the inverse's copy of R (n x n)
    double* w2;      // T^T W: n x kNB (column jt at w2 + jt*kNB)
    int* bad;        // trtri: first zero diagonal (INT_MAX if none)
};

__host__ __device__ inline int64_t qr_nsplit(int64_t k) {
    const int64_t s = (k + 511) / 512;
    return s < 1 ? 1 : (s > 8 ? 8 : s);
}

__global__ void qr_set_desc_kernel(QrDesc* d, QrDesc v) { *d = v; }

constexpr int kPT = 1024;  // panel-step block

// Column c of the panel starting at c0: block b applies the reflector of
// column c to panel column j = c + 1 + b (block 0 also records diag, tau
// and the reflector scale).  One pass over the rows computes both
// |x(c+1:)|^2 and x(c+1:)^T A(c+1:, j) (v^T A(:,j) = A(c,j) + inv_v0 * that
// dot, since v = [1, x(c+1:) * inv_v0]), a second applies the update.
__global__ __launch_bounds__(kPT) void qr_panel_step_kernel(const QrDesc* __restrict__ D, int64_t c0,
                                                            int64_t c) {
    const QrDesc& d = D[blockIdx.z];
    if (c >= d.n) return;
    const int64_t c_end = min(c0 + kNB, d.n);
    const int64_t j = c + 1 + blockIdx.x;
    const bool has_col = j < c_end;
    if (blockIdx.x > 0 && !has_col) return;
    __shared__ double red[2][kPT / 64];
    const double* col = d.a + c * d.lda;
    double* cj = d.a + (has_col ? j : c) * d.lda;
    const double x0 = col[c];
    const double xc = has_col ? cj[c] : 0.0;  // read before the barriers: thread 0 rewrites it
    double ss = 0.0, dot = 0.0;
    for (int64_t i = c + 1 + threadIdx.x; i < d.k; i += kPT) {
        const double xi = col[i];
        ss += xi * xi;
        if (has_col) dot += xi * cj[i];
    }
    for (int o = 32; o >= 1; o >>= 1) {
        ss += __shfl_xor(ss, o);
        dot += __shfl_xor(dot, o);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wave] = ss;
        red[1][wave] = dot;
    }
    __syncthreads();
    double sigma = 0.0, vdot = 0.0;
    for (int w = 0; w < kPT / 64; ++w) {  // same order in every thread
        sigma += red[0][w];
        vdot += red[1][w];
    }
    const double nrm = sqrt(x0 * x0 + sigma);
    const bool active = nrm != 0.0;
    const double beta = x0 >= 0.0 ? -nrm : nrm;
    const double tau = active ? (beta - x0) / beta : 0.0;
    const double inv_v0 = active ? 1.0 / (x0 - beta) : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        d.diag[c] = active ? beta : x0;
        d.tau[c] = tau;
        d.inv_v0[c] = inv_v0;
    }
    if (!active || !has_col) return;
    const double w = tau * (xc + inv_v0 * vdot);
    const double wv = w * inv_v0;
    for (int64_t i = c + 1 + threadIdx.x; i < d.k; i += kPT) cj[i] -= wv * col[i];
    if (threadIdx.x == 0) cj[c] = xc - w;
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
        double acc = 0.0;
        for (int ch = 0; ch < nchunks; ++ch) acc += d.gp[static_cast<int64_t>(ch) * kNB * kNB + e];
        g[e] = acc;
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
        double acc = 0.0;
        if (jt < nt && p < nbp)
            for (int sp = 0; sp < nsplit; ++sp) acc += d.wp[(static_cast<int64_t>(sp) * nt + jt) * kNB + p];
        sw[p][jj] = acc;
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

// A(c0:k, c_end+jt) -= V W2 on 64 x 64 tiles.
__global__ __launch_bounds__(kT) void qr_update_kernel(const QrDesc* __restrict__ D, int64_t c0) {
    const QrDesc& d = D[blockIdx.z];
    const int64_t nbp64 = min<int64_t>(kNB, d.n - c0);
    if (nbp64 <= 0) return;
    const int nbp = static_cast<int>(nbp64);
    const int64_t j0 = c0 + nbp, nt = d.n - j0;
    const int64_t r0 = c0 + static_cast<int64_t>(blockIdx.x) * 64;
    const int64_t jt0 = static_cast<int64_t>(blockIdx.y) * 64;
    if (nt <= 0 || r0 >= d.k || jt0 >= nt) return;
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
        sv[rr][q] = (i < d.k && q < nbp) ? d.v[i + q * d.k] : 0.0;
    }
    __syncthreads();
    // 256 threads: 16 row groups x 16 column groups, 4 x 4 outputs each
    const int rg = threadIdx.x % 16, cgp = threadIdx.x / 16;
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
            if (i < d.k) colp[i] -= acc[x][y];
        }
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

// Column j of R^-1 (problem blockIdx.z) from the copy in wp into d.r.
__global__ __launch_bounds__(kT) void trtri_col_kernel(const QrDesc* __restrict__ D) {
    extern __shared__ double b[];
    const QrDesc& d = D[blockIdx.z];
    const int64_t n = d.n;
    const int64_t j = blockIdx.x;
    if (j >= n) return;
    const double* rc = d.wp;
    for (int64_t i = threadIdx.x; i <= j; i += kT) b[i] = (i == j) ? 1.0 : 0.0;
    __syncthreads();
    for (int64_t i = j; i >= 0; --i) {
        const double xi = b[i] / rc[i + i * n];
        __syncthreads();
        for (int64_t q = threadIdx.x; q < i; q += kT) b[q] -= xi * rc[q + i * n];
        if (threadIdx.x == 0) b[i] = xi;
        __syncthreads();
    }
    for (int64_t i = threadIdx.x; i < n; i += kT) d.r[i + j * d.ldr] = (i <= j) ? b[i] : 0.0;
}

size_t a256(size_t v) { return (v + 255) / 256 * 256; }

// per-problem scratch: diag, tau, inv_v0 | V | Gram partials | T |
// W partials (also the inverse's n x n copy of R) | bad flag
struct Sizes {
    size_t vec3, v, gp, t, wp, w2, total;
};

Sizes sizes(int64_t k, int64_t n) {
    const int64_t nn = std::max<int64_t>(n, 1), kk = std::max<int64_t>(k, 1);
    const int64_t nch = (kk + kGR - 1) / kGR;
    Sizes z;
    z.vec3 = a256(static_cast<size_t>(3 * nn) * 8);
    z.v = a256(static_cast<size_t>(kk * kNB) * 8);
    z.gp = a256(static_cast<size_t>(nch * kNB * kNB) * 8);
    z.t = a256(static_cast<size_t>(kNB * kNB) * 8);
    z.wp = a256(static_cast<size_t>(std::max(qr_nsplit(kk) * nn * kNB, nn * nn)) * 8);
    z.w2 = a256(static_cast<size_t>(nn * kNB) * 8);
    z.total = z.vec3 + z.v + z.gp + z.t + z.wp + z.w2 + 256;
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
    d.bad = reinterpret_cast<int*>(p);
    p += 256;
    return d;
}

size_t batch_bytes(int nprob, const QrProblem* probs) {
    size_t b = a256(static_cast<size_t>(std::max(nprob, 1)) * sizeof(QrDesc));
    for (int i = 0; i < nprob; ++i) b += sizes(probs[i].k, probs[i].n).total;
    return b;
}

// Descriptor table at the start of scratch, then the problems' regions.
struct Batch {
    QrDesc* dev = nullptr;
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
    char* p = base + a256(static_cast<size_t>(std::max(nprob, 1)) * sizeof(QrDesc));
    for (int i = 0; i < nprob; ++i) {
        const QrDesc d = carve(p, probs[i]);
        qr_set_desc_kernel<<<1, 1, 0, s>>>(B.dev + i, d);  // by value: no host buffer to keep alive
        B.kmax = std::max(B.kmax, probs[i].k);
        B.nmax = std::max(B.nmax, probs[i].n);
    }
    TOMOS_LAUNCH_CHECK("qr_set_desc_kernel");
    return B;
}

void factor(const Batch& B, hipStream_t s) {
    const unsigned P = static_cast<unsigned>(B.nprob);
    for (int64_t c0 = 0; c0 < B.nmax; c0 += kNB) {
        const int64_t c_end = std::min<int64_t>(c0 + kNB, B.nmax);
        for (int64_t c = c0; c < c_end; ++c) {
            const unsigned blocks = static_cast<unsigned>(std::max<int64_t>(1, c_end - c - 1));
            qr_panel_step_kernel<<<dim3(blocks, 1, P), kPT, 0, s>>>(B.dev, c0, c);
        }
        TOMOS_LAUNCH_CHECK("qr_panel_step_kernel");
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
                           kT, 0, s>>>(B.dev, c0);
        TOMOS_LAUNCH_CHECK("qr trailing update");
    }
    extract_r_kernel<<<dim3(static_cast<unsigned>(ceil_div(B.nmax * B.nmax, kT)), 1, P), kT, 0, s>>>(B.dev);
    TOMOS_LAUNCH_CHECK("extract_r_kernel");
}

// R -> R^-1 in place for every problem; throws on a zero diagonal.
void invert(const Batch& B, hipStream_t s) {
    const unsigned P = static_cast<unsigned>(B.nprob);
    require(B.nmax * 8 <= 128 * 1024, "trtri: n too large (max 16384)");
    reset_bad_kernel<<<static_cast<unsigned>(ceil_div(B.nprob, 64)), 64, 0, s>>>(B.dev, B.nprob);
    diag_zero_kernel<<<dim3(static_cast<unsigned>(ceil_div(B.nmax, kT)), 1, P), kT, 0, s>>>(B.dev);
    TOMOS_LAUNCH_CHECK("diag_zero_kernel");
    std::vector<QrDesc> h(B.nprob);
    TOMOS_HIP_CHECK(hipMemcpyAsync(h.data(), B.dev, sizeof(QrDesc) * B.nprob, hipMemcpyDeviceToHost, s));
    TOMOS_HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < B.nprob; ++i) {
        if (h[i].n == 0) continue;
        int bad = 0x7FFFFFFF;
        TOMOS_HIP_CHECK(hipMemcpyAsync(&bad, h[i].bad, sizeof(int), hipMemcpyDeviceToHost, s));
        TOMOS_HIP_CHECK(hipStreamSynchronize(s));
        if (bad != 0x7FFFFFFF)
            throw Error(kErrSingular, "trtri: R is singular (zero diagonal at " + std::to_string(bad) +
                                          (B.nprob > 1 ? ", problem " + std::to_string(i) : std::string()) + ")");
    }
    copy_r_kernel<<<dim3(static_cast<unsigned>(ceil_div(B.nmax * B.nmax, kT)), 1, P), kT, 0, s>>>(B.dev);
    const size_t lds = static_cast<size_t>(std::max<int64_t>(B.nmax, 1)) * 8;
    TOMOS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(trtri_col_kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
    trtri_col_kernel<<<dim3(static_cast<unsigned>(B.nmax), 1, P), kT, lds, s>>>(B.dev);
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
                     size_t scratch_bytes, hipStream_t s) {
    require(nprob >= 0 && (nprob == 0 || probs), "geqrf: bad problem list");
    int64_t nmax = 0;
    for (int i = 0; i < nprob; ++i) {
        check_problem(probs[i]);
        nmax = std::max(nmax, probs[i].n);
    }
    if (nprob == 0 || nmax == 0) return;
    const Batch B = setup(nprob, probs, scratch, scratch_bytes, s);
    factor(B, s);
    if (invert_r) invert(B, s);
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
    invert(B, s);
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
    geqrf_r_batched(nprob, probs, true, p, rest, s);
    for (int i = 0; i < nprob; ++i)
        if (probs[i].n > 0)
            TOMOS_HIP_CHECK(hipMemcpy2DAsync(probs[i].a, probs[i].lda * 8, probs[i].r, probs[i].ldr * 8,
                                             probs[i].n * 8, probs[i].n, hipMemcpyDeviceToDevice, s));
    TOMOS_HIP_CHECK(hipStreamSynchronize(s));
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
