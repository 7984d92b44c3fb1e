// FP64 Householder QR of a tall k x n matrix (the CountSketch output) and
// the inverse of its triangular factor, for the sketch-QR preconditioner
// (src/core/precond/sketch_qr.cpp:38-64: geqrf, zero the reflectors,
// trtri('U','N')).
//
// QR, blocked by panels of NB = 32 columns (LAPACK dgeqrf's structure):
//  * panel: one launch per column c.  Every block recomputes the reflector
//    of column c (norm of k-c doubles, L2-resident) and applies it to the
//    remaining panel column it owns: w = v^T A(:,j) as a block reduction,
//    then A(:,j) -= tau*w*v.  Column c itself is never written during its
//    own step (other blocks are reading it); R's diagonal, tau and the
//    reflector scale go to separate arrays.
//  * V and T: the panel's reflectors are written as an explicit unit lower
//    trapezoidal V, and T (upper triangular, H_1...H_nb = I - V T V^T) is
//    formed from the Gram matrix V^T V (dlarft's recurrence).
//  * trailing update A <- (I - V T^T V^T) A over the columns right of the
//    panel, as two tiled FP64 products: W = V^T A (split over rows, partials
//    summed in a fixed order) and A -= V (T^T W).
// All reductions run in a fixed order (deterministic).
//
// Triangular inverse: one block per column j of R^-1 solves R x = e_j by
// column-oriented back substitution with x kept in LDS.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

constexpr int kT = 256;

__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// Reflector of column c: beta = -sign(x0)*||x||, v = [1, x(1:)/(x0-beta)],
// tau = (beta - x0)/beta (LAPACK dlarfg convention).  Returns false for a
// zero column (no reflection).
struct Reflector {
    double beta, tau, inv_v0;
    bool active;
};

__device__ Reflector make_reflector(const double* __restrict__ col, int64_t c, int64_t k,
                                    double* red) {
    double s = 0.0;
    for (int64_t i = c + 1 + threadIdx.x; i < k; i += kT) s += col[i] * col[i];
    const double sigma = block_sum(s, red);
    const double x0 = col[c];
    Reflector R;
    const double nrm = sqrt(x0 * x0 + sigma);
    if (nrm == 0.0) {
        R.active = false;
        R.beta = 0.0;
        R.tau = 0.0;
        R.inv_v0 = 0.0;
        return R;
    }
    R.active = true;
    R.beta = (x0 >= 0.0) ? -nrm : nrm;
    R.tau = (R.beta - x0) / R.beta;
    R.inv_v0 = 1.0 / (x0 - R.beta);
    return R;
}

constexpr int kNB = 32;

// Column c of the panel [c, c_end): reflector + its application to the
// panel columns c+1 .. c_end-1 (one block per column).
__global__ __launch_bounds__(kT) void qr_panel_step_kernel(int64_t k, double* __restrict__ a,
                                                           int64_t lda, int64_t c, int64_t c_end,
                                                           double* __restrict__ diag,
                                                           double* __restrict__ tau,
                                                           double* __restrict__ inv_v0) {
    __shared__ double red[4];
    const double* col = a + c * lda;
    const Reflector R = make_reflector(col, c, k, red);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        diag[c] = R.active ? R.beta : col[c];
        tau[c] = R.active ? R.tau : 0.0;
        inv_v0[c] = R.active ? R.inv_v0 : 0.0;
    }
    if (!R.active) return;
    for (int64_t j = c + 1 + blockIdx.x; j < c_end; j += gridDim.x) {
        double* cj = a + j * lda;
        // read A(c,j) before the reduction's barriers: thread 0 rewrites it below
        const double xc = cj[c];
        double part = 0.0;
        for (int64_t i = c + 1 + threadIdx.x; i < k; i += kT) part += (col[i] * R.inv_v0) * cj[i];
        const double w = R.tau * (xc + block_sum(part, red));
        for (int64_t i = c + 1 + threadIdx.x; i < k; i += kT) cj[i] -= w * (col[i] * R.inv_v0);
        if (threadIdx.x == 0) cj[c] -= w;
        __syncthreads();
    }
}

// V (rows c0..k-1 used, ld = k): unit lower trapezoidal reflector columns.
__global__ __launch_bounds__(kT) void qr_form_v_kernel(int64_t k, const double* __restrict__ a,
                                                       int64_t lda, int64_t c0, int nbp,
                                                       const double* __restrict__ inv_v0,
                                                       double* __restrict__ v) {
    const int64_t rows = k - c0;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (t >= rows * nbp) return;
    const int64_t jj = t / rows, i = c0 + (t - jj * rows);
    const int64_t c = c0 + jj;
    v[i + jj * k] = i < c ? 0.0 : (i == c ? 1.0 : a[i + c * lda] * inv_v0[c]);
}

constexpr int kGR = 128;  // rows per Gram chunk

// Partial Gram V^T V over a chunk of kGR rows: gp[chunk][p + q*kNB].
__global__ __launch_bounds__(kT) void qr_gram_kernel(int64_t k, int64_t c0, int nbp,
                                                     const double* __restrict__ v,
                                                     double* __restrict__ gp) {
    __shared__ double sv[kGR][kNB + 1];
    const int64_t r0 = c0 + static_cast<int64_t>(blockIdx.x) * kGR;
    for (int e = threadIdx.x; e < kGR * kNB; e += kT) {
        const int rr = e % kGR, q = e / kGR;
        const int64_t i = r0 + rr;
        sv[rr][q] = (i < k && q < nbp) ? v[i + q * k] : 0.0;
    }
    __syncthreads();
    const int rows = static_cast<int>(min<int64_t>(kGR, k - r0));
    for (int e = threadIdx.x; e < kNB * kNB; e += kT) {
        const int p = e % kNB, q = e / kNB;
        double acc = 0.0;
        if (p < nbp && q < nbp)
            for (int r = 0; r < rows; ++r) acc += sv[r][p] * sv[r][q];
        gp[static_cast<int64_t>(blockIdx.x) * kNB * kNB + e] = acc;
    }
}

// T (kNB x kNB, column-major) from the Gram partials and tau (dlarft,
// forward columnwise): T(0:i,i) = -tau_i T(0:i,0:i) G(0:i,i), T(i,i) = tau_i.
__global__ __launch_bounds__(64) void qr_form_t_kernel(int nchunks, int nbp,
                                                       const double* __restrict__ gp,
                                                       const double* __restrict__ tau,
                                                       double* __restrict__ t_out) {
    __shared__ double g[kNB * kNB];
    __shared__ double t[kNB * kNB];
    __shared__ double z[kNB];
    for (int e = threadIdx.x; e < kNB * kNB; e += 64) {
        double acc = 0.0;
        for (int ch = 0; ch < nchunks; ++ch) acc += gp[static_cast<int64_t>(ch) * kNB * kNB + e];
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
    for (int e = threadIdx.x; e < kNB * kNB; e += 64) t_out[e] = t[e];
}

// W partials: wp[split][jt][p] = sum over the split's rows of V(i,p) A(i, j0+jt).
// grid (ceil(nt/32), nsplit); 256 threads = 32 p x 8 groups of 4 columns.
__global__ __launch_bounds__(kT) void qr_w_kernel(int64_t k, const double* __restrict__ a,
                                                  int64_t lda, int64_t c0, int nbp, int64_t j0,
                                                  int64_t nt, const double* __restrict__ v,
                                                  int64_t rows_per_split,
                                                  double* __restrict__ wp) {
    __shared__ double sv[64][kNB + 1];
    __shared__ double sa[64][kNB + 1];
    const int p = threadIdx.x % kNB, qg = threadIdx.x / kNB;
    const int64_t jt0 = static_cast<int64_t>(blockIdx.x) * kNB;
    const int64_t ra = c0 + static_cast<int64_t>(blockIdx.y) * rows_per_split;
    const int64_t rb = min(k, ra + rows_per_split);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t r = ra; r < rb; r += 64) {
        for (int e = threadIdx.x; e < 64 * kNB; e += kT) {
            const int rr = e % 64, q = e / 64;
            const int64_t i = r + rr;
            sv[rr][q] = (i < rb && q < nbp) ? v[i + q * k] : 0.0;
            sa[rr][q] = (i < rb && jt0 + q < nt) ? a[i + (j0 + jt0 + q) * lda] : 0.0;
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
        if (jt < nt && p < nbp) wp[(static_cast<int64_t>(blockIdx.y) * nt + jt) * kNB + p] = acc[u];
    }
}

// A(c0:k, j0+jt) -= V (T^T W) on 64 x 64 tiles; every block forms its
// 32 x 64 slice of T^T W from the split partials (fixed order).
__global__ __launch_bounds__(kT) void qr_update_kernel(int64_t k, double* __restrict__ a,
                                                       int64_t lda, int64_t c0, int nbp,
                                                       int64_t j0, int64_t nt,
                                                       const double* __restrict__ v,
                                                       const double* __restrict__ t,
                                                       const double* __restrict__ wp, int nsplit) {
    __shared__ double sw[kNB][64 + 1];
    __shared__ double sw2[kNB][64 + 1];
    __shared__ double sv[64][kNB + 1];
    __shared__ double st[kNB * kNB];
    const int64_t r0 = c0 + static_cast<int64_t>(blockIdx.x) * 64;
    const int64_t jt0 = static_cast<int64_t>(blockIdx.y) * 64;
    for (int e = threadIdx.x; e < kNB * kNB; e += kT) st[e] = t[e];
    for (int e = threadIdx.x; e < kNB * 64; e += kT) {
        const int p = e % kNB, jj = e / kNB;
        const int64_t jt = jt0 + jj;
        double acc = 0.0;
        if (jt < nt && p < nbp)
            for (int sp = 0; sp < nsplit; ++sp) acc += wp[(static_cast<int64_t>(sp) * nt + jt) * kNB + p];
        sw[p][jj] = acc;
    }
    for (int e = threadIdx.x; e < 64 * kNB; e += kT) {
        const int rr = e % 64, q = e / 64;
        const int64_t i = r0 + rr;
        sv[rr][q] = (i < k && q < nbp) ? v[i + q * k] : 0.0;
    }
    __syncthreads();
    // W2 = T^T W: W2(p, j) = sum_{q <= p} T(q, p) W(q, j)
    for (int e = threadIdx.x; e < kNB * 64; e += kT) {
        const int p = e % kNB, jj = e / kNB;
        double acc = 0.0;
        for (int q = 0; q <= p; ++q) acc += st[q + p * kNB] * sw[q][jj];
        sw2[p][jj] = acc;
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
        double* colp = a + (j0 + jt) * lda;
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            const int64_t i = r0 + rg + 16 * x;
            if (i < k) colp[i] -= acc[x][y];
        }
    }
}

struct QrPlan {
    int64_t nchunks, nsplit, rows_per_split;
};

QrPlan qr_plan(int64_t k, int64_t n) {
    QrPlan p;
    p.nchunks = ceil_div(std::max<int64_t>(k, 1), kGR);
    p.nsplit = std::max<int64_t>(1, std::min<int64_t>(8, ceil_div(k, 512)));
    p.rows_per_split = ceil_div(std::max<int64_t>(k, 1), p.nsplit);
    (void)n;
    return p;
}

__global__ void extract_r_kernel(int64_t n, const double* __restrict__ a, int64_t lda,
                                 const double* __restrict__ diag, double* __restrict__ r,
                                 int64_t ldr) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (t >= n * n) return;
    const int64_t j = t / n, i = t - j * n;
    r[i + j * ldr] = (i < j) ? a[i + j * lda] : (i == j ? diag[j] : 0.0);
}

// x = R^-1 e_j, column j of the inverse, written into out (ldo).
__global__ __launch_bounds__(kT) void trtri_col_kernel(int64_t n, const double* __restrict__ r,
                                                       int64_t ldr, double* __restrict__ out,
                                                       int64_t ldo) {
    extern __shared__ double b[];
    const int64_t j = blockIdx.x;
    for (int64_t i = threadIdx.x; i <= j; i += kT) b[i] = (i == j) ? 1.0 : 0.0;
    __syncthreads();
    for (int64_t i = j; i >= 0; --i) {
        const double xi = b[i] / r[i + i * ldr];
        __syncthreads();
        for (int64_t q = threadIdx.x; q < i; q += kT) b[q] -= xi * r[q + i * ldr];
        if (threadIdx.x == 0) b[i] = xi;
        __syncthreads();
    }
    for (int64_t i = threadIdx.x; i < n; i += kT) out[i + j * ldo] = (i <= j) ? b[i] : 0.0;
}

__global__ void diag_zero_kernel(int64_t n, const double* __restrict__ r, int64_t ldr,
                                 int* __restrict__ first_zero) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (i < n && r[i + i * ldr] == 0.0) atomicMin(first_zero, static_cast<int>(i));
}

}  // namespace

// scratch: diag, tau, inv_v0 (n each) | V (k x NB) | Gram partials | T | W partials
size_t geqrf_scratch_bytes(int64_t k, int64_t n) {
    const QrPlan p = qr_plan(k, n);
    const int64_t nn = std::max<int64_t>(n, 1), kk = std::max<int64_t>(k, 1);
    const int64_t doubles = 3 * nn + kk * kNB + p.nchunks * kNB * kNB + kNB * kNB + p.nsplit * nn * kNB;
    return static_cast<size_t>(doubles) * 8;
}

void geqrf_r(int64_t k, int64_t n, double* a, int64_t lda, double* r, int64_t ldr, void* scratch,
             size_t scratch_bytes, hipStream_t s) {
    require(n >= 0 && k >= n, "geqrf: need k >= n (generate_sketch_qr_precond contract)");
    require(lda >= std::max<int64_t>(k, 1) && ldr >= std::max<int64_t>(n, 1), "geqrf: bad ld");
    if (n == 0) return;
    require(a && r, "geqrf: null pointer");
    const size_t need = geqrf_scratch_bytes(k, n);
    if (!scratch) scratch = fallback_scratch(need);
    else require(scratch_bytes >= need, "geqrf: scratch too small");
    const QrPlan P = qr_plan(k, n);
    double* diag = static_cast<double*>(scratch);
    double* tau = diag + n;
    double* inv_v0 = tau + n;
    double* v = inv_v0 + n;
    double* gp = v + k * kNB;
    double* tm = gp + P.nchunks * kNB * kNB;
    double* wp = tm + kNB * kNB;
    for (int64_t c0 = 0; c0 < n; c0 += kNB) {
        const int nbp = static_cast<int>(std::min<int64_t>(kNB, n - c0));
        const int64_t c_end = c0 + nbp;
        for (int64_t c = c0; c < c_end; ++c) {
            const unsigned blocks = static_cast<unsigned>(std::max<int64_t>(1, c_end - c - 1));
            qr_panel_step_kernel<<<blocks, kT, 0, s>>>(k, a, lda, c, c_end, diag, tau, inv_v0);
            TOMOS_LAUNCH_CHECK("qr_panel_step_kernel");
        }
        const int64_t nt = n - c_end;
        if (nt == 0) break;
        const int64_t rows = k - c0;
        qr_form_v_kernel<<<static_cast<unsigned>(ceil_div(rows * nbp, kT)), kT, 0, s>>>(
            k, a, lda, c0, nbp, inv_v0, v);
        const int64_t nch = ceil_div(rows, kGR);
        qr_gram_kernel<<<static_cast<unsigned>(nch), kT, 0, s>>>(k, c0, nbp, v, gp);
        qr_form_t_kernel<<<1, 64, 0, s>>>(static_cast<int>(nch), nbp, gp, tau + c0, tm);
        const int64_t rps = ceil_div(rows, P.nsplit);
        const int64_t nsp = ceil_div(rows, rps);
        qr_w_kernel<<<dim3(static_cast<unsigned>(ceil_div(nt, kNB)), static_cast<unsigned>(nsp)), kT, 0, s>>>(
            k, a, lda, c0, nbp, c_end, nt, v, rps, wp);
        qr_update_kernel<<<dim3(static_cast<unsigned>(ceil_div(rows, 64)), static_cast<unsigned>(ceil_div(nt, 64))), kT, 0, s>>>(
            k, a, lda, c0, nbp, c_end, nt, v, tm, wp, static_cast<int>(nsp));
        TOMOS_LAUNCH_CHECK("qr trailing update");
    }
    extract_r_kernel<<<static_cast<unsigned>(ceil_div(n * n, kT)), kT, 0, s>>>(n, a, lda, diag, r, ldr);
    TOMOS_LAUNCH_CHECK("extract_r_kernel");
    TOMOS_HIP_CHECK(hipStreamSynchronize(s));
}

void trtri_upper(int64_t n, double* r, int64_t ldr, hipStream_t s) {
    require(n >= 0 && ldr >= std::max<int64_t>(n, 1), "trtri: bad sizes");
    if (n == 0) return;
    require(r != nullptr, "trtri: null R");
    require(n * 8 <= 128 * 1024, "trtri: n too large (max 16384)");
    const size_t copy_bytes = static_cast<size_t>(n) * n * 8;
    char* base = static_cast<char*>(fallback_scratch(copy_bytes + 256));
    double* rc = reinterpret_cast<double*>(base);
    int* flag = reinterpret_cast<int*>(base + copy_bytes);
    TOMOS_HIP_CHECK(hipMemcpy2DAsync(rc, n * 8, r, ldr * 8, n * 8, n, hipMemcpyDeviceToDevice, s));
    const int big = 0x7FFFFFFF;
    TOMOS_HIP_CHECK(hipMemcpyAsync(flag, &big, sizeof(int), hipMemcpyHostToDevice, s));
    diag_zero_kernel<<<static_cast<unsigned>(ceil_div(n, kT)), kT, 0, s>>>(n, rc, n, flag);
    TOMOS_LAUNCH_CHECK("diag_zero_kernel");
    int first = big;
    TOMOS_HIP_CHECK(hipMemcpyAsync(&first, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    TOMOS_HIP_CHECK(hipStreamSynchronize(s));
    if (first != big)
        throw Error(kErrSingular, "trtri: R is singular (zero diagonal at " + std::to_string(first) + ")");
    const size_t lds = static_cast<size_t>(n) * 8;
    TOMOS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(trtri_col_kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
    trtri_col_kernel<<<static_cast<unsigned>(n), kT, lds, s>>>(n, rc, n, r, ldr);
    TOMOS_LAUNCH_CHECK("trtri_col_kernel");
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
    return guarded([&] {
        require(n >= 0 && k >= n, "sketch_qr_precond: sketched matrix must have k >= n");
        if (n == 0) return;
        // R goes to a separate buffer, then overwrites the leading block.
        double* rbuf = nullptr;
        const size_t qs = geqrf_scratch_bytes(k, n);
        TOMOS_HIP_CHECK(hipMalloc(&rbuf, static_cast<size_t>(n) * n * 8 + qs));
        try {
            geqrf_r(k, n, a, lda, rbuf, n, rbuf + n * n, qs, s);
            trtri_upper(n, rbuf, n, s);
            TOMOS_HIP_CHECK(hipMemcpy2DAsync(a, lda * 8, rbuf, n * 8, n * 8, n,
                                             hipMemcpyDeviceToDevice, s));
            TOMOS_HIP_CHECK(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipFree(rbuf);
            throw;
        }
        TOMOS_HIP_CHECK(hipFree(rbuf));
    });
}

}  // extern "C"
