// FP64 Householder QR of a tall k x n matrix (the CountSketch output) and
// the inverse of its triangular factor, for the sketch-QR preconditioner
// (src/core/precond/sketch_qr.cpp:38-64: geqrf, zero the reflectors,
// trtri('U','N')).
//
// QR: one launch per column c.  Every block recomputes the reflector of
// column c (norm of k-c doubles, L2-resident) and applies it to the
// trailing columns it owns, one column at a time: w = v^T A(:,j) as a block
// reduction, then A(:,j) -= tau*w*v.  Column c itself is never written
// during its own step (other blocks are reading it); R's diagonal goes to a
// separate array.  All reductions run in a fixed order (deterministic).
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

__global__ __launch_bounds__(kT) void qr_step_kernel(int64_t k, int64_t n, double* __restrict__ a,
                                                     int64_t lda, int64_t c,
                                                     double* __restrict__ diag) {
    __shared__ double red[4];
    const double* col = a + c * lda;
    const Reflector R = make_reflector(col, c, k, red);
    if (blockIdx.x == 0 && threadIdx.x == 0) diag[c] = R.active ? R.beta : col[c];
    if (!R.active) return;
    for (int64_t j = c + 1 + blockIdx.x; j < n; j += gridDim.x) {
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

size_t geqrf_scratch_bytes(int64_t k, int64_t n) {
    (void)k;
    return static_cast<size_t>(std::max<int64_t>(n, 1)) * 8;
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
    double* diag = static_cast<double*>(scratch);
    for (int64_t c = 0; c < n; ++c) {
        const int64_t trailing = n - c - 1;
        const unsigned blocks = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(trailing, 1024)));
        qr_step_kernel<<<blocks, kT, 0, s>>>(k, n, a, lda, c, diag);
        TOMOS_LAUNCH_CHECK("qr_step_kernel");
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
        TOMOS_HIP_CHECK(hipMalloc(&rbuf, static_cast<size_t>(n) * n * 8 + static_cast<size_t>(n) * 8));
        try {
            geqrf_r(k, n, a, lda, rbuf, n, rbuf + n * n, static_cast<size_t>(n) * 8, s);
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
