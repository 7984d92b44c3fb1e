// CSR sparse matrix-vector products (the reference's mv_csr slots,
// src/dispatch/sparse_registries.cpp:225-243; CPU semantics
// src/kernels/cpu/matrix/matrix_cpu_kernels.cpp:92-131) and the CSR
// transpose that lets LSQR form A^T u without atomics.
//
// A*x: L lanes of a wavefront share one row (L in {1, 4, 16, 64}, from the
// mean row length), each lane sums a lane-strided part of the row in FP64
// and the group folds its partial sums with a butterfly of lane shuffles.
// The order of every sum depends only on (m, nnz), so results repeat bit
// for bit.  Bound: HBM, nnz*(value + index bytes) + (m+1) index bytes + the
// y traffic, with x gathered through the caches.
//
// A^T*x: the reference scatters into y (matrix_cpu_kernels.cpp:116-130);
// here A^T is built once as a CSR matrix of its own (a stable radix sort of
// the column indices keeps A's row order inside each row of A^T) and the
// A*x kernel runs on it: no atomics, deterministic, coalesced.
#include <type_traits>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "codec.hpp"
#include "common.hpp"
#include "csr.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

constexpr int kCsrThreads = 256;
constexpr int64_t kCsrMaxBlocks = 8192;  // also bounds the sum-of-squares partials

template <int VT>
__device__ __forceinline__ double load_val(const void* v, int64_t k) {
    if constexpr (VT == TOMOS_F64) return static_cast<const double*>(v)[k];
    else if constexpr (VT == TOMOS_F32) return static_cast<const float*>(v)[k];
    else if constexpr (VT == TOMOS_F16) return codec::half_bits_to_f32(static_cast<const uint16_t*>(v)[k]);
    else return codec::u2f(static_cast<uint32_t>(static_cast<const uint16_t*>(v)[k]) << 16);
}

template <class I>
__device__ __forceinline__ int64_t clamp_ptr(I p, int64_t nnz) {
    const int64_t q = static_cast<int64_t>(p);
    return q < 0 ? 0 : (q > nnz ? nnz : q);
}

struct MvArgs {
    int64_t m, n, nnz;
    const void* vals;
    const void* rp;
    const void* ci;
    const void* x;
    void* y;
    bool x64, y64;
    bool xlds;  // x staged in LDS as FP64 (n <= kXLdsMax)
    Coef c;
};

// x of at most this many entries is copied to LDS by every block: gathers
// from LDS instead of cache lines
constexpr int64_t kXLdsMax = 8192;

extern __shared__ double csr_xs[];

__device__ __forceinline__ double load_x(const MvArgs& a, uint64_t col) {
    if (a.xlds) return csr_xs[col];
    return a.x64 ? static_cast<const double*>(a.x)[col] : static_cast<double>(static_cast<const float*>(a.x)[col]);
}

template <int VT, class I, int L>
__global__ __launch_bounds__(kCsrThreads) void csr_mv_kernel(MvArgs a) {
    if (a.c.skip && *a.c.skip) return;
    if (a.xlds) {
        for (int64_t j = threadIdx.x; j < a.n; j += kCsrThreads)
            csr_xs[j] = a.x64 ? static_cast<const double*>(a.x)[j]
                              : static_cast<double>(static_cast<const float*>(a.x)[j]);
        __syncthreads();
    }
    const double alpha = a.c.dalpha ? *a.c.dalpha * a.c.alpha_mul : a.c.alpha;
    const double beta = a.c.dbeta ? *a.c.dbeta * a.c.beta_mul : a.c.beta;
    const I* rp = static_cast<const I*>(a.rp);
    const I* ci = static_cast<const I*>(a.ci);
    const int lane = threadIdx.x % L;
    const int64_t groups = static_cast<int64_t>(gridDim.x) * (kCsrThreads / L);
    double sq = 0.0;
    for (int64_t row = static_cast<int64_t>(blockIdx.x) * (kCsrThreads / L) + threadIdx.x / L; row < a.m;
         row += groups) {
        const int64_t k0 = clamp_ptr(rp[row], a.nnz);
        int64_t k1 = clamp_ptr(rp[row + 1], a.nnz);
        k1 = k1 < k0 ? k0 : k1;
        // four independent chains keep four entries' loads in flight per lane
        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
        int64_t k = k0 + lane;
        for (; k + 3 * L < k1; k += 4 * L) {
            const uint64_t c0 = static_cast<uint64_t>(static_cast<int64_t>(ci[k]));
            const uint64_t c1 = static_cast<uint64_t>(static_cast<int64_t>(ci[k + L]));
            const uint64_t c2 = static_cast<uint64_t>(static_cast<int64_t>(ci[k + 2 * L]));
            const uint64_t c3 = static_cast<uint64_t>(static_cast<int64_t>(ci[k + 3 * L]));
            const double v0 = load_val<VT>(a.vals, k), v1 = load_val<VT>(a.vals, k + L);
            const double v2 = load_val<VT>(a.vals, k + 2 * L), v3 = load_val<VT>(a.vals, k + 3 * L);
            if (c0 < static_cast<uint64_t>(a.n)) acc0 = fma(v0, load_x(a, c0), acc0);
            if (c1 < static_cast<uint64_t>(a.n)) acc1 = fma(v1, load_x(a, c1), acc1);
            if (c2 < static_cast<uint64_t>(a.n)) acc2 = fma(v2, load_x(a, c2), acc2);
            if (c3 < static_cast<uint64_t>(a.n)) acc3 = fma(v3, load_x(a, c3), acc3);
        }
        for (; k < k1; k += L) {
            const uint64_t c0 = static_cast<uint64_t>(static_cast<int64_t>(ci[k]));
            if (c0 < static_cast<uint64_t>(a.n)) acc0 = fma(load_val<VT>(a.vals, k), load_x(a, c0), acc0);
        }
        double acc = (acc0 + acc1) + (acc2 + acc3);
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, L);
        if (lane == 0) {
            double out = alpha * acc;
            if (beta != 0.0)
                out += beta * (a.y64 ? static_cast<double*>(a.y)[row]
                                     : static_cast<double>(static_cast<float*>(a.y)[row]));
            if (a.y64) {
                static_cast<double*>(a.y)[row] = out;
                sq += out * out;
            } else {
                const float o32 = static_cast<float>(out);
                static_cast<float*>(a.y)[row] = o32;
                sq += static_cast<double>(o32) * o32;
            }
        }
    }
    if (a.c.sumsq) {
        __shared__ double part[kCsrThreads / 64];
        for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
        if (threadIdx.x % 64 == 0) part[threadIdx.x / 64] = sq;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < kCsrThreads / 64; ++w) t += part[w];
            a.c.sumsq[blockIdx.x] = t;
        }
    }
}

// ---- long rows: each row split into S segments ------------------------------
// A matrix of few, long rows (A^T of a tall LSQR matrix: n rows of nnz/n
// entries each) leaves most of the chip idle with one row group per row, so
// block (s, r) sums the s-th of S equal segments of row r into
// part[r * S + s] and a second kernel adds each row's S partials in order.
struct SegArgs {
    int64_t rows, n, nnz, S;
    const void* vals;
    const void* rp;
    const void* ci;
    const void* x;
    bool x64;
    const int* skip;
    double* part;
};

template <int VT, class I>
__global__ __launch_bounds__(kCsrThreads) void csr_seg_kernel(SegArgs a) {
    if (a.skip && *a.skip) return;
    const int64_t r = blockIdx.y, sgi = blockIdx.x;
    const I* rp = static_cast<const I*>(a.rp);
    const I* ci = static_cast<const I*>(a.ci);
    const int64_t k0 = clamp_ptr(rp[r], a.nnz);
    int64_t k1 = clamp_ptr(rp[r + 1], a.nnz);
    k1 = k1 < k0 ? k0 : k1;
    const int64_t seg = (k1 - k0 + a.S - 1) / a.S;
    const int64_t b = k0 + sgi * seg;
    const int64_t e = b + seg < k1 ? b + seg : k1;
    double acc = 0.0;
    for (int64_t k = b + threadIdx.x; k < e; k += kCsrThreads) {
        const uint64_t col = static_cast<uint64_t>(static_cast<int64_t>(ci[k]));
        if (col < static_cast<uint64_t>(a.n)) {
            const double xv = a.x64 ? static_cast<const double*>(a.x)[col]
                                    : static_cast<double>(static_cast<const float*>(a.x)[col]);
            acc = fma(load_val<VT>(a.vals, k), xv, acc);
        }
    }
    __shared__ double part[kCsrThreads / 64];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (threadIdx.x % 64 == 0) part[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kCsrThreads / 64; ++w) t += part[w];
        a.part[r * a.S + sgi] = t;
    }
}

__global__ __launch_bounds__(kCsrThreads) void csr_seg_finish_kernel(int64_t rows, int64_t S, const double* part,
                                                                     int64_t rs, int64_t ss, void* y, bool y64,
                                                                     Coef c) {
    if (c.skip && *c.skip) return;
    const double alpha = c.dalpha ? *c.dalpha * c.alpha_mul : c.alpha;
    const double beta = c.dbeta ? *c.dbeta * c.beta_mul : c.beta;
    const int64_t r = static_cast<int64_t>(blockIdx.x) * kCsrThreads + threadIdx.x;
    double sq = 0.0;
    if (r < rows) {
        double acc = 0.0;
        for (int64_t s = 0; s < S; ++s) acc += part[r * rs + s * ss];
        double out = alpha * acc;
        if (beta != 0.0)
            out += beta * (y64 ? static_cast<double*>(y)[r] : static_cast<double>(static_cast<float*>(y)[r]));
        if (y64) {
            static_cast<double*>(y)[r] = out;
            sq = out * out;
        } else {
            const float o32 = static_cast<float>(out);
            static_cast<float*>(y)[r] = o32;
            sq = static_cast<double>(o32) * o32;
        }
    }
    if (c.sumsq) {
        __shared__ double p[kCsrThreads / 64];
        for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
        if (threadIdx.x % 64 == 0) p[threadIdx.x / 64] = sq;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < kCsrThreads / 64; ++w) t += p[w];
            c.sumsq[blockIdx.x] = t;
        }
    }
}

// ---- A^T x by panels, the panel's slice of x in LDS -------------------------
// Block (q, p) loads x[p*PR, p*PR + len) into LDS and forms, for columns c
// of chunk q, part[p*nout + c] = sum over row c of panel p of the blocked
// transpose (entries index A's rows, all inside the panel's slice).
constexpr int kPanelThreads = 1024;
constexpr int64_t kPanelLdsBytes = 65536;

struct PanelArgs {
    int64_t panels, nout, nnz, panel_rows, mrows, chunk;
    const void* vals;
    const void* rp;
    const void* ci;
    const void* x;
    bool x64;
    const int* skip;
    double* part;
};

template <int VT, class I, bool X64>
__global__ __launch_bounds__(kPanelThreads) void csr_panel_tmv_kernel(PanelArgs a) {
    if (a.skip && *a.skip) return;
    using XT = typename std::conditional<X64, double, float>::type;
    XT* xs = reinterpret_cast<XT*>(csr_xs);
    const int64_t p = blockIdx.y;
    const int64_t base = p * a.panel_rows;
    const int64_t len = std::min<int64_t>(a.panel_rows, a.mrows - base);
    const XT* xg = static_cast<const XT*>(a.x) + base;
    for (int64_t i = threadIdx.x; i < len; i += kPanelThreads) xs[i] = xg[i];
    __syncthreads();
    const I* rp = static_cast<const I*>(a.rp);
    const I* ci = static_cast<const I*>(a.ci);
    const int lane = threadIdx.x % 64;
    const int64_t c0 = static_cast<int64_t>(blockIdx.x) * a.chunk;
    const int64_t c1 = std::min<int64_t>(c0 + a.chunk, a.nout);
    for (int64_t c = c0 + threadIdx.x / 64; c < c1; c += kPanelThreads / 64) {
        const int64_t r = p * a.nout + c;
        const int64_t k0 = clamp_ptr(rp[r], a.nnz);
        int64_t k1 = clamp_ptr(rp[r + 1], a.nnz);
        k1 = k1 < k0 ? k0 : k1;
        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
        const uint64_t ulen = static_cast<uint64_t>(len);
        int64_t k = k0 + lane;
        for (; k + 192 < k1; k += 256) {
            const uint64_t o0 = static_cast<uint64_t>(static_cast<int64_t>(ci[k]) - base);
            const uint64_t o1 = static_cast<uint64_t>(static_cast<int64_t>(ci[k + 64]) - base);
            const uint64_t o2 = static_cast<uint64_t>(static_cast<int64_t>(ci[k + 128]) - base);
            const uint64_t o3 = static_cast<uint64_t>(static_cast<int64_t>(ci[k + 192]) - base);
            const double v0 = load_val<VT>(a.vals, k), v1 = load_val<VT>(a.vals, k + 64);
            const double v2 = load_val<VT>(a.vals, k + 128), v3 = load_val<VT>(a.vals, k + 192);
            if (o0 < ulen) acc0 = fma(v0, static_cast<double>(xs[o0]), acc0);
            if (o1 < ulen) acc1 = fma(v1, static_cast<double>(xs[o1]), acc1);
            if (o2 < ulen) acc2 = fma(v2, static_cast<double>(xs[o2]), acc2);
            if (o3 < ulen) acc3 = fma(v3, static_cast<double>(xs[o3]), acc3);
        }
        for (; k < k1; k += 64) {
            const uint64_t o0 = static_cast<uint64_t>(static_cast<int64_t>(ci[k]) - base);
            if (o0 < ulen) acc0 = fma(load_val<VT>(a.vals, k), static_cast<double>(xs[o0]), acc0);
        }
        double acc = (acc0 + acc1) + (acc2 + acc3);
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) a.part[r] = acc;
    }
}

template <int VT, class I>
void launch_panel(const PanelArgs& g, dim3 grid, size_t lds, hipStream_t s) {
    if (g.x64) csr_panel_tmv_kernel<VT, I, true><<<grid, kPanelThreads, lds, s>>>(g);
    else csr_panel_tmv_kernel<VT, I, false><<<grid, kPanelThreads, lds, s>>>(g);
}

// y[r] = alpha * sum_p part[p*nout + r] + beta * y[r]: one wavefront per
// output, lanes over the panels, a fixed butterfly order
__global__ __launch_bounds__(kCsrThreads) void csr_panel_finish_kernel(int64_t nout, int64_t panels,
                                                                       const double* part, void* y, bool y64,
                                                                       Coef c) {
    if (c.skip && *c.skip) return;
    const double alpha = c.dalpha ? *c.dalpha * c.alpha_mul : c.alpha;
    const double beta = c.dbeta ? *c.dbeta * c.beta_mul : c.beta;
    const int lane = threadIdx.x % 64;
    double sq = 0.0;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * (kCsrThreads / 64) + threadIdx.x / 64; r < nout;
         r += static_cast<int64_t>(gridDim.x) * (kCsrThreads / 64)) {
        double acc = 0.0;
        for (int64_t p = lane; p < panels; p += 64) acc += part[p * nout + r];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) {
            double out = alpha * acc;
            if (beta != 0.0)
                out += beta * (y64 ? static_cast<double*>(y)[r] : static_cast<double>(static_cast<float*>(y)[r]));
            if (y64) {
                static_cast<double*>(y)[r] = out;
                sq += out * out;
            } else {
                const float o32 = static_cast<float>(out);
                static_cast<float*>(y)[r] = o32;
                sq += static_cast<double>(o32) * o32;
            }
        }
    }
    if (c.sumsq) {
        __shared__ double p4[kCsrThreads / 64];
        if (lane == 0) p4[threadIdx.x / 64] = sq;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < kCsrThreads / 64; ++w) t += p4[w];
            c.sumsq[blockIdx.x] = t;
        }
    }
}

// segments per row for the split path, or 0 for the row-group kernel
int64_t seg_count(int64_t m, int64_t nnz) {
    if (m <= 0 || m > 65535) return 0;
    const double mean = static_cast<double>(nnz) / static_cast<double>(m);
    if (mean < 1024.0) return 0;
    const int64_t want = ceil_div(8192, m);                                 // about 8192 blocks in all
    const int64_t most = static_cast<int64_t>(mean / 256.0);                // segments of >= 256 entries
    return std::max<int64_t>(1, std::min<int64_t>(std::min(want, most), 65535));
}

// about eight entries per lane (two rounds of the four chains)
int lanes_for(int64_t m, int64_t nnz) {
    const double mean = m > 0 ? static_cast<double>(nnz) / static_cast<double>(m) : 0.0;
    return mean <= 8.0 ? 1 : mean <= 32.0 ? 4 : mean <= 64.0 ? 8 : mean <= 128.0 ? 16 : mean <= 256.0 ? 32 : 64;
}

int64_t mv_blocks(int64_t m, int lanes) {
    const int64_t b = ceil_div(std::max<int64_t>(m, 1) * lanes, kCsrThreads);
    return std::min<int64_t>(b, kCsrMaxBlocks);
}

template <int VT, class I>
void launch_mv_vt(const MvArgs& a, int lanes, unsigned blocks, hipStream_t s) {
    const size_t lds = a.xlds ? static_cast<size_t>(a.n) * sizeof(double) : 0;
    switch (lanes) {
        case 1: csr_mv_kernel<VT, I, 1><<<blocks, kCsrThreads, lds, s>>>(a); break;
        case 4: csr_mv_kernel<VT, I, 4><<<blocks, kCsrThreads, lds, s>>>(a); break;
        case 8: csr_mv_kernel<VT, I, 8><<<blocks, kCsrThreads, lds, s>>>(a); break;
        case 16: csr_mv_kernel<VT, I, 16><<<blocks, kCsrThreads, lds, s>>>(a); break;
        case 32: csr_mv_kernel<VT, I, 32><<<blocks, kCsrThreads, lds, s>>>(a); break;
        default: csr_mv_kernel<VT, I, 64><<<blocks, kCsrThreads, lds, s>>>(a); break;
    }
}

template <class I>
void launch_mv(int vt, const MvArgs& a, int lanes, unsigned blocks, hipStream_t s) {
    switch (vt) {
        case TOMOS_F64: launch_mv_vt<TOMOS_F64, I>(a, lanes, blocks, s); break;
        case TOMOS_F32: launch_mv_vt<TOMOS_F32, I>(a, lanes, blocks, s); break;
        case TOMOS_F16: launch_mv_vt<TOMOS_F16, I>(a, lanes, blocks, s); break;
        default: launch_mv_vt<TOMOS_BF16, I>(a, lanes, blocks, s); break;
    }
}

// ---- validation -------------------------------------------------------------
// flag bits: 1 row_ptrs[0] != 0, 2 decreasing row_ptrs, 4 row_ptrs[m] != nnz,
// 8 column index out of range
template <class I>
__global__ void csr_check_kernel(int64_t m, int64_t n, int64_t nnz, const I* rp, const I* ci, unsigned* flag) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    unsigned f = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i <= m; i += stride) {
        const int64_t p = static_cast<int64_t>(rp[i]);
        if (i == 0 && p != 0) f |= 1u;
        if (i < m && static_cast<int64_t>(rp[i + 1]) < p) f |= 2u;
        if (i == m && p != nnz) f |= 4u;
    }
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < nnz; k += stride) {
        const int64_t c = static_cast<int64_t>(ci[k]);
        if (c < 0 || c >= n) f |= 8u;
    }
    if (f) atomicOr(flag, f);
}

// ---- transpose ----------------------------------------------------------------
// row index of every stored entry (rows[k] = i for rp[i] <= k < rp[i+1])
// 16 lanes per row: consecutive lanes write consecutive entries
template <class I>
__global__ void csr_row_of_kernel(int64_t m, const I* rp, I* rows) {
    const int lane = threadIdx.x % 16;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * (blockDim.x / 16);
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * (blockDim.x / 16) + threadIdx.x / 16; i < m; i += stride)
        for (int64_t k = static_cast<int64_t>(rp[i]) + lane; k < static_cast<int64_t>(rp[i + 1]); k += 16)
            rows[k] = static_cast<I>(i);
}

// 4-byte and 2-byte values with 32-bit indices travel through the sort
// packed with their row index (row in the high word), so the transpose
// needs no gather through the permutation
template <int VT>
__global__ void csr_pack_rowval_kernel(int64_t nnz, const uint32_t* rows, const void* vals, uint64_t* out) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < nnz; k += stride) {
        uint32_t v;
        if constexpr (VT == TOMOS_F32) v = static_cast<const uint32_t*>(vals)[k];
        else v = static_cast<const uint16_t*>(vals)[k];
        out[k] = (static_cast<uint64_t>(rows[k]) << 32) | v;
    }
}

template <int VT>
__global__ void csr_unpack_rowval_kernel(int64_t nnz, const uint64_t* in, void* vals_t, uint32_t* idx_t) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < nnz; k += stride) {
        const uint64_t w = in[k];
        idx_t[k] = static_cast<uint32_t>(w >> 32);
        if constexpr (VT == TOMOS_F32) static_cast<uint32_t*>(vals_t)[k] = static_cast<uint32_t>(w);
        else static_cast<uint16_t*>(vals_t)[k] = static_cast<uint16_t>(w);
    }
}

// ptr_t[c] = first sorted position with key >= c, from the boundaries of
// the sorted keys (every c in (keys[k-1], keys[k]] gets k).
template <class U>
__global__ void csr_ptr_from_keys_kernel(int64_t nnz, int64_t n, const U* keys, U* ptr_t) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k <= nnz; k += stride) {
        const int64_t prev = k == 0 ? -1 : static_cast<int64_t>(keys[k - 1]);
        int64_t cur = k == nnz ? n : static_cast<int64_t>(keys[k]);
        cur = cur > n ? n : cur;
        for (int64_t c = prev + 1; c <= cur; ++c) ptr_t[c] = static_cast<U>(k);
    }
}

template <int VT, class U>
__global__ void csr_gather_kernel(int64_t nnz, const U* perm, const void* vals, const U* rows, void* vals_t,
                                  U* idx_t) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < nnz; k += stride) {
        const U p = perm[k];
        idx_t[k] = rows[p];
        if constexpr (VT == TOMOS_F64) static_cast<double*>(vals_t)[k] = static_cast<const double*>(vals)[p];
        else if constexpr (VT == TOMOS_F32) static_cast<float*>(vals_t)[k] = static_cast<const float*>(vals)[p];
        else static_cast<uint16_t*>(vals_t)[k] = static_cast<const uint16_t*>(vals)[p];
    }
}

unsigned grid_1d(int64_t work) {
    return static_cast<unsigned>(std::min<int64_t>(std::max<int64_t>(ceil_div(work, 256), 1), 8192));
}

unsigned key_bits(int64_t n) {
    unsigned b = 1;
    while (b < 63 && (int64_t(1) << b) < n) ++b;
    return b;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

template <class U>
size_t sort_temp_bytes(int64_t nnz, int64_t n) {
    size_t bytes = 0;
    rocprim::counting_iterator<U> iota(U(0));
    TOMOS_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, static_cast<const U*>(nullptr),
                                              static_cast<U*>(nullptr), iota, static_cast<U*>(nullptr),
                                              static_cast<size_t>(nnz), 0, key_bits(n)));
    return bytes;
}

// panel keys: (row / panel_rows) * n + column
template <class U>
__global__ void csr_panel_key_kernel(int64_t nnz, int64_t n, int64_t panel_rows, const U* rows, const U* ci,
                                     U* keys) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < nnz; k += stride)
        keys[k] = static_cast<U>((static_cast<int64_t>(rows[k]) / panel_rows) * n + static_cast<int64_t>(ci[k]));
}

size_t sort_temp_bytes_packed(int64_t nnz, int64_t n) {
    size_t bytes = 0;
    TOMOS_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, static_cast<const uint32_t*>(nullptr),
                                              static_cast<uint32_t*>(nullptr), static_cast<const uint64_t*>(nullptr),
                                              static_cast<uint64_t*>(nullptr), static_cast<size_t>(nnz), 0,
                                              key_bits(n)));
    return bytes;
}

// values of 4 or 2 bytes with 32-bit indices: sorted packed with the row
bool packed_sort(const CsrMat& A) { return A.idx_type == TOMOS_INDEX_INT32 && A.val_type != TOMOS_F64; }

template <class U>
size_t transpose_bytes(int64_t nnz, int64_t keys) {
    const size_t e = static_cast<size_t>(std::max<int64_t>(nnz, 1)) * sizeof(U);
    return 4 * align256(e) + align256(sort_temp_bytes<U>(nnz, keys));
}

size_t transpose_bytes_packed(int64_t nnz, int64_t keys) {
    const size_t e = static_cast<size_t>(std::max<int64_t>(nnz, 1));
    // keys_out, keys_in (4 B), rows (4 B), packed in/out (8 B each)
    return 3 * align256(4 * e) + 2 * align256(8 * e) + align256(sort_temp_bytes_packed(nnz, keys));
}

// transpose_impl for packed_sort(A): one sort carries (row, value)
void transpose_packed(const CsrMat& A, int64_t panel_rows, void* vals_t, void* ptr_t, void* idx_t, void* scratch,
                      hipStream_t s) {
    const int64_t panels = A.m > 0 ? ceil_div(A.m, panel_rows) : 1;
    const int64_t nkeys = panels * A.n;
    const size_t e = static_cast<size_t>(std::max<int64_t>(A.nnz, 1));
    char* sc = static_cast<char*>(scratch);
    uint32_t* keys_out = reinterpret_cast<uint32_t*>(sc);
    uint32_t* keys_in = reinterpret_cast<uint32_t*>(sc + align256(4 * e));
    uint32_t* rows = reinterpret_cast<uint32_t*>(sc + 2 * align256(4 * e));
    uint64_t* pin = reinterpret_cast<uint64_t*>(sc + 3 * align256(4 * e));
    uint64_t* pout = reinterpret_cast<uint64_t*>(sc + 3 * align256(4 * e) + align256(8 * e));
    void* temp = sc + 3 * align256(4 * e) + 2 * align256(8 * e);
    size_t temp_bytes = sort_temp_bytes_packed(A.nnz, nkeys);
    if (A.nnz > 0) {
        csr_row_of_kernel<uint32_t><<<grid_1d(A.m * 16), 256, 0, s>>>(A.m, static_cast<const uint32_t*>(A.row_ptrs),
                                                                       rows);
        TOMOS_LAUNCH_CHECK("csr_row_of_kernel");
        const uint32_t* keys = static_cast<const uint32_t*>(A.col_idxs);
        if (panels > 1) {
            csr_panel_key_kernel<uint32_t><<<grid_1d(A.nnz), 256, 0, s>>>(A.nnz, A.n, panel_rows, rows, keys,
                                                                          keys_in);
            TOMOS_LAUNCH_CHECK("csr_panel_key_kernel");
            keys = keys_in;
        }
        const unsigned g = grid_1d(A.nnz);
        if (A.val_type == TOMOS_F32) csr_pack_rowval_kernel<TOMOS_F32><<<g, 256, 0, s>>>(A.nnz, rows, A.vals, pin);
        else csr_pack_rowval_kernel<TOMOS_F16><<<g, 256, 0, s>>>(A.nnz, rows, A.vals, pin);
        TOMOS_LAUNCH_CHECK("csr_pack_rowval_kernel");
        TOMOS_HIP_CHECK(rocprim::radix_sort_pairs(temp, temp_bytes, keys, keys_out, pin, pout,
                                                  static_cast<size_t>(A.nnz), 0, key_bits(nkeys), s));
        if (A.val_type == TOMOS_F32)
            csr_unpack_rowval_kernel<TOMOS_F32><<<g, 256, 0, s>>>(A.nnz, pout, vals_t, static_cast<uint32_t*>(idx_t));
        else
            csr_unpack_rowval_kernel<TOMOS_F16><<<g, 256, 0, s>>>(A.nnz, pout, vals_t, static_cast<uint32_t*>(idx_t));
        TOMOS_LAUNCH_CHECK("csr_unpack_rowval_kernel");
    }
    csr_ptr_from_keys_kernel<uint32_t><<<grid_1d(A.nnz + 1), 256, 0, s>>>(A.nnz, nkeys, keys_out,
                                                                          static_cast<uint32_t*>(ptr_t));
    TOMOS_LAUNCH_CHECK("csr_ptr_from_keys_kernel");
}

// CSR of the panel-blocked transpose: row p*n + c holds column c of A's
// rows [p*panel_rows, (p+1)*panel_rows) in A's row order (panels = 1: A^T).
template <class U>
void transpose_impl(const CsrMat& A, int64_t panel_rows, void* vals_t, void* ptr_t, void* idx_t, void* scratch,
                    hipStream_t s) {
    const int64_t panels = A.m > 0 ? ceil_div(A.m, panel_rows) : 1;
    const int64_t nkeys = panels * A.n;
    const size_t e = static_cast<size_t>(std::max<int64_t>(A.nnz, 1)) * sizeof(U);
    char* sc = static_cast<char*>(scratch);
    U* keys_out = reinterpret_cast<U*>(sc);
    U* perm = reinterpret_cast<U*>(sc + align256(e));
    U* rows = reinterpret_cast<U*>(sc + 2 * align256(e));
    U* keys_in = reinterpret_cast<U*>(sc + 3 * align256(e));
    void* temp = sc + 4 * align256(e);
    size_t temp_bytes = sort_temp_bytes<U>(A.nnz, nkeys);
    if (A.nnz > 0) {
        csr_row_of_kernel<U><<<grid_1d(A.m * 16), 256, 0, s>>>(A.m, static_cast<const U*>(A.row_ptrs), rows);
        TOMOS_LAUNCH_CHECK("csr_row_of_kernel");
        const U* keys = static_cast<const U*>(A.col_idxs);
        if (panels > 1) {
            csr_panel_key_kernel<U><<<grid_1d(A.nnz), 256, 0, s>>>(A.nnz, A.n, panel_rows, rows, keys, keys_in);
            TOMOS_LAUNCH_CHECK("csr_panel_key_kernel");
            keys = keys_in;
        }
        rocprim::counting_iterator<U> iota(U(0));
        TOMOS_HIP_CHECK(rocprim::radix_sort_pairs(temp, temp_bytes, keys, keys_out, iota, perm,
                                                  static_cast<size_t>(A.nnz), 0, key_bits(nkeys), s));
    }
    csr_ptr_from_keys_kernel<U><<<grid_1d(A.nnz + 1), 256, 0, s>>>(A.nnz, nkeys, keys_out, static_cast<U*>(ptr_t));
    TOMOS_LAUNCH_CHECK("csr_ptr_from_keys_kernel");
    if (A.nnz == 0) return;
    const unsigned g = grid_1d(A.nnz);
    U* it = static_cast<U*>(idx_t);
    switch (A.val_type) {
        case TOMOS_F64: csr_gather_kernel<TOMOS_F64, U><<<g, 256, 0, s>>>(A.nnz, perm, A.vals, rows, vals_t, it); break;
        case TOMOS_F32: csr_gather_kernel<TOMOS_F32, U><<<g, 256, 0, s>>>(A.nnz, perm, A.vals, rows, vals_t, it); break;
        default: csr_gather_kernel<TOMOS_F16, U><<<g, 256, 0, s>>>(A.nnz, perm, A.vals, rows, vals_t, it); break;
    }
    TOMOS_LAUNCH_CHECK("csr_gather_kernel");
}

}  // namespace

size_t csr_value_bytes(int t) { return t == TOMOS_F64 ? 8 : t == TOMOS_F32 ? 4 : 2; }
size_t csr_index_bytes(int t) { return t == TOMOS_INDEX_INT64 ? 8 : 4; }

void csr_require_valid_types(const CsrMat& A, const char* who) {
    const std::string w(who);
    require(A.val_type == TOMOS_F64 || A.val_type == TOMOS_F32 || A.val_type == TOMOS_F16 ||
                A.val_type == TOMOS_BF16,
            w + ": CSR values must be F64, F32, F16 or BF16");
    require(A.idx_type == TOMOS_INDEX_INT32 || A.idx_type == TOMOS_INDEX_INT64,
            w + ": CSR index type must be INT32 or INT64");
    require(A.m >= 0 && A.n >= 0 && A.nnz >= 0, w + ": negative size");
    require(A.idx_type == TOMOS_INDEX_INT64 || (A.nnz <= 0x7FFFFFFF && A.m <= 0x7FFFFFFF && A.n <= 0x7FFFFFFF),
            w + ": sizes exceed INT32 indices");
    require(A.row_ptrs != nullptr || A.m == 0, w + ": null row_ptrs");
    require((A.vals && A.col_idxs) || A.nnz == 0, w + ": null values or column indices");
}

int64_t csr_sumsq_blocks_max(int64_t m) {
    return std::max<int64_t>(std::min<int64_t>(ceil_div(std::max<int64_t>(m, 1), 4), kCsrMaxBlocks),
                             ceil_div(std::max<int64_t>(m, 1), kCsrThreads));
}

size_t csr_mv_scratch_bytes(const CsrMat& A) {
    const int64_t S = seg_count(A.m, A.nnz);
    return S ? align256(static_cast<size_t>(A.m * S) * sizeof(double)) : 0;
}

void csr_mv(const CsrMat& A, const Coef& c, int x_type, const void* x, int y_type, void* y, void* scratch,
            size_t scratch_bytes, hipStream_t s, int* sumsq_blocks) {
    require(x_type == TOMOS_F32 || x_type == TOMOS_F64, "csr_mv: x must be F32 or F64");
    require(y_type == TOMOS_F32 || y_type == TOMOS_F64, "csr_mv: y must be F32 or F64");
    if (A.m == 0) {
        if (sumsq_blocks) *sumsq_blocks = 0;
        return;
    }
    require(y != nullptr && (x != nullptr || A.n == 0), "csr_mv: null x or y");
    const int64_t S = seg_count(A.m, A.nnz);
    if (S) {
        require(scratch && scratch_bytes >= csr_mv_scratch_bytes(A), "csr_mv: scratch too small");
        SegArgs g{A.m, A.n, A.nnz, S, A.vals, A.row_ptrs, A.col_idxs, x, x_type == TOMOS_F64, c.skip,
                  static_cast<double*>(scratch)};
        const dim3 grid(static_cast<unsigned>(S), static_cast<unsigned>(A.m));
        if (A.idx_type == TOMOS_INDEX_INT64) {
            switch (A.val_type) {
                case TOMOS_F64: csr_seg_kernel<TOMOS_F64, int64_t><<<grid, kCsrThreads, 0, s>>>(g); break;
                case TOMOS_F32: csr_seg_kernel<TOMOS_F32, int64_t><<<grid, kCsrThreads, 0, s>>>(g); break;
                case TOMOS_F16: csr_seg_kernel<TOMOS_F16, int64_t><<<grid, kCsrThreads, 0, s>>>(g); break;
                default: csr_seg_kernel<TOMOS_BF16, int64_t><<<grid, kCsrThreads, 0, s>>>(g); break;
            }
        } else {
            switch (A.val_type) {
                case TOMOS_F64: csr_seg_kernel<TOMOS_F64, int32_t><<<grid, kCsrThreads, 0, s>>>(g); break;
                case TOMOS_F32: csr_seg_kernel<TOMOS_F32, int32_t><<<grid, kCsrThreads, 0, s>>>(g); break;
                case TOMOS_F16: csr_seg_kernel<TOMOS_F16, int32_t><<<grid, kCsrThreads, 0, s>>>(g); break;
                default: csr_seg_kernel<TOMOS_BF16, int32_t><<<grid, kCsrThreads, 0, s>>>(g); break;
            }
        }
        TOMOS_LAUNCH_CHECK("csr_seg_kernel");
        const unsigned fb = static_cast<unsigned>(ceil_div(A.m, kCsrThreads));
        csr_seg_finish_kernel<<<fb, kCsrThreads, 0, s>>>(A.m, S, static_cast<const double*>(scratch), S, 1, y,
                                                         y_type == TOMOS_F64, c);
        TOMOS_LAUNCH_CHECK("csr_seg_finish_kernel");
        if (sumsq_blocks) *sumsq_blocks = static_cast<int>(fb);
        return;
    }
    MvArgs a{A.m,   A.n, A.nnz, A.vals, A.row_ptrs, A.col_idxs, x, y, x_type == TOMOS_F64, y_type == TOMOS_F64,
             A.n > 0 && A.n <= kXLdsMax && A.nnz >= 4 * A.n, c};
    const int lanes = lanes_for(A.m, A.nnz);
    const unsigned blocks = static_cast<unsigned>(mv_blocks(A.m, lanes));
    if (A.idx_type == TOMOS_INDEX_INT64) launch_mv<int64_t>(A.val_type, a, lanes, blocks, s);
    else launch_mv<int32_t>(A.val_type, a, lanes, blocks, s);
    TOMOS_LAUNCH_CHECK("csr_mv_kernel");
    if (sumsq_blocks) *sumsq_blocks = static_cast<int>(blocks);
}

size_t csr_check_scratch_bytes() { return 256; }

void csr_check(const CsrMat& A, void* scratch, size_t scratch_bytes, hipStream_t s) {
    require(scratch_bytes >= csr_check_scratch_bytes(), "csr_check: scratch too small");
    unsigned* flag = static_cast<unsigned*>(scratch);
    TOMOS_HIP_CHECK(hipMemsetAsync(flag, 0, sizeof(unsigned), s));
    const unsigned g = grid_1d(std::max<int64_t>(A.m + 1, A.nnz));
    if (A.idx_type == TOMOS_INDEX_INT64)
        csr_check_kernel<int64_t><<<g, 256, 0, s>>>(A.m, A.n, A.nnz, static_cast<const int64_t*>(A.row_ptrs),
                                                    static_cast<const int64_t*>(A.col_idxs), flag);
    else
        csr_check_kernel<int32_t><<<g, 256, 0, s>>>(A.m, A.n, A.nnz, static_cast<const int32_t*>(A.row_ptrs),
                                                    static_cast<const int32_t*>(A.col_idxs), flag);
    TOMOS_LAUNCH_CHECK("csr_check_kernel");
    unsigned f = 0;
    TOMOS_HIP_CHECK(hipMemcpyAsync(&f, flag, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    TOMOS_HIP_CHECK(hipStreamSynchronize(s));
    require(!(f & 1u), "CSR: row_ptrs[0] must be 0");
    require(!(f & 2u), "CSR: row_ptrs must be non-decreasing");
    require(!(f & 4u), "CSR: row_ptrs[m] must equal nnz");
    require(!(f & 8u), "CSR: column index out of range");
}

int64_t csr_panel_rows(const CsrMat& A, int x_type) {
    const int64_t pr = kPanelLdsBytes / (x_type == TOMOS_F64 ? 8 : 4);
    if (A.m <= pr || A.n == 0) return std::max<int64_t>(A.m, 1);
    int64_t panels = ceil_div(A.m, pr);
    // keys p*n + c must fit the index type
    const int64_t key_max = A.idx_type == TOMOS_INDEX_INT64 ? (int64_t(1) << 62) : int64_t(0x7FFFFFFF);
    panels = std::max<int64_t>(1, std::min<int64_t>(panels, key_max / A.n));
    const int64_t rows = ceil_div(A.m, panels);
    require(rows <= pr, "CSR transpose: too many panels for the index type");
    return rows;
}

int64_t csr_panels(const CsrMat& A, int64_t panel_rows) { return A.m > 0 ? ceil_div(A.m, panel_rows) : 1; }

size_t csr_transpose_scratch_bytes(const CsrMat& A, int64_t panel_rows) {
    const int64_t nkeys = csr_panels(A, panel_rows) * A.n;
    if (packed_sort(A)) return transpose_bytes_packed(A.nnz, nkeys);
    return A.idx_type == TOMOS_INDEX_INT64 ? transpose_bytes<uint64_t>(A.nnz, nkeys)
                                           : transpose_bytes<uint32_t>(A.nnz, nkeys);
}

void csr_transpose(const CsrMat& A, int64_t panel_rows, void* vals_t, void* ptr_t, void* idx_t, void* scratch,
                   size_t scratch_bytes, hipStream_t s) {
    require(panel_rows >= 1, "csr_transpose: bad panel size");
    require(scratch_bytes >= csr_transpose_scratch_bytes(A, panel_rows), "csr_transpose: scratch too small");
    require(ptr_t != nullptr && ((vals_t && idx_t) || A.nnz == 0), "csr_transpose: null output");
    // the index buffers hold non-negative values: sorted as unsigned
    if (packed_sort(A)) transpose_packed(A, panel_rows, vals_t, ptr_t, idx_t, scratch, s);
    else if (A.idx_type == TOMOS_INDEX_INT64) transpose_impl<uint64_t>(A, panel_rows, vals_t, ptr_t, idx_t, scratch, s);
    else transpose_impl<uint32_t>(A, panel_rows, vals_t, ptr_t, idx_t, scratch, s);
}

size_t csr_tmv_scratch_bytes(const CsrMat& Tb, int64_t panels, int64_t nout) {
    (void)Tb;
    return align256(static_cast<size_t>(std::max<int64_t>(panels * nout, 1)) * sizeof(double));
}

void csr_tmv(const CsrMat& Tb, int64_t panels, int64_t panel_rows, int64_t nout, const Coef& c, int x_type,
             const void* x, int y_type, void* y, void* scratch, size_t scratch_bytes, hipStream_t s,
             int* sumsq_blocks) {
    require(x_type == TOMOS_F32 || x_type == TOMOS_F64, "csr_tmv: x must be F32 or F64");
    require(panel_rows * (x_type == TOMOS_F64 ? 8 : 4) <= kPanelLdsBytes, "csr_tmv: panel does not fit LDS");
    require(scratch_bytes >= csr_tmv_scratch_bytes(Tb, panels, nout), "csr_tmv: scratch too small");
    if (nout == 0) {
        if (sumsq_blocks) *sumsq_blocks = 0;
        return;
    }
    // per-panel partial products, then each output's panels added in order
    double* part = static_cast<double*>(scratch);
    // about 1024 blocks in all; a block's columns share its LDS slice
    const int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(1024, panels), ceil_div(nout, 16)));
    PanelArgs g{panels, nout, Tb.nnz, panel_rows, Tb.n, ceil_div(nout, chunks), Tb.vals, Tb.row_ptrs,
                Tb.col_idxs, x, x_type == TOMOS_F64, c.skip, part};
    const dim3 grid(static_cast<unsigned>(ceil_div(nout, g.chunk)), static_cast<unsigned>(panels));
    const size_t lds = static_cast<size_t>(std::min<int64_t>(panel_rows, std::max<int64_t>(Tb.n, 1))) *
                       (x_type == TOMOS_F64 ? 8 : 4);
    if (Tb.idx_type == TOMOS_INDEX_INT64) {
        switch (Tb.val_type) {
            case TOMOS_F64: launch_panel<TOMOS_F64, int64_t>(g, grid, lds, s); break;
            case TOMOS_F32: launch_panel<TOMOS_F32, int64_t>(g, grid, lds, s); break;
            case TOMOS_F16: launch_panel<TOMOS_F16, int64_t>(g, grid, lds, s); break;
            default: launch_panel<TOMOS_BF16, int64_t>(g, grid, lds, s); break;
        }
    } else {
        switch (Tb.val_type) {
            case TOMOS_F64: launch_panel<TOMOS_F64, int32_t>(g, grid, lds, s); break;
            case TOMOS_F32: launch_panel<TOMOS_F32, int32_t>(g, grid, lds, s); break;
            case TOMOS_F16: launch_panel<TOMOS_F16, int32_t>(g, grid, lds, s); break;
            default: launch_panel<TOMOS_BF16, int32_t>(g, grid, lds, s); break;
        }
    }
    TOMOS_LAUNCH_CHECK("csr_panel_tmv_kernel");
    const unsigned fb = static_cast<unsigned>(std::min<int64_t>(ceil_div(nout, kCsrThreads / 64), kCsrMaxBlocks));
    csr_panel_finish_kernel<<<fb, kCsrThreads, 0, s>>>(nout, panels, part, y, y_type == TOMOS_F64, c);
    TOMOS_LAUNCH_CHECK("csr_panel_finish_kernel");
    if (sumsq_blocks) *sumsq_blocks = static_cast<int>(fb);
}

namespace {

// Y (m x ncols) = alpha * A X + beta * Y for CSR A (m x n) and dense X
// (n x ncols, ldx): one thread per output row, columns strided over
// blockIdx.y; each output is one FP64 sum over the row's entries in
// storage order (deterministic).  Clamped row ranges and skipped
// out-of-range columns keep every access in bounds whatever the indices.
template <int VT, class I>
__global__ __launch_bounds__(kCsrThreads) void csr_mm_kernel(int64_t m, int64_t n, int64_t nnz, int64_t ncols,
                                                             const void* vals, const I* __restrict__ rp,
                                                             const I* __restrict__ ci, const void* X, bool x64,
                                                             int64_t ldx, void* Y, bool y64, int64_t ldy,
                                                             double alpha, double beta) {
    const int64_t row = static_cast<int64_t>(blockIdx.x) * kCsrThreads + threadIdx.x;
    if (row >= m) return;
    const int64_t k0 = clamp_ptr(rp[row], nnz);
    int64_t k1 = clamp_ptr(rp[row + 1], nnz);
    k1 = k1 < k0 ? k0 : k1;
    for (int64_t j = blockIdx.y; j < ncols; j += gridDim.y) {
        double acc = 0.0;
        for (int64_t k = k0; k < k1; ++k) {
            const uint64_t c = static_cast<uint64_t>(static_cast<int64_t>(ci[k]));
            if (c >= static_cast<uint64_t>(n)) continue;
            const int64_t at = static_cast<int64_t>(c) + j * ldx;
            const double xv = x64 ? static_cast<const double*>(X)[at]
                                  : static_cast<double>(static_cast<const float*>(X)[at]);
            acc = fma(load_val<VT>(vals, k), xv, acc);
        }
        const int64_t o = row + j * ldy;
        double out = alpha * acc;
        if (beta != 0.0)
            out += beta * (y64 ? static_cast<double*>(Y)[o] : static_cast<double>(static_cast<float*>(Y)[o]));
        if (y64) static_cast<double*>(Y)[o] = out;
        else static_cast<float*>(Y)[o] = static_cast<float>(out);
    }
}

template <int VT, class I>
void launch_csr_mm(const CsrMat& A, int64_t ncols, const void* X, bool x64, int64_t ldx, void* Y, bool y64,
                   int64_t ldy, double alpha, double beta, hipStream_t s) {
    const dim3 grid(static_cast<unsigned>(ceil_div(A.m, kCsrThreads)),
                    static_cast<unsigned>(std::min<int64_t>(ncols, 65535)));
    csr_mm_kernel<VT, I><<<grid, kCsrThreads, 0, s>>>(A.m, A.n, A.nnz, ncols, A.vals, static_cast<const I*>(A.row_ptrs),
                                                      static_cast<const I*>(A.col_idxs), X, x64, ldx, Y, y64, ldy,
                                                      alpha, beta);
}

}  // namespace

void csr_mm(const CsrMat& A, int64_t ncols, int x_type, const void* X, int64_t ldx, int y_type, void* Y, int64_t ldy,
            double alpha, double beta, hipStream_t s) {
    csr_require_valid_types(A, "csr_mm");
    require(x_type == TOMOS_F32 || x_type == TOMOS_F64, "csr_mm: X must be F32 or F64");
    require(y_type == TOMOS_F32 || y_type == TOMOS_F64, "csr_mm: Y must be F32 or F64");
    require(ncols >= 0, "csr_mm: negative column count");
    require(ldx >= std::max<int64_t>(A.n, 1) && ldy >= std::max<int64_t>(A.m, 1), "csr_mm: leading dimension too small");
    require(ceil_div(A.m, kCsrThreads) <= 0x7FFFFFFF, "csr_mm: too many rows");
    if (A.m == 0 || ncols == 0) return;
    require(A.row_ptrs && Y && (A.n == 0 || X) && (A.nnz == 0 || (A.vals && A.col_idxs)), "csr_mm: null pointer");
    const bool x64 = x_type == TOMOS_F64, y64 = y_type == TOMOS_F64;
#define TOMOS_CSR_MM(I)                                                                                   \
    switch (A.val_type) {                                                                                 \
        case TOMOS_F64: launch_csr_mm<TOMOS_F64, I>(A, ncols, X, x64, ldx, Y, y64, ldy, alpha, beta, s); break; \
        case TOMOS_F32: launch_csr_mm<TOMOS_F32, I>(A, ncols, X, x64, ldx, Y, y64, ldy, alpha, beta, s); break; \
        case TOMOS_F16: launch_csr_mm<TOMOS_F16, I>(A, ncols, X, x64, ldx, Y, y64, ldy, alpha, beta, s); break; \
        default: launch_csr_mm<TOMOS_BF16, I>(A, ncols, X, x64, ldx, Y, y64, ldy, alpha, beta, s); break;      \
    }
    if (A.idx_type == TOMOS_INDEX_INT64) {
        TOMOS_CSR_MM(int64_t)
    } else {
        TOMOS_CSR_MM(int32_t)
    }
#undef TOMOS_CSR_MM
    TOMOS_LAUNCH_CHECK("csr_mm_kernel");
}

}  // namespace tomos_hip

using namespace tomos_hip;

namespace {
CsrMat make_csr(int val_type, int idx_type, int64_t m, int64_t n, int64_t nnz, const void* vals, const void* rp,
                const void* ci) {
    CsrMat A;
    A.val_type = val_type;
    A.idx_type = idx_type;
    A.m = m;
    A.n = n;
    A.nnz = nnz;
    A.vals = vals;
    A.row_ptrs = rp;
    A.col_idxs = ci;
    return A;
}
}  // namespace

extern "C" {

int tomos_csr_check(int val_type, int idx_type, int64_t m, int64_t n, int64_t nnz, const void* row_ptrs,
                    const void* col_idxs, tomos_stream_t st) {
    return guarded([&] {
        const CsrMat A = make_csr(val_type, idx_type, m, n, nnz, nullptr, row_ptrs, col_idxs);
        CsrMat T = A;
        T.vals = nnz ? col_idxs : nullptr;  // values are not read
        csr_require_valid_types(T, "tomos_csr_check");
        csr_check(A, fallback_scratch(csr_check_scratch_bytes()), csr_check_scratch_bytes(), st);
    });
}

size_t tomos_csr_transpose_scratch_size(int val_type, int idx_type, int64_t m, int64_t n, int64_t nnz) {
    const CsrMat A = make_csr(val_type, idx_type, m, n, nnz, nullptr, nullptr, nullptr);
    if (m < 0 || n < 0 || nnz < 0) return 0;
    try {
        return csr_transpose_scratch_bytes(A, std::max<int64_t>(m, 1));
    } catch (...) {
        return 0;
    }
}

int tomos_csr_transpose(int val_type, int idx_type, int64_t m, int64_t n, int64_t nnz, const void* vals,
                        const void* row_ptrs, const void* col_idxs, void* vals_t, void* row_ptrs_t, void* col_idxs_t,
                        void* scratch, size_t scratch_bytes, tomos_stream_t st) {
    return guarded([&] {
        const CsrMat A = make_csr(val_type, idx_type, m, n, nnz, vals, row_ptrs, col_idxs);
        csr_require_valid_types(A, "tomos_csr_transpose");
        csr_check(A, fallback_scratch(csr_check_scratch_bytes()), csr_check_scratch_bytes(), st);
        const size_t need = csr_transpose_scratch_bytes(A, std::max<int64_t>(m, 1));
        if (!scratch) {
            scratch = fallback_scratch(need);
            scratch_bytes = need;
        }
        csr_transpose(A, std::max<int64_t>(m, 1), vals_t, row_ptrs_t, col_idxs_t, scratch, scratch_bytes, st);
    });
}

size_t tomos_csr_mv_scratch_size(int val_type, int idx_type, int trans, int64_t m, int64_t n, int64_t nnz) {
    if (m < 0 || n < 0 || nnz < 0) return 0;
    const CsrMat A = make_csr(val_type, idx_type, m, n, nnz, nullptr, nullptr, nullptr);
    if (!trans) return std::max<size_t>(256, csr_mv_scratch_bytes(A));
    try {
        // sized for FP64 x: the most panels
        const int64_t pr = csr_panel_rows(A, TOMOS_F64), panels = csr_panels(A, pr);
        const CsrMat T = make_csr(val_type, idx_type, panels * n, m, nnz, nullptr, nullptr, nullptr);
        const size_t e = static_cast<size_t>(std::max<int64_t>(nnz, 1));
        return align256(e * csr_value_bytes(val_type)) + align256(e * csr_index_bytes(idx_type)) +
               align256(static_cast<size_t>(panels * n + 1) * csr_index_bytes(idx_type)) +
               std::max(csr_transpose_scratch_bytes(A, pr), csr_tmv_scratch_bytes(T, panels, n));
    } catch (...) {
        return 0;
    }
}

int tomos_csr_mv(int val_type, int idx_type, int trans, double alpha, int64_t m, int64_t n, int64_t nnz,
                 const void* vals, const void* row_ptrs, const void* col_idxs, int x_type, const void* x, double beta,
                 int y_type, void* y, void* scratch, size_t scratch_bytes, tomos_stream_t st) {
    return guarded([&] {
        const CsrMat A = make_csr(val_type, idx_type, m, n, nnz, vals, row_ptrs, col_idxs);
        csr_require_valid_types(A, "tomos_csr_mv");
        require(trans == TOMOS_NO_TRANSPOSE || trans == TOMOS_TRANSPOSE, "tomos_csr_mv: bad trans");
        Coef c;
        c.alpha = alpha;
        c.beta = beta;
        const size_t need = tomos_csr_mv_scratch_size(val_type, idx_type, trans, m, n, nnz);
        if (!scratch) {
            scratch = fallback_scratch(need);
            scratch_bytes = need;
        }
        require(scratch_bytes >= need, "tomos_csr_mv: scratch too small");
        if (trans == TOMOS_NO_TRANSPOSE) {
            csr_mv(A, c, x_type, x, y_type, y, scratch, scratch_bytes, st);
            return;
        }
        // A^T x: transpose into scratch, then the row kernels on A^T
        csr_check(A, scratch, scratch_bytes, st);
        require(x_type == TOMOS_F32 || x_type == TOMOS_F64, "tomos_csr_mv: x must be F32 or F64");
        const int64_t pr = csr_panel_rows(A, x_type), panels = csr_panels(A, pr);
        const size_t e = static_cast<size_t>(std::max<int64_t>(nnz, 1));
        char* sc = static_cast<char*>(scratch);
        void* vt = sc;
        sc += align256(e * csr_value_bytes(val_type));
        void* it = sc;
        sc += align256(e * csr_index_bytes(idx_type));
        void* pt = sc;
        sc += align256(static_cast<size_t>(panels * n + 1) * csr_index_bytes(idx_type));
        csr_transpose(A, pr, vt, pt, it, sc, csr_transpose_scratch_bytes(A, pr), st);
        const CsrMat T = make_csr(val_type, idx_type, panels * n, m, nnz, vt, pt, it);
        csr_tmv(T, panels, pr, n, c, x_type, x, y_type, y, sc, csr_tmv_scratch_bytes(T, panels, n), st);
    });
}

int tomos_csr_mm(int val_type, int idx_type, double alpha, int64_t m, int64_t n, int64_t nnz, const void* vals,
                 const void* row_ptrs, const void* col_idxs, int64_t ncols, int x_type, const void* X, int64_t ldx,
                 double beta, int y_type, void* Y, int64_t ldy, tomos_stream_t st) {
    return guarded([&] {
        const CsrMat A = make_csr(val_type, idx_type, m, n, nnz, vals, row_ptrs, col_idxs);
        csr_mm(A, ncols, x_type, X, ldx, y_type, Y, ldy, alpha, beta, st);
    });
}

}  // extern "C"
