// Dense GEMV on plain and packed column-major matrices, for gfx950.
//
// Storage formats: F32, F64, FP32_16 and FP32_BF16 (planar wh/wr uint16
// planes).  Tiers (see tomos_hip.h): DEFAULT decodes wh only and
// accumulates in FP32; FAST decodes wh only and accumulates in FP64; EXACT
// rebuilds the exact FP32 value from wh+wr and accumulates in FP64.
//
// y = A x   (NO_TRANSPOSE): each thread owns one 16-byte vector of rows
//           (8 packed / 4 fp32 / 2 fp64 rows) and streams every column of
//           its block's column range, 8 columns per batch so 8 independent
//           16-B loads are in flight per lane.  x[j] is wave-uniform and is
//           read through the scalar cache.  When m is too small to fill the
//           chip, columns are split across blocks and a second pass sums the
//           per-split partials in a fixed order.
// z = A^T y (TRANSPOSE): a block owns a chunk of rows and a group of CB
//           columns; each lane keeps CB column accumulators in VGPRs and
//           walks the chunk in 16-B row vectors (loads coalesced along the
//           column), the wave reduces the CB accumulators with a transpose
//           butterfly (CB-1 shuffles per lane instead of 6*CB), the four
//           waves combine through LDS, and the per-chunk partials are summed
//           in a fixed order by a second pass.  No atomics anywhere: results
//           are bit-reproducible run to run.
#include <algorithm>

#include "codec.hpp"
#include "common.hpp"
#include "gemv.hpp"
#include "tomos_hip.h"

namespace tomos_hip {

namespace {

// FMT_PBF with DEC_HI also serves native BF16 storage (the hi plane is a
// bf16 array); FMT_H16 is native IEEE binary16 storage.
enum { FMT_F32 = 0, FMT_F64 = 1, FMT_P16 = 2, FMT_PBF = 3, FMT_H16 = 4 };
enum { DEC_NATIVE = 0, DEC_HI = 1, DEC_EXACT = 2 };

constexpr int kThreads = 256;

template <int FMT>
constexpr int vec_full() {
    return FMT == FMT_F64 ? 2 : (FMT == FMT_F32 ? 4 : 8);
}
template <int FMT>
constexpr bool is16() {
    return FMT == FMT_P16 || FMT == FMT_PBF || FMT == FMT_H16;
}

// Raw (undecoded) bits of VEC consecutive elements.
template <int FMT, int DEC, int VEC>
struct Raw {
    uint32_t h[4];
    uint32_t r[4];
};

template <int FMT, int DEC, int VEC, bool NTL>
__device__ __forceinline__ void load_raw(const void* __restrict__ a,
                                         const uint16_t* __restrict__ wr, int64_t idx,
                                         Raw<FMT, DEC, VEC>& R) {
    if constexpr (is16<FMT>()) {
        const uint16_t* p = static_cast<const uint16_t*>(a) + idx;
        if constexpr (VEC == 8) {
            const uint4 t = load_a16<NTL>(p);
            R.h[0] = t.x; R.h[1] = t.y; R.h[2] = t.z; R.h[3] = t.w;
            if constexpr (DEC == DEC_EXACT) {
                const uint4 q = load_a16<NTL>(wr + idx);
                R.r[0] = q.x; R.r[1] = q.y; R.r[2] = q.z; R.r[3] = q.w;
            }
        } else {
            R.h[0] = *p;
            if constexpr (DEC == DEC_EXACT) R.r[0] = wr[idx];
        }
    } else if constexpr (FMT == FMT_F32) {
        const float* p = static_cast<const float*>(a) + idx;
        if constexpr (VEC == 4) {
            const uint4 t = load_a16<NTL>(p);
            R.h[0] = t.x; R.h[1] = t.y; R.h[2] = t.z; R.h[3] = t.w;
        } else {
            R.h[0] = __float_as_uint(*p);
        }
    } else {
        const double* p = static_cast<const double*>(a) + idx;
        if constexpr (VEC == 2) {
            const uint4 t = load_a16<NTL>(p);
            R.h[0] = t.x; R.h[1] = t.y; R.h[2] = t.z; R.h[3] = t.w;
        } else {
            const uint2 t = *reinterpret_cast<const uint2*>(p);
            R.h[0] = t.x; R.h[1] = t.y;
        }
    }
}

__device__ __forceinline__ float half_bits_to_float(uint32_t hb) {
    return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(hb)));
}

template <int FMT, int DEC>
__device__ __forceinline__ float decode16(uint32_t h, uint32_t r) {
    if constexpr (FMT == FMT_H16) {
        return half_bits_to_float(h);
    } else if constexpr (FMT == FMT_P16) {
        if constexpr (DEC == DEC_EXACT) return codec::fp32_16_decode_exact(h, r);
        else return half_bits_to_float(codec::fp32_16_half_bits(h));
    } else {
        if constexpr (DEC == DEC_EXACT) return __uint_as_float((h << 16) | r);
        else return __uint_as_float(h << 16);
    }
}

template <int FMT, int DEC, int VEC, class ACC>
__device__ __forceinline__ void decode(const Raw<FMT, DEC, VEC>& R, ACC (&out)[VEC]) {
    if constexpr (is16<FMT>()) {
        if constexpr (VEC == 8) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t h = R.h[k];
                const uint32_t r = (DEC == DEC_EXACT) ? R.r[k] : 0u;
                out[2 * k] = static_cast<ACC>(decode16<FMT, DEC>(h & 0xFFFFu, r & 0xFFFFu));
                out[2 * k + 1] = static_cast<ACC>(decode16<FMT, DEC>(h >> 16, r >> 16));
            }
        } else {
            out[0] = static_cast<ACC>(decode16<FMT, DEC>(R.h[0], (DEC == DEC_EXACT) ? R.r[0] : 0u));
        }
    } else if constexpr (FMT == FMT_F32) {
#pragma unroll
        for (int k = 0; k < VEC; ++k) out[k] = static_cast<ACC>(__uint_as_float(R.h[k]));
    } else {
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const uint64_t b = (static_cast<uint64_t>(R.h[2 * k + 1]) << 32) | R.h[2 * k];
            out[k] = static_cast<ACC>(__longlong_as_double(static_cast<long long>(b)));
        }
    }
}

__device__ __forceinline__ double read_coef(double host, const double* dev, double mul) {
    return dev ? mul * (*dev) : host;
}

template <class ACC>
__device__ __forceinline__ ACC load_x(const void* __restrict__ x, bool x64, int64_t j) {
    return x64 ? static_cast<ACC>(static_cast<const double*>(x)[j])
               : static_cast<ACC>(static_cast<const float*>(x)[j]);
}

// Stores alpha*acc + beta*y (y not read when beta == 0); returns the
// stored value.
__device__ __forceinline__ double store_y(void* __restrict__ y, bool y64, int64_t i, double alpha,
                                          double acc, double beta) {
    if (y64) {
        double* p = static_cast<double*>(y) + i;
        double r = alpha * acc;
        if (beta != 0.0) r += beta * (*p);
        *p = r;
        return r;
    } else {
        float* p = static_cast<float*>(y) + i;
        double r = alpha * acc;
        if (beta != 0.0) r += beta * static_cast<double>(*p);
        const float f = static_cast<float>(r);
        *p = f;
        return static_cast<double>(f);
    }
}

// Block-wide FP64 sum, fixed order; every thread of the block must call it.
__device__ __forceinline__ void block_sumsq_out(double v, double* out) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    __shared__ double part[kThreads / 64];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) *out = (part[0] + part[1]) + (part[2] + part[3]);
}

// ------------------------------------------------------------ NO_TRANSPOSE
template <int FMT, int DEC, int VEC, class ACC, bool NTL>
__global__ __launch_bounds__(kThreads) void gemv_n_kernel(
    int64_t m, int64_t n, const void* __restrict__ a, const uint16_t* __restrict__ wr,
    int64_t lda, const void* __restrict__ x, bool x64, void* __restrict__ y, bool y64, Coef coef,
    int64_t col_chunk, ACC* __restrict__ partial) {
    if (coef.skip && *coef.skip) return;
    const int64_t i0 = (static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x) * VEC;
    const int64_t j0 = static_cast<int64_t>(blockIdx.y) * col_chunk;
    const int64_t j1 = min(n, j0 + col_chunk);
    ACC acc[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[v] = ACC(0);
    if (i0 < m) {
        constexpr int U = 8;
        int64_t j = j0;
        for (; j + U <= j1; j += U) {
            Raw<FMT, DEC, VEC> R[U];
#pragma unroll
            for (int u = 0; u < U; ++u) load_raw<FMT, DEC, VEC, NTL>(a, wr, i0 + (j + u) * lda, R[u]);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ACC av[VEC];
                decode<FMT, DEC, VEC, ACC>(R[u], av);
                const ACC xv = load_x<ACC>(x, x64, j + u);
#pragma unroll
                for (int v = 0; v < VEC; ++v) acc[v] = fma(av[v], xv, acc[v]);
            }
        }
        for (; j < j1; ++j) {
            Raw<FMT, DEC, VEC> R;
            load_raw<FMT, DEC, VEC, NTL>(a, wr, i0 + j * lda, R);
            ACC av[VEC];
            decode<FMT, DEC, VEC, ACC>(R, av);
            const ACC xv = load_x<ACC>(x, x64, j);
#pragma unroll
            for (int v = 0; v < VEC; ++v) acc[v] = fma(av[v], xv, acc[v]);
        }
    }
    if (partial) {
        if (i0 >= m) return;
#pragma unroll
        for (int v = 0; v < VEC; ++v)
            if (i0 + v < m) partial[static_cast<int64_t>(blockIdx.y) * m + i0 + v] = acc[v];
        return;
    }
    const double al = read_coef(coef.alpha, coef.dalpha, coef.alpha_mul);
    const double be = read_coef(coef.beta, coef.dbeta, coef.beta_mul);
    double sq = 0.0;
    if (i0 < m) {
#pragma unroll
        for (int v = 0; v < VEC; ++v)
            if (i0 + v < m) {
                const double r = store_y(y, y64, i0 + v, al, static_cast<double>(acc[v]), be);
                sq += r * r;
            }
    }
    if (coef.sumsq) block_sumsq_out(sq, coef.sumsq + blockIdx.x);
}

template <class ACC>
__global__ __launch_bounds__(kThreads) void gemv_n_finish(int64_t m, int splits,
                                                          const ACC* __restrict__ partial,
                                                          void* __restrict__ y, bool y64,
                                                          Coef coef) {
    if (coef.skip && *coef.skip) return;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    double sq = 0.0;
    if (i < m) {
        // four chains: the partial sums' loads overlap
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int c = 0;
        for (; c + 3 < splits; c += 4) {
            s0 += static_cast<double>(partial[static_cast<int64_t>(c) * m + i]);
            s1 += static_cast<double>(partial[static_cast<int64_t>(c + 1) * m + i]);
            s2 += static_cast<double>(partial[static_cast<int64_t>(c + 2) * m + i]);
            s3 += static_cast<double>(partial[static_cast<int64_t>(c + 3) * m + i]);
        }
        for (; c < splits; ++c) s0 += static_cast<double>(partial[static_cast<int64_t>(c) * m + i]);
        const double s = (s0 + s1) + (s2 + s3);
        const double r = store_y(y, y64, i, read_coef(coef.alpha, coef.dalpha, coef.alpha_mul), s,
                                 read_coef(coef.beta, coef.dbeta, coef.beta_mul));
        sq = r * r;
    }
    if (coef.sumsq) block_sumsq_out(sq, coef.sumsq + blockIdx.x);
}

// --------------------------------------------------------------- TRANSPOSE
// Transpose-reduce across a wave: L values per lane in, after log2(L)
// exchange stages lane l holds the wave sum of column l / (64 / L).
template <int L, int OFF, class T>
struct WaveTR {
    static __device__ __forceinline__ void run(T* v, int lane) {
        const bool upper = (lane & OFF) != 0;
#pragma unroll
        for (int t = 0; t < L / 2; ++t) {
            const T send = upper ? v[t] : v[t + L / 2];
            const T keep = upper ? v[t + L / 2] : v[t];
            v[t] = keep + __shfl_xor(send, OFF);
        }
        WaveTR<L / 2, OFF / 2, T>::run(v, lane);
    }
};
template <int OFF, class T>
struct WaveTR<1, OFF, T> {
    static __device__ __forceinline__ void run(T* v, int) {
#pragma unroll
        for (int o = OFF; o >= 1; o >>= 1) v[0] += __shfl_xor(v[0], o);
    }
};

template <int FMT, int DEC, int VEC, class ACC, int CB, bool NTL>
__global__ __launch_bounds__(kThreads) void gemv_t_kernel(
    int64_t m, int64_t n, const void* __restrict__ a, const uint16_t* __restrict__ wr,
    int64_t lda, const void* __restrict__ x, bool x64, int64_t rows_per_chunk, int ncg,
    ACC* __restrict__ partial, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int bid = blockIdx.x;
    const int chunk = bid / ncg;
    const int cg = bid - chunk * ncg;
    const int64_t r0 = static_cast<int64_t>(chunk) * rows_per_chunk;
    const int64_t r1 = min(m, r0 + rows_per_chunk);
    const int64_t c0 = static_cast<int64_t>(cg) * CB;
    const int ncols = static_cast<int>(min<int64_t>(CB, n - c0));

    ACC acc[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[c] = ACC(0);

    constexpr int UB = (sizeof(ACC) == 8 || DEC == DEC_EXACT) ? 8 : 16;
    for (int64_t i = r0 + static_cast<int64_t>(threadIdx.x) * VEC; i < r1;
         i += static_cast<int64_t>(kThreads) * VEC) {
        const int nvalid = static_cast<int>(min<int64_t>(VEC, m - i));
        ACC xv[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) xv[v] = (v < nvalid) ? load_x<ACC>(x, x64, i + v) : ACC(0);
        const int64_t base = i + c0 * lda;
#pragma unroll
        for (int cb = 0; cb < CB; cb += UB) {
            Raw<FMT, DEC, VEC> R[UB];
#pragma unroll
            for (int u = 0; u < UB; ++u)
                if (cb + u < ncols) load_raw<FMT, DEC, VEC, NTL>(a, wr, base + (cb + u) * lda, R[u]);
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                if (cb + u < ncols) {
                    ACC av[VEC];
                    decode<FMT, DEC, VEC, ACC>(R[u], av);
                    if (nvalid < VEC) {
#pragma unroll
                        for (int v = 0; v < VEC; ++v)
                            if (v >= nvalid) av[v] = ACC(0);
                    }
                    ACC s = acc[cb + u];
#pragma unroll
                    for (int v = 0; v < VEC; ++v) s = fma(av[v], xv[v], s);
                    acc[cb + u] = s;
                }
            }
        }
    }

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    WaveTR<CB, 32, ACC>::run(acc, lane);
    __shared__ ACC red[kThreads / 64][CB];
    constexpr int kGroup = 64 / CB;
    if ((lane % kGroup) == 0) red[wave][lane / kGroup] = acc[0];
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < ncols) {
        const int c = threadIdx.x;
        const ACC s = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
        partial[static_cast<int64_t>(chunk) * n + c0 + c] = s;
    }
}

template <class ACC>
__global__ __launch_bounds__(kThreads) void gemv_t_finish(int64_t n, int chunks,
                                                          const ACC* __restrict__ partial,
                                                          void* __restrict__ y, bool y64,
                                                          Coef coef) {
    if (coef.skip && *coef.skip) return;
    const int64_t j = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (j >= n) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int c = 0;
    for (; c + 3 < chunks; c += 4) {
        s0 += static_cast<double>(partial[static_cast<int64_t>(c) * n + j]);
        s1 += static_cast<double>(partial[static_cast<int64_t>(c + 1) * n + j]);
        s2 += static_cast<double>(partial[static_cast<int64_t>(c + 2) * n + j]);
        s3 += static_cast<double>(partial[static_cast<int64_t>(c + 3) * n + j]);
    }
    for (; c < chunks; ++c) s0 += static_cast<double>(partial[static_cast<int64_t>(c) * n + j]);
    const double s = (s0 + s1) + (s2 + s3);
    store_y(y, y64, j, read_coef(coef.alpha, coef.dalpha, coef.alpha_mul), s,
            read_coef(coef.beta, coef.dbeta, coef.beta_mul));
}

// ------------------------------------------------------------------ plans
struct Kind {
    int fmt, dec;
    bool acc64;
};

Kind classify(int a_type, int tier) {
    require(tier == TOMOS_TIER_DEFAULT || tier == TOMOS_TIER_FAST || tier == TOMOS_TIER_EXACT,
            "gemv: unknown tier");
    switch (a_type) {
        case TOMOS_F32: return {FMT_F32, DEC_NATIVE, tier != TOMOS_TIER_DEFAULT};
        case TOMOS_F64: return {FMT_F64, DEC_NATIVE, true};
        case TOMOS_FP32_16:
            return {FMT_P16, tier == TOMOS_TIER_EXACT ? DEC_EXACT : DEC_HI, tier != TOMOS_TIER_DEFAULT};
        case TOMOS_FP32_BF16:
            return {FMT_PBF, tier == TOMOS_TIER_EXACT ? DEC_EXACT : DEC_HI, tier != TOMOS_TIER_DEFAULT};
        // native 16-bit storage: the stored value is the exact one, so EXACT
        // (like FAST) decodes it and accumulates in FP64
        case TOMOS_BF16: return {FMT_PBF, DEC_HI, tier != TOMOS_TIER_DEFAULT};
        case TOMOS_F16: return {FMT_H16, DEC_HI, tier != TOMOS_TIER_DEFAULT};
        default:
            throw Error(kErrUnsupported, "gemv: unsupported matrix type " + std::to_string(a_type));
    }
}

int vec_full_of(int fmt) { return fmt == FMT_F64 ? 2 : (fmt == FMT_F32 ? 4 : 8); }
int cb_of(const Kind& k) { return k.acc64 ? 32 : 64; }
constexpr int64_t kTargetBlocks = 1024;

struct PlanN {
    int64_t bx;
    int splits;
    int64_t col_chunk;
};
PlanN plan_n(const Kind& k, int64_t m, int64_t n) {
    PlanN p;
    p.bx = ceil_div(m, kThreads * vec_full_of(k.fmt));
    p.splits = 1;
    p.col_chunk = std::max<int64_t>(n, 1);
    if (p.bx < kTargetBlocks / 2 && n > 16) {
        int64_t want = ceil_div(kTargetBlocks, p.bx);
        // at most 32 column splits: the finishing pass adds them per row in
        // sequence, and more blocks than that on a small matrix cost more
        // there than they save in the pass over A
        int64_t max_splits = std::min<int64_t>(ceil_div(n, 8), 32);
        int64_t s = std::min(want, max_splits);
        p.col_chunk = ceil_div(n, s);
        p.splits = static_cast<int>(ceil_div(n, p.col_chunk));
    }
    return p;
}

struct PlanT {
    int ncg;
    int64_t rows_per_chunk;
    int chunks;
};
PlanT plan_t(const Kind& k, int64_t m, int64_t n) {
    PlanT p;
    const int64_t cb = cb_of(k);
    p.ncg = static_cast<int>(ceil_div(std::max<int64_t>(n, 1), cb));
    const int64_t step = static_cast<int64_t>(kThreads) * vec_full_of(k.fmt);
    const int64_t row_steps = std::max<int64_t>(ceil_div(m, step), 1);
    const int64_t want_chunks = std::max<int64_t>(1, ceil_div(kTargetBlocks, p.ncg));
    const int64_t kr = std::min<int64_t>(std::max<int64_t>(ceil_div(row_steps, want_chunks), 1), 64);
    p.rows_per_chunk = kr * step;
    p.chunks = static_cast<int>(std::max<int64_t>(ceil_div(m, p.rows_per_chunk), 1));
    return p;
}

// The kind a GEMV runs with: F32 GEMV^T accumulates in FP64 in every tier
// (the FP64-sum kernel streams A faster on gfx950 than the FP32-sum one,
// and it is the more accurate; the reference's tiled FP32 GEMV accumulates
// in FP64 too, mv_cuda_kernels.cu:863-990).
Kind kind_for(int a_type, int tier, int trans) {
    Kind k = classify(a_type, tier);
    if (trans == TOMOS_TRANSPOSE && k.fmt == FMT_F32) k.acc64 = true;
    return k;
}

size_t scratch_for(const Kind& k, int trans, int64_t m, int64_t n) {
    const size_t acc = k.acc64 ? 8 : 4;
    if (trans == TOMOS_NO_TRANSPOSE) {
        PlanN p = plan_n(k, m, n);
        return p.splits > 1 ? static_cast<size_t>(p.splits) * static_cast<size_t>(m) * acc : 0;
    }
    PlanT p = plan_t(k, m, n);
    return static_cast<size_t>(p.chunks) * static_cast<size_t>(n) * acc;
}

// Matrix bytes one GEMV pass reads.
template <int FMT, int DEC>
int64_t matrix_bytes(int64_t m, int64_t n) {
    const int64_t esz = FMT == FMT_F64 ? 8 : (FMT == FMT_F32 ? 4 : (DEC == DEC_EXACT ? 4 : 2));
    return m * n * esz;
}

template <int FMT, int DEC, class ACC>
int run_n(int64_t m, int64_t n, const void* a, const uint16_t* wr, int64_t lda, const void* x,
          bool x64, void* y, bool y64, const Coef& coef, void* scratch, bool vec,
          const PlanN& p, hipStream_t s) {
    constexpr int VF = vec_full<FMT>();
    ACC* part = p.splits > 1 ? static_cast<ACC*>(scratch) : nullptr;
    const int width = vec ? VF : 1;
    const dim3 grid(static_cast<unsigned>(ceil_div(m, static_cast<int64_t>(kThreads) * width)),
                    static_cast<unsigned>(p.splits));
    if (width == VF && stream_loads(matrix_bytes<FMT, DEC>(m, n)))
        gemv_n_kernel<FMT, DEC, VF, ACC, true><<<grid, kThreads, 0, s>>>(m, n, a, wr, lda, x, x64, y, y64,
                                                                         coef, p.col_chunk, part);
    else if (width == VF)
        gemv_n_kernel<FMT, DEC, VF, ACC, false><<<grid, kThreads, 0, s>>>(m, n, a, wr, lda, x, x64, y, y64,
                                                                          coef, p.col_chunk, part);
    else
        gemv_n_kernel<FMT, DEC, 1, ACC, false><<<grid, kThreads, 0, s>>>(m, n, a, wr, lda, x, x64, y, y64,
                                                                         coef, p.col_chunk, part);
    TOMOS_LAUNCH_CHECK("gemv_n_kernel");
    if (part) {
        gemv_n_finish<ACC><<<dim3(static_cast<unsigned>(ceil_div(m, kThreads))), kThreads, 0, s>>>(
            m, p.splits, part, y, y64, coef);
        TOMOS_LAUNCH_CHECK("gemv_n_finish");
        return static_cast<int>(ceil_div(m, kThreads));
    }
    return static_cast<int>(grid.x);
}

template <int FMT, int DEC, class ACC>
void run_t(int64_t m, int64_t n, const void* a, const uint16_t* wr, int64_t lda, const void* x,
           bool x64, void* y, bool y64, const Coef& coef, void* scratch, bool vec,
           const PlanT& p, hipStream_t s, bool finish = true) {
    constexpr int VF = vec_full<FMT>();
    constexpr int CB = sizeof(ACC) == 8 ? 32 : 64;
    ACC* part = static_cast<ACC*>(scratch);
    const int64_t blocks = static_cast<int64_t>(p.chunks) * p.ncg;
    require(blocks < (int64_t(1) << 31), "gemv: grid too large");
    if (m > 0) {
        const dim3 grid(static_cast<unsigned>(blocks));
        // the EXACT tier reads A with cacheable loads: its non-temporal form
        // needs 256 VGPRs (one wave per SIMD), the cacheable one ~140
        if (vec && DEC != DEC_EXACT && stream_loads(matrix_bytes<FMT, DEC>(m, n)))
            gemv_t_kernel<FMT, DEC, VF, ACC, CB, true><<<grid, kThreads, 0, s>>>(
                m, n, a, wr, lda, x, x64, p.rows_per_chunk, p.ncg, part, coef.skip);
        else if (vec)
            gemv_t_kernel<FMT, DEC, VF, ACC, CB, false><<<grid, kThreads, 0, s>>>(
                m, n, a, wr, lda, x, x64, p.rows_per_chunk, p.ncg, part, coef.skip);
        else
            gemv_t_kernel<FMT, DEC, 1, ACC, CB, false><<<grid, kThreads, 0, s>>>(
                m, n, a, wr, lda, x, x64, p.rows_per_chunk, p.ncg, part, coef.skip);
        TOMOS_LAUNCH_CHECK("gemv_t_kernel");
    }
    if (!finish) return;
    gemv_t_finish<ACC><<<dim3(static_cast<unsigned>(ceil_div(n, kThreads))), kThreads, 0, s>>>(
        n, m > 0 ? p.chunks : 0, part, y, y64, coef);
    TOMOS_LAUNCH_CHECK("gemv_t_finish");
}

}  // namespace

size_t gemv_scratch_bytes(int a_type, int tier, int trans, int64_t m, int64_t n) {
    return scratch_for(kind_for(a_type, tier, trans), trans, m, n);
}

int64_t gemv_sumsq_blocks_max(int64_t m) { return ceil_div(std::max<int64_t>(m, 1), kThreads); }

void gemv(int a_type, int tier, int trans, const Coef& coef, int64_t m, int64_t n, const void* a,
          const uint16_t* wr, int64_t lda, int x_type, const void* x, int y_type, void* y,
          void* scratch, size_t scratch_bytes, hipStream_t s, int* sumsq_blocks) {
    require(trans == TOMOS_NO_TRANSPOSE || trans == TOMOS_TRANSPOSE, "gemv: bad trans");
    const Kind k = kind_for(a_type, tier, trans);
    require(m >= 0 && n >= 0, "gemv: negative size");
    require(lda >= std::max<int64_t>(m, 1), "gemv: lda < m");
    require(x_type == TOMOS_F32 || x_type == TOMOS_F64, "gemv: x must be F32 or F64");
    require(y_type == TOMOS_F32 || y_type == TOMOS_F64, "gemv: y must be F32 or F64");
    const int64_t rows_out = trans == TOMOS_NO_TRANSPOSE ? m : n;
    require(coef.sumsq == nullptr || trans == TOMOS_NO_TRANSPOSE, "gemv: sumsq needs NO_TRANSPOSE");
    if (sumsq_blocks) *sumsq_blocks = 0;
    if (rows_out == 0) return;
    require(y != nullptr, "gemv: null y");
    const int64_t inner = trans == TOMOS_NO_TRANSPOSE ? n : m;
    if (inner > 0) require(a != nullptr && x != nullptr, "gemv: null A or x");
    if (k.dec == DEC_EXACT) require(wr != nullptr || inner == 0, "gemv: EXACT tier needs the wr plane");
    const size_t need = scratch_for(k, trans, m, n);
    if (need > 0) {
        if (scratch == nullptr) {
            scratch = fallback_scratch(need);
        } else {
            require(scratch_bytes >= need, "gemv: scratch too small (need " + std::to_string(need) +
                                               " bytes)");
        }
    }
    const bool x64 = x_type == TOMOS_F64, y64 = y_type == TOMOS_F64;
    const int vf = vec_full_of(k.fmt);
    const size_t vb = 16;
    bool vec = (lda % vf == 0) && (reinterpret_cast<uintptr_t>(a) % vb == 0) &&
               (k.dec != DEC_EXACT || reinterpret_cast<uintptr_t>(wr) % vb == 0);

    if (trans == TOMOS_NO_TRANSPOSE) {
        // n == 0 falls through with an empty column range: y = beta*y.
        const PlanN p = plan_n(k, m, n);
int nb = 0;
#define RUN_N(F, D, A) nb = run_n<F, D, A>(m, n, a, wr, lda, x, x64, y, y64, coef, scratch, vec, p, s)
        if (k.fmt == FMT_F32) { if (k.acc64) RUN_N(FMT_F32, DEC_NATIVE, double); else RUN_N(FMT_F32, DEC_NATIVE, float); }
        else if (k.fmt == FMT_F64) RUN_N(FMT_F64, DEC_NATIVE, double);
        else if (k.fmt == FMT_H16) { if (k.acc64) RUN_N(FMT_H16, DEC_HI, double); else RUN_N(FMT_H16, DEC_HI, float); }
        else if (k.fmt == FMT_P16) {
            if (k.dec == DEC_EXACT) RUN_N(FMT_P16, DEC_EXACT, double);
            else if (k.acc64) RUN_N(FMT_P16, DEC_HI, double);
            else RUN_N(FMT_P16, DEC_HI, float);
        } else {
            if (k.dec == DEC_EXACT) RUN_N(FMT_PBF, DEC_EXACT, double);
            else if (k.acc64) RUN_N(FMT_PBF, DEC_HI, double);
            else RUN_N(FMT_PBF, DEC_HI, float);
        }
#undef RUN_N
        if (sumsq_blocks) *sumsq_blocks = nb;
    } else {
        const PlanT p = plan_t(k, m, n);
#define RUN_T(F, D, A) run_t<F, D, A>(m, n, a, wr, lda, x, x64, y, y64, coef, scratch, vec, p, s)
        if (k.fmt == FMT_F32) { if (k.acc64) RUN_T(FMT_F32, DEC_NATIVE, double); else RUN_T(FMT_F32, DEC_NATIVE, float); }
        else if (k.fmt == FMT_F64) RUN_T(FMT_F64, DEC_NATIVE, double);
        else if (k.fmt == FMT_H16) { if (k.acc64) RUN_T(FMT_H16, DEC_HI, double); else RUN_T(FMT_H16, DEC_HI, float); }
        else if (k.fmt == FMT_P16) {
            if (k.dec == DEC_EXACT) RUN_T(FMT_P16, DEC_EXACT, double);
            else if (k.acc64) RUN_T(FMT_P16, DEC_HI, double);
            else RUN_T(FMT_P16, DEC_HI, float);
        } else {
            if (k.dec == DEC_EXACT) RUN_T(FMT_PBF, DEC_EXACT, double);
            else if (k.acc64) RUN_T(FMT_PBF, DEC_HI, double);
            else RUN_T(FMT_PBF, DEC_HI, float);
        }
#undef RUN_T
    }
}

int gemv_t_partials(int a_type, int tier, int64_t m, int64_t n, const void* a, const uint16_t* wr,
                    int64_t lda, int x_type, const void* x, void* scratch, size_t scratch_bytes,
                    const int* skip, hipStream_t s) {
    const Kind k = classify(a_type, tier);
    require(k.acc64, "gemv_t_partials: needs an FP64-accumulating tier");
    require(m >= 1 && n >= 1 && lda >= m, "gemv_t_partials: bad sizes");
    const size_t need = scratch_for(k, TOMOS_TRANSPOSE, m, n);
    require(scratch != nullptr && scratch_bytes >= need, "gemv_t_partials: scratch too small");
    const PlanT p = plan_t(k, m, n);
    const bool x64 = x_type == TOMOS_F64;
    const int vf = vec_full_of(k.fmt);
    const bool vec = (lda % vf == 0) && (reinterpret_cast<uintptr_t>(a) % 16 == 0) &&
                     (k.dec != DEC_EXACT || reinterpret_cast<uintptr_t>(wr) % 16 == 0);
    Coef c;
    c.skip = skip;
#define RUN_TP(F, D) run_t<F, D, double>(m, n, a, wr, lda, x, x64, nullptr, false, c, scratch, vec, p, s, false)
    if (k.fmt == FMT_F32) RUN_TP(FMT_F32, DEC_NATIVE);
    else if (k.fmt == FMT_F64) RUN_TP(FMT_F64, DEC_NATIVE);
    else if (k.fmt == FMT_H16) RUN_TP(FMT_H16, DEC_HI);
    else if (k.fmt == FMT_P16) { if (k.dec == DEC_EXACT) RUN_TP(FMT_P16, DEC_EXACT); else RUN_TP(FMT_P16, DEC_HI); }
    else { if (k.dec == DEC_EXACT) RUN_TP(FMT_PBF, DEC_EXACT); else RUN_TP(FMT_PBF, DEC_HI); }
#undef RUN_TP
    return p.chunks;
}

}  // namespace tomos_hip

using namespace tomos_hip;

extern "C" {

size_t tomos_gemv_scratch_size(int a_type, int tier, int trans, int64_t m, int64_t n) {
    size_t out = 0;
    int rc = guarded([&] {
        require(m >= 0 && n >= 0, "tomos_gemv_scratch_size: negative size");
        out = gemv_scratch_bytes(a_type, tier, trans, m, n);
    });
    return rc == 0 ? out : 0;
}

int tomos_gemv(int a_type, int tier, int trans, double alpha, int64_t m, int64_t n, const void* a,
               const uint16_t* wr, int64_t lda, int x_type, const void* x, double beta,
               int y_type, void* y, void* scratch, size_t scratch_bytes, tomos_stream_t s) {
    return guarded([&] {
        Coef c;
        c.alpha = alpha;
        c.beta = beta;
        gemv(a_type, tier, trans, c, m, n, a, wr, lda, x_type, x, y_type, y, scratch,
             scratch_bytes, s);
    });
}

}  // extern "C"
