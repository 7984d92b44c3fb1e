// Hashed CountSketch Y = S * A on the GPU.
//
// S is the s x m sketch with one +-1 per column: row i of A goes to bucket
// h(i) with sign g(i), both from splitmix64 of i xor a fixed seed
// (src/kernels/cpu/sketch/countsketch_kernels.cpp:59-85).
//
// Apply: A (m x n, column-major, any storage format/tier) is streamed once
// in its natural order.  A block (1024 threads) owns (row split r, column
// group of CG columns) and keeps its CG x s slice of Y in LDS (column-major,
// so the random buckets of a wave spread over the LDS banks).  Each lane
// takes one 16-byte vector of consecutive rows (8 packed / 4 FP32 / 2 FP64
// rows), loads their packed bucket|sign words once, then for each of the CG
// columns loads the same rows (hi plane, plus lo plane for the exact tier)
// and adds sign*A[i,j] into LDS: CG (or 2 CG) independent 16-byte loads per
// lane are in flight, and the index words are read once per column group.
// The LDS slice is written to a per-split partial, and a second pass sums
// the splits in a fixed order in FP64.  Bucket/sign indices are bit-exact;
// the values are FP32 sums within a split (order depends on LDS atomic
// order), FP64 across splits.  A, lda or the split start not 16-byte
// aligned takes the element-wise kernel.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "codec.hpp"
#include "common.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

constexpr uint64_t kBucketSeed = 0x12345678abcdef00ull;
constexpr uint64_t kSignSeed = 0xf00dd00d42babe55ull;

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// packed word: bucket in bits 0..30, bit 31 set for sign -1
__global__ __launch_bounds__(256) void cs_index_kernel(int64_t m, int64_t s,
                                                       int32_t* __restrict__ bucket,
                                                       int8_t* __restrict__ sign,
                                                       uint32_t* __restrict__ packed) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= m) return;
    const uint64_t hb = splitmix64(static_cast<uint64_t>(i) ^ kBucketSeed);
    const uint64_t hs = splitmix64(static_cast<uint64_t>(i) ^ kSignSeed);
    const uint32_t b = static_cast<uint32_t>(hb % static_cast<uint64_t>(s));
    const bool pos = (hs & 1ull) != 0;
    if (bucket) bucket[i] = static_cast<int32_t>(b);
    if (sign) sign[i] = pos ? 1 : -1;
    if (packed) packed[i] = b | (pos ? 0u : 0x80000000u);
}

enum { FMT_F32 = 0, FMT_F64 = 1, FMT_P16 = 2, FMT_PBF = 3 };

template <int FMT, bool EXACT>
__device__ __forceinline__ float load_elem(const void* __restrict__ a,
                                           const uint16_t* __restrict__ wr, int64_t idx) {
    if constexpr (FMT == FMT_F32) return static_cast<const float*>(a)[idx];
    else if constexpr (FMT == FMT_F64) return static_cast<float>(static_cast<const double*>(a)[idx]);
    else {
        const uint32_t h = static_cast<const uint16_t*>(a)[idx];
        if constexpr (FMT == FMT_P16) {
            if constexpr (EXACT) return codec::fp32_16_decode_exact(h, wr[idx]);
            else return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(codec::fp32_16_half_bits(h))));
        } else {
            if constexpr (EXACT) return __uint_as_float((h << 16) | wr[idx]);
            else return __uint_as_float(h << 16);
        }
    }
}

template <int FMT, bool EXACT, int CG>
__global__ __launch_bounds__(256) void cs_apply_kernel(int64_t m, int64_t n,
                                                       const void* __restrict__ a,
                                                       const uint16_t* __restrict__ wr,
                                                       int64_t lda, int64_t s,
                                                       const uint32_t* __restrict__ packed,
                                                       int64_t rows_per_split, int ncg,
                                                       float* __restrict__ partial) {
    extern __shared__ float ylds[];  // [s][CG]
    const int split = blockIdx.x / ncg;
    const int cg = blockIdx.x - split * ncg;
    const int64_t c0 = static_cast<int64_t>(cg) * CG;
    const int ncols = static_cast<int>(min<int64_t>(CG, n - c0));
    for (int64_t t = threadIdx.x; t < s * CG; t += 256) ylds[t] = 0.0f;
    __syncthreads();

    const int64_t r0 = static_cast<int64_t>(split) * rows_per_split;
    const int64_t r1 = min(m, r0 + rows_per_split);
    for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) {
        const uint32_t w = packed[i];
        const int b = static_cast<int>(w & 0x7FFFFFFFu);
        const float sg = (w & 0x80000000u) ? -1.0f : 1.0f;
#pragma unroll
        for (int c = 0; c < CG; ++c) {
            if (c < ncols) {
                const float v = load_elem<FMT, EXACT>(a, wr, i + (c0 + c) * lda);
                atomicAdd(&ylds[b * CG + c], sg * v);
            }
        }
    }
    __syncthreads();
    float* out = partial + static_cast<int64_t>(split) * s * n;
    for (int64_t t = threadIdx.x; t < s * ncols; t += 256) {
        const int64_t c = t / s;
        const int64_t b = t - c * s;
        out[(c0 + c) * s + b] = ylds[b * CG + c];
    }
}

template <int FMT>
constexpr int vec_rows() {
    return FMT == FMT_F64 ? 2 : (FMT == FMT_F32 ? 4 : 8);
}

template <int FMT, bool EXACT>
__device__ __forceinline__ void decode_rows(const uint4& h, const uint4& l,
                                            float (&v)[vec_rows<FMT>()]) {
    if constexpr (FMT == FMT_F64) {
        v[0] = static_cast<float>(__builtin_bit_cast(double, (static_cast<uint64_t>(h.y) << 32) | h.x));
        v[1] = static_cast<float>(__builtin_bit_cast(double, (static_cast<uint64_t>(h.w) << 32) | h.z));
    } else if constexpr (FMT == FMT_F32) {
        v[0] = __uint_as_float(h.x);
        v[1] = __uint_as_float(h.y);
        v[2] = __uint_as_float(h.z);
        v[3] = __uint_as_float(h.w);
    } else {
        const uint32_t hw[4] = {h.x, h.y, h.z, h.w};
        const uint32_t lw[4] = {l.x, l.y, l.z, l.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const uint32_t hh = half ? (hw[k] >> 16) : (hw[k] & 0xFFFFu);
                const uint32_t ll = half ? (lw[k] >> 16) : (lw[k] & 0xFFFFu);
                float x;
                if constexpr (FMT == FMT_P16) {
                    if constexpr (EXACT) x = codec::fp32_16_decode_exact(hh, ll);
                    else x = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(codec::fp32_16_half_bits(hh))));
                } else {
                    x = EXACT ? __uint_as_float((hh << 16) | ll) : __uint_as_float(hh << 16);
                }
                v[2 * k + half] = x;
            }
        }
    }
}

constexpr int kVecThreads = 1024;

// Vector path: rows_per_split, lda and the base pointers are multiples of
// the 16-byte row vector (checked by the host).
template <int FMT, bool EXACT, int CG>
__global__ __launch_bounds__(kVecThreads) void cs_apply_vec_kernel(
    int64_t m, int64_t n, const void* __restrict__ a, const uint16_t* __restrict__ wr, int64_t lda,
    int64_t s, const uint32_t* __restrict__ packed, int64_t rows_per_split, int ncg,
    float* __restrict__ partial) {
    constexpr int V = vec_rows<FMT>();
    constexpr int EB = 16 / V;  // bytes per element of the a plane
    extern __shared__ float ylds[];  // [CG][s]
    const int split = blockIdx.x / ncg;
    const int cg = blockIdx.x - split * ncg;
    const int64_t c0 = static_cast<int64_t>(cg) * CG;
    const int ncols = static_cast<int>(min<int64_t>(CG, n - c0));
    for (int64_t t = threadIdx.x; t < s * CG; t += kVecThreads) ylds[t] = 0.0f;
    __syncthreads();

    const int64_t r0 = static_cast<int64_t>(split) * rows_per_split;
    const int64_t r1 = min(m, r0 + rows_per_split);
    const char* abase = static_cast<const char*>(a);
    for (int64_t i0 = r0 + static_cast<int64_t>(threadIdx.x) * V; i0 < r1;
         i0 += static_cast<int64_t>(kVecThreads) * V) {
        if (i0 + V <= r1) {
            uint32_t w[V];
            if constexpr (V == 8) {
                const uint4 p0 = *reinterpret_cast<const uint4*>(packed + i0);
                const uint4 p1 = *reinterpret_cast<const uint4*>(packed + i0 + 4);
                w[0] = p0.x; w[1] = p0.y; w[2] = p0.z; w[3] = p0.w;
                w[4] = p1.x; w[5] = p1.y; w[6] = p1.z; w[7] = p1.w;
            } else if constexpr (V == 4) {
                const uint4 p0 = *reinterpret_cast<const uint4*>(packed + i0);
                w[0] = p0.x; w[1] = p0.y; w[2] = p0.z; w[3] = p0.w;
            } else {
                const uint2 p0 = *reinterpret_cast<const uint2*>(packed + i0);
                w[0] = p0.x; w[1] = p0.y;
            }
            uint4 h[CG], l[CG];
#pragma unroll
            for (int c = 0; c < CG; ++c) {
                if (c < ncols) {
                    const int64_t e = i0 + (c0 + c) * lda;
                    h[c] = *reinterpret_cast<const uint4*>(abase + e * EB);
                    if constexpr (EXACT && (FMT == FMT_P16 || FMT == FMT_PBF))
                        l[c] = *reinterpret_cast<const uint4*>(wr + e);
                    else
                        l[c] = make_uint4(0, 0, 0, 0);
                }
            }
#pragma unroll
            for (int c = 0; c < CG; ++c) {
                if (c < ncols) {
                    float v[V];
                    decode_rows<FMT, EXACT>(h[c], l[c], v);
#pragma unroll
                    for (int r = 0; r < V; ++r) {
                        const int b = static_cast<int>(w[r] & 0x7FFFFFFFu);
                        atomicAdd(&ylds[c * s + b], (w[r] & 0x80000000u) ? -v[r] : v[r]);
                    }
                }
            }
        } else {
            for (int64_t i = i0; i < r1; ++i) {
                const uint32_t w = packed[i];
                const int b = static_cast<int>(w & 0x7FFFFFFFu);
                for (int c = 0; c < ncols; ++c) {
                    const float v = load_elem<FMT, EXACT>(a, wr, i + (c0 + c) * lda);
                    atomicAdd(&ylds[c * s + b], (w & 0x80000000u) ? -v : v);
                }
            }
        }
    }
    __syncthreads();
    float* out = partial + static_cast<int64_t>(split) * s * n;
    for (int64_t t = threadIdx.x; t < s * ncols; t += kVecThreads) {
        const int64_t c = t / s;
        out[(c0 + c) * s + (t - c * s)] = ylds[t];
    }
}

__global__ __launch_bounds__(256) void cs_finish_kernel(int64_t s, int64_t n, int splits,
                                                        const float* __restrict__ partial,
                                                        double* __restrict__ y, int64_t ldy) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= s * n) return;
    const int64_t c = t / s;
    const int64_t b = t - c * s;
    double acc = 0.0;
    for (int r = 0; r < splits; ++r) acc += static_cast<double>(partial[static_cast<int64_t>(r) * s * n + t]);
    y[b + c * ldy] = acc;
}

// ---- fixed-point path ------------------------------------------------------
// LDS float atomics are the bottleneck of the scatter (about 7x slower than
// 64-bit integer LDS atomics on gfx950), so the vector path accumulates in
// fixed point instead: a first pass finds each column's largest hi-plane
// magnitude, giving an exponent e_j with |A[i,j]| < 2^e_j for every row
// (one bit of margin covers the exact value vs its hi word); each value is
// then added as the 64-bit integer round(v * 2^(shift - e_j)), with
// shift = 62 - H and 2^H > rows per split, so a bucket's sum cannot
// overflow.  Integer sums are exact and order-independent, so the sketch is
// deterministic; the resolution is 2^-shift of the column's magnitude
// (2^-42 at 2e6 rows and 4 splits).  Columns whose max is not finite (Inf
// or NaN from saturated FP32_16 words) take the float path for the call.

// colmax_bits[j] = max over rows of |value| as float bits (monotone for
// non-negative floats) of the values the sketch sums, or (BOUND_HI) of the
// hi words, which bound the exact FP32_BF16 values within a factor 2.
// grid (n, segments): segment y covers rows [y stride, y stride + len)
// (stride == len: every row; stride > len: a sample); 16-byte row vectors.
template <int FMT, bool EXACT, bool NTL>
__global__ __launch_bounds__(256) void cs_colmax_kernel(int64_t m, const void* __restrict__ a,
                                                        const uint16_t* __restrict__ wr, int64_t lda,
                                                        int64_t stride, int64_t len,
                                                        unsigned* __restrict__ colmax_bits) {
    constexpr int V = vec_rows<FMT>();
    constexpr int EB = 16 / V;
    const int64_t j = blockIdx.x;
    const int64_t r0 = static_cast<int64_t>(blockIdx.y) * stride;
    const int64_t r1 = min(m, r0 + len);
    const char* col = static_cast<const char*>(a) + j * lda * EB;
    float mx = 0.0f;
    for (int64_t i0 = r0 + static_cast<int64_t>(threadIdx.x) * V; i0 < r1; i0 += 256 * V) {
        float v[V];
        if (i0 + V <= r1) {
            const uint4 h = load_a16<NTL>(col + i0 * EB);
            uint4 l = make_uint4(0, 0, 0, 0);
            if constexpr (EXACT) l = load_a16<NTL>(wr + i0 + j * lda);
            decode_rows<FMT, EXACT>(h, l, v);
        } else {
            for (int r = 0; r < V; ++r) v[r] = (i0 + r < r1) ? load_elem<FMT, EXACT>(a, wr, i0 + r + j * lda) : 0.0f;
        }
#pragma unroll
        for (int r = 0; r < V; ++r) {
            const float av = fabsf(v[r]);
            mx = (av > mx || av != av) ? av : mx;  // NaN wins
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        const float other = __shfl_xor(mx, o);
        mx = (other > mx || other != other) ? other : mx;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(colmax_bits + j, __float_as_uint(mx));
}

// e_j from the max bits, plus `margin` bits when the max is a sample's;
// flag = 1 if any column is not finite
__global__ __launch_bounds__(256) void cs_colexp_kernel(int64_t n, const unsigned* __restrict__ colmax_bits,
                                                        int margin, int* __restrict__ colexp,
                                                        int* __restrict__ flag) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (j >= n) return;
    const unsigned bits = colmax_bits[j];
    if (bits >= 0x7F800000u) {
        atomicOr(flag, 1);
        colexp[j] = 0;
        return;
    }
    int e = -149;
    if (bits != 0) {
        frexpf(__uint_as_float(bits), &e);
        e += 1;  // margin: an exact FP32_BF16 value is below twice its hi word
    }
    colexp[j] = e + margin;
}

template <int FMT, bool EXACT, int CG, bool NTL>
__global__ __launch_bounds__(kVecThreads) void cs_apply_fx_kernel(
    int64_t m, int64_t n, const void* __restrict__ a, const uint16_t* __restrict__ wr, int64_t lda,
    int64_t s, const uint32_t* __restrict__ packed, int64_t rows_per_split, int ncg,
    const int* __restrict__ colexp, int shift, double* __restrict__ partial, int* __restrict__ over) {
    constexpr int V = vec_rows<FMT>();
    constexpr int EB = 16 / V;
    extern __shared__ unsigned long long yfx[];  // [CG][s]
    const int split = blockIdx.x / ncg;
    const int cg = blockIdx.x - split * ncg;
    const int64_t c0 = static_cast<int64_t>(cg) * CG;
    const int ncols = static_cast<int>(min<int64_t>(CG, n - c0));
    for (int64_t t = threadIdx.x; t < s * CG; t += kVecThreads) yfx[t] = 0ull;
    int sc[CG];
    float bnd[CG];  // |value| < 2^e_c keeps every term below 2^shift (no overflow)
#pragma unroll
    for (int c = 0; c < CG; ++c) {
        sc[c] = c < ncols ? shift - colexp[c0 + c] : 0;
        bnd[c] = c < ncols ? ldexpf(1.0f, colexp[c0 + c]) : 0.0f;
    }
    bool beyond = false;  // a value at or above its column's bound (or not finite)
    __syncthreads();

    const int64_t r0 = static_cast<int64_t>(split) * rows_per_split;
    const int64_t r1 = min(m, r0 + rows_per_split);
    const char* abase = static_cast<const char*>(a);
    for (int64_t i0 = r0 + static_cast<int64_t>(threadIdx.x) * V; i0 < r1;
         i0 += static_cast<int64_t>(kVecThreads) * V) {
        if (i0 + V <= r1) {
            uint32_t w[V];
            if constexpr (V == 8) {
                const uint4 p0 = *reinterpret_cast<const uint4*>(packed + i0);
                const uint4 p1 = *reinterpret_cast<const uint4*>(packed + i0 + 4);
                w[0] = p0.x; w[1] = p0.y; w[2] = p0.z; w[3] = p0.w;
                w[4] = p1.x; w[5] = p1.y; w[6] = p1.z; w[7] = p1.w;
            } else if constexpr (V == 4) {
                const uint4 p0 = *reinterpret_cast<const uint4*>(packed + i0);
                w[0] = p0.x; w[1] = p0.y; w[2] = p0.z; w[3] = p0.w;
            } else {
                const uint2 p0 = *reinterpret_cast<const uint2*>(packed + i0);
                w[0] = p0.x; w[1] = p0.y;
            }
            uint4 h[CG], l[CG];
#pragma unroll
            for (int c = 0; c < CG; ++c) {
                if (c < ncols) {
                    const int64_t e = i0 + (c0 + c) * lda;
                    h[c] = load_a16<NTL>(abase + e * EB);
                    if constexpr (EXACT && (FMT == FMT_P16 || FMT == FMT_PBF))
                        l[c] = load_a16<NTL>(wr + e);
                    else
                        l[c] = make_uint4(0, 0, 0, 0);
                }
            }
#pragma unroll
            for (int c = 0; c < CG; ++c) {
                if (c < ncols) {
                    float v[V];
                    decode_rows<FMT, EXACT>(h[c], l[c], v);
#pragma unroll
                    for (int r = 0; r < V; ++r) {
                        const int b = static_cast<int>(w[r] & 0x7FFFFFFFu);
                        beyond = beyond || !(fabsf(v[r]) < bnd[c]);
                        const long long q = __float2ll_rn(ldexpf(v[r], sc[c]));
                        atomicAdd(&yfx[c * s + b],
                                  static_cast<unsigned long long>((w[r] & 0x80000000u) ? -q : q));
                    }
                }
            }
        } else {
            for (int64_t i = i0; i < r1; ++i) {
                const uint32_t w = packed[i];
                const int b = static_cast<int>(w & 0x7FFFFFFFu);
                for (int c = 0; c < ncols; ++c) {
                    const float v = load_elem<FMT, EXACT>(a, wr, i + (c0 + c) * lda);
                    beyond = beyond || !(fabsf(v) < ldexpf(1.0f, colexp[c0 + c]));
                    const long long q = __float2ll_rn(ldexpf(v, sc[c]));
                    atomicAdd(&yfx[c * s + b], static_cast<unsigned long long>((w & 0x80000000u) ? -q : q));
                }
            }
        }
    }
    if (__syncthreads_or(beyond) && threadIdx.x == 0) atomicOr(over, 1);
    double* out = partial + static_cast<int64_t>(split) * s * n;
    for (int64_t t = threadIdx.x; t < s * ncols; t += kVecThreads) {
        const int64_t c = t / s;
        const long long q = static_cast<long long>(yfx[t]);
        out[(c0 + c) * s + (t - c * s)] = ldexp(static_cast<double>(q), colexp[c0 + c] - shift);
    }
}

__global__ __launch_bounds__(256) void cs_finish_fx_kernel(int64_t s, int64_t n, int splits,
                                                           const double* __restrict__ partial,
                                                           double* __restrict__ y, int64_t ldy) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= s * n) return;
    const int64_t c = t / s;
    const int64_t b = t - c * s;
    double acc = 0.0;
    for (int r = 0; r < splits; ++r) acc += partial[static_cast<int64_t>(r) * s * n + t];
    y[b + c * ldy] = acc;
}

constexpr int64_t kLdsBytes = 128 * 1024;

int cg_for(int64_t s) {
    int cg = 8;  // 16-byte loads in flight per lane per plane
    while (cg > 1 && s * cg * 4 > kLdsBytes) cg >>= 1;
    return cg;
}

struct CsPlan {
    int cg, ncg, splits;
    int64_t rows_per_split;
};

CsPlan cs_plan(int64_t m, int64_t n, int64_t s) {
    CsPlan p;
    p.cg = cg_for(s);
    p.ncg = static_cast<int>(ceil_div(std::max<int64_t>(n, 1), p.cg));
    // about 4 blocks per CU over the grid, each split at least 8192 rows;
    // split starts stay multiples of 8 rows (the widest row vector)
    const int64_t want = std::max<int64_t>(1, ceil_div(4 * 256, p.ncg));
    const int64_t max_splits = std::max<int64_t>(1, ceil_div(m, 8192));
    p.splits = static_cast<int>(std::min(want, max_splits));
    p.rows_per_split = ceil_div(ceil_div(std::max<int64_t>(m, 1), p.splits), 8) * 8;
    p.splits = static_cast<int>(ceil_div(std::max<int64_t>(m, 1), p.rows_per_split));
    return p;
}

// Fixed-point path: int64 LDS slices (CG columns x s buckets x 8 bytes).
CsPlan cs_plan_fx(int64_t m, int64_t n, int64_t s) {
    CsPlan p;
    p.cg = 8;
    while (p.cg > 1 && s * p.cg * 8 > kLdsBytes) p.cg >>= 1;
    p.ncg = static_cast<int>(ceil_div(std::max<int64_t>(n, 1), p.cg));
    const int64_t want = std::max<int64_t>(1, ceil_div(4 * 256, p.ncg));
    const int64_t max_splits = std::max<int64_t>(1, ceil_div(m, 8192));
    p.splits = static_cast<int>(std::min(want, max_splits));
    p.rows_per_split = ceil_div(ceil_div(std::max<int64_t>(m, 1), p.splits), 8) * 8;
    p.splits = static_cast<int>(ceil_div(std::max<int64_t>(m, 1), p.rows_per_split));
    return p;
}

// 62 - H with 2^H > rows per split (plus one bit): a bucket sum of values
// below 2^shift each stays below 2^62
int fx_shift(int64_t rows_per_split) {
    int h = 1;
    while ((int64_t(1) << h) <= rows_per_split) ++h;
    return 62 - (h + 1);
}

size_t a256(size_t v) { return (v + 255) / 256 * 256; }

struct CsLayout {
    size_t idx, colmax, colexp, flag, partial, total;
};

CsLayout cs_layout(int64_t m, int64_t n, int64_t s) {
    const CsPlan pf = cs_plan(m, n, s), px = cs_plan_fx(m, n, s);
    CsLayout L;
    L.idx = 0;
    L.colmax = a256(static_cast<size_t>(std::max<int64_t>(m, 1)) * 4);
    L.colexp = L.colmax + a256(static_cast<size_t>(std::max<int64_t>(n, 1)) * 4);
    L.flag = L.colexp + a256(static_cast<size_t>(std::max<int64_t>(n, 1)) * 4);
    L.partial = L.flag + 256;
    const size_t pb = std::max(static_cast<size_t>(pf.splits) * s * n * 4, static_cast<size_t>(px.splits) * s * n * 8);
    L.total = L.partial + pb;
    return L;
}

size_t cs_scratch(int64_t m, int64_t n, int64_t s) { return cs_layout(m, n, s).total; }

template <int FMT>
bool vec_ok(const void* a, const uint16_t* wr, int64_t lda, bool exact) {
    constexpr int V = vec_rows<FMT>();
    const bool lo = !(exact && (FMT == FMT_P16 || FMT == FMT_PBF)) ||
                    (reinterpret_cast<uintptr_t>(wr) % 16 == 0);
    return reinterpret_cast<uintptr_t>(a) % 16 == 0 && lda % V == 0 && lo;
}

template <int FMT, bool EXACT, int CG>
void launch_apply(const CsPlan& p, int64_t m, int64_t n, const void* a, const uint16_t* wr,
                  int64_t lda, int64_t s, const uint32_t* packed, float* partial,
                  hipStream_t st) {
    const size_t lds = static_cast<size_t>(s) * CG * 4;
    const int64_t blocks = static_cast<int64_t>(p.splits) * p.ncg;
    require(blocks < (int64_t(1) << 31), "countsketch_apply: grid too large");
    if (vec_ok<FMT>(a, wr, lda, EXACT)) {
        auto* k = cs_apply_vec_kernel<FMT, EXACT, CG>;
        TOMOS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            static_cast<int>(lds)));
        k<<<dim3(static_cast<unsigned>(blocks)), kVecThreads, lds, st>>>(
            m, n, a, wr, lda, s, packed, p.rows_per_split, p.ncg, partial);
        TOMOS_LAUNCH_CHECK("cs_apply_vec_kernel");
        return;
    }
    auto* k = cs_apply_kernel<FMT, EXACT, CG>;
    TOMOS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
    k<<<dim3(static_cast<unsigned>(blocks)), 256, lds, st>>>(m, n, a, wr, lda, s, packed,
                                                            p.rows_per_split, p.ncg, partial);
    TOMOS_LAUNCH_CHECK("cs_apply_kernel");
}

template <int FMT, bool EXACT>
void launch_apply_cg(const CsPlan& p, int64_t m, int64_t n, const void* a, const uint16_t* wr,
                     int64_t lda, int64_t s, const uint32_t* packed, float* partial,
                     hipStream_t st) {
    switch (p.cg) {
        case 8: launch_apply<FMT, EXACT, 8>(p, m, n, a, wr, lda, s, packed, partial, st); break;
        case 4: launch_apply<FMT, EXACT, 4>(p, m, n, a, wr, lda, s, packed, partial, st); break;
        case 2: launch_apply<FMT, EXACT, 2>(p, m, n, a, wr, lda, s, packed, partial, st); break;
        default: launch_apply<FMT, EXACT, 1>(p, m, n, a, wr, lda, s, packed, partial, st); break;
    }
}

// Fixed-point apply on the vector path.  The column exponents come from a
// sample of the rows (kSampleSegs segments spread over A, each 1/256 of
// the rows long, 64..1024 rows) plus kSampleMargin bits when A has at least
// kSampleMinRows rows, else from every row.  The apply kernel checks every
// value against its column's bound; a value beyond it (or not finite)
// makes the call redo the exponents from every row and apply again.
constexpr int64_t kSampleSegs = 32;
constexpr int64_t kSampleMinRows = 16384;  // the sample is at most 1/8 of the rows
constexpr int kSampleMargin = 12;           // 2^-31 of the sampled max per term at 2e6 rows, 4 splits

// Enqueue the fixed-point path with the exponents from a sample (sample)
// or from every row, reading nothing back: flags[0] is set if a column's
// max is not finite, flags[1] if a value is beyond its column's bound.
// The caller reads the flags once its work is enqueued.
template <int FMT, bool EXACT>
void fx_enqueue(int64_t m, int64_t n, const void* a, const uint16_t* wr, int64_t lda, int64_t s,
                const uint32_t* packed, char* scratch, const CsLayout& L, double* y, int64_t ldy, bool sample,
                int* flags, hipStream_t st) {
    const CsPlan p = cs_plan_fx(m, n, s);
    unsigned* colmax = reinterpret_cast<unsigned*>(scratch + L.colmax);
    int* colexp = reinterpret_cast<int*>(scratch + L.colexp);
    double* partial = reinterpret_cast<double*>(scratch + L.partial);
    require(n < (int64_t(1) << 31), "countsketch_apply: n too large");
    // FP32_BF16's exact values are bounded by 2x their hi words: the hi plane
    // suffices; FP32_16's exact values need their own decode
    constexpr bool kExactMax = EXACT && FMT == FMT_P16;
    constexpr int64_t kElemBytes = (FMT == FMT_F64 ? 8 : FMT == FMT_F32 ? 4 : 2) * (EXACT && (FMT == FMT_P16 || FMT == FMT_PBF) ? 2 : 1);
    const bool ntl = stream_loads(m * n * kElemBytes);
    TOMOS_HIP_CHECK(hipMemsetAsync(colmax, 0, static_cast<size_t>(n) * 4, st));
    TOMOS_HIP_CHECK(hipMemsetAsync(flags, 0, 8, st));
    int64_t stride, len;
    unsigned segs;
    if (sample) {
        stride = (m / kSampleSegs) / 8 * 8;
        len = std::max<int64_t>(64, std::min<int64_t>(1024, (m / 256) / 8 * 8));
        segs = static_cast<unsigned>(kSampleSegs);
    } else {
        const int64_t csplits = std::max<int64_t>(1, std::min<int64_t>(64, ceil_div(m, 256 * 8 * 64)));
        stride = len = ceil_div(ceil_div(m, csplits), 8) * 8;
        segs = static_cast<unsigned>(ceil_div(m, len));
    }
    const dim3 cgrid(static_cast<unsigned>(n), segs);
    if (ntl && !sample)
        cs_colmax_kernel<FMT, kExactMax, true><<<cgrid, 256, 0, st>>>(m, a, wr, lda, stride, len, colmax);
    else
        cs_colmax_kernel<FMT, kExactMax, false><<<cgrid, 256, 0, st>>>(m, a, wr, lda, stride, len, colmax);
    cs_colexp_kernel<<<dim3(static_cast<unsigned>(ceil_div(n, 256))), 256, 0, st>>>(
        n, colmax, sample ? kSampleMargin : 0, colexp, flags);
    TOMOS_LAUNCH_CHECK("cs_colmax_kernel");
    const int shift = fx_shift(p.rows_per_split);
    const size_t lds = static_cast<size_t>(s) * p.cg * 8;
    const int64_t blocks = static_cast<int64_t>(p.splits) * p.ncg;
    require(blocks < (int64_t(1) << 31), "countsketch_apply: grid too large");
    auto launch = [&](auto* k) {
        TOMOS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        k<<<dim3(static_cast<unsigned>(blocks)), kVecThreads, lds, st>>>(m, n, a, wr, lda, s, packed,
                                                                        p.rows_per_split, p.ncg, colexp, shift,
                                                                        partial, flags + 1);
    };
    if (ntl) {
        switch (p.cg) {
            case 8: launch(cs_apply_fx_kernel<FMT, EXACT, 8, true>); break;
            case 4: launch(cs_apply_fx_kernel<FMT, EXACT, 4, true>); break;
            case 2: launch(cs_apply_fx_kernel<FMT, EXACT, 2, true>); break;
            default: launch(cs_apply_fx_kernel<FMT, EXACT, 1, true>); break;
        }
    } else {
        switch (p.cg) {
            case 8: launch(cs_apply_fx_kernel<FMT, EXACT, 8, false>); break;
            case 4: launch(cs_apply_fx_kernel<FMT, EXACT, 4, false>); break;
            case 2: launch(cs_apply_fx_kernel<FMT, EXACT, 2, false>); break;
            default: launch(cs_apply_fx_kernel<FMT, EXACT, 1, false>); break;
        }
    }
    TOMOS_LAUNCH_CHECK("cs_apply_fx_kernel");
    cs_finish_fx_kernel<<<dim3(static_cast<unsigned>(ceil_div(s * n, 256))), 256, 0, st>>>(s, n, p.splits, partial,
                                                                                          y, ldy);
    TOMOS_LAUNCH_CHECK("cs_finish_fx_kernel");
}

// The fixed-point path for one call; returns false (the float path runs)
// when it does not apply (unaligned A) or a column's max over every row is
// not finite.
template <int FMT, bool EXACT>
bool try_apply_fx(int64_t m, int64_t n, const void* a, const uint16_t* wr, int64_t lda, int64_t s,
                  const uint32_t* packed, char* scratch, const CsLayout& L, double* y, int64_t ldy,
                  hipStream_t st) {
    if (!vec_ok<FMT>(a, wr, lda, EXACT)) return false;
    int* flags = reinterpret_cast<int*>(scratch + L.flag);
    int h[2] = {0, 0};
    auto read_flags = [&]() {
        TOMOS_HIP_CHECK(hipMemcpyAsync(h, flags, 8, hipMemcpyDeviceToHost, st));
        TOMOS_HIP_CHECK(hipStreamSynchronize(st));
    };
    if (m >= kSampleMinRows) {  // sampled exponents: one pass over A when they hold
        fx_enqueue<FMT, EXACT>(m, n, a, wr, lda, s, packed, scratch, L, y, ldy, true, flags, st);
        read_flags();
        if (!h[0] && !h[1]) return true;
    }
    fx_enqueue<FMT, EXACT>(m, n, a, wr, lda, s, packed, scratch, L, y, ldy, false, flags, st);
    read_flags();
    return !h[0] && !h[1];  // the float path redoes a call whose full-row exponents are not finite
}

}  // namespace

size_t countsketch_scratch_bytes(int64_t m, int64_t n, int64_t s) { return cs_scratch(m, n, s); }

void countsketch_apply(int a_type, int tier, int64_t m, int64_t n, const void* a,
                       const uint16_t* wr, int64_t lda, int64_t s, double* y, int64_t ldy,
                       void* scratch, size_t scratch_bytes, hipStream_t st) {
    require(m >= 0 && n >= 0 && s >= 1, "countsketch_apply: bad sizes");
    require(s <= kLdsBytes / 4, "countsketch_apply: s too large (max 32768)");
    require(m < (int64_t(1) << 31), "countsketch_apply: m must be < 2^31");
    require(lda >= std::max<int64_t>(m, 1) && ldy >= s, "countsketch_apply: bad leading dimension");
    require(tier == TOMOS_TIER_DEFAULT || tier == TOMOS_TIER_FAST || tier == TOMOS_TIER_EXACT,
            "countsketch_apply: bad tier");
    if (n == 0) return;
    require(y != nullptr, "countsketch_apply: null y");
    const size_t need = cs_scratch(m, n, s);
    if (scratch == nullptr) scratch = fallback_scratch(need);
    else require(scratch_bytes >= need, "countsketch_apply: scratch too small");
    const CsPlan p = cs_plan(m, n, s);
    const CsLayout L = cs_layout(m, n, s);
    char* sc = static_cast<char*>(scratch);
    uint32_t* packed = reinterpret_cast<uint32_t*>(sc + L.idx);
    float* partial = reinterpret_cast<float*>(sc + L.partial);
    if (m > 0) {
        require(a != nullptr, "countsketch_apply: null A");
        cs_index_kernel<<<dim3(static_cast<unsigned>(ceil_div(m, 256))), 256, 0, st>>>(
            m, s, nullptr, nullptr, packed);
        TOMOS_LAUNCH_CHECK("cs_index_kernel");
        const bool exact = tier == TOMOS_TIER_EXACT;
        if (exact && (a_type == TOMOS_FP32_16 || a_type == TOMOS_FP32_BF16))
            require(wr != nullptr, "countsketch_apply: EXACT needs wr");
        bool done = false;
        switch (a_type) {
            case TOMOS_F32: done = try_apply_fx<FMT_F32, false>(m, n, a, wr, lda, s, packed, sc, L, y, ldy, st); break;
            case TOMOS_F64: done = try_apply_fx<FMT_F64, false>(m, n, a, wr, lda, s, packed, sc, L, y, ldy, st); break;
            case TOMOS_FP32_16:
                done = exact ? try_apply_fx<FMT_P16, true>(m, n, a, wr, lda, s, packed, sc, L, y, ldy, st)
                             : try_apply_fx<FMT_P16, false>(m, n, a, wr, lda, s, packed, sc, L, y, ldy, st);
                break;
            case TOMOS_FP32_BF16:
                done = exact ? try_apply_fx<FMT_PBF, true>(m, n, a, wr, lda, s, packed, sc, L, y, ldy, st)
                             : try_apply_fx<FMT_PBF, false>(m, n, a, wr, lda, s, packed, sc, L, y, ldy, st);
                break;
            default: break;
        }
        if (done) return;
        switch (a_type) {
            case TOMOS_F32: launch_apply_cg<FMT_F32, false>(p, m, n, a, wr, lda, s, packed, partial, st); break;
            case TOMOS_F64: launch_apply_cg<FMT_F64, false>(p, m, n, a, wr, lda, s, packed, partial, st); break;
            case TOMOS_FP32_16:
                require(!exact || wr, "countsketch_apply: EXACT needs wr");
                if (exact) launch_apply_cg<FMT_P16, true>(p, m, n, a, wr, lda, s, packed, partial, st);
                else launch_apply_cg<FMT_P16, false>(p, m, n, a, wr, lda, s, packed, partial, st);
                break;
            case TOMOS_FP32_BF16:
                require(!exact || wr, "countsketch_apply: EXACT needs wr");
                if (exact) launch_apply_cg<FMT_PBF, true>(p, m, n, a, wr, lda, s, packed, partial, st);
                else launch_apply_cg<FMT_PBF, false>(p, m, n, a, wr, lda, s, packed, partial, st);
                break;
            default: throw Error(kErrUnsupported, "countsketch_apply: unsupported a_type");
        }
    }
    const int64_t total = s * n;
    cs_finish_kernel<<<dim3(static_cast<unsigned>(ceil_div(total, 256))), 256, 0, st>>>(
        s, n, m > 0 ? p.splits : 0, partial, y, ldy);
    TOMOS_LAUNCH_CHECK("cs_finish_kernel");
}

// Many sketches in one call, no host synchronisation between them: every
// problem's fixed-point path is enqueued in turn on `st` (sharing one
// scratch region, which the stream order makes safe), each with its own two
// flag words after that region; one read of all flags at the end sends the
// problems whose sampled exponents did not hold (or that the fixed-point
// path does not take) through the one-problem call.
size_t countsketch_batched_scratch_bytes(int nprob, const tomos_cs_problem* pr) {
    size_t b = 256;
    for (int i = 0; i < nprob; ++i) b = std::max(b, cs_scratch(pr[i].m, pr[i].n, pr[i].s));
    return a256(b) + a256(static_cast<size_t>(std::max(nprob, 1)) * 8);
}

void countsketch_apply_batched(int a_type, int tier, int nprob, const tomos_cs_problem* pr, void* scratch,
                               size_t scratch_bytes, hipStream_t st) {
    require(nprob >= 0 && (nprob == 0 || pr), "countsketch_apply_batched: bad problem list");
    require(tier == TOMOS_TIER_DEFAULT || tier == TOMOS_TIER_FAST || tier == TOMOS_TIER_EXACT,
            "countsketch_apply_batched: bad tier");
    for (int i = 0; i < nprob; ++i) {
        const tomos_cs_problem& q = pr[i];
        require(q.m >= 0 && q.n >= 0 && q.s >= 1 && q.s <= kLdsBytes / 4 && q.m < (int64_t(1) << 31),
                "countsketch_apply_batched: bad sizes");
        require(q.lda >= std::max<int64_t>(q.m, 1) && q.ldy >= q.s, "countsketch_apply_batched: bad leading dimension");
        require(q.n == 0 || q.y, "countsketch_apply_batched: null y");
        require(q.n == 0 || q.m == 0 || q.a, "countsketch_apply_batched: null A");
    }
    if (nprob == 0) return;
    const size_t need = countsketch_batched_scratch_bytes(nprob, pr);
    if (scratch == nullptr) scratch = fallback_scratch(need);
    else require(scratch_bytes >= need, "countsketch_apply_batched: scratch too small");
    char* sc = static_cast<char*>(scratch);
    int* flags = reinterpret_cast<int*>(sc + (need - a256(static_cast<size_t>(nprob) * 8)));
    const bool exact = tier == TOMOS_TIER_EXACT;
    std::vector<char> queued(static_cast<size_t>(nprob), 0);
    for (int i = 0; i < nprob; ++i) {
        const tomos_cs_problem& q = pr[i];
        if (q.n == 0 || q.m == 0) continue;
        if (exact && (a_type == TOMOS_FP32_16 || a_type == TOMOS_FP32_BF16) && !q.wr) continue;
        const CsLayout L = cs_layout(q.m, q.n, q.s);
        uint32_t* packed = reinterpret_cast<uint32_t*>(sc + L.idx);
        const bool sample = q.m >= kSampleMinRows;
        auto go = [&](auto fmt_tag, auto exact_tag) {
            constexpr int F = decltype(fmt_tag)::value;
            constexpr bool E = decltype(exact_tag)::value;
            if (!vec_ok<F>(q.a, q.wr, q.lda, E)) return;
            cs_index_kernel<<<dim3(static_cast<unsigned>(ceil_div(q.m, 256))), 256, 0, st>>>(q.m, q.s, nullptr,
                                                                                             nullptr, packed);
            TOMOS_LAUNCH_CHECK("cs_index_kernel");
            fx_enqueue<F, E>(q.m, q.n, q.a, q.wr, q.lda, q.s, packed, sc, L, q.y, q.ldy, sample, flags + 2 * i, st);
            queued[i] = 1;
        };
        using T = std::true_type;
        using Fl = std::false_type;
        switch (a_type) {
            case TOMOS_F32: go(std::integral_constant<int, FMT_F32>{}, Fl{}); break;
            case TOMOS_F64: go(std::integral_constant<int, FMT_F64>{}, Fl{}); break;
            case TOMOS_FP32_16:
                if (exact) go(std::integral_constant<int, FMT_P16>{}, T{});
                else go(std::integral_constant<int, FMT_P16>{}, Fl{});
                break;
            case TOMOS_FP32_BF16:
                if (exact) go(std::integral_constant<int, FMT_PBF>{}, T{});
                else go(std::integral_constant<int, FMT_PBF>{}, Fl{});
                break;
            default: break;
        }
    }
    std::vector<int> h(2 * static_cast<size_t>(nprob), 0);
    TOMOS_HIP_CHECK(hipMemcpyAsync(h.data(), flags, h.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    TOMOS_HIP_CHECK(hipStreamSynchronize(st));
    for (int i = 0; i < nprob; ++i) {
        const tomos_cs_problem& q = pr[i];
        if (q.n == 0) continue;
        if (queued[i] && !h[2 * i] && !h[2 * i + 1]) continue;
        countsketch_apply(a_type, tier, q.m, q.n, q.a, q.wr, q.lda, q.s, q.y, q.ldy, sc, need, st);
    }
}

}  // namespace tomos_hip

using namespace tomos_hip;

extern "C" {

size_t tomos_countsketch_apply_batched_scratch_size(int nprob, const tomos_cs_problem* probs) {
    if (nprob < 0 || (nprob > 0 && !probs)) return 0;
    return countsketch_batched_scratch_bytes(nprob, probs);
}

int tomos_countsketch_apply_batched(int a_type, int tier, int nprob, const tomos_cs_problem* probs, void* scratch,
                                    size_t scratch_bytes, tomos_stream_t s) {
    return guarded([&] { countsketch_apply_batched(a_type, tier, nprob, probs, scratch, scratch_bytes, s); });
}

int tomos_countsketch_indices(int64_t m, int64_t s, int32_t* bucket, int8_t* sign,
                              tomos_stream_t st) {
    return guarded([&] {
        require(m >= 0 && s >= 1, "tomos_countsketch_indices: bad sizes");
        require(s <= 0x7FFFFFFF, "tomos_countsketch_indices: s too large");
        if (m == 0) return;
        require(bucket || sign, "tomos_countsketch_indices: no output");
        cs_index_kernel<<<dim3(static_cast<unsigned>(ceil_div(m, 256))), 256, 0, st>>>(
            m, s, bucket, sign, nullptr);
        TOMOS_LAUNCH_CHECK("cs_index_kernel");
    });
}

size_t tomos_countsketch_scratch_size(int64_t m, int64_t n, int64_t s) {
    if (m < 0 || n < 0 || s < 1) return 0;
    return countsketch_scratch_bytes(m, n, s);
}

int tomos_countsketch_apply(int a_type, int tier, int64_t m, int64_t n, const void* a,
                            const uint16_t* wr, int64_t lda, int64_t s, double* y, int64_t ldy,
                            void* scratch, size_t scratch_bytes, tomos_stream_t st) {
    return guarded([&] {
        countsketch_apply(a_type, tier, m, n, a, wr, lda, s, y, ldy, scratch, scratch_bytes, st);
    });
}

}  // extern "C"
