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
// flag = 1 if the column is not finite
__device__ __forceinline__ void cs_set_exponent(unsigned bits, int margin, int* __restrict__ colexp,
                                                int* __restrict__ flag) {
    if (bits >= 0x7F800000u) {
        atomicOr(flag, 1);
        *colexp = 0;
        return;
    }
    int e = -149;
    if (bits != 0) {
        frexpf(__uint_as_float(bits), &e);
        e += 1;  // margin: an exact FP32_BF16 value is below twice its hi word
    }
    *colexp = e + margin;
}

__global__ __launch_bounds__(256) void cs_colexp_kernel(int64_t n, const unsigned* __restrict__ colmax_bits,
                                                        int margin, int* __restrict__ colexp,
                                                        int* __restrict__ flag) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (j >= n) return;
    cs_set_exponent(colmax_bits[j], margin, colexp + j, flag);
}

// Sampled exponents in one pass, one block per column: the max over `segs`
// segments of `len` rows (starts `stride` apart; len a multiple of 8) reduced
// in the block, then e_j + margin written directly.  The whole sample of a
// column is one block's work, so no column max crosses blocks (no atomics
// to one address from many blocks, no cleared buffer).  segs = 1, stride =
// 0, len = m rounded up to 8 and margin 0 give the exact exponents.
template <int FMT, bool EXACT>
__device__ __forceinline__ void cs_sample_exp_body(int64_t j, int64_t m, const void* __restrict__ a,
                                                   const uint16_t* __restrict__ wr, int64_t lda, int64_t stride,
                                                   int len, int segs, int margin, int* __restrict__ colexp,
                                                   int* __restrict__ flag) {
    constexpr int V = vec_rows<FMT>();
    constexpr int EB = 16 / V;
    __shared__ float wmax[4];
    const char* col = static_cast<const char*>(a) + j * lda * EB;
    const unsigned vps = static_cast<unsigned>(len / V);
    const unsigned total = vps * static_cast<unsigned>(segs);
    float mx = 0.0f;
#pragma unroll 4
    for (unsigned t = threadIdx.x; t < total; t += 256) {
        const unsigned sg = t / vps;
        const int64_t i0 = static_cast<int64_t>(sg) * stride + static_cast<int64_t>(t - sg * vps) * V;
        float v[V];
        if (i0 + V <= m) {
            const uint4 h = load_a16<false>(col + i0 * EB);
            uint4 l = make_uint4(0, 0, 0, 0);
            if constexpr (EXACT) l = load_a16<false>(wr + i0 + j * lda);
            decode_rows<FMT, EXACT>(h, l, v);
        } else {
            for (int r = 0; r < V; ++r) v[r] = (i0 + r < m) ? load_elem<FMT, EXACT>(a, wr, i0 + r + j * lda) : 0.0f;
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
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) mx = (wmax[w] > mx || wmax[w] != wmax[w]) ? wmax[w] : mx;
        cs_set_exponent(__float_as_uint(mx), margin, colexp + j, flag);
    }
}

template <int FMT, bool EXACT>
__global__ __launch_bounds__(256) void cs_sample_exp_kernel(int64_t m, const void* __restrict__ a,
                                                            const uint16_t* __restrict__ wr, int64_t lda,
                                                            int64_t stride, int len, int segs, int margin,
                                                            int* __restrict__ colexp, int* __restrict__ flag) {
    cs_sample_exp_body<FMT, EXACT>(blockIdx.x, m, a, wr, lda, stride, len, segs, margin, colexp, flag);
}

// One problem of a grouped CountSketch: its operands, its fixed-point plan
// (splits of rows_per_split rows x column groups of the launch's CG), the
// sample that sets its exponents and where its scratch and result lie.
struct CsGroupDesc {
    const void* a;
    const uint16_t* wr;
    uint32_t* packed;
    int* colexp;
    unsigned long long* partial;
    int* flags;  // [0] a column max not finite, [1] a value beyond its bound
    double* y;
    int64_t m, n, lda, s, ldy, rows_per_split, stride;
    int ncg, splits, shift, len, segs, margin;
    int block0;  // first block within its apply launch
    int local;   // index within its apply launch's descriptors
    int owner0;  // block0 plus the launch's first entry in the owner table
    int pad;
};

// Descriptors travel as kernel arguments, kCsChunk a launch (under 4 KB)
constexpr int kCsChunk = 24;
struct CsDescChunk {
    CsGroupDesc d[kCsChunk];
};
static_assert(sizeof(CsDescChunk) + 16 <= 4096, "descriptor chunk exceeds the kernel-argument limit");

__global__ __launch_bounds__(64) void cs_set_desc_kernel(CsGroupDesc* __restrict__ dst, CsDescChunk c, int count) {
    if (static_cast<int>(threadIdx.x) < count) dst[threadIdx.x] = c.d[threadIdx.x];
}

// owner[b] = the launch-local descriptor of apply block b; grid (problems)
__global__ __launch_bounds__(256) void cs_owner_kernel(const CsGroupDesc* __restrict__ d, int* __restrict__ owner) {
    const CsGroupDesc& q = d[blockIdx.x];
    const int nb = q.splits * q.ncg;
    for (int b = threadIdx.x; b < nb; b += 256) owner[q.owner0 + b] = q.local;
}

// grid (ceil(max m / 256), problems)
__global__ __launch_bounds__(256) void cs_group_index_kernel(const CsGroupDesc* __restrict__ d) {
    const CsGroupDesc& q = d[blockIdx.y];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= q.m) return;
    const uint64_t hb = splitmix64(static_cast<uint64_t>(i) ^ kBucketSeed);
    const uint64_t hs = splitmix64(static_cast<uint64_t>(i) ^ kSignSeed);
    const uint32_t b = static_cast<uint32_t>(hb % static_cast<uint64_t>(q.s));
    q.packed[i] = b | ((hs & 1ull) ? 0u : 0x80000000u);
}

// grid (max n, problems)
template <int FMT, bool EXACT>
__global__ __launch_bounds__(256) void cs_group_exp_kernel(const CsGroupDesc* __restrict__ d) {
    const CsGroupDesc& q = d[blockIdx.y];
    if (static_cast<int64_t>(blockIdx.x) >= q.n) return;
    cs_sample_exp_body<FMT, EXACT>(blockIdx.x, q.m, q.a, q.wr, q.lda, q.stride, q.len, q.segs, q.margin, q.colexp,
                                   q.flags);
}

template <int FMT, bool EXACT, int CG, bool NTL>
__device__ __forceinline__ void cs_apply_fx_body(
    int split, int cg, int64_t m, int64_t n, const void* __restrict__ a, const uint16_t* __restrict__ wr,
    int64_t lda, int64_t s, const uint32_t* __restrict__ packed, int64_t rows_per_split,
    const int* __restrict__ colexp, int shift, unsigned long long* __restrict__ partial, int* __restrict__ over) {
    constexpr int V = vec_rows<FMT>();
    constexpr int EB = 16 / V;
    extern __shared__ unsigned long long yfx[];  // [CG][s]
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
    unsigned long long* out = partial + static_cast<int64_t>(split) * s * n + c0 * s;
    for (int64_t t = threadIdx.x; t < s * ncols; t += kVecThreads) out[t] = yfx[t];
}

template <int FMT, bool EXACT, int CG, bool NTL>
__global__ __launch_bounds__(kVecThreads) void cs_apply_fx_kernel(
    int64_t m, int64_t n, const void* __restrict__ a, const uint16_t* __restrict__ wr, int64_t lda,
    int64_t s, const uint32_t* __restrict__ packed, int64_t rows_per_split, int ncg,
    const int* __restrict__ colexp, int shift, unsigned long long* __restrict__ partial,
    int* __restrict__ over) {
    const int split = blockIdx.x / ncg;
    cs_apply_fx_body<FMT, EXACT, CG, NTL>(split, blockIdx.x - split * ncg, m, n, a, wr, lda, s, packed,
                                          rows_per_split, colexp, shift, partial, over);
}

// The blocks of every problem in one launch: block b belongs to problem
// owner[b] and is its block b - block0 (split-major over its column groups)
template <int FMT, bool EXACT, int CG, bool NTL>
__global__ __launch_bounds__(kVecThreads) void cs_group_apply_kernel(const CsGroupDesc* __restrict__ d,
                                                                     const int* __restrict__ owner) {
    const CsGroupDesc& q = d[owner[blockIdx.x]];
    const int lb = static_cast<int>(blockIdx.x) - q.block0;
    const int split = lb / q.ncg;
    cs_apply_fx_body<FMT, EXACT, CG, NTL>(split, lb - split * q.ncg, q.m, q.n, q.a, q.wr, q.lda, q.s, q.packed,
                                          q.rows_per_split, q.colexp, q.shift, q.partial, q.flags + 1);
}

// y = the splits' integer sums added as integers (exact: |sum| < m 2^shift
// < 2^61), then rounded once to FP64 and scaled by 2^(e_c - shift).  The
// result depends on the rows and the exponents only, not on how the rows
// were split.
__global__ __launch_bounds__(256) void cs_finish_fx_kernel(int64_t s, int64_t n, int splits,
                                                           const unsigned long long* __restrict__ partial,
                                                           const int* __restrict__ colexp, int shift,
                                                           double* __restrict__ y, int64_t ldy) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= s * n) return;
    const int64_t c = t / s;
    const int64_t b = t - c * s;
    unsigned long long acc = 0ull;
    for (int r = 0; r < splits; ++r) acc += partial[static_cast<int64_t>(r) * s * n + t];
    y[b + c * ldy] = ldexp(static_cast<double>(static_cast<long long>(acc)), colexp[c] - shift);
}

// grid (ceil(max s n / 256), problems)
__global__ __launch_bounds__(256) void cs_group_finish_kernel(const CsGroupDesc* __restrict__ d) {
    const CsGroupDesc& q = d[blockIdx.y];
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= q.s * q.n) return;
    const int64_t c = t / q.s;
    unsigned long long acc = 0ull;
    for (int r = 0; r < q.splits; ++r) acc += q.partial[static_cast<int64_t>(r) * q.s * q.n + t];
    q.y[(t - c * q.s) + c * q.ldy] = ldexp(static_cast<double>(static_cast<long long>(acc)), q.colexp[c] - q.shift);
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

// 62 - H with 2^H > m (plus one bit): a bucket sum over all m rows of
// values below 2^shift each stays below 2^61, so split sums add exactly
int fx_shift(int64_t m) {
    int h = 1;
    while ((int64_t(1) << h) <= m) ++h;
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
                int* flags, bool clear_flags, hipStream_t st) {
    const CsPlan p = cs_plan_fx(m, n, s);
    unsigned* colmax = reinterpret_cast<unsigned*>(scratch + L.colmax);
    int* colexp = reinterpret_cast<int*>(scratch + L.colexp);
    unsigned long long* partial = reinterpret_cast<unsigned long long*>(scratch + L.partial);
    require(n < (int64_t(1) << 31), "countsketch_apply: n too large");
    // FP32_BF16's exact values are bounded by 2x their hi words: the hi plane
    // suffices; FP32_16's exact values need their own decode
    constexpr bool kExactMax = EXACT && FMT == FMT_P16;
    constexpr int64_t kElemBytes = (FMT == FMT_F64 ? 8 : FMT == FMT_F32 ? 4 : 2) * (EXACT && (FMT == FMT_P16 || FMT == FMT_PBF) ? 2 : 1);
    const bool ntl = stream_loads(m * n * kElemBytes);
    if (clear_flags) TOMOS_HIP_CHECK(hipMemsetAsync(flags, 0, 8, st));
    if (sample) {
        // segments start (m / 32) rounded down to 8 rows apart and are
        // m / 256 rows (64..1024, a multiple of 8) long: inside A for m >= 16384
        const int64_t stride = (m / kSampleSegs) / 8 * 8;
        const int64_t len = std::max<int64_t>(64, std::min<int64_t>(1024, (m / 256) / 8 * 8));
        require(m >= kSampleMinRows && (kSampleSegs - 1) * stride + len <= m, "countsketch_apply: bad sample");
        cs_sample_exp_kernel<FMT, kExactMax><<<dim3(static_cast<unsigned>(n)), 256, 0, st>>>(
            m, a, wr, lda, stride, static_cast<int>(len), static_cast<int>(kSampleSegs), kSampleMargin, colexp,
            flags);
        TOMOS_LAUNCH_CHECK("cs_sample_exp_kernel");
    } else {
        TOMOS_HIP_CHECK(hipMemsetAsync(colmax, 0, static_cast<size_t>(n) * 4, st));
        const int64_t csplits = std::max<int64_t>(1, std::min<int64_t>(64, ceil_div(m, 256 * 8 * 64)));
        const int64_t stride = ceil_div(ceil_div(m, csplits), 8) * 8;
        const dim3 cgrid(static_cast<unsigned>(n), static_cast<unsigned>(ceil_div(m, stride)));
        if (ntl)
            cs_colmax_kernel<FMT, kExactMax, true><<<cgrid, 256, 0, st>>>(m, a, wr, lda, stride, stride, colmax);
        else
            cs_colmax_kernel<FMT, kExactMax, false><<<cgrid, 256, 0, st>>>(m, a, wr, lda, stride, stride, colmax);
        cs_colexp_kernel<<<dim3(static_cast<unsigned>(ceil_div(n, 256))), 256, 0, st>>>(n, colmax, 0, colexp,
                                                                                        flags);
        TOMOS_LAUNCH_CHECK("cs_colmax_kernel");
    }
    const int shift = fx_shift(m);
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
                                                                                          colexp, shift, y, ldy);
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
        fx_enqueue<FMT, EXACT>(m, n, a, wr, lda, s, packed, scratch, L, y, ldy, true, flags, true, st);
        read_flags();
        if (!h[0] && !h[1]) return true;
    }
    fx_enqueue<FMT, EXACT>(m, n, a, wr, lda, s, packed, scratch, L, y, ldy, false, flags, true, st);
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

// Many sketches in one call.  Every problem that the fixed-point path takes
// (16-byte aligned rows, s <= 16384) runs in grouped launches: one index
// launch, one exponent launch, one apply launch per column-group width
// (8, 4, 2, 1 columns, whichever keep a block's int64 buckets within
// 128 KB of LDS) whose blocks cover every such problem's splits of up to
// 65536 rows, and one finish launch.  The flags of all problems are read
// once at the end; the problems they reject (sampled exponents that do not
// hold, a column max that is not finite) and the ones not grouped go
// through the one-problem call.  Results equal the one-problem call's bit
// for bit: a fixed-point sketch depends on the rows and the exponents only.
namespace {

constexpr int64_t kGroupRowsPerSplit = 65536;
constexpr int64_t kGroupLdsBytes = 128 * 1024;

int group_cg(int64_t s) {
    int cg = 8;
    while (cg > 1 && s * cg * 8 > kGroupLdsBytes) cg >>= 1;
    return cg;
}

bool group_candidate(const tomos_cs_problem& q) {
    return q.n > 0 && q.m > 0 && q.s * 8 <= kGroupLdsBytes;
}

bool group_takes(int a_type, bool exact, const tomos_cs_problem& q) {
    if (!group_candidate(q)) return false;
    const bool packed = a_type == TOMOS_FP32_16 || a_type == TOMOS_FP32_BF16;
    const int v = a_type == TOMOS_F64 ? 2 : a_type == TOMOS_F32 ? 4 : 8;
    if (!packed && a_type != TOMOS_F32 && a_type != TOMOS_F64) return false;
    if (exact && packed && (!q.wr || reinterpret_cast<uintptr_t>(q.wr) % 16 != 0)) return false;
    return reinterpret_cast<uintptr_t>(q.a) % 16 == 0 && q.lda % v == 0;
}

struct GroupLayout {
    size_t flags, desc, owner, total;
    std::vector<size_t> base;  // per problem: packed, then colexp, then partial
    int64_t blocks;
};

// Offsets for every candidate problem (an upper bound of what a call uses)
GroupLayout group_layout(int nprob, const tomos_cs_problem* pr) {
    GroupLayout L;
    L.blocks = 0;
    int nc = 0;
    L.base.assign(static_cast<size_t>(nprob), 0);
    size_t off = 0;
    L.flags = off;
    off += a256(static_cast<size_t>(std::max(nprob, 1)) * 8);
    for (int i = 0; i < nprob; ++i) {
        const tomos_cs_problem& q = pr[i];
        if (!group_candidate(q)) continue;
        ++nc;
        const int64_t splits = ceil_div(q.m, kGroupRowsPerSplit);
        L.blocks += splits * ceil_div(q.n, group_cg(q.s));
    }
    L.desc = off;
    off += a256(static_cast<size_t>(std::max(nc, 1)) * sizeof(CsGroupDesc));
    L.owner = off;
    off += a256(static_cast<size_t>(std::max<int64_t>(L.blocks, 1)) * 4);
    for (int i = 0; i < nprob; ++i) {
        const tomos_cs_problem& q = pr[i];
        if (!group_candidate(q)) continue;
        L.base[i] = off;
        const int64_t splits = ceil_div(q.m, kGroupRowsPerSplit);
        off += a256(static_cast<size_t>(q.m) * 4) + a256(static_cast<size_t>(q.n) * 4) +
               a256(static_cast<size_t>(splits * q.s * q.n) * 8);
    }
    L.total = off;
    return L;
}

template <int FMT, bool EXACT>
void group_launch(const std::vector<CsGroupDesc>& hd, const int* cg_first,
                  const int* cg_count, const int64_t* blk_first, const int64_t* blk_count, CsGroupDesc* dd,
                  int* downer, int64_t max_m, int64_t max_n, int64_t max_sn, bool ntl, hipStream_t st) {
    const int nd = static_cast<int>(hd.size());
    for (int i0 = 0; i0 < nd; i0 += kCsChunk) {
        CsDescChunk c{};
        const int cnt = std::min(kCsChunk, nd - i0);
        for (int k = 0; k < cnt; ++k) c.d[k] = hd[static_cast<size_t>(i0 + k)];
        cs_set_desc_kernel<<<1, 64, 0, st>>>(dd + i0, c, cnt);
    }
    TOMOS_LAUNCH_CHECK("cs_set_desc_kernel");
    cs_owner_kernel<<<dim3(static_cast<unsigned>(nd)), 256, 0, st>>>(dd, downer);
    TOMOS_LAUNCH_CHECK("cs_owner_kernel");
    cs_group_index_kernel<<<dim3(static_cast<unsigned>(ceil_div(max_m, 256)), static_cast<unsigned>(nd)), 256, 0,
                            st>>>(dd);
    TOMOS_LAUNCH_CHECK("cs_group_index_kernel");
    constexpr bool kExactMax = EXACT && FMT == FMT_P16;
    cs_group_exp_kernel<FMT, kExactMax><<<dim3(static_cast<unsigned>(max_n), static_cast<unsigned>(nd)), 256, 0, st>>>(
        dd);
    TOMOS_LAUNCH_CHECK("cs_group_exp_kernel");
    const int widths[4] = {8, 4, 2, 1};
    for (int w = 0; w < 4; ++w) {
        if (cg_count[w] == 0) continue;
        int64_t smax = 0;
        for (int k = cg_first[w]; k < cg_first[w] + cg_count[w]; ++k) smax = std::max(smax, hd[k].s);
        const size_t lds = static_cast<size_t>(smax) * widths[w] * 8;
        const unsigned blocks = static_cast<unsigned>(blk_count[w]);
        const CsGroupDesc* d0 = dd + cg_first[w];
        const int* o0 = downer + blk_first[w];
        auto launch = [&](auto* k) {
            TOMOS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
            k<<<dim3(blocks), kVecThreads, lds, st>>>(d0, o0);
        };
        switch (widths[w] * 2 + (ntl ? 1 : 0)) {
            case 17: launch(cs_group_apply_kernel<FMT, EXACT, 8, true>); break;
            case 16: launch(cs_group_apply_kernel<FMT, EXACT, 8, false>); break;
            case 9: launch(cs_group_apply_kernel<FMT, EXACT, 4, true>); break;
            case 8: launch(cs_group_apply_kernel<FMT, EXACT, 4, false>); break;
            case 5: launch(cs_group_apply_kernel<FMT, EXACT, 2, true>); break;
            case 4: launch(cs_group_apply_kernel<FMT, EXACT, 2, false>); break;
            case 3: launch(cs_group_apply_kernel<FMT, EXACT, 1, true>); break;
            default: launch(cs_group_apply_kernel<FMT, EXACT, 1, false>); break;
        }
        TOMOS_LAUNCH_CHECK("cs_group_apply_kernel");
    }
    cs_group_finish_kernel<<<dim3(static_cast<unsigned>(ceil_div(max_sn, 256)), static_cast<unsigned>(nd)), 256, 0,
                             st>>>(dd);
    TOMOS_LAUNCH_CHECK("cs_group_finish_kernel");
}

}  // namespace

size_t countsketch_batched_scratch_bytes(int nprob, const tomos_cs_problem* pr) {
    size_t b = group_layout(nprob, pr).total;
    for (int i = 0; i < nprob; ++i) b = std::max(b, cs_scratch(pr[i].m, pr[i].n, pr[i].s));
    return b;
}

void countsketch_apply_batched(int a_type, int tier, int nprob, const tomos_cs_problem* pr, void* scratch,
                               size_t scratch_bytes, hipStream_t st) {
    require(nprob >= 0 && nprob <= 65535 && (nprob == 0 || pr), "countsketch_apply_batched: bad problem list");
    require(tier == TOMOS_TIER_DEFAULT || tier == TOMOS_TIER_FAST || tier == TOMOS_TIER_EXACT,
            "countsketch_apply_batched: bad tier");
    for (int i = 0; i < nprob; ++i) {
        const tomos_cs_problem& q = pr[i];
        require(q.m >= 0 && q.n >= 0 && q.s >= 1 && q.s <= kLdsBytes / 4 && q.m < (int64_t(1) << 31) &&
                    q.n < (int64_t(1) << 31),
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
    const GroupLayout L = group_layout(nprob, pr);
    const bool exact = tier == TOMOS_TIER_EXACT;
    int* flags = reinterpret_cast<int*>(sc + L.flags);
    // grouped problems ordered by column-group width (one apply launch each)
    std::vector<CsGroupDesc> hd;
    std::vector<int> which;
    int64_t nblocks = 0;
    int cg_first[4] = {0, 0, 0, 0}, cg_count[4] = {0, 0, 0, 0};
    int64_t blk_first[4] = {0, 0, 0, 0}, blk_count[4] = {0, 0, 0, 0};
    int64_t max_m = 0, max_n = 0, max_sn = 0, bytes = 0;
    const int64_t eb = (a_type == TOMOS_F64 ? 8 : a_type == TOMOS_F32 ? 4 : 2) *
                       (exact && (a_type == TOMOS_FP32_16 || a_type == TOMOS_FP32_BF16) ? 2 : 1);
    const int widths[4] = {8, 4, 2, 1};
    for (int w = 0; w < 4; ++w) {
        cg_first[w] = static_cast<int>(hd.size());
        blk_first[w] = nblocks;
        for (int i = 0; i < nprob; ++i) {
            const tomos_cs_problem& q = pr[i];
            if (!group_takes(a_type, exact, q) || group_cg(q.s) != widths[w]) continue;
            CsGroupDesc d{};
            d.a = q.a;
            d.wr = q.wr;
            char* pb = sc + L.base[i];
            d.packed = reinterpret_cast<uint32_t*>(pb);
            d.colexp = reinterpret_cast<int*>(pb + a256(static_cast<size_t>(q.m) * 4));
            d.partial = reinterpret_cast<unsigned long long*>(pb + a256(static_cast<size_t>(q.m) * 4) +
                                                              a256(static_cast<size_t>(q.n) * 4));
            d.flags = flags + 2 * i;
            d.y = q.y;
            d.m = q.m;
            d.n = q.n;
            d.lda = q.lda;
            d.s = q.s;
            d.ldy = q.ldy;
            d.splits = static_cast<int>(ceil_div(q.m, kGroupRowsPerSplit));
            d.rows_per_split = ceil_div(ceil_div(q.m, d.splits), 8) * 8;
            d.splits = static_cast<int>(ceil_div(q.m, d.rows_per_split));
            d.ncg = static_cast<int>(ceil_div(q.n, widths[w]));
            d.shift = fx_shift(q.m);
            if (q.m >= kSampleMinRows) {
                d.stride = (q.m / kSampleSegs) / 8 * 8;
                d.len = static_cast<int>(std::max<int64_t>(64, std::min<int64_t>(1024, (q.m / 256) / 8 * 8)));
                d.segs = static_cast<int>(kSampleSegs);
                d.margin = kSampleMargin;
            } else {  // every row: one segment
                d.stride = 0;
                d.len = static_cast<int>(ceil_div(q.m, 8) * 8);
                d.segs = 1;
                d.margin = 0;
            }
            d.block0 = static_cast<int>(nblocks - blk_first[w]);
            d.local = static_cast<int>(hd.size()) - cg_first[w];
            d.owner0 = static_cast<int>(nblocks);
            nblocks += static_cast<int64_t>(d.splits) * d.ncg;
            hd.push_back(d);
            which.push_back(i);
            max_m = std::max(max_m, q.m);
            max_n = std::max(max_n, q.n);
            max_sn = std::max(max_sn, q.s * q.n);
            bytes += q.m * q.n * eb;
        }
        cg_count[w] = static_cast<int>(hd.size()) - cg_first[w];
        blk_count[w] = nblocks - blk_first[w];
    }
    require(nblocks <= L.blocks && nblocks < (int64_t(1) << 31), "countsketch_apply_batched: grid too large");
    TOMOS_HIP_CHECK(hipMemsetAsync(flags, 0, static_cast<size_t>(nprob) * 8, st));
    std::vector<char> grouped(static_cast<size_t>(nprob), 0);
    if (!hd.empty()) {
        CsGroupDesc* dd = reinterpret_cast<CsGroupDesc*>(sc + L.desc);
        int* downer = reinterpret_cast<int*>(sc + L.owner);
        const bool ntl = stream_loads(bytes);
        switch (a_type) {
            case TOMOS_F32:
                group_launch<FMT_F32, false>(hd, cg_first, cg_count, blk_first, blk_count, dd, downer, max_m,
                                             max_n, max_sn, ntl, st);
                break;
            case TOMOS_F64:
                group_launch<FMT_F64, false>(hd, cg_first, cg_count, blk_first, blk_count, dd, downer, max_m,
                                             max_n, max_sn, ntl, st);
                break;
            case TOMOS_FP32_16:
                if (exact)
                    group_launch<FMT_P16, true>(hd, cg_first, cg_count, blk_first, blk_count, dd, downer,
                                                max_m, max_n, max_sn, ntl, st);
                else
                    group_launch<FMT_P16, false>(hd, cg_first, cg_count, blk_first, blk_count, dd, downer,
                                                 max_m, max_n, max_sn, ntl, st);
                break;
            default:
                if (exact)
                    group_launch<FMT_PBF, true>(hd, cg_first, cg_count, blk_first, blk_count, dd, downer,
                                                max_m, max_n, max_sn, ntl, st);
                else
                    group_launch<FMT_PBF, false>(hd, cg_first, cg_count, blk_first, blk_count, dd, downer,
                                                 max_m, max_n, max_sn, ntl, st);
                break;
        }
        for (int i : which) grouped[static_cast<size_t>(i)] = 1;
    }
    std::vector<int> h(2 * static_cast<size_t>(nprob), 0);
    TOMOS_HIP_CHECK(hipMemcpyAsync(h.data(), flags, h.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    TOMOS_HIP_CHECK(hipStreamSynchronize(st));
    for (int i = 0; i < nprob; ++i) {
        const tomos_cs_problem& q = pr[i];
        if (q.n == 0) continue;
        if (grouped[static_cast<size_t>(i)] && !h[2 * i] && !h[2 * i + 1]) continue;
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
