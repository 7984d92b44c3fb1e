// Dense Gaussian sketch: fill an (rows x cols, ld) column-major matrix with
// i.i.d. N(0,1) values (create_gaussian_sketch(Dense&),
// src/dispatch/gaussian_sketch_registries.cpp:16-30,
// src/kernels/cuda/sketch/gaussian_sketch_cuda_kernels.cu:21-45).
//
// The reference seeds its generator from the wall clock, so its sketches
// are not reproducible; this one is counter-based and takes the seed as an
// argument, so a (seed, shape) pair always gives the same matrix, whatever
// ld, stream or launch shape.
//
// Generator.  Entries are produced in pairs: pair p covers rows 2q and
// 2q+1 of column j, p = j * ceil(rows / 2) + q.  Its 64 random bits are
// the p-th output of SplitMix64 started at `seed` (Steele, Lea & Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014):
// z = mix(seed + (p + 1) * 0x9e3779b97f4a7c15).  The top 23 bits give
// u1 = (b + 1/2) 2^-23 in (0, 1), bits 8..31 give u2 = b 2^-24 in [0, 1),
// and the Box-Muller transform (Box & Muller 1958) turns them into two
// independent normals r cos(2 pi u2), r sin(2 pi u2), r = sqrt(-2 ln u1).
// The transform runs in FP32 (device logf, sqrtf, sincospif), the way the
// reference computes its values in float before the cast.  The output
// type's conversion rounds to nearest even.
//
// Roofline: each entry is written once and nothing is read, so the kernel
// is bounded by HBM writes (sizeof(T) bytes per entry) or by the VALU
// work of the mixer and the transform (about 100 operations per pair).
#include "codec.hpp"
#include "common.hpp"
#include "tomos_hip.h"

namespace tomos_hip {
namespace {

constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ull;
constexpr int kGaussThreads = 256;

__device__ __forceinline__ uint64_t sm64_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float2 box_muller(uint64_t z) {
    // b + 1/2 with b < 2^23 is exact in FP32 (with 24-bit b it would round
    // and move u1 near 1, where r is most sensitive to it)
    const float u1 = (static_cast<float>(static_cast<uint32_t>(z >> 41)) + 0.5f) * 0x1p-23f;
    const float u2 = static_cast<float>(static_cast<uint32_t>(z >> 8) & 0xFFFFFFu) * 0x1p-24f;
    // logf keeps its relative accuracy for u1 near 1, where r is small
    // (the bare v_log_f32 does not: its absolute error there, amplified by
    // the square root, reached 2.6e-5); sincospif takes 2 u2 exactly, with
    // no multiplication by pi to round.
    const float r = sqrtf(-2.0f * logf(u1));
    float s, c;
    sincospif(2.0f * u2, &s, &c);
    return make_float2(r * c, r * s);
}

// Two adjacent entries of one column; VEC: the pair is stored with one
// aligned store (base and ld even, checked on the host).
template <int TYPE, bool VEC>
__device__ __forceinline__ void store_pair(void* col, int64_t row, float a, float b, bool two) {
    if constexpr (TYPE == TOMOS_F64) {
        double* p = static_cast<double*>(col) + row;
        if (VEC && two) *reinterpret_cast<double2*>(p) = make_double2(a, b);
        else { p[0] = a; if (two) p[1] = b; }
    } else if constexpr (TYPE == TOMOS_F32) {
        float* p = static_cast<float*>(col) + row;
        if (VEC && two) *reinterpret_cast<float2*>(p) = make_float2(a, b);
        else { p[0] = a; if (two) p[1] = b; }
    } else {
        uint32_t ha, hb;
        if constexpr (TYPE == TOMOS_F16) {
            const _Float16 x = static_cast<_Float16>(a), y = static_cast<_Float16>(b);
            ha = __builtin_bit_cast(uint16_t, x);
            hb = __builtin_bit_cast(uint16_t, y);
        } else {
            ha = codec::f32_to_bf16_bits(a);
            hb = codec::f32_to_bf16_bits(b);
        }
        uint16_t* p = static_cast<uint16_t*>(col) + row;
        if (VEC && two) *reinterpret_cast<uint32_t*>(p) = ha | (hb << 16);
        else { p[0] = static_cast<uint16_t>(ha); if (two) p[1] = static_cast<uint16_t>(hb); }
    }
}

// grid.x covers the row pairs of a column (kGaussThreads per block),
// grid.y strides over the columns.
template <int TYPE, bool VEC>
__global__ __launch_bounds__(kGaussThreads) void gaussian_fill_kernel(int64_t rows, int64_t cols,
                                                                      void* out, int64_t ld,
                                                                      uint64_t seed) {
    constexpr int64_t esz = TYPE == TOMOS_F64 ? 8 : TYPE == TOMOS_F32 ? 4 : 2;
    const int64_t half = (rows + 1) >> 1;
    const int64_t q = static_cast<int64_t>(blockIdx.x) * kGaussThreads + threadIdx.x;
    if (q >= half) return;
    const int64_t row = 2 * q;
    const bool two = row + 1 < rows;
    for (int64_t j = blockIdx.y; j < cols; j += gridDim.y) {
        const uint64_t p = static_cast<uint64_t>(j) * static_cast<uint64_t>(half) + static_cast<uint64_t>(q);
        const float2 g = box_muller(sm64_mix(seed + (p + 1) * kGolden));
        store_pair<TYPE, VEC>(static_cast<char*>(out) + j * ld * esz, row, g.x, g.y, two);
    }
}

template <int TYPE>
void launch_fill(int64_t rows, int64_t cols, void* out, int64_t ld, uint64_t seed, hipStream_t st) {
    constexpr int64_t esz = TYPE == TOMOS_F64 ? 8 : TYPE == TOMOS_F32 ? 4 : 2;
    const int64_t half = (rows + 1) >> 1;
    const bool vec = (reinterpret_cast<uintptr_t>(out) % (2 * esz) == 0) && (ld % 2 == 0);
    const dim3 grid(static_cast<unsigned>(ceil_div(half, kGaussThreads)),
                    static_cast<unsigned>(cols < 65535 ? cols : 65535));
    if (vec) gaussian_fill_kernel<TYPE, true><<<grid, kGaussThreads, 0, st>>>(rows, cols, out, ld, seed);
    else gaussian_fill_kernel<TYPE, false><<<grid, kGaussThreads, 0, st>>>(rows, cols, out, ld, seed);
    TOMOS_LAUNCH_CHECK("gaussian_fill_kernel");
}

// One entry per sketch row (the reference's CSR sketches,
// generate_countsketch_csr_idxs, countsketch_kernels.cpp:34-53): row i
// holds column z % cols, z the i-th SplitMix64 output from seed ^ kColSalt;
// kind 0 stores the sign of z's top bit (+-1), kind 1 leaves the values to
// the Gaussian fill.
constexpr uint64_t kColSalt = 0xc2b2ae3d27d4eb4full;

template <class I, int VT>
__global__ void csr_sketch_pattern_kernel(int64_t rows, int64_t cols, uint64_t seed, I* rp, I* ci, void* vals,
                                          bool signs, bool pattern) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i <= rows; i += stride) {
        if (pattern) rp[i] = static_cast<I>(i);
        if (i == rows) break;
        const uint64_t z = sm64_mix((seed ^ kColSalt) + (static_cast<uint64_t>(i) + 1) * kGolden);
        if (pattern) ci[i] = static_cast<I>(z % static_cast<uint64_t>(cols));
        if (signs) {
            const float v = (z >> 63) ? 1.0f : -1.0f;
            if constexpr (VT == TOMOS_F64) static_cast<double*>(vals)[i] = v;
            else if constexpr (VT == TOMOS_F32) static_cast<float*>(vals)[i] = v;
            else if constexpr (VT == TOMOS_F16) static_cast<uint16_t*>(vals)[i] = (z >> 63) ? 0x3C00 : 0xBC00;
            else static_cast<uint16_t*>(vals)[i] = (z >> 63) ? 0x3F80 : 0xBF80;
        }
    }
}

template <class I>
void launch_pattern(int vt, int64_t rows, int64_t cols, uint64_t seed, void* rp, void* ci, void* vals, bool signs,
                    bool pattern, hipStream_t st) {
    const unsigned g = static_cast<unsigned>(std::min<int64_t>(std::max<int64_t>(ceil_div(rows + 1, 256), 1), 8192));
    I* r = static_cast<I*>(rp);
    I* c = static_cast<I*>(ci);
    switch (vt) {
        case TOMOS_F64: csr_sketch_pattern_kernel<I, TOMOS_F64><<<g, 256, 0, st>>>(rows, cols, seed, r, c, vals, signs, pattern); break;
        case TOMOS_F32: csr_sketch_pattern_kernel<I, TOMOS_F32><<<g, 256, 0, st>>>(rows, cols, seed, r, c, vals, signs, pattern); break;
        case TOMOS_F16: csr_sketch_pattern_kernel<I, TOMOS_F16><<<g, 256, 0, st>>>(rows, cols, seed, r, c, vals, signs, pattern); break;
        default: csr_sketch_pattern_kernel<I, TOMOS_BF16><<<g, 256, 0, st>>>(rows, cols, seed, r, c, vals, signs, pattern); break;
    }
    TOMOS_LAUNCH_CHECK("csr_sketch_pattern_kernel");
}

}  // namespace
}  // namespace tomos_hip

using namespace tomos_hip;

extern "C" int tomos_csr_sketch(int kind, int val_type, int idx_type, int64_t rows, int64_t cols, void* vals,
                                void* row_ptrs, void* col_idxs, uint64_t seed, tomos_stream_t st) {
    return guarded([&] {
        require(kind >= 0 && kind <= 3,
                "tomos_csr_sketch: kind must be 0 (+-1), 1 (Gaussian), 2 (pattern only) or 3 (+-1 values only)");
        require(val_type == TOMOS_F64 || val_type == TOMOS_F32 || val_type == TOMOS_F16 || val_type == TOMOS_BF16,
                "tomos_csr_sketch: values must be F64, F32, F16 or BF16");
        require(idx_type == TOMOS_INDEX_INT32 || idx_type == TOMOS_INDEX_INT64, "tomos_csr_sketch: bad index type");
        require(rows >= 0 && cols >= 1, "tomos_csr_sketch: need rows >= 0 and cols >= 1");
        require(idx_type == TOMOS_INDEX_INT64 || (rows <= 0x7FFFFFFF && cols <= 0x7FFFFFFF),
                "tomos_csr_sketch: sizes exceed INT32 indices");
        const bool pattern = kind != 3, signs = kind == 0 || kind == 3;
        require((!pattern || row_ptrs) && (rows == 0 || ((vals || kind == 2) && (col_idxs || kind == 3))),
                "tomos_csr_sketch: null output");
        if (idx_type == TOMOS_INDEX_INT64)
            launch_pattern<int64_t>(val_type, rows, cols, seed, row_ptrs, col_idxs, vals, signs, pattern, st);
        else
            launch_pattern<int32_t>(val_type, rows, cols, seed, row_ptrs, col_idxs, vals, signs, pattern, st);
        if (kind == 1 && rows > 0) {
            switch (val_type) {
                case TOMOS_F64: launch_fill<TOMOS_F64>(rows, 1, vals, rows, seed, st); break;
                case TOMOS_F32: launch_fill<TOMOS_F32>(rows, 1, vals, rows, seed, st); break;
                case TOMOS_F16: launch_fill<TOMOS_F16>(rows, 1, vals, rows, seed, st); break;
                default: launch_fill<TOMOS_BF16>(rows, 1, vals, rows, seed, st); break;
            }
        }
    });
}

extern "C" int tomos_gaussian_fill(int type, int64_t rows, int64_t cols, void* out, int64_t ld,
                                   uint64_t seed, tomos_stream_t st) {
    return guarded([&] {
        require(rows >= 0 && cols >= 0, "tomos_gaussian_fill: negative size");
        require(ld >= (rows > 1 ? rows : 1), "tomos_gaussian_fill: ld < rows");
        if (rows == 0 || cols == 0) return;
        require(out != nullptr, "tomos_gaussian_fill: null output");
        require(ceil_div((rows + 1) >> 1, kGaussThreads) <= 0x7FFFFFFF, "tomos_gaussian_fill: too many rows");
        switch (type) {
            case TOMOS_F64: launch_fill<TOMOS_F64>(rows, cols, out, ld, seed, st); break;
            case TOMOS_F32: launch_fill<TOMOS_F32>(rows, cols, out, ld, seed, st); break;
            case TOMOS_F16: launch_fill<TOMOS_F16>(rows, cols, out, ld, seed, st); break;
            case TOMOS_BF16: launch_fill<TOMOS_BF16>(rows, cols, out, ld, seed, st); break;
            default: throw Error(kErrUnsupported, "tomos_gaussian_fill: type must be F64, F32, F16 or BF16");
        }
    });
}
