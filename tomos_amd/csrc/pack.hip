// Pack (FP32 -> planar wh/wr) and unpack/decode kernels.
//
// Layout: column-major m x n source with leading dimension ld_src, planar
// uint16 wh/wr destinations with leading dimension ld_dst.  One thread
// handles a run of 8 consecutive rows of one column: 8 fp32 loads (32 B)
// and two 16-B stores when the rows are contiguous and aligned, scalar
// otherwise.  Grid = (row runs, columns): fully coalesced along rows,
// HBM-bound (8 B moved per element).
#include "codec.hpp"
#include "common.hpp"
#include "tomos_hip.h"

namespace tomos_hip {

namespace {

template <int VARIANT>
__device__ __forceinline__ void encode(float v, uint16_t& h, uint16_t& r) {
    if constexpr (VARIANT == TOMOS_PACK_FP32_16) codec::fp32_16_encode<0>(v, h, r);
    else if constexpr (VARIANT == TOMOS_PACK_FP32_16_CLAMPED) codec::fp32_16_encode<1>(v, h, r);
    else if constexpr (VARIANT == TOMOS_PACK_FP32_16_TRUNCATED) codec::fp32_16_encode<2>(v, h, r);
    else if constexpr (VARIANT == TOMOS_PACK_FP32_BF16) codec::fp32_bf16_encode<false>(v, h, r);
    else codec::fp32_bf16_encode<true>(v, h, r);
}

constexpr int kRun = 8;  // rows per thread

template <int VARIANT, bool VECTOR>
__global__ __launch_bounds__(256) void pack_kernel(int64_t m, const float* __restrict__ src,
                                                   int64_t ld_src, uint16_t* __restrict__ wh,
                                                   uint16_t* __restrict__ wr, int64_t ld_dst) {
    const int64_t j = blockIdx.y;
    const int64_t i0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * kRun;
    if (i0 >= m) return;
    const float* s = src + j * ld_src + i0;
    uint16_t* h = wh + j * ld_dst + i0;
    uint16_t* r = wr + j * ld_dst + i0;
    if (VECTOR && i0 + kRun <= m) {
        const float4 a = reinterpret_cast<const float4*>(s)[0];
        const float4 b = reinterpret_cast<const float4*>(s)[1];
        const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        uint16_t hh[8], rr[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) encode<VARIANT>(v[k], hh[k], rr[k]);
        uint4 ph, pr;
        ph.x = hh[0] | (uint32_t(hh[1]) << 16);
        ph.y = hh[2] | (uint32_t(hh[3]) << 16);
        ph.z = hh[4] | (uint32_t(hh[5]) << 16);
        ph.w = hh[6] | (uint32_t(hh[7]) << 16);
        pr.x = rr[0] | (uint32_t(rr[1]) << 16);
        pr.y = rr[2] | (uint32_t(rr[3]) << 16);
        pr.z = rr[4] | (uint32_t(rr[5]) << 16);
        pr.w = rr[6] | (uint32_t(rr[7]) << 16);
        *reinterpret_cast<uint4*>(h) = ph;
        *reinterpret_cast<uint4*>(r) = pr;
    } else {
        const int cnt = static_cast<int>(min<int64_t>(kRun, m - i0));
        for (int k = 0; k < cnt; ++k) encode<VARIANT>(s[k], h[k], r[k]);
    }
}

template <int DEC>
__device__ __forceinline__ float decode_f32(uint32_t h, uint32_t r) {
    if constexpr (DEC == TOMOS_DECODE_FP32_16_EXACT) return codec::fp32_16_decode_exact(h, r);
    else if constexpr (DEC == TOMOS_DECODE_FP32_16_HALF)
        return codec::half_bits_to_f32(codec::fp32_16_half_bits(h));
    else if constexpr (DEC == TOMOS_DECODE_FP32_16_HALF_FAST)
        return codec::half_bits_to_f32(codec::fp32_16_half_bits_fast(h));
    else if constexpr (DEC == TOMOS_DECODE_FP32_16_BF16)
        return codec::u2f(codec::fp32_16_bf16_bits(h) << 16);
    else if constexpr (DEC == TOMOS_DECODE_FP32_BF16_EXACT) return codec::fp32_bf16_decode_exact(h, r);
    else if constexpr (DEC == TOMOS_DECODE_FP32_BF16_HI) return codec::fp32_bf16_decode_hi(h);
    else if constexpr (DEC == TOMOS_DECODE_FP32_BF16_EXACT_ROUNDED)
        return codec::fp32_bf16_decode_exact_rounded(h, r);
    else return codec::fp32_16_decode_exact_truncated(h, r);
}

template <int DEC>
constexpr bool needs_wr() {
    return DEC == TOMOS_DECODE_FP32_16_EXACT || DEC == TOMOS_DECODE_FP32_BF16_EXACT ||
           DEC == TOMOS_DECODE_FP32_BF16_EXACT_ROUNDED ||
           DEC == TOMOS_DECODE_FP32_16_EXACT_TRUNCATED;
}

template <int DEC>
__global__ __launch_bounds__(256) void unpack_kernel(int64_t m, const uint16_t* __restrict__ wh,
                                                     const uint16_t* __restrict__ wr,
                                                     int64_t ld_src, float* __restrict__ dst,
                                                     int64_t ld_dst) {
    const int64_t j = blockIdx.y;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t h = wh[i + j * ld_src];
    const uint32_t r = needs_wr<DEC>() ? wr[i + j * ld_src] : 0u;
    dst[i + j * ld_dst] = decode_f32<DEC>(h, r);
}

template <int DEC>
__global__ __launch_bounds__(256) void decode_wh_kernel(int64_t count,
                                                        const uint16_t* __restrict__ wh,
                                                        uint16_t* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t h = wh[i];
    uint32_t o;
    if constexpr (DEC == TOMOS_DECODE_FP32_16_HALF) o = codec::fp32_16_half_bits(h);
    else if constexpr (DEC == TOMOS_DECODE_FP32_16_HALF_FAST) o = codec::fp32_16_half_bits_fast(h);
    else if constexpr (DEC == TOMOS_DECODE_FP32_16_BF16) o = codec::fp32_16_bf16_bits(h);
    else o = h;  // FP32_BF16 hi-only: wh already is the bf16 pattern
    out[i] = static_cast<uint16_t>(o);
}

template <int VARIANT>
void launch_pack(int64_t m, int64_t n, const float* src, int64_t ld_src, uint16_t* wh,
                 uint16_t* wr, int64_t ld_dst, hipStream_t s) {
    const bool vec = (ld_src % 4 == 0) && (ld_dst % 8 == 0) &&
                     (reinterpret_cast<uintptr_t>(src) % 16 == 0) &&
                     (reinterpret_cast<uintptr_t>(wh) % 16 == 0) &&
                     (reinterpret_cast<uintptr_t>(wr) % 16 == 0);
    const dim3 grid(static_cast<unsigned>(ceil_div(m, 256 * kRun)), static_cast<unsigned>(n));
    if (vec)
        pack_kernel<VARIANT, true><<<grid, 256, 0, s>>>(m, src, ld_src, wh, wr, ld_dst);
    else
        pack_kernel<VARIANT, false><<<grid, 256, 0, s>>>(m, src, ld_src, wh, wr, ld_dst);
    TOMOS_LAUNCH_CHECK("pack_kernel");
}

template <int DEC>
void launch_unpack(int64_t m, int64_t n, const uint16_t* wh, const uint16_t* wr, int64_t ld_src,
                   float* dst, int64_t ld_dst, hipStream_t s) {
    const dim3 grid(static_cast<unsigned>(ceil_div(m, 256)), static_cast<unsigned>(n));
    unpack_kernel<DEC><<<grid, 256, 0, s>>>(m, wh, wr, ld_src, dst, ld_dst);
    TOMOS_LAUNCH_CHECK("unpack_kernel");
}

}  // namespace

}  // namespace tomos_hip

using namespace tomos_hip;

extern "C" {

int tomos_pack(int variant, int64_t m, int64_t n, const float* src, int64_t ld_src, uint16_t* wh,
               uint16_t* wr, int64_t ld_dst, tomos_stream_t s) {
    return guarded([&] {
        require(m >= 0 && n >= 0, "tomos_pack: negative size");
        require(ld_src >= m && ld_dst >= m, "tomos_pack: leading dimension < m");
        require(n <= 2147483647, "tomos_pack: too many columns");
        if (m == 0 || n == 0) return;
        require(src && wh && wr, "tomos_pack: null pointer");
        switch (variant) {
            case TOMOS_PACK_FP32_16:
                launch_pack<TOMOS_PACK_FP32_16>(m, n, src, ld_src, wh, wr, ld_dst, s);
                break;
            case TOMOS_PACK_FP32_16_CLAMPED:
                launch_pack<TOMOS_PACK_FP32_16_CLAMPED>(m, n, src, ld_src, wh, wr, ld_dst, s);
                break;
            case TOMOS_PACK_FP32_16_TRUNCATED:
                launch_pack<TOMOS_PACK_FP32_16_TRUNCATED>(m, n, src, ld_src, wh, wr, ld_dst, s);
                break;
            case TOMOS_PACK_FP32_BF16:
                launch_pack<TOMOS_PACK_FP32_BF16>(m, n, src, ld_src, wh, wr, ld_dst, s);
                break;
            case TOMOS_PACK_FP32_BF16_ROUNDED:
                launch_pack<TOMOS_PACK_FP32_BF16_ROUNDED>(m, n, src, ld_src, wh, wr, ld_dst, s);
                break;
            default:
                throw Error(kErrInvalidArgument, "tomos_pack: unknown variant");
        }
    });
}

int tomos_unpack(int dec, int64_t m, int64_t n, const uint16_t* wh, const uint16_t* wr,
                 int64_t ld_src, float* dst, int64_t ld_dst, tomos_stream_t s) {
    return guarded([&] {
        require(m >= 0 && n >= 0, "tomos_unpack: negative size");
        require(ld_src >= m && ld_dst >= m, "tomos_unpack: leading dimension < m");
        require(n <= 2147483647, "tomos_unpack: too many columns");
        if (m == 0 || n == 0) return;
        require(wh && dst, "tomos_unpack: null pointer");
        const bool exact = dec == TOMOS_DECODE_FP32_16_EXACT || dec == TOMOS_DECODE_FP32_BF16_EXACT ||
                           dec == TOMOS_DECODE_FP32_BF16_EXACT_ROUNDED ||
                           dec == TOMOS_DECODE_FP32_16_EXACT_TRUNCATED;
        require(!exact || wr != nullptr, "tomos_unpack: exact decoders need the wr plane");
        switch (dec) {
#define TOMOS_UNPACK_CASE(D) \
    case D: launch_unpack<D>(m, n, wh, wr, ld_src, dst, ld_dst, s); break;
            TOMOS_UNPACK_CASE(TOMOS_DECODE_FP32_16_EXACT)
            TOMOS_UNPACK_CASE(TOMOS_DECODE_FP32_16_HALF)
            TOMOS_UNPACK_CASE(TOMOS_DECODE_FP32_16_HALF_FAST)
            TOMOS_UNPACK_CASE(TOMOS_DECODE_FP32_16_BF16)
            TOMOS_UNPACK_CASE(TOMOS_DECODE_FP32_BF16_EXACT)
            TOMOS_UNPACK_CASE(TOMOS_DECODE_FP32_BF16_HI)
            TOMOS_UNPACK_CASE(TOMOS_DECODE_FP32_BF16_EXACT_ROUNDED)
            TOMOS_UNPACK_CASE(TOMOS_DECODE_FP32_16_EXACT_TRUNCATED)
#undef TOMOS_UNPACK_CASE
            default:
                throw Error(kErrInvalidArgument, "tomos_unpack: unknown decoder");
        }
    });
}

int tomos_decode_wh(int dec, int64_t count, const uint16_t* wh, uint16_t* out, tomos_stream_t s) {
    return guarded([&] {
        require(count >= 0, "tomos_decode_wh: negative count");
        if (count == 0) return;
        require(wh && out, "tomos_decode_wh: null pointer");
        const dim3 grid(static_cast<unsigned>(ceil_div(count, 256)));
        switch (dec) {
            case TOMOS_DECODE_FP32_16_HALF:
                decode_wh_kernel<TOMOS_DECODE_FP32_16_HALF><<<grid, 256, 0, s>>>(count, wh, out);
                break;
            case TOMOS_DECODE_FP32_16_HALF_FAST:
                decode_wh_kernel<TOMOS_DECODE_FP32_16_HALF_FAST><<<grid, 256, 0, s>>>(count, wh, out);
                break;
            case TOMOS_DECODE_FP32_16_BF16:
                decode_wh_kernel<TOMOS_DECODE_FP32_16_BF16><<<grid, 256, 0, s>>>(count, wh, out);
                break;
            case TOMOS_DECODE_FP32_BF16_HI:
                decode_wh_kernel<TOMOS_DECODE_FP32_BF16_HI><<<grid, 256, 0, s>>>(count, wh, out);
                break;
            default:
                throw Error(kErrInvalidArgument,
                            "tomos_decode_wh: decoder must produce 16-bit values");
        }
        TOMOS_LAUNCH_CHECK("decode_wh_kernel");
    });
}

}  // extern "C"
