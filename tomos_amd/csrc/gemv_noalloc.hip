// tomos_gemv_noalloc: the GEMV of tomos_gemv for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  The partial sums that tomos_gemv keeps
// in a caller- or library-provided scratch buffer (the chunk sums of a
// TRANSPOSE GEMV, the column-split sums of a short NO_TRANSPOSE GEMV) go to
// `scratch`, a fixed array that the code object defines: every GPU the
// library runs on holds its own copy until the process ends.
//
// The kernels write every partial they read within the same call, so the
// array needs no clearing between calls.  Calls that overlap on one GPU
// (two streams, or two host threads) share the array and race.
#include <mutex>
#include <string>

#include "common.hpp"
#include "gemv.hpp"
#include "tomos_hip.h"

static __device__ __align__(256) unsigned char scratch[TOMOS_GEMV_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of the current GPU's copy of `scratch`, looked up once per
// device with that device current.  The host-side name of the array is not
// a device address: only the pointer returned here is passed to kernels.
void* device_scratch() {
    int dev = 0;
    TOMOS_HIP_CHECK(hipGetDevice(&dev));
    tomos_hip::require(dev >= 0 && dev < kMaxDevices, "tomos_gemv_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[dev] == nullptr) {
        void* p = nullptr;
        TOMOS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        tomos_hip::require(p != nullptr, "tomos_gemv_noalloc: no device address for the scratch array");
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

}  // namespace

extern "C" {

int tomos_gemv_noalloc(int a_type, int tier, int trans, double alpha, int64_t m, int64_t n,
                       const void* a, const uint16_t* wr, int64_t lda, int x_type, const void* x,
                       double beta, int y_type, void* y, tomos_stream_t s) {
    return tomos_hip::guarded([&] {
        tomos_hip::require(m >= 0 && n >= 0, "tomos_gemv_noalloc: negative size");
        const size_t need = tomos_hip::gemv_scratch_bytes(a_type, tier, trans, m, n);
        if (need > static_cast<size_t>(TOMOS_GEMV_NOALLOC_SCRATCH_BYTES))
            throw tomos_hip::Error(tomos_hip::kErrWorkspace,
                                   "tomos_gemv_noalloc: this shape needs " + std::to_string(need) +
                                       " scratch bytes, above the fixed " +
                                       std::to_string(static_cast<size_t>(TOMOS_GEMV_NOALLOC_SCRATCH_BYTES)) +
                                       "; call tomos_gemv with a scratch buffer");
        void* p = device_scratch();
        tomos_hip::Coef c;
        c.alpha = alpha;
        c.beta = beta;
        tomos_hip::gemv(a_type, tier, trans, c, m, n, a, wr, lda, x_type, x, y_type, y, p,
                        static_cast<size_t>(TOMOS_GEMV_NOALLOC_SCRATCH_BYTES), s);
    });
}

}  // extern "C"
