// Shared runtime helpers for libtomos_hip: error capture for the C ABI,
// launch checking, device properties, and the per-device scratch cache
// used when a caller passes no scratch buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace tomos_hip {

// Thrown inside the library, converted to a status code + message at the
// extern "C" boundary (exceptions never cross the ABI).
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum : int {
    kOk = 0,
    kErrInvalidArgument = 1,
    kErrHip = 2,
    kErrUnsupported = 3,
    kErrWorkspace = 4,
    kErrNotConverged = 5,
    kErrSingular = 6,
    kErrInternal = 7,
};

// 16-byte loads of matrix data.  A matrix far larger than the caches is
// streamed with non-temporal loads: on gfx950 a 4 GB read stream runs at
// ~6.9 TB/s that way against ~6.0 TB/s with plain loads.  A matrix that
// fits the 256 MB MALL is better left cacheable, because an iterative
// solver reads it again a few microseconds later (the batched problems).
__device__ __forceinline__ uint4 load_stream16(const void* p) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 v = __builtin_nontemporal_load(static_cast<const u32x4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}
template <bool NTL>
__device__ __forceinline__ uint4 load_a16(const void* p) {
    if constexpr (NTL) return load_stream16(p);
    else return *static_cast<const uint4*>(p);
}

// Load policy for a pass that reads `bytes` of matrix data: true = use
// non-temporal loads.  Auto mode picks them above kStreamLoadBytes; the
// process-wide override (tomos_stream_loads_set) forces either kind.
constexpr int64_t kStreamLoadBytes = int64_t(192) << 20;
bool stream_loads(int64_t bytes);
void set_stream_load_mode(int mode);  // 0 auto, 1 always non-temporal, 2 never

void set_last_error(const std::string& msg);
const char* last_error();

inline void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw Error(kErrHip, std::string(what) + ": " + hipGetErrorString(e));
}

#define TOMOS_HIP_CHECK(expr) ::tomos_hip::check_hip((expr), #expr)
#define TOMOS_LAUNCH_CHECK(name) ::tomos_hip::check_hip(hipGetLastError(), name)

inline void require(bool cond, const std::string& msg) {
    if (!cond) throw Error(kErrInvalidArgument, msg);
}

// Number of compute units of the current device (cached per device).
int cu_count();

// A second stream of the current device (created on first use, kept until
// the process ends) and events to fork work onto it and join it back.
struct SideStream {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr;
    hipEvent_t join[2] = {nullptr, nullptr};
};
const SideStream& side_stream();

// Per-device grow-only scratch for callers that pass no scratch buffer.
// Not thread-safe; kept until process exit.
void* fallback_scratch(size_t bytes);

// Wrap a C-ABI body: returns kOk, or an error code with the message stored
// for tomos_last_error().
template <class F>
int guarded(F&& body) {
    try {
        body();
        return kOk;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return kErrInternal;
    } catch (...) {
        set_last_error("unknown error");
        return kErrInternal;
    }
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace tomos_hip
