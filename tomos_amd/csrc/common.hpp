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
