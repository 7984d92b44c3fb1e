// Internal GEMV interface shared by the C ABI, LSQR and CountSketch code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace tomos_hip {

// alpha/beta either from the host values or, when the device pointer is
// non-null, read on the device (LSQR keeps its scalars on the GPU).
struct Coef {
    double alpha = 1.0;
    double beta = 0.0;
    const double* dalpha = nullptr;
    const double* dbeta = nullptr;
    double alpha_mul = 1.0;  // multiplies *dalpha
    double beta_mul = 1.0;   // multiplies *dbeta
    const int* skip = nullptr;  // device flag: when non-zero the launch is a no-op
};

size_t gemv_scratch_bytes(int a_type, int tier, int trans, int64_t m, int64_t n);

// y = alpha*op(A)*x + beta*y.  Types/tiers as in tomos_hip.h.  scratch must
// hold gemv_scratch_bytes(); throws tomos_hip::Error on bad arguments.
void gemv(int a_type, int tier, int trans, const Coef& coef, int64_t m, int64_t n,
          const void* a, const uint16_t* wr, int64_t lda, int x_type, const void* x,
          int y_type, void* y, void* scratch, size_t scratch_bytes, hipStream_t s);

}  // namespace tomos_hip
