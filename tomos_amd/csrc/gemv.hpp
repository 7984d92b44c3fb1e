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
    double* sumsq = nullptr;    // NO_TRANSPOSE: per-block sums of y^2 (as stored) go here
};

size_t gemv_scratch_bytes(int a_type, int tier, int trans, int64_t m, int64_t n);

// y = alpha*op(A)*x + beta*y.  Types/tiers as in tomos_hip.h.  scratch must
// hold gemv_scratch_bytes(); throws tomos_hip::Error on bad arguments.
// With coef.sumsq set (NO_TRANSPOSE only), *sumsq_blocks receives the
// number of partial sums written (their total is |y|^2 in FP64).
void gemv(int a_type, int tier, int trans, const Coef& coef, int64_t m, int64_t n,
          const void* a, const uint16_t* wr, int64_t lda, int x_type, const void* x,
          int y_type, void* y, void* scratch, size_t scratch_bytes, hipStream_t s,
          int* sumsq_blocks = nullptr);

// Upper bound on *sumsq_blocks for an m-row NO_TRANSPOSE GEMV.
int64_t gemv_sumsq_blocks_max(int64_t m);

// TRANSPOSE partials only: chunk sums of A^T x written to scratch as
// [chunks][n] doubles (the tier must accumulate in FP64: FAST or EXACT, or
// an F64 matrix); returns the number of chunks.
int gemv_t_partials(int a_type, int tier, int64_t m, int64_t n, const void* a,
                    const uint16_t* wr, int64_t lda, int x_type, const void* x, void* scratch,
                    size_t scratch_bytes, const int* skip, hipStream_t s);

}  // namespace tomos_hip
