// Single-pass LSQR matvec pair (A p and A^T u from one read of A).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace tomos_hip {

struct FusedPlan {
    int nblocks = 0;
    int64_t tiles_per_block = 0;
    int splits = 1;
    size_t tpart_bytes = 0;  // [nblocks][n] float block partials of A^T u_raw
    size_t tsum_bytes = 0;   // [splits][n] double level-1 sums
};

// A in a_type with the per-iteration (DEFAULT) tier, column-major,
// 16-byte aligned, lda a multiple of the 16-byte vector, n <= 1536.
bool lsqr_fused_supported(int a_type, int tier, int64_t m, int64_t n, int64_t lda, const void* a);
FusedPlan lsqr_fused_plan(int a_type, int64_t m, int64_t n);

// u_raw <- A p + (*u_coef) u_raw for all m rows; sumsq[b] = sum of u_raw^2
// over block b's rows (plan.nblocks entries); tsum[s][j] = sum over the
// blocks of split s of (A^T u_raw)_j.  Every launch is a no-op when *skip.
void lsqr_fused_launch(int a_type, int64_t m, int64_t n, const void* a, int64_t lda, const void* p,
                       bool p64, void* u, bool u64, const double* u_coef, double* sumsq,
                       float* tpart, double* tsum, const FusedPlan& plan, const int* skip,
                       hipStream_t s);

}  // namespace tomos_hip
