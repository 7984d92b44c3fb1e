// Sparse (CSR) matrix products shared by the C ABI and LSQR.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "gemv.hpp"

namespace tomos_hip {

// A CSR matrix on the device: m rows, n columns, nnz stored entries.
// val_type: TOMOS_F64, TOMOS_F32, TOMOS_F16 or TOMOS_BF16; idx_type:
// TOMOS_INDEX_INT32 or TOMOS_INDEX_INT64 (row_ptrs and col_idxs alike).
struct CsrMat {
    int val_type = -1, idx_type = -1;
    int64_t m = 0, n = 0, nnz = 0;
    const void* vals = nullptr;
    const void* row_ptrs = nullptr;
    const void* col_idxs = nullptr;
};

void csr_require_valid_types(const CsrMat& A, const char* who);

// y = alpha*A*x + beta*y over A's rows, FP64 accumulation, Coef contract as
// gemv (device alpha/beta, skip flag, per-block sums of y^2 into
// coef.sumsq, their count into *sumsq_blocks).  Memory-safe for any index
// contents: row ranges are clamped to [0, nnz] and out-of-range column
// indices contribute nothing (csr_check reports both).
void csr_mv(const CsrMat& A, const Coef& c, int x_type, const void* x, int y_type, void* y,
            hipStream_t s, int* sumsq_blocks = nullptr);
int64_t csr_sumsq_blocks_max(int64_t m);

// Device-side validation: row_ptrs[0] = 0, non-decreasing, row_ptrs[m] =
// nnz, 0 <= col_idxs < n.  Synchronous; throws Error(kErrInvalidArgument).
void csr_check(const CsrMat& A, void* scratch, size_t scratch_bytes, hipStream_t s);
size_t csr_check_scratch_bytes();

// CSR of A^T (n x m) into (vals_t, ptr_t, idx_t): within each row of A^T
// the entries keep A's row order, so products with it are deterministic.
// A must be valid (csr_check).  scratch: csr_transpose_scratch_bytes().
size_t csr_transpose_scratch_bytes(const CsrMat& A);
void csr_transpose(const CsrMat& A, void* vals_t, void* ptr_t, void* idx_t, void* scratch,
                   size_t scratch_bytes, hipStream_t s);

size_t csr_value_bytes(int val_type);
size_t csr_index_bytes(int idx_type);

}  // namespace tomos_hip
