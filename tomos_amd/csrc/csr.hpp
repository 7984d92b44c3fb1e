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
// scratch: csr_mv_scratch_bytes(A) bytes (row-segment partial sums of a
// matrix with few, long rows; 0 otherwise).
void csr_mv(const CsrMat& A, const Coef& c, int x_type, const void* x, int y_type, void* y, void* scratch,
            size_t scratch_bytes, hipStream_t s, int* sumsq_blocks = nullptr);
size_t csr_mv_scratch_bytes(const CsrMat& A);
int64_t csr_sumsq_blocks_max(int64_t m);

// Device-side validation: row_ptrs[0] = 0, non-decreasing, row_ptrs[m] =
// nnz, 0 <= col_idxs < n.  Synchronous; throws Error(kErrInvalidArgument).
void csr_check(const CsrMat& A, void* scratch, size_t scratch_bytes, hipStream_t s);
size_t csr_check_scratch_bytes();

// Panel-blocked transpose of A: A's rows are cut into panels of
// panel_rows rows and row p*n + c of the result (a CSR matrix of
// panels*n rows over A's m columns) holds column c of panel p, in A's row
// order; panel_rows >= m gives plain A^T (n x m).  Products with it are
// deterministic.  A must be valid (csr_check).  csr_panel_rows(A, x_type)
// picks panels whose slice of the gathered vector (of x_type) fits the
// 64 KB of LDS a product block stages it in.
int64_t csr_panel_rows(const CsrMat& A, int x_type);
int64_t csr_panels(const CsrMat& A, int64_t panel_rows);
size_t csr_transpose_scratch_bytes(const CsrMat& A, int64_t panel_rows);
void csr_transpose(const CsrMat& A, int64_t panel_rows, void* vals_t, void* ptr_t, void* idx_t, void* scratch,
                   size_t scratch_bytes, hipStream_t s);

// y (nout) = alpha * A^T x + beta * y from the blocked transpose Tb
// (panels * nout rows): per-panel products, then each output's panel
// partials summed in panel order.  Coef contract as csr_mv.
size_t csr_tmv_scratch_bytes(const CsrMat& Tb, int64_t panels, int64_t nout);
void csr_tmv(const CsrMat& Tb, int64_t panels, int64_t panel_rows, int64_t nout, const Coef& c, int x_type,
             const void* x, int y_type, void* y, void* scratch, size_t scratch_bytes, hipStream_t s,
             int* sumsq_blocks = nullptr);

// Y (m x ncols, ldy) = alpha * A X + beta * Y for dense X (n x ncols, ldx);
// X, Y F32 or F64, FP64 sums in storage order; beta == 0 does not read Y.
// Memory-safe for any index contents (as csr_mv).
void csr_mm(const CsrMat& A, int64_t ncols, int x_type, const void* X, int64_t ldx, int y_type, void* Y, int64_t ldy,
            double alpha, double beta, hipStream_t s);

size_t csr_value_bytes(int val_type);
size_t csr_index_bytes(int idx_type);

}  // namespace tomos_hip
