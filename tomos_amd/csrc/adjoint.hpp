// Internal interface between the adjoint's single-problem code and its
// grouped (many problems per launch) form.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "tomos_hip.h"

namespace tomos_hip {

// Relative drop of the inner residual that ends a restart's inner CG
// iterations early: when the iterations multiply by a 16-bit hi word of A
// (per-iteration tier below the restart tier on a packed format), the
// operator they use differs from A^T A by about the hi word's relative
// rounding, so one restart cannot reduce the true residual by much more
// than that; iterating further until tol only spends matrix passes that
// the next restart's exact residual discards.  The drop is set just below
// that rounding (bf16 hi word 2^-8, fp16 hi word 2^-11); 0 (no early end)
// when the iterations are exact.
inline double adjoint_inner_drop(const tomos_lsqr_options& o) {
    if (o.mv_tier == o.restart_tier || o.mv_tier == TOMOS_TIER_EXACT) return 0.0;
    if (o.a_type == TOMOS_FP32_BF16) return 1.0 / 256.0;
    if (o.a_type == TOMOS_FP32_16) return 1.0 / 2048.0;
    return 0.0;
}

// grad_b = A w and grad_A = r w^T - (A w) x^T (r = b - A x) at `tier`;
// scratch: grad_scratch_bytes() bytes.
size_t grad_scratch_bytes(int a_type, int tier, int64_t m, int64_t n);
void lsq_grad(int a_type, int tier, int64_t m, int64_t n, const void* a, const uint16_t* wr, int64_t lda,
              const double* b, const double* x, const double* w, double* grad_b, float* grad_a, int64_t ldg,
              void* scratch, size_t scratch_bytes, hipStream_t s);

// The whole batch on the device: every problem's restarted, R-preconditioned
// CG in grouped launches (adjoint_batched.hip), then the gradients.
size_t adjoint_grouped_scratch_bytes(const tomos_lsqr_options& o, int nprob, const tomos_lsq_problem* probs);
void adjoint_grouped(const tomos_lsqr_options& o, int nprob, const tomos_lsq_problem* probs, const double* const* g,
                     double* const* w, double* const* grad_b, float* const* grad_a, const int64_t* ldg,
                     void* scratch, size_t scratch_bytes, int* iterations, hipStream_t s);

}  // namespace tomos_hip
