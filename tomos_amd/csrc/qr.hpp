// Batched FP64 Householder QR / triangular inverse (internal interface).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace tomos_hip {

// One k x n problem (k >= n): A (lda) is overwritten, R goes to r (ldr).
struct QrProblem {
    int64_t k, n, lda, ldr;
    double* a;
    double* r;
};

size_t geqrf_scratch_bytes(int64_t k, int64_t n);
size_t geqrf_batched_scratch_bytes(int nprob, const QrProblem* probs);
// R of every problem (R^-1 when invert_r; copy_back: also written over the
// leading n x n block of A); synchronous.
void geqrf_r_batched(int nprob, const QrProblem* probs, bool invert_r, void* scratch,
                     size_t scratch_bytes, hipStream_t s, bool copy_back = false);
// R^-1 of each sketch written over its leading n x n block (ldr/r are
// filled in); scratch: sketch_precond_batched_bytes().
size_t sketch_precond_batched_bytes(int nprob, const QrProblem* probs);
void sketch_precond_batched(int nprob, QrProblem* probs, void* scratch, size_t scratch_bytes,
                            hipStream_t s);

}  // namespace tomos_hip
