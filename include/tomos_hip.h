/*
 * tomos_hip.h -- C ABI of libtomos_hip.so, the MI355X (gfx950) kernels
 * behind the TOMOS LSQR hot path.
 *
 * Conventions (all entry points):
 *   - plain pointers and sizes only; every device buffer is owned by the
 *     caller; matrices are column-major with an explicit leading dimension;
 *   - the packed formats are PLANAR: a wh plane and a wr plane of uint16,
 *     each lda*n long (the reference stores wr right after wh in one
 *     allocation, src/core/matrix/fp32_16_pack.hpp:861-874; a caller with
 *     that layout passes wr = wh + lda*n);
 *   - functions return 0 on success or a positive TOMOS_ERR_* code, and the
 *     message is then available from tomos_last_error(); nothing throws;
 *   - work is enqueued on the given stream (NULL = the default stream) and
 *     is asynchronous unless a comment says otherwise.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the reference checkout).
 */
#ifndef TOMOS_HIP_H
#define TOMOS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* tomos_stream_t; /* == hipStream_t */
typedef struct ihipEvent_t* tomos_event_t;   /* == hipEvent_t */

/* Status codes. */
enum {
    TOMOS_OK = 0,
    TOMOS_ERR_INVALID_ARGUMENT = 1,
    TOMOS_ERR_HIP = 2,
    TOMOS_ERR_UNSUPPORTED = 3,
    TOMOS_ERR_WORKSPACE = 4,
    TOMOS_ERR_NOT_CONVERGED = 5,
    TOMOS_ERR_SINGULAR = 6,
    TOMOS_ERR_INTERNAL = 7
};

/* Element types; values equal the reference's NumericType
 * (include/core/precision.hpp:43-50) with FP32_BF16 appended. */
enum {
    TOMOS_F64 = 0,
    TOMOS_F32 = 1,
    TOMOS_F16 = 2,
    TOMOS_FP32_16 = 3,
    TOMOS_BF16 = 4,
    TOMOS_FP32_BF16 = 5
};

enum { TOMOS_NO_TRANSPOSE = 0, TOMOS_TRANSPOSE = 1 };

/* Encoders (FP32 -> planar wh/wr). */
enum {
    TOMOS_PACK_FP32_16 = 0,           /* fp32_16_split          fp32_16_pack.hpp:129-190 */
    TOMOS_PACK_FP32_16_CLAMPED = 1,   /* fp32_16_split_clamped  fp32_16_pack.hpp:200-248 */
    TOMOS_PACK_FP32_16_TRUNCATED = 2, /* fp32_16_split_truncated fp32_16_pack.hpp:263-299 */
    TOMOS_PACK_FP32_BF16 = 3,         /* fp32_bf16_split         fp32_bf16_pack.hpp:68-73 */
    TOMOS_PACK_FP32_BF16_ROUNDED = 4  /* fp32_bf16_split_rounded fp32_bf16_pack.hpp:117-145 */
};

/* Decoders (planar wh[/wr] -> FP32 or 16-bit bits). */
enum {
    TOMOS_DECODE_FP32_16_EXACT = 0,      /* fp32_16_combine            fp32_16_pack.hpp:301-359 */
    TOMOS_DECODE_FP32_16_HALF = 1,       /* fp32_16_half_only          fp32_16_pack.hpp:553-571 */
    TOMOS_DECODE_FP32_16_HALF_FAST = 2,  /* fp32_16_half_only_fast     fp32_16_pack.hpp:586-595 */
    TOMOS_DECODE_FP32_16_BF16 = 3,       /* fp32_16_bf16_only          fp32_16_pack.hpp:812-849 */
    TOMOS_DECODE_FP32_BF16_EXACT = 4,    /* fp32_bf16_combine          fp32_bf16_pack.hpp:78-81 */
    TOMOS_DECODE_FP32_BF16_HI = 5,       /* fp32_bf16_bf16_only        fp32_bf16_pack.hpp:91-93 */
    TOMOS_DECODE_FP32_BF16_EXACT_ROUNDED = 6, /* fp32_bf16_combine_rounded fp32_bf16_pack.hpp:180-186 */
    TOMOS_DECODE_FP32_16_EXACT_TRUNCATED = 7  /* fp32_16_combine_truncated fp32_16_pack.hpp:498-522 */
};

/* GEMV/GEMM compute tiers (reference dispatch::Fp32_16 slots,
 * src/dispatch/fp32_16_registries.hpp:40-54):
 *   DEFAULT: the per-iteration tier -- FP32_16: half_only decode of wh
 *            ("pure_half"), FP32_BF16: wh read as bf16 ("hi-only"); both
 *            accumulate in FP32.  F32: FP32 accumulate.  F64: FP64.
 *   FAST:    FP32_16 half_only decode / FP32_BF16 hi-only, FP64 accumulate
 *            ("fast", fp32_16_kernels.cpp:90-113).  F32: FP64 accumulate.
 *   EXACT:   wh+wr exact reconstruction, FP64 accumulate ("accurate",
 *            fp32_16_kernels.cpp:28-55).  F32/F64: FP64 accumulate. */
enum { TOMOS_TIER_DEFAULT = 0, TOMOS_TIER_FAST = 1, TOMOS_TIER_EXACT = 2 };
/* GEMM only: the reference's materialize slot (exact decode, then an FP32
 * GEMM: dispatch::Fp32_16::mm_workspace, fp32_16_registries.cpp:432-449),
 * computed as a 3-term split on 16-bit matrix cores. */
enum { TOMOS_TIER_MATERIALIZE = 3 };

/* ---- runtime helpers -------------------------------------------------- */
const char* tomos_last_error(void);
int tomos_version(void);
int tomos_set_device(int device);
int tomos_get_device_info(int* cu_count, int* arch_major, int* arch_minor,
                          size_t* total_mem);
int tomos_malloc(void** ptr, size_t bytes);
int tomos_free(void* ptr);
int tomos_memcpy_htod(void* dst, const void* src, size_t bytes, tomos_stream_t stream);
int tomos_memcpy_dtoh(void* dst, const void* src, size_t bytes, tomos_stream_t stream);
int tomos_memcpy_dtod(void* dst, const void* src, size_t bytes, tomos_stream_t stream);
int tomos_memset(void* ptr, int value, size_t bytes, tomos_stream_t stream);
int tomos_stream_synchronize(tomos_stream_t stream);
int tomos_device_synchronize(void);
int tomos_stream_create(tomos_stream_t* stream);
int tomos_stream_destroy(tomos_stream_t stream);
int tomos_event_create(tomos_event_t* ev);
int tomos_event_destroy(tomos_event_t ev);
int tomos_event_record(tomos_event_t ev, tomos_stream_t stream);
int tomos_event_elapsed_ms(float* ms, tomos_event_t start, tomos_event_t stop);

/* Cache policy of the matrix loads in the streaming kernels (GEMV, the
 * single-pass LSQR kernel, CountSketch, the batched residual):
 * 0 = auto (non-temporal loads when a pass reads more than 192 MiB of
 * matrix data, cacheable below, so matrices that fit the MALL stay there
 * between iterations), 1 = always non-temporal, 2 = always cacheable.
 * Process-wide; results do not depend on it.  No reference counterpart. */
int tomos_stream_loads_set(int mode);

/* Device-to-device copy through a plain vectorized kernel (16-byte loads
 * and stores); exists so profilers can calibrate HBM byte counters against
 * a kernel of known traffic.  No reference counterpart. */
int tomos_copy_kernel(void* dst, const void* src, size_t bytes, tomos_stream_t stream);
/* Read-only stream over bytes (rounded down to 16-byte vectors) of src with
 * non-temporal 16-byte loads: the achievable HBM read rate, next to the
 * published peak.  *flag (device) is left unchanged in practice.  No
 * reference counterpart. */
int tomos_read_kernel(const void* src, size_t bytes, unsigned* flag, tomos_stream_t stream);

/* Deterministic synthetic data on the device (no reference counterpart;
 * used by benchmarks to build the 1e6..2e6-row problems in HBM).
 * out[i + j*ld] for i<m, j<n: N(0,1) from a counter-based hash of
 * (seed, i + j*m) and Box-Muller, times col_scale[j] (nullable) and times
 * 2^k with k uniform in [exp_lo, exp_hi] drawn per element (exp_lo ==
 * exp_hi == 0 disables it). */
int tomos_fill_normal_f32(float* out, int64_t m, int64_t n, int64_t ld, uint64_t seed,
                          const float* col_scale, int exp_lo, int exp_hi,
                          tomos_stream_t stream);

/* dst (rows x cols, ldd) <- src (rows x cols, lds), column-major, types
 * TOMOS_F32 / TOMOS_F64 in either direction (FP64 -> FP32 rounds to
 * nearest).  The reference casts a host copy element by element where a
 * matrix changes precision (lsqr_precision_runner.cu:620-628); here the
 * cast runs on the device.  Asynchronous. */
int tomos_convert(int src_type, int dst_type, int64_t rows, int64_t cols, const void* src, int64_t lds,
                  void* dst, int64_t ldd, tomos_stream_t stream);

/* ---- pack / decode ---------------------------------------------------- */
/* Replaces cuda::split_dense_fp32_16 (fp32_16_cuda_kernels.cu:271-299) and
 * cpu::split_dense_fp32_16 (fp32_16_kernels.cpp:8-16), for every encoder
 * variant above.  src is m x n (ld_src); wh/wr are m x n (ld_dst). */
int tomos_pack(int variant, int64_t m, int64_t n, const float* src, int64_t ld_src,
               uint16_t* wh, uint16_t* wr, int64_t ld_dst, tomos_stream_t stream);

/* Replaces cuda::combine_dense_fp32_16 (fp32_16_cuda_kernels.cu:301-323)
 * for every FP32-producing decoder above.  wr may be NULL for the wh-only
 * decoders. */
int tomos_unpack(int decoder, int64_t m, int64_t n, const uint16_t* wh, const uint16_t* wr,
                 int64_t ld_src, float* dst, int64_t ld_dst, tomos_stream_t stream);

/* Element-wise wh -> 16-bit decode, returning raw half (HALF, HALF_FAST) or
 * bfloat16 (BF16, FP32_BF16_HI) bit patterns: the device form of
 * fp32_16_half_only / _fast (fp32_16_pack.hpp:553-595) and
 * fp32_16_bf16_only (:812-849), for bit-exact checks. */
int tomos_decode_wh(int decoder, int64_t count, const uint16_t* wh, uint16_t* out,
                    tomos_stream_t stream);

/* ---- GEMV ------------------------------------------------------------- */
/* y = alpha*op(A)*x + beta*y (y is not read when beta == 0).
 * Replaces the mv_dense_fn slots of dispatch::Dense / dispatch::Fp32_16
 * (src/dispatch/dense_registries.hpp:29-31, fp32_16_registries.hpp:27-29):
 * mv_pure_half_impl (fp32_16_registries.cpp:65-263), mv_fast_impl
 * (:265-285), mv_accurate_impl (:287-368), the FP32_BF16 kernels
 * (fp32_16_cuda_kernels.cu:28785, :29636, :7587) and mv_fp32_tiled
 * (mv_cuda_kernels.cu:863-990).
 *   a_type: TOMOS_F32 | TOMOS_F64 | TOMOS_FP32_16 | TOMOS_FP32_BF16 |
 *           TOMOS_F16 | TOMOS_BF16 (native 16-bit storage, the reference's
 *           mv_fp16_native_mixed, mv_cuda_kernels.cu:1898-1926; FAST and
 *           EXACT both accumulate the stored values in FP64);
 *   a: the F32, F64 or 16-bit matrix, or the wh plane for packed types; wr:
 *      the wr plane (used by the EXACT tier only; may be NULL otherwise);
 *   x_type, y_type: TOMOS_F32 | TOMOS_F64.
 * scratch: device buffer of at least tomos_gemv_scratch_size() bytes, or
 * NULL to use a library-owned buffer (grown with hipMalloc on demand). */
size_t tomos_gemv_scratch_size(int a_type, int tier, int trans, int64_t m, int64_t n);
int tomos_gemv(int a_type, int tier, int trans, double alpha, int64_t m, int64_t n,
               const void* a, const uint16_t* wr, int64_t lda, int x_type, const void* x,
               double beta, int y_type, void* y, void* scratch, size_t scratch_bytes,
               tomos_stream_t stream);

/* tomos_gemv with no scratch argument, allocating nothing.  It has no
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results
 * as tomos_gemv.  Its partial sums live in a fixed device array of
 * TOMOS_GEMV_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends;
 * a shape whose tomos_gemv_scratch_size() exceeds it is rejected with
 * TOMOS_ERR_WORKSPACE.  All calls on one GPU share that array: calls that
 * overlap on one GPU (different streams or host threads) race. */
#define TOMOS_GEMV_NOALLOC_SCRATCH_BYTES (32u << 20)
int tomos_gemv_noalloc(int a_type, int tier, int trans, double alpha, int64_t m, int64_t n,
                       const void* a, const uint16_t* wr, int64_t lda, int x_type, const void* x,
                       double beta, int y_type, void* y, tomos_stream_t stream);

/* ---- GEMM ------------------------------------------------------------- */
/* C = alpha*op(A)*op(B) + beta*C with packed or plain A, FP32 B and C.
 * Replaces the mm_dense_fn slots (dense_registries.hpp:32-34) and
 * dispatch::Fp32_16::mm_{pure_half,fast,accurate,workspace}
 * (fp32_16_registries.cpp:370-453).  M x N result, K inner dimension.
 * Packed A (FP32_16, FP32_BF16):
 *   DEFAULT      16-bit matrix cores (the reference's tensor-core GEMMs on
 *                wh, mm_cuda_kernels.cu:35372, :37826): A's hi word as the
 *                16-bit tiers decode it (FP32_BF16: bf16; FP32_16:
 *                fp32_16_half_only, fp16), B rounded to nearest-even
 *                bf16 / fp16, FP32 accumulate;
 *   FAST, EXACT  as for GEMV (FP64 accumulate; FAST: hi word, EXACT: wh+wr);
 *   MATERIALIZE  A_hi B_hi + A_hi B_lo + A_lo B_hi of the exact values on
 *                16-bit matrix cores, FP32 accumulate.
 * DEFAULT and MATERIALIZE stage both operands as K-contiguous 16-bit planes
 * in the library-owned scratch buffer (grown with hipMalloc on demand; not
 * thread-safe).  F32 / F64 A: DEFAULT accumulates in FP32, FAST/EXACT FP64. */
int tomos_gemm(int a_type, int tier, int trans_a, int trans_b, double alpha, int64_t M,
               int64_t N, int64_t K, const void* a, const uint16_t* wr, int64_t lda,
               const float* b, int64_t ldb, double beta, float* c, int64_t ldc,
               tomos_stream_t stream);

/* Split-precision packed GEMM on the matrix cores:
 *   C = alpha * A B + beta * C,  A (M x K, lda) and B (K x N, ldb) both
 *   packed as a_type (FP32_16 or FP32_BF16 planes, column-major), C FP32.
 * Each operand is decoded exactly and split into hi = rn16(a) and
 * lo = rn16(a - hi) 16-bit planes (bf16 for FP32_BF16; fp16 with a
 * power-of-two operand scale for FP32_16); the product is
 * A_hi B_hi + A_hi B_lo + A_lo B_hi with FP32 accumulation (bf16/fp16
 * MFMA 32x32x16).  No reference counterpart (the reference's 16-bit packed
 * GEMMs use the hi word only, mm_cuda_kernels.cu:35372, :37826).
 * scratch: tomos_gemm_split_scratch_size() bytes or NULL. */
size_t tomos_gemm_split_scratch_size(int64_t M, int64_t N, int64_t K);
int tomos_gemm_split(int a_type, int64_t M, int64_t N, int64_t K, double alpha,
                     const uint16_t* a_wh, const uint16_t* a_wr, int64_t lda,
                     const uint16_t* b_wh, const uint16_t* b_wr, int64_t ldb, double beta,
                     float* c, int64_t ldc, void* scratch, size_t scratch_bytes,
                     tomos_stream_t stream);

/* ---- LSQR's matvec pair from one read of A ---------------------------- */
/* u <- A p + u_coef * u  (all m rows), then t <- A^T u and *unorm2 <- |u|^2,
 * with A read from HBM once: the two mv_apply calls of one LSQR step
 * (src/core/solver/lsqr.cpp:560-577, A*(P v) - alpha*u followed by A^T u)
 * with the 1/beta scaling left to the caller (A^T (u/beta) = (A^T u)/beta).
 * A: F32, FP32_16 or FP32_BF16 (hi plane only: the per-iteration tier),
 * column-major, 16-byte aligned, lda a multiple of 4 (F32) / 8 (packed),
 * n <= 2048; otherwise TOMOS_ERR_UNSUPPORTED.  p, t: FP64 (n); u: FP32 (m);
 * unorm2: one FP64 in device memory, or NULL.  Deterministic (fixed
 * reduction order).  scratch: tomos_matvec_pair_scratch_size() bytes or NULL. */
size_t tomos_matvec_pair_scratch_size(int a_type, int64_t m, int64_t n);
int tomos_matvec_pair(int a_type, int64_t m, int64_t n, const void* a, int64_t lda,
                      const double* p, double u_coef, float* u, double* t, double* unorm2,
                      void* scratch, size_t scratch_bytes, tomos_stream_t stream);

/* Tuned launch shape of the single-pass kernel behind tomos_matvec_pair and
 * the fused LSQR iteration, per (a_type, n): row_groups (2, 4 or 8 lanes
 * share one column of a row tile; 0 = default), threads per block (256 or
 * 512; 0 = 256 with an explicit row_groups) and blocks_per_cu (grid size;
 * 0 = as many as are resident at once).  Process-wide; consulted at
 * every launch.  Scratch sizes do not depend on it.  Set by the autotuner
 * (tomos_amd/autotune.py); no reference counterpart.  Setting all three to
 * 0 removes the (a_type, n) entry; clear removes every entry. */
int tomos_fused_config_set(int a_type, int64_t n, int row_groups, int blocks_per_cu, int threads);
int tomos_fused_config_clear(void);
/* Installed shape for (a_type, n): all three outputs 0 when none is set. */
int tomos_fused_config_get(int a_type, int64_t n, int* row_groups, int* blocks_per_cu, int* threads);

/* Kernel timer (no reference counterpart: the reference times whole solves
 * with std::chrono, lsqr_precision_runner.cu:653-659).  Between begin and
 * end, every launch of the single-pass kernel (tomos_matvec_pair, the fused
 * LSQR iteration, the adjoint's fused products) made by the calling host
 * thread is bracketed by HIP events on its stream, up to `capacity`
 * launches (1..65536; later launches are not timed).  end waits for the
 * recorded launches, writes their device durations (ms, launch order) to
 * durations_ms[0 .. min(count, capacity)), the count to *launches, and
 * frees the events.  A launch enqueued after the solve's stop flag was set
 * exits at once and shows as a duration of a few microseconds. */
int tomos_kernel_timer_begin(int capacity);
int tomos_kernel_timer_end(float* durations_ms, int capacity, int* launches);

/* ---- CountSketch ------------------------------------------------------ */
/* bucket[i] = splitmix64(i ^ 0x12345678abcdef00) mod s,
 * sign[i]   = bit0(splitmix64(i ^ 0xf00dd00d42babe55)) ? +1 : -1,
 * bit-exact with cpu::countsketch_apply_impl's hashing
 * (src/kernels/cpu/sketch/countsketch_kernels.cpp:59-85). */
int tomos_countsketch_indices(int64_t m, int64_t s, int32_t* bucket, int8_t* sign,
                              tomos_stream_t stream);

/* ---- Sparse (CSR) ------------------------------------------------------ */
/* Index types of CSR row_ptrs / col_idxs (IntType, precision.hpp:42). */
enum { TOMOS_INDEX_INT32 = 0, TOMOS_INDEX_INT64 = 1 };

/* A CSR matrix: m x n, nnz entries; vals (val_type TOMOS_F64, TOMOS_F32,
 * TOMOS_F16 or TOMOS_BF16), row_ptrs (m+1) and col_idxs (nnz) of idx_type
 * (CsrStorage, src/core/matrix/sparse_storage.hpp:16-47).
 * tomos_csr_check validates row_ptrs[0] = 0, non-decreasing row_ptrs,
 * row_ptrs[m] = nnz and 0 <= col_idxs < n on the device (synchronous;
 * TOMOS_ERR_INVALID_ARGUMENT with the reason). */
int tomos_csr_check(int val_type, int idx_type, int64_t m, int64_t n, int64_t nnz, const void* row_ptrs,
                    const void* col_idxs, tomos_stream_t stream);

/* y = alpha*op(A)*x + beta*y, x/y TOMOS_F32 or TOMOS_F64, FP64
 * accumulation; beta == 0 does not read y.  Replaces the mv_csr slots
 * (sparse_registries.cpp:225-243; CPU semantics
 * matrix_cpu_kernels.cpp:92-131).  NO_TRANSPOSE is asynchronous and
 * memory-safe for any index contents (out-of-range entries contribute
 * nothing); TRANSPOSE validates A, builds A^T in scratch
 * (tomos_csr_mv_scratch_size bytes, or NULL for a library buffer) and
 * applies it: deterministic, no atomics. */
size_t tomos_csr_mv_scratch_size(int val_type, int idx_type, int trans, int64_t m, int64_t n, int64_t nnz);
int tomos_csr_mv(int val_type, int idx_type, int trans, double alpha, int64_t m, int64_t n, int64_t nnz,
                 const void* vals, const void* row_ptrs, const void* col_idxs, int x_type, const void* x,
                 double beta, int y_type, void* y, void* scratch, size_t scratch_bytes, tomos_stream_t stream);

/* Y (m x ncols, ldy) = alpha * A X + beta * Y, A CSR (m x n), X dense
 * (n x ncols, ldx), X / Y TOMOS_F32 or TOMOS_F64, each output one FP64 sum
 * over its row's entries in storage order; beta == 0 does not read Y.
 * Replaces cpu::mm_csr with both operands untransposed
 * (src/kernels/cpu/matrix/mm_cpu_kernels.cpp:85-108), the kernel behind
 * apply_count_sketch(Sparse&, Dense&, Dense&) (sketch_registries.cpp:27-41).
 * One launch, memory-safe for any index contents.  Asynchronous. */
int tomos_csr_mm(int val_type, int idx_type, double alpha, int64_t m, int64_t n, int64_t nnz, const void* vals,
                 const void* row_ptrs, const void* col_idxs, int64_t ncols, int x_type, const void* X,
                 int64_t ldx, double beta, int y_type, void* Y, int64_t ldy, tomos_stream_t stream);

/* CSR of A^T (n x m: vals_t nnz, row_ptrs_t n+1, col_idxs_t nnz, same
 * types) with each row of A^T in A's row order.  Validates A first.  No
 * reference counterpart (the reference's transposed product scatters,
 * matrix_cpu_kernels.cpp:116-130). */
size_t tomos_csr_transpose_scratch_size(int val_type, int idx_type, int64_t m, int64_t n, int64_t nnz);
int tomos_csr_transpose(int val_type, int idx_type, int64_t m, int64_t n, int64_t nnz, const void* vals,
                        const void* row_ptrs, const void* col_idxs, void* vals_t, void* row_ptrs_t,
                        void* col_idxs_t, void* scratch, size_t scratch_bytes, tomos_stream_t stream);

/* Dense Gaussian sketch: out (rows x cols, ld, column-major; type
 * TOMOS_F64, TOMOS_F32, TOMOS_F16 or TOMOS_BF16) <- i.i.d. N(0,1).
 * Replaces cuda::generate_gaussian_values
 * (src/kernels/cuda/sketch/gaussian_sketch_cuda_kernels.cu:36-45) and the
 * CPU generator (src/kernels/cpu/sketch/gaussian_sketch_kernels.cpp:10-18),
 * which seed themselves from the clock; here the seed is an argument and
 * the values depend only on (seed, rows, cols): entries (2q, j) and
 * (2q+1, j) are the Box-Muller pair of the p-th SplitMix64 output,
 * p = j * ceil(rows/2) + q.  Entries between rows and ld are not written.
 * Asynchronous on `stream`. */
int tomos_gaussian_fill(int type, int64_t rows, int64_t cols, void* out, int64_t ld, uint64_t seed,
                        tomos_stream_t stream);

/* The reference's CSR sketch form (rows x cols, one entry per row: row i
 * holds column h(i); generate_countsketch_csr_idxs,
 * countsketch_kernels.cpp:34-53, which draws h from rand() seeded by the
 * clock): row_ptrs[i] = i, col_idxs[i] = z_i % cols with z_i the i-th
 * SplitMix64 output from seed ^ 0xc2b2ae3d27d4eb4f; values +-1 from z_i's
 * top bit (kind 0, the count sketch) or N(0,1) as tomos_gaussian_fill of a
 * rows x 1 column with `seed` (kind 1, create_gaussian_sketch(Sparse&),
 * gaussian_sketch_registries.cpp:33-57).  nnz = rows.  Kind 2 writes the
 * pattern only (row_ptrs, col_idxs; vals untouched and may be NULL) and
 * kind 3 the +-1 values only (row_ptrs and col_idxs untouched, may be NULL):
 * the separate generate_count_sketch_idxs / _values registry slots
 * (sketch_registries.hpp:14-15, called in that order by
 * generate_count_sketch, countsketch.cpp:15-18).  Asynchronous. */
int tomos_csr_sketch(int kind, int val_type, int idx_type, int64_t rows, int64_t cols, void* vals,
                     void* row_ptrs, void* col_idxs, uint64_t seed, tomos_stream_t stream);

/* Y (s x n, FP64, ldy) = S * A, S the hashed s x m CountSketch above,
 * A m x n in any a_type/tier (packed A is decoded per tier).  Y is
 * overwritten.  Replaces countsketch_apply_matrix_hashed
 * (countsketch_cuda_kernels.cu:253-285, host :435-450) and the CPU apply.
 * scratch: tomos_countsketch_scratch_size() bytes or NULL. */
size_t tomos_countsketch_scratch_size(int64_t m, int64_t n, int64_t s);
int tomos_countsketch_apply(int a_type, int tier, int64_t m, int64_t n, const void* a,
                            const uint16_t* wr, int64_t lda, int64_t s, double* y,
                            int64_t ldy, void* scratch, size_t scratch_bytes,
                            tomos_stream_t stream);

/* One CountSketch of a batch: Y (s x n, ldy) = S A, A m x n (lda; wr the
 * lo plane of a packed A at the EXACT tier, else unused). */
typedef struct {
    int64_t m, n, lda, s, ldy;
    const void* a;
    const uint16_t* wr;
    double* y;
} tomos_cs_problem;

/* tomos_countsketch_apply over nprob problems of one a_type and tier, with
 * one host synchronisation for the whole batch instead of one per problem
 * (the fixed-point path's exponent checks are read together at the end;
 * the problems they reject are redone one by one).  Results equal the
 * one-problem call's bit for bit.  No counterpart in the reference, whose
 * batched preconditioner calls countsketch_apply_matrix_hashed once per
 * problem; it is for batched-solve callers.  scratch:
 * tomos_countsketch_apply_batched_scratch_size() bytes or NULL. */
size_t tomos_countsketch_apply_batched_scratch_size(int nprob, const tomos_cs_problem* probs);
int tomos_countsketch_apply_batched(int a_type, int tier, int nprob, const tomos_cs_problem* probs,
                                    void* scratch, size_t scratch_bytes, tomos_stream_t stream);

/* ---- Householder QR / triangular inverse (FP64) ---------------------- */
/* R (n x n upper, ldr) of the Householder QR of the k x n FP64 matrix A
 * (lda; A is overwritten with intermediate data).  k >= n required.
 * Replaces the geqrf step of generate_sketch_qr_precond
 * (src/core/precond/sketch_qr.cpp:38-57, CUDA sketch_qr_cuda_kernels.cu:65-151).
 * Synchronous. */
size_t tomos_geqrf_scratch_size(int64_t k, int64_t n);
int tomos_geqrf_r_f64(int64_t k, int64_t n, double* a, int64_t lda, double* r, int64_t ldr,
                      void* scratch, size_t scratch_bytes, tomos_stream_t stream);

/* In-place inverse of the n x n upper-triangular R (ldr); strictly lower
 * part set to 0.  Returns TOMOS_ERR_SINGULAR for a zero diagonal.
 * Replaces dtrtri('U','N') (sketch_qr.cpp:59-63).  Synchronous. */
int tomos_trtri_upper_f64(int64_t n, double* r, int64_t ldr, tomos_stream_t stream);

/* Panel factorization of the QR entry points above: 0 (default) = each
 * 32-column panel by CholeskyQR2 with its Householder vectors
 * reconstructed (no column-by-column pass over the rows), and the whole
 * factorization redone with column-by-column Householder panels for a
 * sketch where a panel is too ill-conditioned for it; 1 = the
 * column-by-column Householder panels only.  Process-wide; returns the
 * status.  No reference counterpart (the reference calls LAPACK dgeqrf). */
int tomos_qr_panels_set(int mode);

/* generate_sketch_qr_precond (src/core/precond/sketch_qr.cpp:68-122): the
 * k x n FP64 sketch (lda) is factored and its leading n x n block is
 * overwritten with R^-1 (upper), the rest of the block zeroed.
 * Synchronous. */
int tomos_sketch_qr_precond_f64(int64_t k, int64_t n, double* a, int64_t lda,
                                tomos_stream_t stream);

/* Many sketches at once (uneven k, n): the same factorization with every
 * launch covering all problems (the launches of the largest one), R^-1 of
 * each written over its leading n x n block.  scratch:
 * tomos_sketch_qr_precond_batched_scratch_size() bytes, or NULL (then the
 * call allocates and frees it).  Synchronous.  The batched form of
 * generate_sketch_qr_precond (src/core/precond/sketch_qr.cpp:68-122). */
typedef struct {
    int64_t k, n, lda;
    double* a;
} tomos_qr_problem;
size_t tomos_sketch_qr_precond_batched_scratch_size(int nprob, const tomos_qr_problem* probs);
int tomos_sketch_qr_precond_batched_f64(int nprob, const tomos_qr_problem* probs, void* scratch,
                                        size_t scratch_bytes, tomos_stream_t stream);

/* ---- LSQR ------------------------------------------------------------- */
/* Restarted, right-preconditioned Golub-Kahan LSQR on one GPU, with the
 * scalar recurrences (alpha, beta, rho, phi, c, s, theta) and all norms in
 * FP64 in device memory.  Semantics follow fused_lsqr_impl
 * (src/core/solver/lsqr.cpp:685-747): restart loop, residual recompute
 * u = b - A*x at each restart on `restart_tier`, inner iterations on
 * `mv_tier`, stop when |phi_bar|/|b| < tol or < max_relres, or when the
 * true restart residual is below them.  Replaces lsqr_dense_impl
 * (src/dispatch/lsqr_registries.cpp:88-159). */
typedef struct tomos_lsqr_options {
    int a_type;        /* TOMOS_F32 | TOMOS_F64 | TOMOS_FP32_16 | TOMOS_FP32_BF16 | TOMOS_F16 | TOMOS_BF16 */
    int mv_tier;       /* per-iteration GEMV tier (TOMOS_TIER_*) */
    int restart_tier;  /* restart residual tier (EXACT = "implicit duplicate") */
    int value_type;    /* Krylov vectors u, v, w: TOMOS_F32 | TOMOS_F64 */
    int refine_type;   /* b and x: TOMOS_F32 | TOMOS_F64 */
    int precond_type;  /* R^-1 element type (TOMOS_F32 | TOMOS_F64), or -1: none */
    double tol;
    double max_relres;
    int max_iters;     /* iterations per restart */
    int restarts;
    int check_every;   /* host polls the device stop flag every N iterations (>=1) */
    int fused;         /* 1: one pass over A per iteration (A*p and A^T*u fused) */
    int confirm;       /* 1: a stop on the |phi_bar| estimate only ends the current restart; the
                          solve stops when the next restart's true residual |b - A x|/|b|
                          (restart tier) is below tol or max_relres.  0: the reference's rule.
                          Zero-initialise the struct so older callers get 0. */
} tomos_lsqr_options;

/* Arena: the Krylov vectors u (m), v, w, temp1 (n, value_type) and
 * precond_v (n, precond_type), carved in the reference's order and sized by
 * the reference's formula (create_lsqr_workspace, lsqr.cpp:389-406).
 * Scratch: device state, residual history and GEMV partial sums (the
 * reference's Workspace::mv_transpose_scratch_, workspace.hpp:170-182). */
size_t tomos_lsqr_arena_size(const tomos_lsqr_options* opt, int64_t m, int64_t n);
size_t tomos_lsqr_scratch_size(const tomos_lsqr_options* opt, int64_t m, int64_t n);

/* x (refine_type) is overwritten.  rinv: n x n upper (ldr) or NULL.
 * arena must hold tomos_lsqr_arena_size() bytes (else
 * TOMOS_ERR_WORKSPACE, as lsqr_registries.cpp:59-63); scratch holds
 * tomos_lsqr_scratch_size() bytes or is NULL (library-owned buffer).
 * history (host, nullable) receives up to history_cap relative residuals,
 * one per stopping check, as lsqr_solve's residual_history.
 * Synchronous: returns after the solve completed. */
int tomos_lsqr_solve(const tomos_lsqr_options* opt, int64_t m, int64_t n, const void* a,
                     const uint16_t* wr, int64_t lda, const void* b, void* x,
                     const void* rinv, int64_t ldr, void* arena, size_t arena_bytes,
                     void* scratch, size_t scratch_bytes, int* iterations, double* history,
                     int history_cap, int* history_len, tomos_stream_t stream);

/* tomos_lsqr_solve with an explicit refine-precision duplicate of A: the
 * restart residuals u = b - A x read a_ref (ref_type TOMOS_F32 or
 * TOMOS_F64, lda_ref) with FP64 accumulation instead of A, while the inner
 * iterations use A at mv_tier (lsqr_solve's mtx_refine_precision,
 * lsqr.cpp:463-469, lsqr_registries.cpp:88-159).  Same arena and scratch
 * sizes as tomos_lsqr_solve. */
int tomos_lsqr_solve_refined(const tomos_lsqr_options* opt, int64_t m, int64_t n, const void* a,
                             const uint16_t* wr, int64_t lda, int ref_type, const void* a_ref,
                             int64_t lda_ref, const void* b, void* x, const void* rinv, int64_t ldr,
                             void* arena, size_t arena_bytes, void* scratch, size_t scratch_bytes,
                             int* iterations, double* history, int history_cap, int* history_len,
                             tomos_stream_t stream);

/* tomos_lsqr_solve over a CSR matrix (lsqr_solve(Context, Sparse&, ...),
 * include/core/solver.hpp:74-78; lsqr_sparse_impl,
 * lsqr_registries.cpp:164-196).  opt->a_type is the CSR value type
 * (TOMOS_F64, TOMOS_F32, TOMOS_F16 or TOMOS_BF16); every product
 * accumulates in FP64, so mv_tier/restart_tier and fused are ignored.  A is
 * validated (tomos_csr_check) and A^T is built in the scratch once per
 * solve, so A^T u needs no atomics and iterates repeat bit for bit.
 * ref_vals (ref_type TOMOS_F32/F64, or NULL): the refine-precision
 * duplicate's values over A's pattern, read only by the restart residuals
 * (mtx_refine_precision, lsqr_registries.cpp:180-183).  The arena is
 * tomos_lsqr_arena_size(opt, m, n) bytes, the scratch
 * tomos_lsqr_csr_scratch_size() bytes or NULL.  Synchronous. */
size_t tomos_lsqr_csr_scratch_size(const tomos_lsqr_options* opt, int idx_type, int64_t m, int64_t n,
                                   int64_t nnz);
int tomos_lsqr_solve_csr(const tomos_lsqr_options* opt, int idx_type, int64_t m, int64_t n, int64_t nnz,
                         const void* vals, const void* row_ptrs, const void* col_idxs, int ref_type,
                         const void* ref_vals, const void* b, void* x, const void* rinv, int64_t ldr,
                         void* arena, size_t arena_bytes, void* scratch, size_t scratch_bytes,
                         int* iterations, double* history, int history_cap, int* history_len,
                         tomos_stream_t stream);

/* ---- Many uneven LSQR problems in one set of launches ----------------- */
/* One problem: A (m x n, lda; for packed storage the hi plane, `wr` the lo
 * plane), b and x FP64, R^-1 (FP64, n x n, ldr) of its sketch-QR
 * preconditioner. */
typedef struct {
    int64_t m, n, lda;
    const void* a;
    const uint16_t* wr;
    const double* b;
    double* x;
    const double* rinv;
    int64_t ldr;
} tomos_lsq_problem;

/* Restarted, R-preconditioned LSQR on every problem of the batch, each with
 * the single-problem solver's control flow, scalars and stop test
 * (fused_lsqr_impl, src/core/solver/lsqr.cpp:685-747, per problem), but one
 * step of all problems is the same few launches whatever their number and
 * shapes: a grouped exact-residual kernel (restarts), the grouped
 * single-pass A p / A^T u kernel and a tail (per-problem scalars, R^-1
 * products spread over many blocks).  With use_graph the whole solve
 * (opt->restarts restarts of 1 + opt->max_iters steps; problems that have
 * stopped exit at once) is captured once per launch configuration and
 * submitted as one graph launch; a NULL stream then runs on a private
 * stream ordered after earlier NULL-stream work.
 * Options: a_type F32/FP32_16/FP32_BF16, mv_tier DEFAULT, value_type F32,
 * refine_type F64, precond_type F64 (R^-1 required), n <= 2048, 16-byte
 * aligned planes with lda a multiple of 4 (F32) / 8 (packed); columns that
 * start on 128-byte lines (lda a multiple of 32 / 64) read at full speed,
 * a 64-byte column offset costs ~1.4x the HBM fetch.
 * row_groups / threads / tiles_per_block: launch shape of the grouped
 * single-pass kernel (0 = the tuned shape for the largest n, else default).  iterations and
 * final_relres (|phi_bar|/|b|, the solver's estimate): one per problem, or
 * NULL.  scratch: tomos_lsqr_batched_scratch_size() bytes or NULL.
 * No reference counterpart (the reference solves one problem per call). */
size_t tomos_lsqr_batched_scratch_size(const tomos_lsqr_options* opt, int nprob,
                                       const tomos_lsq_problem* probs, int row_groups, int threads,
                                       int tiles_per_block);
int tomos_lsqr_batched_solve(const tomos_lsqr_options* opt, int nprob, const tomos_lsq_problem* probs,
                             int row_groups, int threads, int tiles_per_block, int use_graph, void* scratch,
                             size_t scratch_bytes, int* iterations, double* final_relres,
                             tomos_stream_t stream);

/* ---- Adjoint of the least-squares solve (implicit differentiation) ----- */
/* x = argmin |A x - b| with upstream gradient g = dL/dx:
 *   w = (A^T A)^-1 g,  grad_b = A w,  grad_A = r w^T - (A w) x^T,  r = b - A x.
 * No reference counterpart (the reference has no adjoint).
 *
 * tomos_lsq_adjoint_solve: w by restarted, R-preconditioned CG on the normal
 * equations with the LSQR solver's kernels and precision split: each
 * restart forms res = g - A^T A w at opt->restart_tier (FP64 accumulation)
 * and stops when |res|/|g| < opt->tol; inner iterations (opt->max_iters per
 * restart, opt->restarts restarts) form A d and A^T (A d) at opt->mv_tier
 * from one read of A (tomos_matvec_pair's kernel, when opt->fused and the
 * shape allow).  rinv: R^-1 (n x n, ldr, opt->precond_type F64/F32) or NULL
 * with precond_type -1.  g, w: FP64 (n).  value_type/refine_type are not
 * used.  history: relative residual per step (restart checks included).
 * scratch: tomos_lsq_adjoint_scratch_size() bytes or NULL. */
size_t tomos_lsq_adjoint_scratch_size(const tomos_lsqr_options* opt, int64_t m, int64_t n);
int tomos_lsq_adjoint_solve(const tomos_lsqr_options* opt, int64_t m, int64_t n, const void* a,
                            const uint16_t* wr, int64_t lda, const double* g, const void* rinv,
                            int64_t ldr, double* w, void* scratch, size_t scratch_bytes,
                            int* iterations, double* history, int history_cap, int* history_len,
                            tomos_stream_t stream);

/* grad_b = A w (FP64, m) at `tier`; when grad_a is non-NULL also
 * r = b - A x (same tier) and grad_a = r w^T - grad_b x^T (FP32, m x n,
 * column-major, ldg >= m) from one rank-2 kernel that reads only r, grad_b,
 * w and x.  b: FP64 (m); x, w: FP64 (n).
 * scratch: tomos_lsq_grad_scratch_size() bytes or NULL. */
size_t tomos_lsq_grad_scratch_size(int a_type, int tier, int64_t m, int64_t n);
int tomos_lsq_grad(int a_type, int tier, int64_t m, int64_t n, const void* a, const uint16_t* wr,
                   int64_t lda, const double* b, const double* x, const double* w, double* grad_b,
                   float* grad_a, int64_t ldg, void* scratch, size_t scratch_bytes,
                   tomos_stream_t stream);

/* Adjoint and gradients of many problems (tomos_lsq_problem: A, wr, lda, b,
 * x, R^-1 per problem, as for tomos_lsqr_batched_solve): for each problem p,
 * w[p] = (A^T A)^-1 g[p] by tomos_lsq_adjoint_solve's control flow and
 * kernels, then grad_b[p] = A w[p] and, when grad_a and grad_a[p] are
 * non-NULL, grad_a[p] (FP32, ldg[p] >= m) as tomos_lsq_grad at
 * opt->restart_tier.
 *   workers == 0: all problems on the device in the same launches: one
 *     captured hipGraph of 16 CG iterations of every problem (the batched
 *     LSQR solver's grouped single-pass kernel for A d and A^T A d, one
 *     block per problem for the n-sized work) replayed until every problem
 *     has stopped, with a captured restart sequence between replays when
 *     some problem restarts (the exact A^T A w of every restarting problem
 *     in one grouped launch over row items when the restart tier decodes
 *     exactly and every n <= 1024, otherwise per problem the restart-tier
 *     GEMV pair).  The gradients follow: for packed problems with every
 *     n <= 512, one launch over the same row items (A w and A x from one
 *     read of A, then grad_b and grad_A); otherwise tomos_lsq_grad per
 *     problem.  Same control flow and stop tests per problem as
 *     tomos_lsq_adjoint_solve; results agree with it to rounding.  Needs
 *     a_type F32/FP32_16/FP32_BF16, mv_tier DEFAULT, precond_type F64 (R^-1),
 *     n <= 2048, 16-byte aligned A with lda a multiple of 4 (F32) / 8.
 *   workers 1..16: the problems (largest first) are taken by that many host
 *     threads, each issuing the one-problem calls on its own HIP stream
 *     (default priority) with its own slice of the scratch; every result is
 *     bit-identical to the one-problem calls.
 * scratch (required): tomos_lsq_adjoint_batched_scratch_size() bytes for the
 * same workers.  iterations: one per problem, or NULL.  Synchronous: waits
 * for `stream` (workers >= 1) or runs on it (workers == 0) and returns when
 * every result is written.  No reference counterpart. */
size_t tomos_lsq_adjoint_batched_scratch_size(const tomos_lsqr_options* opt, int nprob,
                                              const tomos_lsq_problem* probs, int workers);
int tomos_lsq_adjoint_batched(const tomos_lsqr_options* opt, int nprob, const tomos_lsq_problem* probs,
                              const double* const* g, double* const* w, double* const* grad_b,
                              float* const* grad_a, const int64_t* ldg, int workers, void* scratch,
                              size_t scratch_bytes, int* iterations, tomos_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* TOMOS_HIP_H */
