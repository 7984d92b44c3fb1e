// Single-pass LSQR matvec pair (A p and A^T u from one read of A).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace tomos_hip {

struct FusedShape {
    int g, threads;
};

struct FusedPlan {
    int g = 8;       // row groups per tile (lanes sharing one column)
    int threads = 256;
    int rounds = 0;  // column rounds per lane
    int nblocks = 0;
    int64_t tiles_per_block = 0;
    int splits = 1;
    size_t tpart_bytes = 0;  // [nblocks][n] float block partials of A^T u_raw
    size_t tsum_bytes = 0;   // [splits][n] double level-1 sums
    bool stream = false;     // non-temporal loads of A (stream_loads of its bytes)
};

// A in a_type with the per-iteration (DEFAULT) tier, column-major,
// 16-byte aligned, lda a multiple of the 16-byte vector, n <= 2048.
bool lsqr_fused_supported(int a_type, int tier, int64_t m, int64_t n, int64_t lda, const void* a);
// g: row groups per tile (2, 4 or 8), threads per block (256 or 512), 0 =
// the tuned or default shape for n; blocks_per_cu: grid size in blocks per
// CU (0 = as many as are resident).  A shape without a kernel for n falls
// back to the default.
FusedPlan lsqr_fused_plan(int a_type, int64_t m, int64_t n, int g = 0, int blocks_per_cu = 0,
                          int threads = 0);
FusedShape lsqr_fused_default_shape(int64_t n);
// upper bounds over every plan for (m, n): blocks and FusedPlan::tpart_bytes
int64_t lsqr_fused_max_blocks(int64_t m);
size_t lsqr_fused_max_tpart_bytes(int64_t m, int64_t n);
size_t lsqr_fused_max_tsum_bytes(int64_t n);
// Process-wide tuned shapes, looked up by lsqr_fused_plan when it is called
// with g = blocks_per_cu = 0.  (0, 0) removes the entry.
void lsqr_fused_config_set(int a_type, int64_t n, int g, int blocks_per_cu, int threads);
void lsqr_fused_config_clear();
bool lsqr_fused_config_get(int a_type, int64_t n, int* g, int* blocks_per_cu, int* threads);

// u_raw <- A p + (*u_coef) u_raw for all m rows; sumsq[b] = sum of u_raw^2
// over block b's rows (plan.nblocks entries); tsum[s][j] = sum over the
// blocks of split s of (A^T u_raw)_j.  Every launch is a no-op when *skip.
void lsqr_fused_launch(int a_type, int64_t m, int64_t n, const void* a, int64_t lda, const void* p,
                       bool p64, void* u, bool u64, const double* u_coef, double* sumsq,
                       float* tpart, double* tsum, const FusedPlan& plan, const int* skip,
                       hipStream_t s);

// ---- restart in one pass ------------------------------------------------
// u_raw = u - A x at the exact tier (u holds b on entry, x FP64, FP64 row
// sums), sumsq[b] per block and tsum = level-1 sums of A_hi^T u_raw, from
// one read of both planes.  xzero: x is known to be zero (the first restart
// of a solve), so u_raw = u and only the hi plane is read.  Supported for
// F32 and FP32_BF16 at the EXACT tier, n <= 1024, aligned as the
// single-pass kernel.
bool lsqr_restart_supported(int a_type, int tier, int64_t m, int64_t n, int64_t lda, const void* a,
                            const uint16_t* wr);
FusedPlan lsqr_restart_plan(int a_type, int64_t m, int64_t n);
void lsqr_restart_launch(int a_type, int64_t m, int64_t n, const void* a, const uint16_t* wr, int64_t lda,
                         const double* x, bool xzero, void* u, bool u64, double* sumsq, float* tpart,
                         double* tsum, const FusedPlan& plan, const int* skip, hipStream_t s);

// ---- many problems in one launch ----------------------------------------
// Device table entry of one problem: A (hi plane / F32), p (FP64, n),
// u_raw (FP32, m) updated in place with *u_coef, per-item partial slots
// (sumsq[slot], tpart[slot][n]) and a skip flag.
struct GroupProb {
    int64_t m, n, lda;
    const void* a;
    const double* p;
    float* u;
    const double* u_coef;
    const int* skip;
    double* sumsq;
    float* tpart;
    int rounds;  // column rounds covering n for the launch's (g, threads)
};

// One block's work: tiles [tile0, tile1) of problem `prob`, partial slot `slot`.
struct GroupItem {
    int prob, slot;
    int64_t tile0, tile1;
};

// Rows per tile for (a_type, g): g * (8 packed / 4 FP32 rows per 16-byte vector).
int64_t fused_tile_rows(int a_type, int g);
// Column rounds of the kernel for (g, threads, n); 0 when no kernel covers n.
int fused_rounds(int g, int threads, int64_t n);
// grid = nitems blocks; rounds must cover the largest n of the problems.
// ntl: non-temporal loads (a batch far larger than the MALL: stream_loads()
// of the batch's bytes per pass)
void grouped_fused_launch(int a_type, int g, int threads, int rounds, const GroupProb* probs,
                          const GroupItem* items, int nitems, hipStream_t s, bool ntl = false);
// Blocks of the grouped kernel resident per CU (occupancy), >= 1.
int grouped_fused_occupancy(int a_type, int g, int threads, int rounds);

}  // namespace tomos_hip
