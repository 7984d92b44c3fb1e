// C++ API mirror of TOMOS's BLAS front-ends, sketch + preconditioner,
// LSQR solver and dispatch registries for the LSQR hot path
// (include/core/blas.hpp, solver.hpp, countsketch.hpp, src/core/precond/
// sketch_qr.hpp, src/dispatch/*_registries.hpp).
//
// Registry contract kept alike (src/dispatch/dense_registries.hpp:29-34,
// fp32_16_registries.hpp:22-70, lsqr_registries.hpp:21-68): singleton tables
// of function pointers built once and immutable, keyed by
// (Device, NumericType...) with the same mixed-radix encode; a lookup of an
// empty slot throws std::runtime_error.  Only the Device::HIP slots are
// populated: there is no CPU implementation in this library.
#pragma once

#include <cstdint>
#include <vector>

#include "core.hpp"

namespace tomos {

// ---------------------------------------------------------------- BLAS
// y = alpha*op(A)*x + beta*y through dispatch::Dense::lookup_mv with the
// default compute rule (FP32_16 -> FP16 "pure_half", FP32_BF16 -> BF16
// hi-only, native types -> their own type).  blas.cpp:23-54
Status mv_apply(Device dev, TransposeType trans, double alpha, Dense& A, Dense& x, double beta,
                Dense& y);
// Explicit compute tier; workspace: optional device scratch (float slots)
// sized by mv_workspace_size().  blas.hpp:157-159
Status mv_apply(Device dev, TransposeType trans, double alpha, Dense& A, Dense& x, double beta,
                Dense& y, NumericType compute_type, float* workspace = nullptr);
std::size_t mv_workspace_size(Dense& A, TransposeType trans, NumericType compute_type);

// precond_apply = mv_apply on the preconditioner.  blas.cpp:88-96
void precond_apply(Device dev, TransposeType trans, double alpha, Dense& P, Dense& x, double beta,
                   Dense& y);

// C = alpha*op(A)*op(B) + beta*C.  use_workspace selects the
// decode-then-GEMM tier for packed A (blas.hpp:212-214).
Status mm_apply(Device dev, TransposeType transA, TransposeType transB, double alpha, Dense& A,
                Dense& B, double beta, Dense& C, bool use_workspace = false);

template <typename value_t>
void set_values(value_t value, Dense& target);

// FP32_16 tier entry points (blas.hpp:259-284)
Status split_fp32_16(Dense& source, Dense& target);
Status combine_fp32_16(Dense& source, Dense& target);
Status mv_apply_fp32_16_pure_half(Device dev, TransposeType trans, double alpha, Dense& A, Dense& x,
                                  double beta, Dense& y);
Status mv_apply_fp32_16_fast(Device dev, TransposeType trans, double alpha, Dense& A, Dense& x,
                             double beta, Dense& y);
Status mv_apply_fp32_16_ext(Device dev, TransposeType trans, double alpha, Dense& A, Dense& x,
                            double beta, Dense& y);
Status mm_apply_fp32_16_pure_half(Device dev, TransposeType transA, TransposeType transB,
                                  double alpha, Dense& A, Dense& B, double beta, Dense& C);
Status mm_apply_fp32_16_fast(Device dev, TransposeType transA, TransposeType transB, double alpha,
                             Dense& A, Dense& B, double beta, Dense& C);
Status mm_apply_fp32_16_ext(Device dev, TransposeType transA, TransposeType transB, double alpha,
                            Dense& A, Dense& B, double beta, Dense& C);

// FP32_BF16 counterparts (the reference has these only as raw kernels,
// fp32_bf16_pack.hpp / fp32_16_cuda_kernels.cu:28785, :29636, :7587).
Status split_fp32_bf16(Dense& source, Dense& target, bool rounded = false);
Status combine_fp32_bf16(Dense& source, Dense& target, bool rounded = false);
Status mv_apply_fp32_bf16_hi(Device dev, TransposeType trans, double alpha, Dense& A, Dense& x,
                             double beta, Dense& y);
Status mv_apply_fp32_bf16_ext(Device dev, TransposeType trans, double alpha, Dense& A, Dense& x,
                              double beta, Dense& y);

// ------------------------------------------------------ sketch + precond
// The hashed CountSketch S (k x m): column i of S has +-1 in row h(i),
// h and the sign from splitmix64 with the reference's fixed seeds
// (countsketch_kernels.cpp:59-85).  Deterministic.
struct CountSketch {
    dim<2> size;  // {k, m}
};
CountSketch create_count_sketch(std::size_t k, std::size_t m);
void generate_count_sketch(CountSketch& sketch);
// target (k x n, FP64 or FP32) = S * source.  countsketch.hpp:56-60
void apply_count_sketch(CountSketch& sketch, Dense& source, Dense& target);
// Householder QR of the k x n sketch (FP64, or FP32 factored in FP64 and
// written back as FP32); the leading n x n block is overwritten with R^-1.
// Throws std::invalid_argument if k < n and std::runtime_error if R is
// singular.  sketch_qr.cpp:68-122
void generate_sketch_qr_precond(CountSketch& sketch, Dense& sketched_mtx);
// The reference's signature (sketch_qr.hpp:19): the sketch is consulted
// only for its device, which must be sketched_mtx's.
void generate_sketch_qr_precond(Sparse& sketch, Dense& sketched_mtx);

// The reference's CSR sketch form (k x m, one entry per row at a hashed
// column; countsketch.hpp:56-60): values +-1.  generate_count_sketch runs
// the registry's generate_idxs slot and then its generate_values slot
// (countsketch.cpp:15-18).  Deterministic per process (a seed sequence, as
// the unseeded Gaussian sketch below); needs nnz == k.
void generate_count_sketch(Sparse& sketch);
// target (k x n, FP32 or FP64) = S * source (m x n, FP32 or FP64) for a CSR
// S, through dispatch::Sketch::lookup_apply_count_sketch_dense (the
// reference registers this on the CPU only, sketch_registries.cpp:25-40).
void apply_count_sketch(Sparse& sketch, Dense& source, Dense& target);
// CSR * CSR into a dense target (countsketch.hpp:62): declared for the
// drop-in, nothing is registered on the accelerator, so the lookup throws
// std::runtime_error.
void apply_count_sketch(Sparse& sketch, Sparse& source, Dense& target);

// Dense Gaussian sketch: S <- i.i.d. N(0,1), S FP64, FP32, FP16 or BF16
// (include/core/sketch.hpp:11-13).  The reference seeds each call from the
// clock; here a call without a seed takes the next seed of a process-wide
// sequence, so successive sketches differ and a program's sketches repeat
// from run to run.  The seeded overload fixes the values: they depend only
// on (seed, rows, cols), not on the lead dimension.
Status create_gaussian_sketch(Dense& S);
Status create_gaussian_sketch(Dense& S, std::uint64_t seed);
// Sparse Gaussian sketch: the CSR form above with N(0,1) values
// (sketch.hpp:15-18); needs nnz == rows.
Status create_gaussian_sketch(Sparse& S);
Status create_gaussian_sketch(Sparse& S, std::uint64_t seed);

// ------------------------------------------------------------ sparse BLAS
// y = alpha*op(A)*x + beta*y for CSR A; x, y FP32 or FP64, FP64
// accumulation.  blas.hpp:127-137.  TRANSPOSE validates A and builds A^T
// (deterministic, no atomics) for the call.
Status mv_apply(Device dev, TransposeType trans, double alpha, Sparse& A, Dense& x, double beta, Dense& y);

// ------------------------------------------------------------------ LSQR
// include/core/solver.hpp:15-72
Workspace create_lsqr_workspace(Context& ctx, dim<2> size, NumericType value_type);
Workspace create_lsqr_workspace(Context& ctx, dim<2> size, NumericType value_type,
                                NumericType precond_type);
void lsqr_solve(Context context, Dense& A, Dense& b, Dense& x, Workspace& workspace, double tol,
                int max_iters, double max_relres, int* iterations = nullptr,
                std::vector<double>* residual_history = nullptr, int restarts = 1,
                Dense* mtx_refine_precision = nullptr);
// CSR A (solver.hpp:74-78): A^T built once per call; mtx_refine_precision,
// when given, must have A's pattern (same row pointers and column indices)
// with values at b/x's precision.
void lsqr_solve(Context context, Sparse& A, Dense& b, Dense& x, Workspace& workspace, double tol,
                int max_iters, double max_relres, int* iterations = nullptr,
                std::vector<double>* residual_history = nullptr, int restarts = 1,
                Sparse* mtx_refine_precision = nullptr);

// ------------------------------------------------------------ registries
namespace dispatch {

constexpr std::size_t num_registry_devices = static_cast<std::size_t>(Device::Count);
constexpr std::size_t num_registry_numeric_types = static_cast<std::size_t>(NumericType::Count);

using mv_dense_fn = void (*)(TransposeType trans, double alpha, tomos::Dense& A, tomos::Dense& x,
                             double beta, tomos::Dense& y, float* workspace);
using mm_dense_fn = void (*)(TransposeType transA, TransposeType transB, double alpha,
                             tomos::Dense& A, tomos::Dense& B, double beta, tomos::Dense& C);
using split_fn = void (*)(tomos::Dense& source, tomos::Dense& target);
using combine_fn = void (*)(tomos::Dense& source, tomos::Dense& target);
using lsqr_dense_fn = void (*)(Context context, tomos::Dense& A, tomos::Dense& b, tomos::Dense& x,
                               Workspace& workspace, double tol, int max_iters, double max_relres,
                               int* iterations, std::vector<double>* history, int restarts,
                               tomos::Dense* mtx_refine);
// sketch_registries.hpp:14-19
using generate_count_sketch_values_fn = Status (*)(tomos::Sparse&);
using generate_count_sketch_idxs_fn = Status (*)(tomos::Sparse&);
using apply_count_sketch_dense_fn = void (*)(tomos::Sparse&, tomos::Dense&, tomos::Dense&);
using apply_count_sketch_sparse_fn = void (*)(tomos::Sparse&, tomos::Sparse&, tomos::Dense&);
using create_gaussian_sketch_dense_fn = Status (*)(tomos::Dense& S, std::uint64_t seed);
using create_gaussian_sketch_sparse_fn = Status (*)(tomos::Sparse& S, std::uint64_t seed);
using mv_sparse_fn = void (*)(TransposeType trans, double alpha, tomos::Sparse& A, tomos::Dense& x, double beta,
                              tomos::Dense& y);
using lsqr_sparse_fn = void (*)(Context context, tomos::Sparse& A, tomos::Dense& b, tomos::Dense& x,
                                Workspace& workspace, double tol, int max_iters, double max_relres,
                                int* iterations, std::vector<double>* history, int restarts,
                                tomos::Sparse* mtx_refine);

struct Dense {
    static const Dense& get();
    mv_dense_fn lookup_mv(Device dev, NumericType A, NumericType x, NumericType y,
                          NumericType compute) const;
    mm_dense_fn lookup_mm(Device dev, NumericType A, NumericType B, NumericType C,
                          bool use_workspace) const;
    static std::size_t encode_mv(Device dev, NumericType A, NumericType x, NumericType y,
                                 NumericType compute);
    static std::size_t encode_mm(Device dev, NumericType A, NumericType B, NumericType C,
                                 bool use_workspace);

private:
    Dense();
    static constexpr std::size_t Nmv = num_registry_devices * num_registry_numeric_types *
                                       num_registry_numeric_types * num_registry_numeric_types *
                                       num_registry_numeric_types;
    static constexpr std::size_t Nmm = num_registry_devices * num_registry_numeric_types *
                                       num_registry_numeric_types * num_registry_numeric_types * 2;
    std::array<mv_dense_fn, Nmv> mv_{};
    std::array<mm_dense_fn, Nmm> mm_{};
};

// Per-device slots of the packed formats.  fp32_16_registries.hpp:37-70
struct Fp32_16 {
    static const Fp32_16& get();
    split_fn lookup_split(Device dev) const;
    combine_fn lookup_combine(Device dev) const;
    mv_dense_fn lookup_mv_pure_half(Device dev) const;
    mv_dense_fn lookup_mv_fast(Device dev) const;
    mv_dense_fn lookup_mv_accurate(Device dev) const;
    mm_dense_fn lookup_mm_pure_half(Device dev) const;
    mm_dense_fn lookup_mm_fast(Device dev) const;
    mm_dense_fn lookup_mm_accurate(Device dev) const;
    mm_dense_fn lookup_mm_workspace(Device dev) const;

private:
    Fp32_16();
    std::array<split_fn, num_registry_devices> split_{};
    std::array<combine_fn, num_registry_devices> combine_{};
    std::array<mv_dense_fn, num_registry_devices> mv_pure_half_{}, mv_fast_{}, mv_accurate_{};
    std::array<mm_dense_fn, num_registry_devices> mm_pure_half_{}, mm_fast_{}, mm_accurate_{},
        mm_workspace_{};
};

struct Fp32Bf16 {
    static const Fp32Bf16& get();
    split_fn lookup_split(Device dev) const;
    split_fn lookup_split_rounded(Device dev) const;
    combine_fn lookup_combine(Device dev) const;
    combine_fn lookup_combine_rounded(Device dev) const;
    mv_dense_fn lookup_mv_hi(Device dev) const;
    mv_dense_fn lookup_mv_accurate(Device dev) const;
    mm_dense_fn lookup_mm_hi(Device dev) const;
    mm_dense_fn lookup_mm_accurate(Device dev) const;

private:
    Fp32Bf16();
    std::array<split_fn, num_registry_devices> split_{}, split_rounded_{};
    std::array<combine_fn, num_registry_devices> combine_{}, combine_rounded_{};
    std::array<mv_dense_fn, num_registry_devices> mv_hi_{}, mv_accurate_{};
    std::array<mm_dense_fn, num_registry_devices> mm_hi_{}, mm_accurate_{};
};

struct Lsqr {
    static const Lsqr& get();
    lsqr_dense_fn lookup_solve_dense(Device dev, NumericType mv_type, NumericType value_type,
                                     NumericType refine_type, NumericType precond_type) const;
    lsqr_sparse_fn lookup_solve_sparse(Device dev, NumericType mv_type, NumericType value_type,
                                       NumericType refine_type, NumericType precond_type) const;
    static std::size_t encode(Device dev, NumericType mv_type, NumericType value_type,
                              NumericType refine_type, NumericType precond_type);

private:
    Lsqr();
    static constexpr std::size_t N = num_registry_devices * num_registry_numeric_types *
                                     num_registry_numeric_types * num_registry_numeric_types *
                                     num_registry_numeric_types;
    std::array<lsqr_dense_fn, N> dense_{};
    std::array<lsqr_sparse_fn, N> sparse_{};
};

// CSR products per (device, A, x, y, index type).  sparse_registries.hpp
struct Sparse {
    static const Sparse& get();
    mv_sparse_fn lookup_mv(Device dev, NumericType A, NumericType x, NumericType y, IntType index) const;
    static std::size_t encode_mv(Device dev, NumericType A, NumericType x, NumericType y, IntType index);

private:
    Sparse();
    static constexpr std::size_t N = num_registry_devices * num_registry_numeric_types *
                                     num_registry_numeric_types * num_registry_numeric_types * 2;
    std::array<mv_sparse_fn, N> mv_{};
};

// sketch_registries.hpp:21-66: slots keyed by (device, value type),
// (device, index type) and (device, sketch, source, target, index types)
// with the same mixed-radix encodes.
constexpr std::size_t num_registry_index_types = static_cast<std::size_t>(IntType::Count);

struct Sketch {
    static const Sketch& get();
    generate_count_sketch_values_fn lookup_generate_count_sketch_values(Device dev, NumericType type) const;
    generate_count_sketch_idxs_fn lookup_generate_count_sketch_idxs(Device dev, IntType type) const;
    apply_count_sketch_dense_fn lookup_apply_count_sketch_dense(Device dev, NumericType sketch_type,
                                                                NumericType source_type,
                                                                NumericType target_type,
                                                                IntType index_type) const;
    apply_count_sketch_sparse_fn lookup_apply_count_sketch_sparse(Device dev, NumericType sketch_type,
                                                                  NumericType source_type,
                                                                  NumericType target_type,
                                                                  IntType index_type) const;
    static std::size_t encode_values(Device dev, NumericType type);
    static std::size_t encode_idxs(Device dev, IntType type);
    static std::size_t encode_apply(Device dev, NumericType sketch_type, NumericType source_type,
                                    NumericType target_type, IntType index_type);

private:
    Sketch();
    static constexpr std::size_t N_values = num_registry_devices * num_registry_numeric_types;
    static constexpr std::size_t N_idxs = num_registry_devices * num_registry_index_types;
    static constexpr std::size_t N_apply = num_registry_devices * num_registry_numeric_types *
                                           num_registry_numeric_types * num_registry_numeric_types *
                                           num_registry_index_types;
    std::array<generate_count_sketch_values_fn, N_values> generate_values_{};
    std::array<generate_count_sketch_idxs_fn, N_idxs> generate_idxs_{};
    std::array<apply_count_sketch_dense_fn, N_apply> apply_dense_{};
    std::array<apply_count_sketch_sparse_fn, N_apply> apply_sparse_{};
};

// gaussian_sketch_registries.hpp: one slot per (device, value type).  The
// slot takes the seed as well (the reference's generates its own).
struct GaussianSketch {
    static const GaussianSketch& get();
    create_gaussian_sketch_dense_fn lookup_create_dense(Device dev, NumericType type) const;
    create_gaussian_sketch_sparse_fn lookup_create_sparse(Device dev, NumericType type, IntType index) const;
    static std::size_t encode_dense(Device dev, NumericType type);

private:
    GaussianSketch();
    std::array<create_gaussian_sketch_dense_fn, num_registry_devices * num_registry_numeric_types>
        dense_{};
    std::array<create_gaussian_sketch_sparse_fn, num_registry_devices * num_registry_numeric_types * 2>
        sparse_{};
};

}  // namespace dispatch
}  // namespace tomos
