// C++ API mirror of TOMOS's core types (include/core/precision.hpp,
// matrix.hpp, workspace.hpp, status.hpp, src/core/context/device.hpp) for
// the LSQR hot path, backed by libtomos_hip.so on MI355X.
//
// Kept alike on purpose (so reference callers compile against it):
//   * NumericType enum values FP64=0, FP32=1, FP16=2, FP32_16=3, BF16=4, with
//     FP32_BF16=5 appended; the pointer variant's alternative index equals
//     the enum value (precision.hpp:26-39);
//   * Device index 2 is the accelerator (the reference's Device::CUDA slot),
//     so registry table sizes and encodings are unchanged;
//   * Dense: column-major strided storage, owning (create_strided) or
//     wrapping (wrap_strided), move-only; packed types hold lead_dim*ncols
//     uint32 slots laid out as the wh plane followed by the wr plane
//     (fp32_16_pack.hpp:861-874);
//   * Workspace fields and add_precond's type check (workspace.hpp:37-120).
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <variant>

struct ihipStream_t;  // hipStream_t == ihipStream_t*

namespace tomos {

enum class Device : std::size_t { CPU = 0, OMP = 1, HIP = 2, Count };

enum class NumericType : std::size_t {
    FP64 = 0,
    FP32 = 1,
    FP16 = 2,
    FP32_16 = 3,
    BF16 = 4,
    FP32_BF16 = 5,
    Count
};

enum class Layout : std::size_t { ROW_MAJOR, COL_MAJOR };
enum class TransposeType : std::size_t { NO_TRANSPOSE = 0, TRANSPOSE = 1 };
enum Status { SUCCESS = 0, FAILURE = 1 };

template <std::size_t N>
using dim = std::array<std::size_t, N>;

// 16-bit storage element types (FP16 / BF16 slots) and the packed-word
// types of FP32_16 / FP32_BF16 (one uint32 slot per element).
struct half_t { uint16_t bits; };
struct bf16_t { uint16_t bits; };
struct fp32_bf16_word { uint32_t bits; };

// Alternative index == NumericType value.
using NumericPtrVariant =
    std::variant<double*, float*, half_t*, uint32_t*, bf16_t*, fp32_bf16_word*>;

constexpr std::size_t numeric_type_size(NumericType t) {
    return t == NumericType::FP64 ? 8
         : (t == NumericType::FP16 || t == NumericType::BF16) ? 2
         : 4;
}

const char* numeric_type_name(NumericType t);

// hipStream_t-compatible handle.
using stream_t = ::ihipStream_t*;

// Execution context: the device (and, for the accelerator, the stream every
// call is enqueued on; nullptr = default stream).  src/core/context/context.hpp:49-67
class Context {
public:
    explicit Context(Device d, stream_t s = nullptr) : device_(d), stream_(s) {}
    Device get_device() const { return device_; }
    stream_t get_stream() const { return stream_; }

private:
    Device device_;
    stream_t stream_;
};

template <typename T>
struct StridedView {
    dim<2> size_;
    std::size_t lead_dim_;
    Layout layout_;
    T* values_;
};

// Dense column-major matrix.  include/core/matrix.hpp:43-121
class Dense {
public:
    static Dense create_strided(Context& ctx, Layout layout, NumericType type, dim<2> size);
    static Dense create_strided(Context& ctx, Layout layout, NumericType type, dim<2> size,
                                std::size_t lead_dim);
    static Dense wrap_strided(Context& ctx, Layout layout, NumericType type, dim<2> size,
                              std::size_t lead_dim, void* values);

    Dense(Dense&& o) noexcept;
    Dense& operator=(Dense&& o) noexcept;
    Dense(const Dense&) = delete;
    Dense& operator=(const Dense&) = delete;
    ~Dense();

    NumericType get_value_type() const { return type_; }
    Device get_device() const { return device_; }
    dim<2> get_size() const { return size_; }
    std::size_t get_lead_dim() const { return lead_dim_; }
    Layout get_layout() const { return layout_; }
    void* data() const { return values_; }
    bool owns() const { return owning_; }
    stream_t stream() const { return stream_; }
    NumericPtrVariant values() const;

    // Packed types: the two planes.
    uint16_t* wh() const;
    uint16_t* wr() const;

    // host <-> device staging of the raw storage (lead_dim * ncols slots)
    void copy_from_host(const void* src, std::size_t bytes);
    void copy_to_host(void* dst, std::size_t bytes) const;
    std::size_t storage_bytes() const;

private:
    Dense() = default;
    Device device_ = Device::HIP;
    NumericType type_ = NumericType::FP32;
    Layout layout_ = Layout::COL_MAJOR;
    dim<2> size_{0, 0};
    std::size_t lead_dim_ = 0;
    void* values_ = nullptr;
    bool owning_ = false;
    stream_t stream_ = nullptr;
};

template <typename T>
StridedView<T> as_strided_view(Dense& d) {
    return StridedView<T>{d.get_size(), d.get_lead_dim(), d.get_layout(),
                          static_cast<T*>(d.data())};
}

// Index type of CSR row pointers and column indices.  precision.hpp:42
enum class IntType : std::size_t { INT32 = 0, INT64 = 1, Count };

// CSR sparse matrix (m x n, nnz stored entries; row_ptrs m+1, col_idxs
// nnz, values nnz, all on the device).  include/core/matrix.hpp:131-180;
// storage as CsrStorage, src/core/matrix/sparse_storage.hpp:16-47.
// Value types FP64, FP32, FP16, BF16.
class Sparse {
public:
    static Sparse create_csr(Context& ctx, NumericType value_type, IntType index_type, dim<2> size,
                             std::size_t nnz);

    Sparse(Sparse&& o) noexcept;
    Sparse& operator=(Sparse&& o) noexcept;
    Sparse(const Sparse&) = delete;
    Sparse& operator=(const Sparse&) = delete;
    ~Sparse();

    NumericType get_value_type() const { return type_; }
    IntType get_index_type() const { return index_; }
    Device get_device() const { return device_; }
    dim<2> get_size() const { return size_; }
    std::size_t get_nnz() const { return nnz_; }
    stream_t stream() const { return stream_; }
    void* values() const { return values_; }
    void* row_ptrs() const { return row_ptrs_; }
    void* col_idxs() const { return col_idxs_; }

    // host -> device staging: row_ptrs (m+1), col_idxs (nnz) in the index
    // type, values (nnz) in the value type; any pointer may be null (skipped)
    void copy_from_host(const void* row_ptrs, const void* col_idxs, const void* values);
    void copy_to_host(void* row_ptrs, void* col_idxs, void* values) const;

private:
    Sparse() = default;
    void release();
    Device device_ = Device::HIP;
    NumericType type_ = NumericType::FP64;
    IntType index_ = IntType::INT32;
    dim<2> size_{0, 0};
    std::size_t nnz_ = 0;
    void* values_ = nullptr;
    void* row_ptrs_ = nullptr;
    void* col_idxs_ = nullptr;
    stream_t stream_ = nullptr;
};

// Device pointers of a CSR matrix.  matrix.hpp CsrView
template <typename T, typename I>
struct CsrView {
    dim<2> size_;
    std::size_t nnz_;
    T* values_;
    I* row_ptrs_;
    I* col_idxs_;
};

template <typename T, typename I>
CsrView<T, I> as_csr_view(Sparse& s) {
    return CsrView<T, I>{s.get_size(), s.get_nnz(), static_cast<T*>(s.values()),
                         static_cast<I*>(s.row_ptrs()), static_cast<I*>(s.col_idxs())};
}

// Identifies the concrete type behind Workspace::precond_.  workspace.hpp:21-35
enum class PrecondKind : std::size_t { None = 0, Dense = 1, Sparse = 2 };

// Byte arena for LSQR's working vectors plus the fields lsqr_solve reads.
// include/core/workspace.hpp:37-106
struct Workspace {
    Device device_;
    std::size_t capacity_ = 0;
    std::byte* data_ = nullptr;

    PrecondKind precond_kind_ = PrecondKind::None;
    void* precond_ = nullptr;

    NumericType value_type_ = NumericType::FP64;
    NumericType precond_type_ = NumericType::FP64;

    // Restart residuals read the packed A at its exact tier (wh+wr).
    bool use_implicit_duplicate_ = false;

    // Per-iteration matvec compute tier override (FP16 = "pure_half" for
    // FP32_16, BF16 = hi-only for FP32_BF16, FP32 = exact).
    NumericType mv_compute_type_ = NumericType::FP16;
    bool mv_compute_type_set_ = false;

    // Device scratch grown at solve time (device state, history, GEMV
    // partial sums); owned, reused across solves.
    float* mv_transpose_scratch_ = nullptr;
    std::size_t mv_transpose_scratch_capacity_ = 0;

    // Iterations between host polls of the device stop flag.
    int check_every_ = 8;

    // Extension (no reference counterpart): run each iteration's A p and
    // A^T u from one read of A (single-pass kernel) and restarts in one
    // pass where supported; the recurrence is the same up to FP32
    // summation order.  On by default (the kernels fall back to separate
    // products where it does not apply); false reads A twice per iteration
    // as the reference's two GEMVs do.
    bool single_pass_ = true;

    // Extension: a stop on the |phi_bar| estimate only ends the current
    // restart, and the solve stops when the next restart's true residual is
    // below tol.  Off by default (the reference stops on the estimate,
    // lsqr.cpp:417-425).
    bool confirm_true_residual_ = false;

    Workspace(Device device, std::size_t nbytes);
    ~Workspace();
    Workspace(Workspace&& o) noexcept;
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;

    void reserve_scratch(std::size_t bytes);
};

template <typename T>
void add_precond(T& precond, Workspace& workspace) {
    if (precond.get_value_type() != workspace.precond_type_)
        throw std::invalid_argument(
            "add_precond: precond's NumericType doesn't match the precond_type the workspace "
            "was created with.");
    workspace.precond_kind_ = PrecondKind::Dense;
    workspace.precond_ = &precond;
}

}  // namespace tomos
