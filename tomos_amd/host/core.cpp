// Dense / Workspace / Context implementation over the libtomos_hip C ABI.
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>

#include "tomos/core.hpp"
#include "tomos_hip.h"

namespace tomos {

namespace {
void check(int rc, const char* what) {
    if (rc != TOMOS_OK)
        throw std::runtime_error(std::string(what) + ": " + tomos_last_error());
}
void require_device(Device d, const char* what) {
    if (d == Device::OMP)
        throw std::invalid_argument(std::string(what) + ": Device::OMP is not supported.");
    if (d == Device::CPU)
        throw std::runtime_error(std::string(what) +
                                 ": Device::CPU has no implementation in this library "
                                 "(the product path is the GPU; the CPU oracle is test-only).");
}
}  // namespace

const char* numeric_type_name(NumericType t) {
    switch (t) {
        case NumericType::FP64: return "FP64";
        case NumericType::FP32: return "FP32";
        case NumericType::FP16: return "FP16";
        case NumericType::FP32_16: return "FP32_16";
        case NumericType::BF16: return "BF16";
        case NumericType::FP32_BF16: return "FP32_BF16";
        default: return "?";
    }
}

Dense Dense::create_strided(Context& ctx, Layout layout, NumericType type, dim<2> size) {
    const std::size_t ld = layout == Layout::COL_MAJOR ? size[0] : size[1];
    return create_strided(ctx, layout, type, size, ld);
}

Dense Dense::create_strided(Context& ctx, Layout layout, NumericType type, dim<2> size,
                            std::size_t lead_dim) {
    require_device(ctx.get_device(), "Dense::create_strided");
    if (layout != Layout::COL_MAJOR)
        throw std::invalid_argument("Dense::create_strided: only COL_MAJOR storage is supported.");
    if (lead_dim < size[0]) throw std::invalid_argument("Dense::create_strided: lead_dim < rows.");
    Dense d;
    d.device_ = ctx.get_device();
    d.type_ = type;
    d.layout_ = layout;
    d.size_ = size;
    d.lead_dim_ = lead_dim;
    d.stream_ = ctx.get_stream();
    // initialize_strided_storage allocates lead_dim * ncols slots
    // (src/core/memory/dense_storage.cpp:19-26)
    const std::size_t bytes = d.storage_bytes();
    check(tomos_malloc(&d.values_, bytes ? bytes : 1), "Dense::create_strided");
    check(tomos_memset(d.values_, 0, bytes ? bytes : 1, d.stream_), "Dense::create_strided");
    d.owning_ = true;
    return d;
}

Dense Dense::wrap_strided(Context& ctx, Layout layout, NumericType type, dim<2> size,
                          std::size_t lead_dim, void* values) {
    require_device(ctx.get_device(), "Dense::wrap_strided");
    if (layout != Layout::COL_MAJOR)
        throw std::invalid_argument("Dense::wrap_strided: only COL_MAJOR storage is supported.");
    Dense d;
    d.device_ = ctx.get_device();
    d.type_ = type;
    d.layout_ = layout;
    d.size_ = size;
    d.lead_dim_ = lead_dim;
    d.values_ = values;
    d.owning_ = false;
    d.stream_ = ctx.get_stream();
    return d;
}

Dense::Dense(Dense&& o) noexcept { *this = std::move(o); }

Dense& Dense::operator=(Dense&& o) noexcept {
    if (this != &o) {
        if (owning_ && values_) tomos_free(values_);
        device_ = o.device_;
        type_ = o.type_;
        layout_ = o.layout_;
        size_ = o.size_;
        lead_dim_ = o.lead_dim_;
        values_ = o.values_;
        owning_ = o.owning_;
        stream_ = o.stream_;
        o.values_ = nullptr;
        o.owning_ = false;
    }
    return *this;
}

Dense::~Dense() {
    if (owning_ && values_) tomos_free(values_);
}

std::size_t Dense::storage_bytes() const {
    return lead_dim_ * size_[1] * numeric_type_size(type_);
}

NumericPtrVariant Dense::values() const {
    switch (type_) {
        case NumericType::FP64: return static_cast<double*>(values_);
        case NumericType::FP32: return static_cast<float*>(values_);
        case NumericType::FP16: return static_cast<half_t*>(values_);
        case NumericType::FP32_16: return static_cast<uint32_t*>(values_);
        case NumericType::BF16: return static_cast<bf16_t*>(values_);
        case NumericType::FP32_BF16: return static_cast<fp32_bf16_word*>(values_);
        default: throw std::invalid_argument("Dense::values: bad NumericType");
    }
}

uint16_t* Dense::wh() const {
    if (type_ != NumericType::FP32_16 && type_ != NumericType::FP32_BF16)
        throw std::invalid_argument("Dense::wh: not a packed matrix");
    return static_cast<uint16_t*>(values_);
}

uint16_t* Dense::wr() const { return wh() + lead_dim_ * size_[1]; }

void Dense::copy_from_host(const void* src, std::size_t bytes) {
    if (bytes > storage_bytes()) throw std::invalid_argument("Dense::copy_from_host: too many bytes");
    check(tomos_memcpy_htod(values_, src, bytes, stream_), "Dense::copy_from_host");
}

void Dense::copy_to_host(void* dst, std::size_t bytes) const {
    if (bytes > storage_bytes()) throw std::invalid_argument("Dense::copy_to_host: too many bytes");
    check(tomos_memcpy_dtoh(dst, values_, bytes, stream_), "Dense::copy_to_host");
}

namespace {
std::size_t index_size(IntType t) { return t == IntType::INT64 ? 8 : 4; }
}  // namespace

Sparse Sparse::create_csr(Context& ctx, NumericType value_type, IntType index_type, dim<2> size,
                          std::size_t nnz) {
    require_device(ctx.get_device(), "Sparse::create_csr");
    if (value_type != NumericType::FP64 && value_type != NumericType::FP32 && value_type != NumericType::FP16 &&
        value_type != NumericType::BF16)
        throw std::invalid_argument("Sparse::create_csr: values must be FP64, FP32, FP16 or BF16.");
    if (index_type != IntType::INT32 && index_type != IntType::INT64)
        throw std::invalid_argument("Sparse::create_csr: bad index type.");
    if (index_type == IntType::INT32 && (nnz > 0x7FFFFFFFu || size[0] > 0x7FFFFFFFu || size[1] > 0x7FFFFFFFu))
        throw std::invalid_argument("Sparse::create_csr: sizes exceed INT32 indices.");
    Sparse sp;
    sp.device_ = ctx.get_device();
    sp.type_ = value_type;
    sp.index_ = index_type;
    sp.size_ = size;
    sp.nnz_ = nnz;
    sp.stream_ = ctx.get_stream();
    const std::size_t is = index_size(index_type);
    const std::size_t pb = (size[0] + 1) * is, ib = std::max<std::size_t>(nnz, 1) * is,
                      vb = std::max<std::size_t>(nnz, 1) * numeric_type_size(value_type);
    try {
        check(tomos_malloc(&sp.row_ptrs_, pb), "Sparse::create_csr");
        check(tomos_malloc(&sp.col_idxs_, ib), "Sparse::create_csr");
        check(tomos_malloc(&sp.values_, vb), "Sparse::create_csr");
        // an all-zero CSR is valid: every row empty (row_ptrs[m] = 0 = nnz only when nnz == 0)
        check(tomos_memset(sp.row_ptrs_, 0, pb, sp.stream_), "Sparse::create_csr");
        check(tomos_memset(sp.col_idxs_, 0, ib, sp.stream_), "Sparse::create_csr");
        check(tomos_memset(sp.values_, 0, vb, sp.stream_), "Sparse::create_csr");
    } catch (...) {
        sp.release();
        throw;
    }
    return sp;
}

void Sparse::release() {
    if (values_) tomos_free(values_);
    if (row_ptrs_) tomos_free(row_ptrs_);
    if (col_idxs_) tomos_free(col_idxs_);
    values_ = row_ptrs_ = col_idxs_ = nullptr;
}

Sparse::Sparse(Sparse&& o) noexcept { *this = std::move(o); }

Sparse& Sparse::operator=(Sparse&& o) noexcept {
    if (this != &o) {
        release();
        device_ = o.device_;
        type_ = o.type_;
        index_ = o.index_;
        size_ = o.size_;
        nnz_ = o.nnz_;
        values_ = o.values_;
        row_ptrs_ = o.row_ptrs_;
        col_idxs_ = o.col_idxs_;
        stream_ = o.stream_;
        o.values_ = o.row_ptrs_ = o.col_idxs_ = nullptr;
    }
    return *this;
}

Sparse::~Sparse() { release(); }

void Sparse::copy_from_host(const void* rp, const void* ci, const void* vals) {
    const std::size_t is = index_size(index_);
    if (rp) check(tomos_memcpy_htod(row_ptrs_, rp, (size_[0] + 1) * is, stream_), "Sparse::copy_from_host");
    if (ci && nnz_) check(tomos_memcpy_htod(col_idxs_, ci, nnz_ * is, stream_), "Sparse::copy_from_host");
    if (vals && nnz_)
        check(tomos_memcpy_htod(values_, vals, nnz_ * numeric_type_size(type_), stream_), "Sparse::copy_from_host");
}

void Sparse::copy_to_host(void* rp, void* ci, void* vals) const {
    const std::size_t is = index_size(index_);
    if (rp) check(tomos_memcpy_dtoh(rp, row_ptrs_, (size_[0] + 1) * is, stream_), "Sparse::copy_to_host");
    if (ci && nnz_) check(tomos_memcpy_dtoh(ci, col_idxs_, nnz_ * is, stream_), "Sparse::copy_to_host");
    if (vals && nnz_)
        check(tomos_memcpy_dtoh(vals, values_, nnz_ * numeric_type_size(type_), stream_), "Sparse::copy_to_host");
}

Workspace::Workspace(Device device, std::size_t nbytes) : device_(device), capacity_(nbytes) {
    require_device(device, "Workspace");
    void* p = nullptr;
    check(tomos_malloc(&p, nbytes ? nbytes : 1), "Workspace");
    data_ = static_cast<std::byte*>(p);
}

Workspace::Workspace(Workspace&& o) noexcept
    : device_(o.device_),
      capacity_(o.capacity_),
      data_(o.data_),
      precond_kind_(o.precond_kind_),
      precond_(o.precond_),
      value_type_(o.value_type_),
      precond_type_(o.precond_type_),
      use_implicit_duplicate_(o.use_implicit_duplicate_),
      mv_compute_type_(o.mv_compute_type_),
      mv_compute_type_set_(o.mv_compute_type_set_),
      mv_transpose_scratch_(o.mv_transpose_scratch_),
      mv_transpose_scratch_capacity_(o.mv_transpose_scratch_capacity_),
      check_every_(o.check_every_),
      single_pass_(o.single_pass_),
      confirm_true_residual_(o.confirm_true_residual_) {
    o.data_ = nullptr;
    o.mv_transpose_scratch_ = nullptr;
    o.mv_transpose_scratch_capacity_ = 0;
}

Workspace::~Workspace() {
    if (data_) tomos_free(data_);
    if (mv_transpose_scratch_) tomos_free(mv_transpose_scratch_);
    data_ = nullptr;
    mv_transpose_scratch_ = nullptr;
}

void Workspace::reserve_scratch(std::size_t bytes) {
    if (bytes <= mv_transpose_scratch_capacity_) return;
    if (mv_transpose_scratch_) {
        check(tomos_device_synchronize(), "Workspace::reserve_scratch");
        tomos_free(mv_transpose_scratch_);
    }
    mv_transpose_scratch_ = nullptr;
    mv_transpose_scratch_capacity_ = 0;
    void* p = nullptr;
    check(tomos_malloc(&p, bytes), "Workspace::reserve_scratch");
    mv_transpose_scratch_ = static_cast<float*>(p);
    mv_transpose_scratch_capacity_ = bytes;
}

}  // namespace tomos
