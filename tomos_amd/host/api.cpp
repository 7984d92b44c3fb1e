// BLAS front-ends, sketch/preconditioner, LSQR and the dispatch registries,
// calling the gfx950 kernels through the libtomos_hip C ABI.
#include <algorithm>
#include <atomic>
#include <stdexcept>
#include <string>
#include <vector>

#include "tomos/api.hpp"
#include "tomos_hip.h"

namespace tomos {

namespace {

void check(int rc, const char* what) {
    if (rc == TOMOS_OK) return;
    const std::string msg = std::string(what) + ": " + tomos_last_error();
    if (rc == TOMOS_ERR_INVALID_ARGUMENT || rc == TOMOS_ERR_WORKSPACE) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

int abi_type(NumericType t) { return static_cast<int>(t); }

bool is_packed(NumericType t) { return t == NumericType::FP32_16 || t == NumericType::FP32_BF16; }

// Compute tier selected by a compute NumericType for storage type A
// (dense_registries.cpp FP32_16 delegation, blas.cpp:33-54).
int tier_for(NumericType a, NumericType compute) {
    if (is_packed(a)) {
        if (compute == NumericType::FP32) return TOMOS_TIER_EXACT;
        return TOMOS_TIER_DEFAULT;  // FP16 (FP32_16) / BF16 (FP32_BF16)
    }
    if (a == NumericType::FP32 && compute == NumericType::FP64) return TOMOS_TIER_EXACT;
    return TOMOS_TIER_DEFAULT;
}

NumericType default_compute(NumericType a) {
    if (a == NumericType::FP32_16) return NumericType::FP16;
    if (a == NumericType::FP32_BF16) return NumericType::BF16;
    return a;
}

std::size_t vec_len(tomos::Dense& v) { return v.get_size()[0] * v.get_size()[1]; }

void gemv_call(int tier, TransposeType trans, double alpha, tomos::Dense& A, tomos::Dense& x,
               double beta, tomos::Dense& y, float* workspace) {
    const auto sz = A.get_size();
    const int64_t m = static_cast<int64_t>(sz[0]), n = static_cast<int64_t>(sz[1]);
    const bool t = trans == TransposeType::TRANSPOSE;
    if (vec_len(x) != static_cast<std::size_t>(t ? m : n) || vec_len(y) != static_cast<std::size_t>(t ? n : m))
        throw std::invalid_argument("mv_apply: operand sizes do not match A");
    const void* a = is_packed(A.get_value_type()) ? static_cast<const void*>(A.wh()) : A.data();
    const uint16_t* wr = is_packed(A.get_value_type()) ? A.wr() : nullptr;
    const size_t need = tomos_gemv_scratch_size(abi_type(A.get_value_type()), tier, t ? 1 : 0, m, n);
    check(tomos_gemv(abi_type(A.get_value_type()), tier, t ? 1 : 0, alpha, m, n, a, wr,
                     static_cast<int64_t>(A.get_lead_dim()), abi_type(x.get_value_type()), x.data(),
                     beta, abi_type(y.get_value_type()), y.data(), workspace,
                     workspace ? need : 0, A.stream()),
          "mv_apply");
}

template <int TIER>
void mv_impl(TransposeType trans, double alpha, tomos::Dense& A, tomos::Dense& x, double beta,
             tomos::Dense& y, float* workspace) {
    gemv_call(TIER, trans, alpha, A, x, beta, y, workspace);
}

void gemm_call(int tier, TransposeType ta, TransposeType tb, double alpha, tomos::Dense& A,
               tomos::Dense& B, double beta, tomos::Dense& C) {
    if (B.get_value_type() != NumericType::FP32 || C.get_value_type() != NumericType::FP32)
        throw std::invalid_argument("mm_apply: B and C must be FP32");
    const bool tA = ta == TransposeType::TRANSPOSE, tB = tb == TransposeType::TRANSPOSE;
    const auto sa = A.get_size(), sb = B.get_size(), sc = C.get_size();
    const std::size_t M = tA ? sa[1] : sa[0], K = tA ? sa[0] : sa[1];
    const std::size_t N = tB ? sb[0] : sb[1];
    if ((tB ? sb[1] : sb[0]) != K || sc[0] != M || sc[1] != N)
        throw std::invalid_argument("mm_apply: operand sizes do not match");
    const void* a = is_packed(A.get_value_type()) ? static_cast<const void*>(A.wh()) : A.data();
    const uint16_t* wr = is_packed(A.get_value_type()) ? A.wr() : nullptr;
    check(tomos_gemm(abi_type(A.get_value_type()), tier, tA ? 1 : 0, tB ? 1 : 0, alpha,
                     static_cast<int64_t>(M), static_cast<int64_t>(N), static_cast<int64_t>(K), a, wr,
                     static_cast<int64_t>(A.get_lead_dim()), static_cast<const float*>(B.data()),
                     static_cast<int64_t>(B.get_lead_dim()), beta, static_cast<float*>(C.data()),
                     static_cast<int64_t>(C.get_lead_dim()), A.stream()),
          "mm_apply");
}

template <int TIER>
void mm_impl(TransposeType ta, TransposeType tb, double alpha, tomos::Dense& A, tomos::Dense& B,
             double beta, tomos::Dense& C) {
    gemm_call(TIER, ta, tb, alpha, A, B, beta, C);
}

// "workspace" tier (fp32_16_registries.cpp:432-449 materialize tier: the
// exact values of A through an FP32 GEMM): here the 3-term split of the
// exact values on 16-bit matrix cores, FP32 accumulate.
void mm_materialize(TransposeType ta, TransposeType tb, double alpha, tomos::Dense& A,
                    tomos::Dense& B, double beta, tomos::Dense& C) {
    gemm_call(TOMOS_TIER_MATERIALIZE, ta, tb, alpha, A, B, beta, C);
    check(tomos_stream_synchronize(A.stream()), "mm_apply (materialize)");
}

template <int VARIANT>
void split_impl(tomos::Dense& source, tomos::Dense& target) {
    if (source.get_value_type() != NumericType::FP32) throw std::invalid_argument("split: source must be FP32");
    const auto s = source.get_size();
    if (target.get_size() != s) throw std::invalid_argument("split: size mismatch");
    check(tomos_pack(VARIANT, static_cast<int64_t>(s[0]), static_cast<int64_t>(s[1]),
                     static_cast<const float*>(source.data()), static_cast<int64_t>(source.get_lead_dim()),
                     target.wh(), target.wr(), static_cast<int64_t>(target.get_lead_dim()),
                     target.stream()),
          "split");
}

template <int DEC>
void combine_impl(tomos::Dense& source, tomos::Dense& target) {
    if (target.get_value_type() != NumericType::FP32) throw std::invalid_argument("combine: target must be FP32");
    const auto s = target.get_size();
    if (source.get_size() != s) throw std::invalid_argument("combine: size mismatch");
    check(tomos_unpack(DEC, static_cast<int64_t>(s[0]), static_cast<int64_t>(s[1]), source.wh(),
                       source.wr(), static_cast<int64_t>(source.get_lead_dim()),
                       static_cast<float*>(target.data()), static_cast<int64_t>(target.get_lead_dim()),
                       source.stream()),
          "combine");
}

void device_guard(Device dev, const char* what) {
    if (dev == Device::OMP) throw std::invalid_argument(std::string(what) + ": Device::OMP is not supported.");
}

}  // namespace

// ============================================================ registries
namespace dispatch {

std::size_t Dense::encode_mv(Device dev, NumericType A, NumericType x, NumericType y, NumericType c) {
    const std::size_t T = num_registry_numeric_types;
    return static_cast<std::size_t>(dev) * T * T * T * T + static_cast<std::size_t>(A) * T * T * T +
           static_cast<std::size_t>(x) * T * T + static_cast<std::size_t>(y) * T +
           static_cast<std::size_t>(c);
}

std::size_t Dense::encode_mm(Device dev, NumericType A, NumericType B, NumericType C, bool ws) {
    const std::size_t T = num_registry_numeric_types;
    return static_cast<std::size_t>(dev) * T * T * T * 2 + static_cast<std::size_t>(A) * T * T * 2 +
           static_cast<std::size_t>(B) * T * 2 + static_cast<std::size_t>(C) * 2 + (ws ? 1 : 0);
}

Dense::Dense() {
    mv_.fill(nullptr);
    mm_.fill(nullptr);
    const Device d = Device::HIP;
    const NumericType vt[] = {NumericType::FP32, NumericType::FP64};
    for (NumericType x : vt)
        for (NumericType y : vt) {
            mv_[encode_mv(d, NumericType::FP32, x, y, NumericType::FP32)] = &mv_impl<TOMOS_TIER_DEFAULT>;
            mv_[encode_mv(d, NumericType::FP32, x, y, NumericType::FP64)] = &mv_impl<TOMOS_TIER_EXACT>;
            mv_[encode_mv(d, NumericType::FP64, x, y, NumericType::FP64)] = &mv_impl<TOMOS_TIER_DEFAULT>;
            mv_[encode_mv(d, NumericType::FP32_16, x, y, NumericType::FP16)] = &mv_impl<TOMOS_TIER_DEFAULT>;
            mv_[encode_mv(d, NumericType::FP32_16, x, y, NumericType::FP32)] = &mv_impl<TOMOS_TIER_EXACT>;
            mv_[encode_mv(d, NumericType::FP32_BF16, x, y, NumericType::BF16)] = &mv_impl<TOMOS_TIER_DEFAULT>;
            mv_[encode_mv(d, NumericType::FP32_BF16, x, y, NumericType::FP32)] = &mv_impl<TOMOS_TIER_EXACT>;
            // native 16-bit storage (mv_fp16_native_mixed, dense_registries.cpp:332-333):
            // FP32 accumulation at the storage's own compute type, FP64 when FP32/FP64 is asked
            mv_[encode_mv(d, NumericType::FP16, x, y, NumericType::FP16)] = &mv_impl<TOMOS_TIER_DEFAULT>;
            mv_[encode_mv(d, NumericType::FP16, x, y, NumericType::FP32)] = &mv_impl<TOMOS_TIER_FAST>;
            mv_[encode_mv(d, NumericType::BF16, x, y, NumericType::BF16)] = &mv_impl<TOMOS_TIER_DEFAULT>;
            mv_[encode_mv(d, NumericType::BF16, x, y, NumericType::FP32)] = &mv_impl<TOMOS_TIER_FAST>;
        }
    const NumericType f = NumericType::FP32;
    mm_[encode_mm(d, NumericType::FP32, f, f, false)] = &mm_impl<TOMOS_TIER_DEFAULT>;
    mm_[encode_mm(d, NumericType::FP64, f, f, false)] = &mm_impl<TOMOS_TIER_DEFAULT>;
    mm_[encode_mm(d, NumericType::FP32_16, f, f, false)] = &mm_impl<TOMOS_TIER_EXACT>;
    mm_[encode_mm(d, NumericType::FP32_16, f, f, true)] = &mm_materialize;
    mm_[encode_mm(d, NumericType::FP32_BF16, f, f, false)] = &mm_impl<TOMOS_TIER_EXACT>;
    mm_[encode_mm(d, NumericType::FP32_BF16, f, f, true)] = &mm_materialize;
}

const Dense& Dense::get() {
    static const Dense instance;
    return instance;
}

mv_dense_fn Dense::lookup_mv(Device dev, NumericType A, NumericType x, NumericType y,
                             NumericType c) const {
    device_guard(dev, "lookup_mv");
    auto fn = mv_[encode_mv(dev, A, x, y, c)];
    if (!fn)
        throw std::runtime_error(std::string("No dense mv kernel registered for device/A/x/y/compute = ") +
                                 std::to_string(static_cast<int>(dev)) + "/" + numeric_type_name(A) + "/" +
                                 numeric_type_name(x) + "/" + numeric_type_name(y) + "/" +
                                 numeric_type_name(c));
    return fn;
}

mm_dense_fn Dense::lookup_mm(Device dev, NumericType A, NumericType B, NumericType C, bool ws) const {
    device_guard(dev, "lookup_mm");
    auto fn = mm_[encode_mm(dev, A, B, C, ws)];
    if (!fn) throw std::runtime_error("No dense mm kernel registered for the requested combination.");
    return fn;
}

#define TOMOS_SLOT(arr, name)                                                                  \
    if (dev == Device::OMP) throw std::invalid_argument(#name ": Device::OMP is not supported."); \
    auto fn = arr[static_cast<std::size_t>(dev)];                                              \
    if (!fn) throw std::runtime_error(#name ": nothing registered for this device.");          \
    return fn;

Fp32_16::Fp32_16() {
    const std::size_t h = static_cast<std::size_t>(Device::HIP);
    split_[h] = &split_impl<TOMOS_PACK_FP32_16>;
    combine_[h] = &combine_impl<TOMOS_DECODE_FP32_16_EXACT>;
    mv_pure_half_[h] = &mv_impl<TOMOS_TIER_DEFAULT>;
    mv_fast_[h] = &mv_impl<TOMOS_TIER_FAST>;
    mv_accurate_[h] = &mv_impl<TOMOS_TIER_EXACT>;
    mm_pure_half_[h] = &mm_impl<TOMOS_TIER_DEFAULT>;
    mm_fast_[h] = &mm_impl<TOMOS_TIER_FAST>;
    mm_accurate_[h] = &mm_impl<TOMOS_TIER_EXACT>;
    mm_workspace_[h] = &mm_materialize;
}
const Fp32_16& Fp32_16::get() {
    static const Fp32_16 instance;
    return instance;
}
split_fn Fp32_16::lookup_split(Device dev) const { TOMOS_SLOT(split_, lookup_split) }
combine_fn Fp32_16::lookup_combine(Device dev) const { TOMOS_SLOT(combine_, lookup_combine) }
mv_dense_fn Fp32_16::lookup_mv_pure_half(Device dev) const { TOMOS_SLOT(mv_pure_half_, lookup_mv_pure_half) }
mv_dense_fn Fp32_16::lookup_mv_fast(Device dev) const { TOMOS_SLOT(mv_fast_, lookup_mv_fast) }
mv_dense_fn Fp32_16::lookup_mv_accurate(Device dev) const { TOMOS_SLOT(mv_accurate_, lookup_mv_accurate) }
mm_dense_fn Fp32_16::lookup_mm_pure_half(Device dev) const { TOMOS_SLOT(mm_pure_half_, lookup_mm_pure_half) }
mm_dense_fn Fp32_16::lookup_mm_fast(Device dev) const { TOMOS_SLOT(mm_fast_, lookup_mm_fast) }
mm_dense_fn Fp32_16::lookup_mm_accurate(Device dev) const { TOMOS_SLOT(mm_accurate_, lookup_mm_accurate) }
mm_dense_fn Fp32_16::lookup_mm_workspace(Device dev) const { TOMOS_SLOT(mm_workspace_, lookup_mm_workspace) }

Fp32Bf16::Fp32Bf16() {
    const std::size_t h = static_cast<std::size_t>(Device::HIP);
    split_[h] = &split_impl<TOMOS_PACK_FP32_BF16>;
    split_rounded_[h] = &split_impl<TOMOS_PACK_FP32_BF16_ROUNDED>;
    combine_[h] = &combine_impl<TOMOS_DECODE_FP32_BF16_EXACT>;
    combine_rounded_[h] = &combine_impl<TOMOS_DECODE_FP32_BF16_EXACT_ROUNDED>;
    mv_hi_[h] = &mv_impl<TOMOS_TIER_DEFAULT>;
    mv_accurate_[h] = &mv_impl<TOMOS_TIER_EXACT>;
    mm_hi_[h] = &mm_impl<TOMOS_TIER_DEFAULT>;
    mm_accurate_[h] = &mm_impl<TOMOS_TIER_EXACT>;
}
const Fp32Bf16& Fp32Bf16::get() {
    static const Fp32Bf16 instance;
    return instance;
}
split_fn Fp32Bf16::lookup_split(Device dev) const { TOMOS_SLOT(split_, lookup_split) }
split_fn Fp32Bf16::lookup_split_rounded(Device dev) const { TOMOS_SLOT(split_rounded_, lookup_split_rounded) }
combine_fn Fp32Bf16::lookup_combine(Device dev) const { TOMOS_SLOT(combine_, lookup_combine) }
combine_fn Fp32Bf16::lookup_combine_rounded(Device dev) const { TOMOS_SLOT(combine_rounded_, lookup_combine_rounded) }
mv_dense_fn Fp32Bf16::lookup_mv_hi(Device dev) const { TOMOS_SLOT(mv_hi_, lookup_mv_hi) }
mv_dense_fn Fp32Bf16::lookup_mv_accurate(Device dev) const { TOMOS_SLOT(mv_accurate_, lookup_mv_accurate) }
mm_dense_fn Fp32Bf16::lookup_mm_hi(Device dev) const { TOMOS_SLOT(mm_hi_, lookup_mm_hi) }
mm_dense_fn Fp32Bf16::lookup_mm_accurate(Device dev) const { TOMOS_SLOT(mm_accurate_, lookup_mm_accurate) }

#undef TOMOS_SLOT

}  // namespace dispatch

// ======================================================================
// LSQR: lsqr_dense_impl (src/dispatch/lsqr_registries.cpp:88-159)
// ======================================================================
namespace {

template <NumericType MV, NumericType VAL, NumericType REF, NumericType PRE>
void lsqr_dense_impl(Context context, tomos::Dense& A, tomos::Dense& b, tomos::Dense& x,
                     Workspace& ws, double tol, int max_iters, double max_relres, int* iterations,
                     std::vector<double>* history, int restarts, tomos::Dense* mtx_refine) {
    if (ws.device_ != context.get_device())
        throw std::invalid_argument("lsqr_solve: workspace and context must be on the same device.");
    const auto sz = A.get_size();
    if (mtx_refine) {
        // explicit duplicate (lsqr.cpp:463-469): the restart residuals read it
        if (mtx_refine->get_value_type() != NumericType::FP32 && mtx_refine->get_value_type() != NumericType::FP64)
            throw std::invalid_argument("lsqr_solve: the refine-precision duplicate must be FP32 or FP64.");
        if (mtx_refine->get_size() != sz || mtx_refine->get_device() != A.get_device())
            throw std::invalid_argument("lsqr_solve: the refine-precision duplicate must match A's size and device.");
        if (ws.use_implicit_duplicate_)
            throw std::invalid_argument(
                "lsqr_solve: an explicit duplicate and Workspace::use_implicit_duplicate_ are mutually exclusive.");
    }
    const int64_t m = static_cast<int64_t>(sz[0]), n = static_cast<int64_t>(sz[1]);
    if (vec_len(b) != sz[0] || vec_len(x) != sz[1])
        throw std::invalid_argument("lsqr_solve: b/x sizes do not match A");
    if (ws.use_implicit_duplicate_ && !is_packed(MV))
        throw std::invalid_argument(
            "lsqr_solve: Workspace::use_implicit_duplicate_ requires A to be stored packed "
            "(FP32_16 or FP32_BF16).");
    const NumericType compute = ws.mv_compute_type_set_ ? ws.mv_compute_type_ : default_compute(MV);
    // the compute type must be one the dense registry knows for A
    (void)dispatch::Dense::get().lookup_mv(context.get_device(), MV, VAL, VAL, compute);
    const bool pre = ws.precond_kind_ == PrecondKind::Dense;
    tomos::Dense* P = pre ? static_cast<tomos::Dense*>(ws.precond_) : nullptr;

    tomos_lsqr_options o{};
    o.a_type = abi_type(MV);
    o.mv_tier = tier_for(MV, compute);
    o.restart_tier = ws.use_implicit_duplicate_ ? TOMOS_TIER_EXACT : o.mv_tier;
    o.value_type = abi_type(VAL);
    o.refine_type = abi_type(REF);
    o.precond_type = pre ? abi_type(PRE) : -1;
    o.tol = tol;
    o.max_relres = max_relres;
    o.max_iters = max_iters;
    o.restarts = restarts;
    o.check_every = std::max(1, ws.check_every_);
    o.fused = ws.single_pass_ ? 1 : 0;
    o.confirm = ws.confirm_true_residual_ ? 1 : 0;
    const size_t arena = tomos_lsqr_arena_size(&o, m, n);
    const size_t scratch = tomos_lsqr_scratch_size(&o, m, n);
    if (arena == 0 || scratch == 0) check(TOMOS_ERR_INVALID_ARGUMENT, "lsqr_solve");
    if (arena > ws.capacity_)
        throw std::invalid_argument("lsqr_solve: workspace is too small for LSQR's working vectors (need " +
                                    std::to_string(arena) + " bytes, have " +
                                    std::to_string(ws.capacity_) + ").");
    // the only allocation on the solve path, once per shape (lsqr_registries.cpp:132-145)
    ws.reserve_scratch(scratch);
    const int cap = std::max(0, restarts) * (std::max(0, max_iters) + 1) + 1;
    std::vector<double> hist(history ? static_cast<std::size_t>(cap) : 0);
    int it = 0, hl = 0;
    const void* a = is_packed(MV) ? static_cast<const void*>(A.wh()) : A.data();
    const uint16_t* wr = is_packed(MV) ? A.wr() : nullptr;
    if (mtx_refine)
        check(tomos_lsqr_solve_refined(&o, m, n, a, wr, static_cast<int64_t>(A.get_lead_dim()),
                                       abi_type(mtx_refine->get_value_type()), mtx_refine->data(),
                                       static_cast<int64_t>(mtx_refine->get_lead_dim()), b.data(), x.data(),
                                       P ? P->data() : nullptr, P ? static_cast<int64_t>(P->get_lead_dim()) : 1,
                                       ws.data_, ws.capacity_, ws.mv_transpose_scratch_,
                                       ws.mv_transpose_scratch_capacity_, &it, history ? hist.data() : nullptr,
                                       history ? cap : 0, &hl, context.get_stream()),
              "lsqr_solve");
    else
        check(tomos_lsqr_solve(&o, m, n, a, wr, static_cast<int64_t>(A.get_lead_dim()), b.data(), x.data(),
                               P ? P->data() : nullptr, P ? static_cast<int64_t>(P->get_lead_dim()) : 1,
                               ws.data_, ws.capacity_, ws.mv_transpose_scratch_,
                               ws.mv_transpose_scratch_capacity_, &it, history ? hist.data() : nullptr,
                               history ? cap : 0, &hl, context.get_stream()),
              "lsqr_solve");
    if (iterations) *iterations = it;
    if (history) history->insert(history->end(), hist.begin(), hist.begin() + std::min(hl, cap));
}

int abi_index(IntType t) { return t == IntType::INT64 ? TOMOS_INDEX_INT64 : TOMOS_INDEX_INT32; }

// lsqr_sparse_impl (src/dispatch/lsqr_registries.cpp:164-196)
template <NumericType MV, NumericType VAL, NumericType REF, NumericType PRE>
void lsqr_sparse_impl(Context context, tomos::Sparse& A, tomos::Dense& b, tomos::Dense& x, Workspace& ws,
                      double tol, int max_iters, double max_relres, int* iterations, std::vector<double>* history,
                      int restarts, tomos::Sparse* mtx_refine) {
    if (ws.device_ != context.get_device())
        throw std::invalid_argument("lsqr_solve: workspace and context must be on the same device.");
    const auto sz = A.get_size();
    const int64_t m = static_cast<int64_t>(sz[0]), n = static_cast<int64_t>(sz[1]);
    if (vec_len(b) != sz[0] || vec_len(x) != sz[1])
        throw std::invalid_argument("lsqr_solve: b/x sizes do not match A");
    if (mtx_refine) {
        if (mtx_refine->get_value_type() != NumericType::FP32 && mtx_refine->get_value_type() != NumericType::FP64)
            throw std::invalid_argument("lsqr_solve: the refine-precision duplicate must be FP32 or FP64.");
        if (mtx_refine->get_size() != sz || mtx_refine->get_nnz() != A.get_nnz() ||
            mtx_refine->get_index_type() != A.get_index_type() || mtx_refine->get_device() != A.get_device())
            throw std::invalid_argument("lsqr_solve: the refine-precision duplicate must match A's size and pattern.");
        // the restarts read its values over A's indices: the patterns must agree
        const std::size_t is = A.get_index_type() == IntType::INT64 ? 8 : 4;
        std::vector<unsigned char> ra((sz[0] + 1) * is), rb(ra.size()), ca(A.get_nnz() * is), cb(ca.size());
        A.copy_to_host(ra.data(), ca.data(), nullptr);
        mtx_refine->copy_to_host(rb.data(), cb.data(), nullptr);
        if (ra != rb || ca != cb)
            throw std::invalid_argument("lsqr_solve: the refine-precision duplicate must have A's sparsity pattern.");
    }
    if (ws.use_implicit_duplicate_)
        throw std::invalid_argument("lsqr_solve: Workspace::use_implicit_duplicate_ needs dense packed A.");
    const bool pre = ws.precond_kind_ == PrecondKind::Dense;
    if (ws.precond_kind_ == PrecondKind::Sparse)
        throw std::invalid_argument("lsqr_solve: a sparse preconditioner is not supported; use a dense R^-1.");
    tomos::Dense* P = pre ? static_cast<tomos::Dense*>(ws.precond_) : nullptr;
    tomos_lsqr_options o{};
    o.a_type = abi_type(MV);
    o.mv_tier = TOMOS_TIER_DEFAULT;
    o.restart_tier = TOMOS_TIER_DEFAULT;
    o.value_type = abi_type(VAL);
    o.refine_type = abi_type(REF);
    o.precond_type = pre ? abi_type(PRE) : -1;
    o.tol = tol;
    o.max_relres = max_relres;
    o.max_iters = max_iters;
    o.restarts = restarts;
    o.check_every = std::max(1, ws.check_every_);
    o.fused = 0;
    o.confirm = ws.confirm_true_residual_ ? 1 : 0;
    const int idx = abi_index(A.get_index_type());
    const int64_t nnz = static_cast<int64_t>(A.get_nnz());
    const size_t arena = tomos_lsqr_arena_size(&o, m, n);
    const size_t scratch = tomos_lsqr_csr_scratch_size(&o, idx, m, n, nnz);
    if (arena == 0 || scratch == 0) check(TOMOS_ERR_INVALID_ARGUMENT, "lsqr_solve");
    if (arena > ws.capacity_)
        throw std::invalid_argument("lsqr_solve: workspace is too small for LSQR's working vectors (need " +
                                    std::to_string(arena) + " bytes, have " + std::to_string(ws.capacity_) + ").");
    ws.reserve_scratch(scratch);
    const int cap = std::max(0, restarts) * (std::max(0, max_iters) + 1) + 1;
    std::vector<double> hist(history ? static_cast<std::size_t>(cap) : 0);
    int it = 0, hl = 0;
    check(tomos_lsqr_solve_csr(&o, idx, m, n, nnz, A.values(), A.row_ptrs(), A.col_idxs(),
                               mtx_refine ? abi_type(mtx_refine->get_value_type()) : -1,
                               mtx_refine ? mtx_refine->values() : nullptr, b.data(), x.data(),
                               P ? P->data() : nullptr, P ? static_cast<int64_t>(P->get_lead_dim()) : 1, ws.data_,
                               ws.capacity_, ws.mv_transpose_scratch_, ws.mv_transpose_scratch_capacity_, &it,
                               history ? hist.data() : nullptr, history ? cap : 0, &hl, context.get_stream()),
          "lsqr_solve");
    if (iterations) *iterations = it;
    if (history) history->insert(history->end(), hist.begin(), hist.begin() + std::min(hl, cap));
}

}  // namespace

namespace dispatch {

std::size_t Lsqr::encode(Device dev, NumericType mv, NumericType val, NumericType ref, NumericType pre) {
    const std::size_t T = num_registry_numeric_types;
    return static_cast<std::size_t>(dev) * T * T * T * T + static_cast<std::size_t>(mv) * T * T * T +
           static_cast<std::size_t>(val) * T * T + static_cast<std::size_t>(ref) * T +
           static_cast<std::size_t>(pre);
}

namespace {
constexpr std::size_t kLsqrSlots = num_registry_devices * num_registry_numeric_types * num_registry_numeric_types *
                                   num_registry_numeric_types * num_registry_numeric_types;
template <NumericType MV, NumericType VAL, NumericType REF>
void register_pre(std::array<lsqr_dense_fn, kLsqrSlots>& t) {
    t[Lsqr::encode(Device::HIP, MV, VAL, REF, NumericType::FP32)] =
        &lsqr_dense_impl<MV, VAL, REF, NumericType::FP32>;
    t[Lsqr::encode(Device::HIP, MV, VAL, REF, NumericType::FP64)] =
        &lsqr_dense_impl<MV, VAL, REF, NumericType::FP64>;
}
template <NumericType MV, NumericType VAL, NumericType REF>
void register_pre(std::array<lsqr_sparse_fn, kLsqrSlots>& t) {
    t[Lsqr::encode(Device::HIP, MV, VAL, REF, NumericType::FP32)] =
        &lsqr_sparse_impl<MV, VAL, REF, NumericType::FP32>;
    t[Lsqr::encode(Device::HIP, MV, VAL, REF, NumericType::FP64)] =
        &lsqr_sparse_impl<MV, VAL, REF, NumericType::FP64>;
}
template <NumericType MV, NumericType VAL, class Arr>
void register_ref(Arr& t) {
    register_pre<MV, VAL, NumericType::FP32>(t);
    register_pre<MV, VAL, NumericType::FP64>(t);
}
template <NumericType MV, class Arr>
void register_val(Arr& t) {
    register_ref<MV, NumericType::FP32>(t);
    register_ref<MV, NumericType::FP64>(t);
}
}  // namespace

Lsqr::Lsqr() {
    dense_.fill(nullptr);
    register_val<NumericType::FP32>(dense_);
    register_val<NumericType::FP64>(dense_);
    register_val<NumericType::FP32_16>(dense_);
    register_val<NumericType::FP32_BF16>(dense_);
    register_val<NumericType::FP16>(dense_);
    register_val<NumericType::BF16>(dense_);
    sparse_.fill(nullptr);
    register_val<NumericType::FP64>(sparse_);
    register_val<NumericType::FP32>(sparse_);
    register_val<NumericType::FP16>(sparse_);
    register_val<NumericType::BF16>(sparse_);
}

const Lsqr& Lsqr::get() {
    static const Lsqr instance;
    return instance;
}

lsqr_sparse_fn Lsqr::lookup_solve_sparse(Device dev, NumericType mv, NumericType val, NumericType ref,
                                         NumericType pre) const {
    if (dev == Device::OMP) throw std::invalid_argument("lookup_solve_sparse: Device::OMP is not supported.");
    auto fn = sparse_[encode(dev, mv, val, ref, pre)];
    if (!fn)
        throw std::runtime_error(
            "No LSQR sparse kernel registered for the requested "
            "device/mv_type/value_type/refine_type/precond_type combination.");
    return fn;
}

// ---- CSR products (sparse_registries.cpp:225-243) ----
std::size_t Sparse::encode_mv(Device dev, NumericType A, NumericType x, NumericType y, IntType index) {
    const std::size_t T = num_registry_numeric_types;
    return (((static_cast<std::size_t>(dev) * T + static_cast<std::size_t>(A)) * T + static_cast<std::size_t>(x)) *
                T +
            static_cast<std::size_t>(y)) *
               2 +
           static_cast<std::size_t>(index);
}

namespace {
void mv_csr_impl(TransposeType trans, double alpha, tomos::Sparse& A, tomos::Dense& x, double beta,
                 tomos::Dense& y) {
    const auto sz = A.get_size();
    const bool t = trans == TransposeType::TRANSPOSE;
    if (vec_len(x) != (t ? sz[0] : sz[1]) || vec_len(y) != (t ? sz[1] : sz[0]))
        throw std::invalid_argument("mv_apply: operand sizes do not match A");
    check(tomos_csr_mv(abi_type(A.get_value_type()), abi_index(A.get_index_type()), t ? 1 : 0, alpha,
                       static_cast<int64_t>(sz[0]), static_cast<int64_t>(sz[1]), static_cast<int64_t>(A.get_nnz()),
                       A.values(), A.row_ptrs(), A.col_idxs(), abi_type(x.get_value_type()), x.data(), beta,
                       abi_type(y.get_value_type()), y.data(), nullptr, 0, A.stream()),
          "mv_apply");
}
}  // namespace

Sparse::Sparse() {
    mv_.fill(nullptr);
    for (NumericType a : {NumericType::FP64, NumericType::FP32, NumericType::FP16, NumericType::BF16})
        for (NumericType xv : {NumericType::FP64, NumericType::FP32})
            for (NumericType yv : {NumericType::FP64, NumericType::FP32})
                for (IntType it : {IntType::INT32, IntType::INT64})
                    mv_[encode_mv(Device::HIP, a, xv, yv, it)] = &mv_csr_impl;
}
const Sparse& Sparse::get() {
    static const Sparse instance;
    return instance;
}
mv_sparse_fn Sparse::lookup_mv(Device dev, NumericType A, NumericType x, NumericType y, IntType index) const {
    if (dev == Device::OMP) throw std::invalid_argument("mv_apply: Device::OMP is not supported.");
    auto fn = mv_[encode_mv(dev, A, x, y, index)];
    if (!fn) throw std::runtime_error("mv_apply: no CSR kernel registered for this device/type combination.");
    return fn;
}

lsqr_dense_fn Lsqr::lookup_solve_dense(Device dev, NumericType mv, NumericType val, NumericType ref,
                                       NumericType pre) const {
    if (dev == Device::OMP) throw std::invalid_argument("lookup_solve_dense: Device::OMP is not supported.");
    auto fn = dense_[encode(dev, mv, val, ref, pre)];
    if (!fn)
        throw std::runtime_error(
            "No LSQR dense kernel registered for the requested "
            "device/mv_type/value_type/refine_type/precond_type combination.");
    return fn;
}

}  // namespace dispatch

// ======================================================================
// Sketch + preconditioner
// ======================================================================
namespace {
void cs_apply_hashed(CountSketch& S, tomos::Dense& source, tomos::Dense& target) {
    if (target.get_value_type() != NumericType::FP64)
        throw std::invalid_argument("apply_count_sketch: target must be FP64");
    const auto ss = source.get_size(), ts = target.get_size();
    if (S.size[1] != ss[0] || ts[0] != S.size[0] || ts[1] != ss[1])
        throw std::invalid_argument("apply_count_sketch: size mismatch");
    const NumericType t = source.get_value_type();
    const void* a = is_packed(t) ? static_cast<const void*>(source.wh()) : source.data();
    const uint16_t* wr = is_packed(t) ? source.wr() : nullptr;
    check(tomos_countsketch_apply(abi_type(t), is_packed(t) ? TOMOS_TIER_EXACT : TOMOS_TIER_DEFAULT,
                                  static_cast<int64_t>(ss[0]), static_cast<int64_t>(ss[1]), a, wr,
                                  static_cast<int64_t>(source.get_lead_dim()), static_cast<int64_t>(S.size[0]),
                                  static_cast<double*>(target.data()),
                                  static_cast<int64_t>(target.get_lead_dim()), nullptr, 0, source.stream()),
          "apply_count_sketch");
}
}  // namespace

// The registry's accelerator slots: CSR sketch pattern / values
// (tomos_csr_sketch kinds 2 and 3) and S * source for a CSR S (tomos_csr_mm).
Status cs_generate_idxs_impl(tomos::Sparse& S);
Status cs_generate_values_impl(tomos::Sparse& S);
void cs_apply_csr_impl(tomos::Sparse& S, tomos::Dense& source, tomos::Dense& target);

namespace dispatch {
std::size_t Sketch::encode_values(Device dev, NumericType type) {
    return static_cast<std::size_t>(dev) * num_registry_numeric_types + static_cast<std::size_t>(type);
}
std::size_t Sketch::encode_idxs(Device dev, IntType type) {
    return static_cast<std::size_t>(dev) * num_registry_index_types + static_cast<std::size_t>(type);
}
std::size_t Sketch::encode_apply(Device dev, NumericType sk, NumericType src, NumericType tgt, IntType idx) {
    constexpr std::size_t T = num_registry_numeric_types, I = num_registry_index_types;
    return static_cast<std::size_t>(dev) * T * T * T * I + static_cast<std::size_t>(sk) * T * T * I +
           static_cast<std::size_t>(src) * T * I + static_cast<std::size_t>(tgt) * I +
           static_cast<std::size_t>(idx);
}
Sketch::Sketch() {
    generate_values_.fill(nullptr);
    generate_idxs_.fill(nullptr);
    apply_dense_.fill(nullptr);
    apply_sparse_.fill(nullptr);
    const NumericType vals[] = {NumericType::FP64, NumericType::FP32, NumericType::FP16, NumericType::BF16};
    const NumericType dense[] = {NumericType::FP64, NumericType::FP32};
    const IntType idxs[] = {IntType::INT32, IntType::INT64};
    for (NumericType v : vals) generate_values_[encode_values(Device::HIP, v)] = &cs_generate_values_impl;
    for (IntType i : idxs) generate_idxs_[encode_idxs(Device::HIP, i)] = &cs_generate_idxs_impl;
    for (NumericType sk : vals)
        for (NumericType src : dense)
            for (NumericType tgt : dense)
                for (IntType i : idxs) apply_dense_[encode_apply(Device::HIP, sk, src, tgt, i)] = &cs_apply_csr_impl;
    // apply_sparse_ (CSR * CSR) stays empty: SpGEMM is not part of this library
}
const Sketch& Sketch::get() {
    static const Sketch instance;
    return instance;
}
namespace {
void no_omp(Device dev, const char* who) {
    if (dev == Device::OMP) throw std::invalid_argument(std::string(who) + ": Device::OMP is not supported.");
}
}  // namespace
generate_count_sketch_values_fn Sketch::lookup_generate_count_sketch_values(Device dev, NumericType type) const {
    no_omp(dev, "generate_count_sketch");
    auto fn = generate_values_[encode_values(dev, type)];
    if (!fn) throw std::runtime_error("generate_count_sketch: no values kernel registered for this device/type.");
    return fn;
}
generate_count_sketch_idxs_fn Sketch::lookup_generate_count_sketch_idxs(Device dev, IntType type) const {
    no_omp(dev, "generate_count_sketch");
    auto fn = generate_idxs_[encode_idxs(dev, type)];
    if (!fn) throw std::runtime_error("generate_count_sketch: no index kernel registered for this device/type.");
    return fn;
}
apply_count_sketch_dense_fn Sketch::lookup_apply_count_sketch_dense(Device dev, NumericType sk, NumericType src,
                                                                    NumericType tgt, IntType idx) const {
    no_omp(dev, "apply_count_sketch");
    auto fn = apply_dense_[encode_apply(dev, sk, src, tgt, idx)];
    if (!fn) throw std::runtime_error("apply_count_sketch: nothing registered for this device/type combination (dense source).");
    return fn;
}
apply_count_sketch_sparse_fn Sketch::lookup_apply_count_sketch_sparse(Device dev, NumericType sk, NumericType src,
                                                                      NumericType tgt, IntType idx) const {
    no_omp(dev, "apply_count_sketch");
    auto fn = apply_sparse_[encode_apply(dev, sk, src, tgt, idx)];
    if (!fn) throw std::runtime_error("apply_count_sketch: nothing registered for this device/type combination (sparse source).");
    return fn;
}
}  // namespace dispatch

namespace {
Status gaussian_dense_impl(tomos::Dense& S, std::uint64_t seed) {
    device_guard(S.get_device(), "create_gaussian_sketch");
    const auto sz = S.get_size();
    check(tomos_gaussian_fill(abi_type(S.get_value_type()), static_cast<int64_t>(sz[0]),
                              static_cast<int64_t>(sz[1]), S.data(), static_cast<int64_t>(S.get_lead_dim()),
                              seed, S.stream()),
          "create_gaussian_sketch");
    return SUCCESS;
}
// Seeds of the unseeded overload: the SplitMix64 increments from a fixed
// start, one per call.
std::atomic<std::uint64_t> g_gauss_calls{0};
Status csr_sketch_impl(tomos::Sparse& S, std::uint64_t seed, int kind, const char* who) {
    device_guard(S.get_device(), who);
    const auto sz = S.get_size();
    if (S.get_nnz() != sz[0])
        throw std::invalid_argument(std::string(who) + ": nnz_per_row == 1 only; the sketch's nnz must equal its rows.");
    check(tomos_csr_sketch(kind, abi_type(S.get_value_type()), abi_index(S.get_index_type()),
                           static_cast<int64_t>(sz[0]), static_cast<int64_t>(sz[1]), S.values(), S.row_ptrs(),
                           S.col_idxs(), seed, S.stream()),
          who);
    return SUCCESS;
}
Status gaussian_sparse_impl(tomos::Sparse& S, std::uint64_t seed) {
    return csr_sketch_impl(S, seed, 1, "create_gaussian_sketch");
}
std::uint64_t next_sketch_seed() {
    const std::uint64_t k = g_gauss_calls.fetch_add(1);
    return 0x6a09e667f3bcc909ull + k * 0x9e3779b97f4a7c15ull;
}
}  // namespace

namespace dispatch {
std::size_t GaussianSketch::encode_dense(Device dev, NumericType type) {
    return static_cast<std::size_t>(dev) * num_registry_numeric_types + static_cast<std::size_t>(type);
}
GaussianSketch::GaussianSketch() {
    dense_.fill(nullptr);
    sparse_.fill(nullptr);
    for (NumericType t : {NumericType::FP64, NumericType::FP32, NumericType::FP16, NumericType::BF16}) {
        dense_[encode_dense(Device::HIP, t)] = &gaussian_dense_impl;
        for (IntType it : {IntType::INT32, IntType::INT64})
            sparse_[encode_dense(Device::HIP, t) * 2 + static_cast<std::size_t>(it)] = &gaussian_sparse_impl;
    }
}
const GaussianSketch& GaussianSketch::get() {
    static const GaussianSketch instance;
    return instance;
}
create_gaussian_sketch_dense_fn GaussianSketch::lookup_create_dense(Device dev, NumericType type) const {
    if (dev == Device::OMP) throw std::invalid_argument("create_gaussian_sketch: Device::OMP is not supported.");
    auto fn = dense_[encode_dense(dev, type)];
    if (!fn) throw std::runtime_error("create_gaussian_sketch: nothing registered for this device/type.");
    return fn;
}
}  // namespace dispatch


dispatch::create_gaussian_sketch_sparse_fn dispatch::GaussianSketch::lookup_create_sparse(Device dev, NumericType type,
                                                                                IntType index) const {
    if (dev == Device::OMP) throw std::invalid_argument("create_gaussian_sketch: Device::OMP is not supported.");
    auto fn = sparse_[encode_dense(dev, type) * 2 + static_cast<std::size_t>(index)];
    if (!fn) throw std::runtime_error("create_gaussian_sketch: nothing registered for this device/type.");
    return fn;
}

Status create_gaussian_sketch(tomos::Sparse& S, std::uint64_t seed) {
    return dispatch::GaussianSketch::get().lookup_create_sparse(S.get_device(), S.get_value_type(),
                                                                S.get_index_type())(S, seed);
}

Status create_gaussian_sketch(tomos::Sparse& S) { return create_gaussian_sketch(S, next_sketch_seed()); }

Status cs_generate_idxs_impl(tomos::Sparse& S) {
    return csr_sketch_impl(S, next_sketch_seed(), 2, "generate_count_sketch");
}

Status cs_generate_values_impl(tomos::Sparse& S) {
    return csr_sketch_impl(S, next_sketch_seed(), 3, "generate_count_sketch");
}

void cs_apply_csr_impl(tomos::Sparse& sketch, tomos::Dense& source, tomos::Dense& target) {
    const auto ss = sketch.get_size(), xs = source.get_size(), ys = target.get_size();
    if (xs[0] != ss[1] || ys[0] != ss[0] || ys[1] != xs[1])
        throw std::invalid_argument("apply_count_sketch: sizes do not match (S k x m, source m x n, target k x n).");
    if (sketch.get_device() != source.get_device() || source.get_device() != target.get_device())
        throw std::invalid_argument("apply_count_sketch: sketch, source and target must be on the same device.");
    const NumericType xt = source.get_value_type(), yt = target.get_value_type();
    check(tomos_csr_mm(abi_type(sketch.get_value_type()), abi_index(sketch.get_index_type()), 1.0,
                       static_cast<int64_t>(ss[0]), static_cast<int64_t>(ss[1]), static_cast<int64_t>(sketch.get_nnz()),
                       sketch.values(), sketch.row_ptrs(), sketch.col_idxs(), static_cast<int64_t>(xs[1]), abi_type(xt),
                       source.data(), static_cast<int64_t>(source.get_lead_dim()), 0.0, abi_type(yt), target.data(),
                       static_cast<int64_t>(target.get_lead_dim()), sketch.stream()),
          "apply_count_sketch");
}

// countsketch.cpp:15-18
void generate_count_sketch(tomos::Sparse& sketch) {
    const auto& reg = dispatch::Sketch::get();
    reg.lookup_generate_count_sketch_idxs(sketch.get_device(), sketch.get_index_type())(sketch);
    reg.lookup_generate_count_sketch_values(sketch.get_device(), sketch.get_value_type())(sketch);
}

// countsketch.cpp:20-26 (sketch_dispatch: lookup by the four types)
void apply_count_sketch(tomos::Sparse& sketch, tomos::Dense& source, tomos::Dense& target) {
    dispatch::Sketch::get().lookup_apply_count_sketch_dense(sketch.get_device(), sketch.get_value_type(),
                                                            source.get_value_type(), target.get_value_type(),
                                                            sketch.get_index_type())(sketch, source, target);
}

void apply_count_sketch(tomos::Sparse& sketch, tomos::Sparse& source, tomos::Dense& target) {
    dispatch::Sketch::get().lookup_apply_count_sketch_sparse(sketch.get_device(), sketch.get_value_type(),
                                                             source.get_value_type(), target.get_value_type(),
                                                             sketch.get_index_type())(sketch, source, target);
}

Status create_gaussian_sketch(tomos::Dense& S, std::uint64_t seed) {
    return dispatch::GaussianSketch::get().lookup_create_dense(S.get_device(), S.get_value_type())(S, seed);
}

Status create_gaussian_sketch(tomos::Dense& S) { return create_gaussian_sketch(S, next_sketch_seed()); }

CountSketch create_count_sketch(std::size_t k, std::size_t m) { return CountSketch{{k, m}}; }

void generate_count_sketch(CountSketch& sketch) {
    // The hashed sketch is fully defined by its size and the fixed seeds.
    if (sketch.size[0] == 0) throw std::invalid_argument("generate_count_sketch: zero rows");
}

// The hashed sketch is an extension outside the reference's Sparse-keyed
// registry: it calls the CountSketch kernel directly.
void apply_count_sketch(CountSketch& sketch, tomos::Dense& source, tomos::Dense& target) {
    device_guard(source.get_device(), "apply_count_sketch");
    const NumericType t = source.get_value_type();
    if (t != NumericType::FP32 && t != NumericType::FP64 && !is_packed(t))
        throw std::runtime_error("apply_count_sketch: nothing registered for this device/type.");
    cs_apply_hashed(sketch, source, target);
}

namespace {
void sketch_qr_f64(double* a, int64_t k, int64_t n, int64_t ld, stream_t st) {
    const int rc = tomos_sketch_qr_precond_f64(k, n, a, ld, st);
    if (rc == TOMOS_ERR_SINGULAR)
        throw std::runtime_error(std::string("generate_sketch_qr_precond: trtri failed -- R is singular (") +
                                 tomos_last_error() + ")");
    check(rc, "generate_sketch_qr_precond");
}
}  // namespace

// sketch_qr.cpp:68-122: FP64 sketches are factored in place; an FP32 sketch
// is widened to FP64 on the device, factored there (the FP64 Householder QR
// and R^-1), and R^-1 is rounded back into its leading n x n block.
void generate_sketch_qr_precond_impl(tomos::Dense& sketched) {
    device_guard(sketched.get_device(), "generate_sketch_qr_precond");
    if (sketched.get_device() != Device::HIP)
        throw std::runtime_error("generate_sketch_qr_precond: unsupported device.");
    const NumericType t = sketched.get_value_type();
    if (t != NumericType::FP64 && t != NumericType::FP32)
        throw std::runtime_error(
            "generate_sketch_qr_precond: unsupported NumericType (only FP64/FP32 are supported).");
    const auto s = sketched.get_size();
    if (s[0] < s[1])
        throw std::invalid_argument(
            "generate_sketch_qr_precond: sketched_mtx must have at least as many rows as columns (k >= n).");
    const int64_t k = static_cast<int64_t>(s[0]), n = static_cast<int64_t>(s[1]),
                  ld = static_cast<int64_t>(sketched.get_lead_dim());
    if (t == NumericType::FP64) {
        sketch_qr_f64(static_cast<double*>(sketched.data()), k, n, ld, sketched.stream());
        return;
    }
    if (n == 0) return;
    void* wide = nullptr;
    check(tomos_malloc(&wide, static_cast<size_t>(k) * static_cast<size_t>(n) * 8), "generate_sketch_qr_precond");
    struct Free {
        void* p;
        ~Free() { tomos_free(p); }
    } guard{wide};
    check(tomos_convert(TOMOS_F32, TOMOS_F64, k, n, sketched.data(), ld, wide, k, sketched.stream()),
          "generate_sketch_qr_precond");
    sketch_qr_f64(static_cast<double*>(wide), k, n, k, sketched.stream());
    check(tomos_convert(TOMOS_F64, TOMOS_F32, n, n, wide, k, sketched.data(), ld, sketched.stream()),
          "generate_sketch_qr_precond");
    check(tomos_stream_synchronize(sketched.stream()), "generate_sketch_qr_precond");
}

void generate_sketch_qr_precond(CountSketch& sketch, tomos::Dense& sketched) {
    (void)sketch;
    generate_sketch_qr_precond_impl(sketched);
}

void generate_sketch_qr_precond(tomos::Sparse& sketch, tomos::Dense& sketched) {
    if (sketch.get_device() != sketched.get_device())
        throw std::invalid_argument(
            "generate_sketch_qr_precond: sketch and sketched_mtx must be on the same device.");
    generate_sketch_qr_precond_impl(sketched);
}

// ======================================================================
// Front-ends (src/core/blas/blas.cpp, src/core/solver/lsqr.cpp:389-411)
// ======================================================================
Status mv_apply(Device dev, TransposeType trans, double alpha, tomos::Dense& A, tomos::Dense& x, double beta,
                tomos::Dense& y, NumericType compute, float* workspace) {
    auto fn = dispatch::Dense::get().lookup_mv(dev, A.get_value_type(), x.get_value_type(),
                                               y.get_value_type(), compute);
    fn(trans, alpha, A, x, beta, y, workspace);
    return SUCCESS;
}

Status mv_apply(Device dev, TransposeType trans, double alpha, tomos::Dense& A, tomos::Dense& x, double beta,
                tomos::Dense& y) {
    return mv_apply(dev, trans, alpha, A, x, beta, y, default_compute(A.get_value_type()));
}

std::size_t mv_workspace_size(tomos::Dense& A, TransposeType trans, NumericType compute) {
    const auto s = A.get_size();
    return tomos_gemv_scratch_size(abi_type(A.get_value_type()), tier_for(A.get_value_type(), compute),
                                   trans == TransposeType::TRANSPOSE ? 1 : 0, static_cast<int64_t>(s[0]),
                                   static_cast<int64_t>(s[1]));
}

void precond_apply(Device dev, TransposeType trans, double alpha, tomos::Dense& P, tomos::Dense& x,
                   double beta, tomos::Dense& y) {
    mv_apply(dev, trans, alpha, P, x, beta, y);
}

Status mm_apply(Device dev, TransposeType ta, TransposeType tb, double alpha, tomos::Dense& A, tomos::Dense& B,
                double beta, tomos::Dense& C, bool use_workspace) {
    auto fn = dispatch::Dense::get().lookup_mm(dev, A.get_value_type(), B.get_value_type(),
                                               C.get_value_type(), use_workspace);
    fn(ta, tb, alpha, A, B, beta, C);
    return SUCCESS;
}

template <typename value_t>
void set_values(value_t value, tomos::Dense& target) {
    const double v = static_cast<double>(value);
    if (v == 0.0) {
        check(tomos_memset(target.data(), 0, target.storage_bytes(), target.stream()), "set_values");
        return;
    }
    const NumericType t = target.get_value_type();
    if (t != NumericType::FP32 && t != NumericType::FP64)
        throw std::invalid_argument("set_values: non-zero fill needs an FP32/FP64 target");
    const std::size_t cnt = target.get_lead_dim() * target.get_size()[1];
    if (t == NumericType::FP64) {
        std::vector<double> h(cnt, v);
        target.copy_from_host(h.data(), cnt * 8);
    } else {
        std::vector<float> h(cnt, static_cast<float>(v));
        target.copy_from_host(h.data(), cnt * 4);
    }
}
template void set_values<double>(double, tomos::Dense&);
template void set_values<float>(float, tomos::Dense&);

Status split_fp32_16(tomos::Dense& source, tomos::Dense& target) {
    dispatch::Fp32_16::get().lookup_split(target.get_device())(source, target);
    return SUCCESS;
}
Status combine_fp32_16(tomos::Dense& source, tomos::Dense& target) {
    dispatch::Fp32_16::get().lookup_combine(source.get_device())(source, target);
    return SUCCESS;
}
Status mv_apply_fp32_16_pure_half(Device dev, TransposeType t, double a, tomos::Dense& A, tomos::Dense& x,
                                  double b, tomos::Dense& y) {
    dispatch::Fp32_16::get().lookup_mv_pure_half(dev)(t, a, A, x, b, y, nullptr);
    return SUCCESS;
}
Status mv_apply_fp32_16_fast(Device dev, TransposeType t, double a, tomos::Dense& A, tomos::Dense& x, double b,
                             tomos::Dense& y) {
    dispatch::Fp32_16::get().lookup_mv_fast(dev)(t, a, A, x, b, y, nullptr);
    return SUCCESS;
}
Status mv_apply_fp32_16_ext(Device dev, TransposeType t, double a, tomos::Dense& A, tomos::Dense& x, double b,
                            tomos::Dense& y) {
    dispatch::Fp32_16::get().lookup_mv_accurate(dev)(t, a, A, x, b, y, nullptr);
    return SUCCESS;
}
Status mm_apply_fp32_16_pure_half(Device dev, TransposeType ta, TransposeType tb, double a, tomos::Dense& A,
                                  tomos::Dense& B, double b, tomos::Dense& C) {
    dispatch::Fp32_16::get().lookup_mm_pure_half(dev)(ta, tb, a, A, B, b, C);
    return SUCCESS;
}
Status mm_apply_fp32_16_fast(Device dev, TransposeType ta, TransposeType tb, double a, tomos::Dense& A,
                             tomos::Dense& B, double b, tomos::Dense& C) {
    dispatch::Fp32_16::get().lookup_mm_fast(dev)(ta, tb, a, A, B, b, C);
    return SUCCESS;
}
Status mm_apply_fp32_16_ext(Device dev, TransposeType ta, TransposeType tb, double a, tomos::Dense& A,
                            tomos::Dense& B, double b, tomos::Dense& C) {
    dispatch::Fp32_16::get().lookup_mm_accurate(dev)(ta, tb, a, A, B, b, C);
    return SUCCESS;
}
Status split_fp32_bf16(tomos::Dense& source, tomos::Dense& target, bool rounded) {
    auto& r = dispatch::Fp32Bf16::get();
    (rounded ? r.lookup_split_rounded(target.get_device()) : r.lookup_split(target.get_device()))(source, target);
    return SUCCESS;
}
Status combine_fp32_bf16(tomos::Dense& source, tomos::Dense& target, bool rounded) {
    auto& r = dispatch::Fp32Bf16::get();
    (rounded ? r.lookup_combine_rounded(source.get_device()) : r.lookup_combine(source.get_device()))(source,
                                                                                                       target);
    return SUCCESS;
}
Status mv_apply_fp32_bf16_hi(Device dev, TransposeType t, double a, tomos::Dense& A, tomos::Dense& x, double b,
                             tomos::Dense& y) {
    dispatch::Fp32Bf16::get().lookup_mv_hi(dev)(t, a, A, x, b, y, nullptr);
    return SUCCESS;
}
Status mv_apply_fp32_bf16_ext(Device dev, TransposeType t, double a, tomos::Dense& A, tomos::Dense& x, double b,
                              tomos::Dense& y) {
    dispatch::Fp32Bf16::get().lookup_mv_accurate(dev)(t, a, A, x, b, y, nullptr);
    return SUCCESS;
}

Workspace create_lsqr_workspace(Context& ctx, dim<2> size, NumericType value_type, NumericType precond_type) {
    // nbytes = m*sz(v) + 3n*sz(v) + (sz(p)-1) + n*sz(p)  (lsqr.cpp:389-406)
    const std::size_t m = size[0], n = size[1];
    const std::size_t vs = numeric_type_size(value_type), ps = numeric_type_size(precond_type);
    Workspace ws(ctx.get_device(), m * vs + 3 * n * vs + (ps - 1) + n * ps);
    ws.value_type_ = value_type;
    ws.precond_type_ = precond_type;
    return ws;
}

Workspace create_lsqr_workspace(Context& ctx, dim<2> size, NumericType value_type) {
    return create_lsqr_workspace(ctx, size, value_type, value_type);
}

Status mv_apply(Device dev, TransposeType trans, double alpha, tomos::Sparse& A, tomos::Dense& x, double beta,
                tomos::Dense& y) {
    device_guard(dev, "mv_apply");
    dispatch::Sparse::get().lookup_mv(dev, A.get_value_type(), x.get_value_type(), y.get_value_type(),
                                      A.get_index_type())(trans, alpha, A, x, beta, y);
    return SUCCESS;
}

void lsqr_solve(Context context, tomos::Sparse& A, tomos::Dense& b, tomos::Dense& x, Workspace& workspace,
                double tol, int max_iters, double max_relres, int* iterations,
                std::vector<double>* residual_history, int restarts, tomos::Sparse* mtx_refine) {
    if (mtx_refine && mtx_refine->get_value_type() != b.get_value_type())
        throw std::invalid_argument(
            "lsqr_solve: mtx_refine_precision's NumericType must match b/x's (refine_type).");
    auto fn = dispatch::Lsqr::get().lookup_solve_sparse(context.get_device(), A.get_value_type(),
                                                        workspace.value_type_, b.get_value_type(),
                                                        workspace.precond_type_);
    fn(context, A, b, x, workspace, tol, max_iters, max_relres, iterations, residual_history, restarts,
       mtx_refine);
}

void lsqr_solve(Context context, tomos::Dense& A, tomos::Dense& b, tomos::Dense& x, Workspace& workspace,
                double tol, int max_iters, double max_relres, int* iterations,
                std::vector<double>* residual_history, int restarts, tomos::Dense* mtx_refine) {
    if (mtx_refine && mtx_refine->get_value_type() != b.get_value_type())
        throw std::invalid_argument(
            "lsqr_solve: mtx_refine_precision's NumericType must match b/x's (refine_type).");
    auto fn = dispatch::Lsqr::get().lookup_solve_dense(context.get_device(), A.get_value_type(),
                                                       workspace.value_type_, b.get_value_type(),
                                                       workspace.precond_type_);
    fn(context, A, b, x, workspace, tol, max_iters, max_relres, iterations, residual_history, restarts,
       mtx_refine);
}

}  // namespace tomos
