// C++ API tests: the drop-in surface (Dense / Workspace / registries /
// lsqr_solve) used the way the reference's own tests use it
// (test/core/solver/test_lsqr.cpp, test/dispatch/test_*_registries.cpp,
// test/core/workspace/test_workspace.cpp, test/core/matrix/test_fp32_16.cpp).
// Runs on the GPU; prints one line per case and exits non-zero on failure.
#include <cmath>
#include <cstring>
#include <cstdio>
#include <functional>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "tomos/api.hpp"

using namespace tomos;

static int g_fail = 0;

#define EXPECT(cond)                                                                    \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            std::printf("    expectation failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__); \
            throw std::runtime_error("expectation failed");                             \
        }                                                                               \
    } while (0)

template <class E, class F>
bool throws(F&& f) {
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

static void run(const char* name, const std::function<void()>& f) {
    try {
        f();
        std::printf("PASS %s\n", name);
    } catch (const std::exception& e) {
        std::printf("FAIL %s: %s\n", name, e.what());
        ++g_fail;
    }
    std::fflush(stdout);
}

static Context ctx{Device::HIP};

template <class T>
static Dense make(NumericType t, std::size_t m, std::size_t n, const std::vector<T>& colmajor) {
    Dense d = Dense::create_strided(ctx, Layout::COL_MAJOR, t, {m, n});
    d.copy_from_host(colmajor.data(), colmajor.size() * sizeof(T));
    return d;
}

template <class T>
static std::vector<T> fetch(const Dense& d, std::size_t count) {
    std::vector<T> h(count);
    d.copy_to_host(h.data(), count * sizeof(T));
    return h;
}

static std::vector<double> rows_to_cols(const std::vector<double>& r, std::size_t m, std::size_t n) {
    std::vector<double> c(m * n);
    for (std::size_t i = 0; i < m; ++i)
        for (std::size_t j = 0; j < n; ++j) c[i + j * m] = r[i * n + j];
    return c;
}

struct Sys {
    const char* name;
    std::size_t m, n;
    std::vector<double> a_rows, b, x;
    double tol, xtol;
};

int main() {
    run("registries_are_singletons", [] {
        EXPECT(&dispatch::Lsqr::get() == &dispatch::Lsqr::get());
        EXPECT(&dispatch::Dense::get() == &dispatch::Dense::get());
        EXPECT(&dispatch::Fp32_16::get() == &dispatch::Fp32_16::get());
    });
    run("omp_device_throws", [] {
        EXPECT(throws<std::invalid_argument>([] { dispatch::Fp32_16::get().lookup_split(Device::OMP); }));
        EXPECT(throws<std::invalid_argument>([] {
            dispatch::Lsqr::get().lookup_solve_dense(Device::OMP, NumericType::FP32, NumericType::FP32,
                                                    NumericType::FP32, NumericType::FP32);
        }));
    });
    run("cpu_slots_are_empty", [] {
        // no CPU implementation is shipped: lookups throw (test_fp32_16_registries.cpp:31-34 style)
        EXPECT(throws<std::runtime_error>([] { dispatch::Fp32_16::get().lookup_mm_workspace(Device::CPU); }));
        EXPECT(throws<std::runtime_error>([] {
            dispatch::Dense::get().lookup_mv(Device::CPU, NumericType::FP32, NumericType::FP32,
                                             NumericType::FP32, NumericType::FP32);
        }));
    });
    run("unregistered_combination_throws", [] {
        EXPECT(throws<std::runtime_error>([] {
            dispatch::Dense::get().lookup_mv(Device::HIP, NumericType::FP32_16, NumericType::FP32,
                                             NumericType::FP32, NumericType::FP32_16);
        }));
        EXPECT(throws<std::runtime_error>([] {
            dispatch::Lsqr::get().lookup_solve_dense(Device::HIP, NumericType::FP32, NumericType::FP16,
                                                    NumericType::FP32, NumericType::FP32);
        }));
    });
    run("workspace_arena_formula_and_precond_check", [] {
        Workspace ws = create_lsqr_workspace(ctx, {100, 7}, NumericType::FP32, NumericType::FP64);
        EXPECT(ws.capacity_ == 100 * 4 + 3 * 7 * 4 + 7 + 7 * 8);
        Dense p32 = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32, {7, 7});
        EXPECT(throws<std::invalid_argument>([&] { add_precond(p32, ws); }));
        Dense p64 = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {7, 7});
        add_precond(p64, ws);
        EXPECT(ws.precond_kind_ == PrecondKind::Dense && ws.precond_ == &p64);
    });

    std::vector<Sys> systems = {
        {"identity3", 3, 3, {1, 0, 0, 0, 1, 0, 0, 0, 1}, {1, 2, 3}, {1, 2, 3}, 1e-10, 1e-8},
        {"diag124", 3, 3, {1, 0, 0, 0, 2, 0, 0, 0, 4}, {1, 4, 8}, {1, 2, 2}, 1e-10, 1e-8},
        {"overdetermined_consistent", 4, 2, {1, 0, 0, 1, 1, 1, 1, -1}, {1, 1, 2, 0}, {1, 1}, 1e-10, 1e-7},
        {"overdetermined_inconsistent", 4, 2, {1, 1, 1, 2, 1, 3, 1, 4}, {6, 5, 7, 10}, {3.5, 1.4}, 1e-10, 1e-6},
    };
    for (const Sys& s : systems) {
        run((std::string("lsqr_solve_fp64_") + s.name).c_str(), [&] {
            Dense A = make(NumericType::FP64, s.m, s.n, rows_to_cols(s.a_rows, s.m, s.n));
            Dense b = make(NumericType::FP64, s.m, 1, s.b);
            Dense x = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {s.n, 1});
            Workspace ws = create_lsqr_workspace(ctx, {s.m, s.n}, NumericType::FP64);
            int iters = 0;
            std::vector<double> hist;
            lsqr_solve(ctx, A, b, x, ws, s.tol, 200, s.tol, &iters, &hist);
            auto sol = fetch<double>(x, s.n);
            for (std::size_t i = 0; i < s.n; ++i) EXPECT(std::fabs(sol[i] - s.x[i]) < s.xtol);
            EXPECT(iters > 0 && hist.size() == static_cast<std::size_t>(iters));
        });
    }
    run("lsqr_solve_fp32_system", [] {
        std::vector<float> a = {1, 0, 0, 0, 2, 0, 0, 0, 3}, bb = {1, 4, 9};
        Dense A = make(NumericType::FP32, 3, 3, a);
        Dense b = make(NumericType::FP32, 3, 1, bb);
        Dense x = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32, {3, 1});
        Workspace ws = create_lsqr_workspace(ctx, {3, 3}, NumericType::FP32);
        lsqr_solve(ctx, A, b, x, ws, 1e-5, 50, 1e-5);
        auto sol = fetch<float>(x, 3);
        EXPECT(std::fabs(sol[0] - 1) < 1e-5 && std::fabs(sol[1] - 2) < 1e-5 && std::fabs(sol[2] - 3) < 1e-5);
    });
    run("lsqr_workspace_too_small_throws", [] {
        std::vector<float> a(12, 1.0f), bb(4, 1.0f);
        Dense A = make(NumericType::FP32, 4, 3, a);
        Dense b = make(NumericType::FP32, 4, 1, bb);
        Dense x = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32, {3, 1});
        Workspace ws = create_lsqr_workspace(ctx, {2, 3}, NumericType::FP32);
        EXPECT(throws<std::invalid_argument>([&] { lsqr_solve(ctx, A, b, x, ws, 1e-6, 10, 1e-6); }));
    });
    run("lsqr_refine_type_mismatch_throws", [] {
        std::vector<float> a(12, 1.0f), bb(4, 1.0f);
        Dense A = make(NumericType::FP32, 4, 3, a);
        Dense b = make(NumericType::FP32, 4, 1, bb);
        Dense x = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32, {3, 1});
        Dense Ar = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {4, 3});
        Workspace ws = create_lsqr_workspace(ctx, {4, 3}, NumericType::FP32);
        EXPECT(throws<std::invalid_argument>(
            [&] { lsqr_solve(ctx, A, b, x, ws, 1e-6, 10, 1e-6, nullptr, nullptr, 1, &Ar); }));
        EXPECT(throws<std::invalid_argument>([&] {
            ws.use_implicit_duplicate_ = true;  // needs packed A (lsqr_registries.cpp:106-112)
            lsqr_solve(ctx, A, b, x, ws, 1e-6, 10, 1e-6);
        }));
    });
    run("sketch_qr_rejects_k_lt_n", [] {
        CountSketch S = create_count_sketch(3, 10);
        Dense Y = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {3, 5});
        EXPECT(throws<std::invalid_argument>([&] { generate_sketch_qr_precond(S, Y); }));
    });
    run("sketch_preconditioned_lsqr_40x5", [] {
        // LsqrSketchPrecondTest (test_lsqr.cpp:461-523): mt19937(42), U(-1,1)
        const std::size_t m = 40, n = 5, k = 30;
        std::mt19937 gen(42);
        std::uniform_real_distribution<double> dist(-1.0, 1.0);
        std::vector<double> a(m * n), xt(n), bb(m, 0.0);
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = 0; i < m; ++i) a[i + j * m] = dist(gen);
        for (auto& v : xt) v = dist(gen);
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = 0; i < m; ++i) bb[i] += a[i + j * m] * xt[j];
        Dense A = make(NumericType::FP64, m, n, a);
        Dense b = make(NumericType::FP64, m, 1, bb);
        CountSketch S = create_count_sketch(k, m);
        generate_count_sketch(S);
        Dense sk = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {k, n});
        apply_count_sketch(S, A, sk);
        generate_sketch_qr_precond(S, sk);
        Dense P = Dense::wrap_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, n}, k, sk.data());
        Dense x = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, 1});
        Workspace ws = create_lsqr_workspace(ctx, {m, n}, NumericType::FP64);
        add_precond(P, ws);
        lsqr_solve(ctx, A, b, x, ws, 1e-10, 100, 1e-10);
        auto sol = fetch<double>(x, n);
        for (std::size_t i = 0; i < n; ++i) EXPECT(std::fabs(sol[i] - xt[i]) < 1e-6);
    });
    run("fp32_16_and_fp32_bf16_round_trip_and_tiers", [] {
        const std::size_t m = 20, n = 15;
        std::mt19937 gen(1);
        std::uniform_real_distribution<float> dist(-8.0f, 8.0f);
        std::vector<float> a(m * n), xv(n);
        for (auto& v : a) v = dist(gen);
        for (auto& v : xv) v = dist(gen);
        Dense src = make(NumericType::FP32, m, n, a);
        Dense x = make(NumericType::FP32, n, 1, xv);
        for (NumericType pt : {NumericType::FP32_16, NumericType::FP32_BF16}) {
            Dense P = Dense::create_strided(ctx, Layout::COL_MAJOR, pt, {m, n});
            Dense back = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32, {m, n});
            if (pt == NumericType::FP32_16) {
                split_fp32_16(src, P);
                combine_fp32_16(P, back);
            } else {
                split_fp32_bf16(src, P);
                combine_fp32_bf16(P, back);
            }
            auto r = fetch<float>(back, m * n);
            for (std::size_t i = 0; i < m * n; ++i) EXPECT(r[i] == a[i]);
            std::vector<double> ref(m, 0.0);
            for (std::size_t j = 0; j < n; ++j)
                for (std::size_t i = 0; i < m; ++i) ref[i] += double(a[i + j * m]) * xv[j];
            Dense y = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32, {m, 1});
            mv_apply(Device::HIP, TransposeType::NO_TRANSPOSE, 1.0, P, x, 0.0, y);  // 16-bit tier
            auto lo = fetch<float>(y, m);
            mv_apply(Device::HIP, TransposeType::NO_TRANSPOSE, 1.0, P, x, 0.0, y, NumericType::FP32);
            auto hi = fetch<float>(y, m);
            for (std::size_t i = 0; i < m; ++i) {
                EXPECT(std::fabs(lo[i] - ref[i]) <= 0.15 * std::fabs(ref[i]) + 0.15);
                EXPECT(std::fabs(hi[i] - ref[i]) <= 1e-3 * std::fabs(ref[i]) + 1e-4);
            }
        }
    });
    run("mm_apply_packed_tiers", [] {
        const std::size_t M = 33, N = 17, K = 29;
        std::mt19937 gen(3);
        std::uniform_real_distribution<float> dist(-8.0f, 8.0f);
        std::vector<float> a(M * K), bm(K * N);
        for (auto& v : a) v = dist(gen);
        for (auto& v : bm) v = dist(gen);
        Dense src = make(NumericType::FP32, M, K, a);
        Dense B = make(NumericType::FP32, K, N, bm);
        Dense P = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32_16, {M, K});
        split_fp32_16(src, P);
        std::vector<double> ref(M * N, 0.0);
        for (std::size_t j = 0; j < N; ++j)
            for (std::size_t l = 0; l < K; ++l)
                for (std::size_t i = 0; i < M; ++i) ref[i + j * M] += double(a[i + l * M]) * bm[l + j * K];
        for (bool ws : {false, true}) {
            Dense C = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32, {M, N});
            mm_apply(Device::HIP, TransposeType::NO_TRANSPOSE, TransposeType::NO_TRANSPOSE, 1.0, P, B, 0.0, C, ws);
            auto c = fetch<float>(C, M * N);
            for (std::size_t i = 0; i < M * N; ++i) EXPECT(std::fabs(c[i] - ref[i]) <= 1e-4 * std::fabs(ref[i]) + 1e-3);
        }
    });
    run("mixed_precision_lsqr_implicit_duplicates", [] {
        // FP32_BF16 A, hi-only per-iteration matvecs, exact restart residual,
        // sketch-QR preconditioner: converges to 1e-6 (runner flags
        // --implicit-duplicates --precond, lsqr_precision_runner.cu:638-656)
        const std::size_t m = 6000, n = 60, k = 240;
        std::mt19937_64 gen(5);
        std::normal_distribution<double> nd(0.0, 1.0);
        std::vector<float> a(m * n);
        std::vector<double> xt(n), bb(m, 0.0);
        for (std::size_t j = 0; j < n; ++j) {
            const double d = std::pow(10.0, -4.0 * double(j) / double(n - 1));
            for (std::size_t i = 0; i < m; ++i) a[i + j * m] = static_cast<float>(nd(gen) * d);
        }
        for (auto& v : xt) v = nd(gen);
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = 0; i < m; ++i) bb[i] += double(a[i + j * m]) * xt[j];
        Dense src = make(NumericType::FP32, m, n, a);
        Dense A = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32_BF16, {m, n});
        split_fp32_bf16(src, A);
        Dense b = make(NumericType::FP64, m, 1, bb);
        CountSketch S = create_count_sketch(k, m);
        Dense sk = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {k, n});
        apply_count_sketch(S, A, sk);
        generate_sketch_qr_precond(S, sk);
        Dense P = Dense::wrap_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, n}, k, sk.data());
        Dense x = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, 1});
        Workspace ws = create_lsqr_workspace(ctx, {m, n}, NumericType::FP32, NumericType::FP64);
        add_precond(P, ws);
        ws.use_implicit_duplicate_ = true;
        ws.mv_compute_type_ = NumericType::BF16;
        ws.mv_compute_type_set_ = true;
        int iters = 0;
        lsqr_solve(ctx, A, b, x, ws, 1e-6, 10, 1e-6, &iters, nullptr, 30);
        auto sol = fetch<double>(x, n);
        double rn = 0, bn = 0;
        for (std::size_t i = 0; i < m; ++i) {
            double r = -bb[i];
            for (std::size_t j = 0; j < n; ++j) r += double(a[i + j * m]) * sol[j];
            rn += r * r;
            bn += bb[i] * bb[i];
        }
        std::printf("    iterations=%d relres=%.3e\n", iters, std::sqrt(rn / bn));
        EXPECT(std::sqrt(rn / bn) < 1.5e-6);
    });
    run("reference_runner_sketch_sequence", [] {
        // The runner's preconditioner build and solve, call for call
        // (lsqr_precision_runner.cu:609-656): Sparse S (k = 2n rows, nnz = k),
        // generate_count_sketch(S), apply_count_sketch(S, A, sketched),
        // generate_sketch_qr_precond(S, sketched), R^-1 copied into its own
        // Dense, then lsqr_solve on packed A with implicit duplicates.
        const std::size_t m = 6000, n = 60, k = 2 * n;
        std::mt19937_64 gen(11);
        std::normal_distribution<double> nd(0.0, 1.0);
        std::vector<float> a(m * n);
        std::vector<double> a64(m * n), xt(n), bb(m, 0.0);
        for (std::size_t j = 0; j < n; ++j) {
            const double d = std::pow(10.0, -3.0 * double(j) / double(n - 1));
            for (std::size_t i = 0; i < m; ++i) {
                a[i + j * m] = static_cast<float>(nd(gen) * d);
                a64[i + j * m] = a[i + j * m];
            }
        }
        for (auto& v : xt) v = nd(gen);
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = 0; i < m; ++i) bb[i] += a64[i + j * m] * xt[j];
        Dense A64 = make(NumericType::FP64, m, n, a64);
        Dense A32 = make(NumericType::FP32, m, n, a);
        // registry shape (sketch_registries.hpp:21-66)
        const auto& reg = dispatch::Sketch::get();
        EXPECT(&reg == &dispatch::Sketch::get());
        EXPECT(throws<std::runtime_error>([&] {
            reg.lookup_apply_count_sketch_sparse(Device::HIP, NumericType::FP64, NumericType::FP64,
                                                 NumericType::FP64, IntType::INT32);
        }));
        EXPECT(throws<std::runtime_error>(
            [&] { reg.lookup_generate_count_sketch_values(Device::CPU, NumericType::FP64); }));
        EXPECT(throws<std::invalid_argument>(
            [&] { reg.lookup_generate_count_sketch_idxs(Device::OMP, IntType::INT32); }));
        EXPECT(throws<std::runtime_error>([&] {
            reg.lookup_apply_count_sketch_dense(Device::HIP, NumericType::FP64, NumericType::FP32_BF16,
                                                NumericType::FP64, IntType::INT32);
        }));

        Sparse S = Sparse::create_csr(ctx, NumericType::FP64, IntType::INT32, {k, m}, k);
        generate_count_sketch(S);
        Dense sketched = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {k, n});
        apply_count_sketch(S, A64, sketched);
        // the product is the signed row gather S A, exactly
        std::vector<int> rp(k + 1), ci(k);
        std::vector<double> sv(k);
        S.copy_to_host(rp.data(), ci.data(), sv.data());
        auto skh = fetch<double>(sketched, k * n);
        double gerr = 0;
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = 0; i < k; ++i) gerr = std::max(gerr, std::fabs(skh[i + j * k] - sv[i] * a64[ci[i] + j * m]));
        EXPECT(gerr == 0.0);
        // FP32 build type (the runner builds at FP32 unless --precond-type fp64):
        // the same S on the FP32 source into an FP32 target, factored in FP64
        Dense sk32 = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32, {k, n});
        apply_count_sketch(S, A32, sk32);
        generate_sketch_qr_precond(S, sketched);  // leading n x n block <- R^-1
        generate_sketch_qr_precond(S, sk32);
        auto r64 = fetch<double>(sketched, k * n);
        auto r32 = fetch<float>(sk32, k * n);
        double dmax = 0, rmax = 0;
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = 0; i < n; ++i) {
                dmax = std::max(dmax, std::fabs(double(r32[i + j * k]) - r64[i + j * k]));
                rmax = std::max(rmax, std::fabs(r64[i + j * k]));
                if (i > j) EXPECT(r64[i + j * k] == 0.0 && r32[i + j * k] == 0.0f);
            }
        std::printf("    R^-1 from the FP32 sketch vs the FP64 one: max |diff| / max |R^-1| = %.2e\n", dmax / rmax);
        EXPECT(dmax <= 1e-5 * rmax);
        Dense precond = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, n});
        std::vector<double> ph(n * n);
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = 0; i < n; ++i) ph[i + j * n] = r64[i + j * k];
        precond.copy_from_host(ph.data(), ph.size() * 8);

        Dense A = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32_BF16, {m, n});
        split_fp32_bf16(A32, A);
        Dense b = make(NumericType::FP64, m, 1, bb);
        Dense x = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, 1});
        Workspace ws = create_lsqr_workspace(ctx, {m, n}, NumericType::FP32, NumericType::FP64);
        add_precond(precond, ws);
        ws.use_implicit_duplicate_ = true;
        ws.mv_compute_type_ = NumericType::BF16;
        ws.mv_compute_type_set_ = true;
        EXPECT(ws.single_pass_);  // the single-pass kernels are the default
        ws.confirm_true_residual_ = true;
        int iters = 0;
        std::vector<double> hist;
        lsqr_solve(ctx, A, b, x, ws, 1e-6, 10, 1e-6, &iters, &hist, 40);
        auto sol = fetch<double>(x, n);
        double rn = 0, bn = 0;
        for (std::size_t i = 0; i < m; ++i) {
            double r = -bb[i];
            for (std::size_t j = 0; j < n; ++j) r += a64[i + j * m] * sol[j];
            rn += r * r;
            bn += bb[i] * bb[i];
        }
        std::printf("    iterations=%d history=%zu true relres=%.3e\n", iters, hist.size(), std::sqrt(rn / bn));
        EXPECT(std::sqrt(rn / bn) <= 1e-6);
        // the two-pass path of the same solve reaches the same tolerance
        Dense x2 = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, 1});
        ws.single_pass_ = false;
        lsqr_solve(ctx, A, b, x2, ws, 1e-6, 10, 1e-6, &iters, nullptr, 40);
        auto sol2 = fetch<double>(x2, n);
        double dx = 0, xn = 0;
        for (std::size_t j = 0; j < n; ++j) {
            dx = std::max(dx, std::fabs(sol2[j] - sol[j]));
            xn = std::max(xn, std::fabs(sol[j]));
        }
        std::printf("    two-pass iterations=%d, max |x_two - x_single| / max |x| = %.2e\n", iters, dx / xn);
        EXPECT(dx <= 1e-2 * xn);
        Sparse S2 = Sparse::create_csr(ctx, NumericType::FP64, IntType::INT32, {3, m}, 3);
        Dense small = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {3, n});
        EXPECT(throws<std::invalid_argument>([&] { generate_sketch_qr_precond(S2, small); }));
    });
    run("native_bf16_lsqr_explicit_duplicate", [] {
        // A stored as bf16 (round to nearest even) plus a copy at the refine
        // precision (FP64, as b and x) passed as mtx_refine (runner flag
        // --explicit-duplicates, lsqr_precision_runner.cu:559-580, 656): the
        // restarts read the copy, so the solve reaches 1e-6 against the
        // unrounded matrix; without it the
        // residual stalls near bf16's rounding of A.
        const std::size_t m = 6000, n = 60, k = 240;
        std::mt19937_64 gen(7);
        std::normal_distribution<double> nd(0.0, 1.0);
        std::vector<float> a(m * n);
        std::vector<uint16_t> a16(m * n);
        std::vector<double> xt(n), bb(m, 0.0);
        for (std::size_t j = 0; j < n; ++j) {
            const double d = std::pow(10.0, -3.0 * double(j) / double(n - 1));
            for (std::size_t i = 0; i < m; ++i) a[i + j * m] = static_cast<float>(nd(gen) * d);
        }
        for (std::size_t i = 0; i < m * n; ++i) {
            uint32_t u;
            std::memcpy(&u, &a[i], 4);
            a16[i] = static_cast<uint16_t>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
        }
        for (auto& v : xt) v = nd(gen);
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = 0; i < m; ++i) bb[i] += double(a[i + j * m]) * xt[j];
        Dense A32 = make(NumericType::FP32, m, n, a);
        Dense A = make(NumericType::BF16, m, n, a16);
        Dense A64 = make(NumericType::FP64, m, n, std::vector<double>(a.begin(), a.end()));
        Dense b = make(NumericType::FP64, m, 1, bb);
        CountSketch S = create_count_sketch(k, m);
        Dense sk = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {k, n});
        apply_count_sketch(S, A32, sk);
        generate_sketch_qr_precond(S, sk);
        Dense P = Dense::wrap_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, n}, k, sk.data());
        auto relres = [&](const Dense& x) {
            auto sol = fetch<double>(x, n);
            double rn = 0, bn = 0;
            for (std::size_t i = 0; i < m; ++i) {
                double r = -bb[i];
                for (std::size_t j = 0; j < n; ++j) r += double(a[i + j * m]) * sol[j];
                rn += r * r;
                bn += bb[i] * bb[i];
            }
            return std::sqrt(rn / bn);
        };
        double res[2];
        for (int dup = 0; dup < 2; ++dup) {
            Dense x = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, 1});
            Workspace ws = create_lsqr_workspace(ctx, {m, n}, NumericType::FP32, NumericType::FP64);
            add_precond(P, ws);
            int iters = 0;
            lsqr_solve(ctx, A, b, x, ws, 1e-6, 8, 1e-6, &iters, nullptr, 30, dup ? &A64 : nullptr);
            res[dup] = relres(x);
            std::printf("    duplicate=%d iterations=%d relres=%.3e\n", dup, iters, res[dup]);
        }
        EXPECT(res[1] < 1.5e-6);
        EXPECT(res[0] > res[1] && res[0] < 4.0 * std::ldexp(1.0, -8));
        Dense x = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, 1});
        Workspace ws = create_lsqr_workspace(ctx, {m, n}, NumericType::FP32, NumericType::FP64);
        Dense small = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {m - 1, n});
        EXPECT(throws<std::invalid_argument>(
            [&] { lsqr_solve(ctx, A, b, x, ws, 1e-6, 8, 1e-6, nullptr, nullptr, 1, &small); }));
        EXPECT(throws<std::invalid_argument>(  // must be at b/x's precision
            [&] { lsqr_solve(ctx, A, b, x, ws, 1e-6, 8, 1e-6, nullptr, nullptr, 1, &A32); }));
    });
    run("gaussian_sketch_registry_and_preconditioner", [] {
        // create_gaussian_sketch (sketch.hpp:11-13): seeded values repeat,
        // unseeded calls differ, 16-bit types are registered, packed are not
        EXPECT(throws<std::runtime_error>(
            [] { dispatch::GaussianSketch::get().lookup_create_dense(Device::HIP, NumericType::FP32_16); }));
        EXPECT(throws<std::runtime_error>(
            [] { dispatch::GaussianSketch::get().lookup_create_dense(Device::CPU, NumericType::FP32); }));
        Dense h = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::BF16, {5, 3});
        create_gaussian_sketch(h);
        const std::size_t m = 6000, n = 60, k = 240;
        Dense S1 = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32, {k, m});
        Dense S2 = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32, {k, m});
        create_gaussian_sketch(S1, 77);
        create_gaussian_sketch(S2, 77);
        auto v1 = fetch<float>(S1, k * m), v2 = fetch<float>(S2, k * m);
        EXPECT(v1 == v2);
        create_gaussian_sketch(S2);
        auto v3 = fetch<float>(S2, k * m);
        create_gaussian_sketch(S2);
        auto v4 = fetch<float>(S2, k * m);
        EXPECT(v3 != v4 && v3 != v1);
        double mean = 0, var = 0;
        for (float v : v1) mean += v;
        mean /= double(v1.size());
        for (float v : v1) var += (v - mean) * (v - mean);
        var /= double(v1.size());
        EXPECT(std::fabs(mean) < 0.005 && std::fabs(var - 1.0) < 0.01);
        // the sketch as a preconditioner: S A by mm_apply, R from its QR
        std::mt19937_64 gen(11);
        std::normal_distribution<double> nd(0.0, 1.0);
        std::vector<float> a(m * n);
        std::vector<double> xt(n), bb(m, 0.0);
        for (std::size_t j = 0; j < n; ++j) {
            const double d = std::pow(10.0, -4.0 * double(j) / double(n - 1));
            for (std::size_t i = 0; i < m; ++i) a[i + j * m] = static_cast<float>(nd(gen) * d);
        }
        for (auto& v : xt) v = nd(gen);
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = 0; i < m; ++i) bb[i] += double(a[i + j * m]) * xt[j];
        Dense A = make(NumericType::FP32, m, n, a);
        Dense Y = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP32, {k, n});
        mm_apply(Device::HIP, TransposeType::NO_TRANSPOSE, TransposeType::NO_TRANSPOSE, 1.0, S1, A, 0.0, Y);
        auto y = fetch<float>(Y, k * n);
        Dense sk = make(NumericType::FP64, k, n, std::vector<double>(y.begin(), y.end()));
        CountSketch unused = create_count_sketch(k, m);
        generate_sketch_qr_precond(unused, sk);
        Dense P = Dense::wrap_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, n}, k, sk.data());
        Dense b = make(NumericType::FP64, m, 1, bb);
        int its[2] = {0, 0};
        for (int pre = 0; pre < 2; ++pre) {
            Dense x = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, 1});
            Workspace ws = create_lsqr_workspace(ctx, {m, n}, NumericType::FP32, NumericType::FP64);
            if (pre) add_precond(P, ws);
            lsqr_solve(ctx, A, b, x, ws, 1e-5, 2000, 1e-5, &its[pre], nullptr, 1);
        }
        std::printf("    iterations: plain %d, Gaussian-sketch preconditioned %d\n", its[0], its[1]);
        EXPECT(its[1] > 0 && its[1] <= 40 && its[1] * 5 < its[0]);
    });
    // ------------------------------------------------------------ CSR
    auto make_csr = [](NumericType vt, std::size_t m, std::size_t n, const std::vector<int>& rp,
                       const std::vector<int>& ci, const std::vector<double>& v) {
        Sparse A = Sparse::create_csr(ctx, vt, IntType::INT32, {m, n}, ci.size());
        std::vector<unsigned char> vb(v.size() * numeric_type_size(vt));
        for (std::size_t k = 0; k < v.size(); ++k) {
            if (vt == NumericType::FP64) std::memcpy(&vb[k * 8], &v[k], 8);
            else {
                const float f = static_cast<float>(v[k]);
                if (vt == NumericType::FP32) std::memcpy(&vb[k * 4], &f, 4);
                else {
                    uint32_t u;
                    std::memcpy(&u, &f, 4);
                    uint16_t h = vt == NumericType::BF16 ? static_cast<uint16_t>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16)
                                                         : static_cast<uint16_t>(0);
                    if (vt == NumericType::FP16) {  // the test values are small integers: exact in half
                        const int e = f == 0.0f ? 0 : static_cast<int>((u >> 23) & 0xFF) - 127 + 15;
                        h = f == 0.0f ? 0 : static_cast<uint16_t>(((u >> 16) & 0x8000u) | (e << 10) | ((u >> 13) & 0x3FFu));
                    }
                    std::memcpy(&vb[k * 2], &h, 2);
                }
            }
        }
        A.copy_from_host(rp.data(), ci.data(), vb.data());
        return A;
    };
    run("sparse_lsqr_known_answers", [&] {
        // SparseDiagonalSystem / SparseOverdeterminedConsistent (test_lsqr.cpp:375-440)
        Sparse D = make_csr(NumericType::FP64, 3, 3, {0, 1, 2, 3}, {0, 1, 2}, {1, 2, 4});
        Dense b = make(NumericType::FP64, 3, 1, std::vector<double>{1, 4, 8});
        Dense x = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {3, 1});
        Workspace ws = create_lsqr_workspace(ctx, {3, 3}, NumericType::FP64);
        lsqr_solve(ctx, D, b, x, ws, 1e-10, 50, 1e-10);
        auto s1 = fetch<double>(x, 3);
        EXPECT(std::fabs(s1[0] - 1) < 1e-8 && std::fabs(s1[1] - 2) < 1e-8 && std::fabs(s1[2] - 2) < 1e-8);
        for (NumericType mv : {NumericType::FP64, NumericType::FP32, NumericType::FP16, NumericType::BF16}) {
            Sparse A = make_csr(mv, 4, 2, {0, 1, 2, 4, 6}, {0, 1, 0, 1, 0, 1}, {1, 1, 1, 1, 1, -1});
            Dense b2 = make(NumericType::FP64, 4, 1, std::vector<double>{1, 1, 2, 0});
            Dense x2 = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {2, 1});
            Dense P = make(NumericType::FP64, 2, 2, std::vector<double>{0.9, 0, 0, 0.9});
            Workspace ws2 = create_lsqr_workspace(ctx, {4, 2}, NumericType::FP32, NumericType::FP64);
            add_precond(P, ws2);
            lsqr_solve(ctx, A, b2, x2, ws2, 1e-8, 100, 1e-8);
            auto s2 = fetch<double>(x2, 2);
            EXPECT(std::fabs(s2[0] - 1) < 1e-6 && std::fabs(s2[1] - 1) < 1e-6);
        }
    });
    run("sparse_mv_apply_and_registry", [&] {
        Sparse A = make_csr(NumericType::FP32, 4, 2, {0, 1, 2, 4, 6}, {0, 1, 0, 1, 0, 1}, {1, 1, 1, 1, 1, -1});
        Dense x = make(NumericType::FP64, 2, 1, std::vector<double>{2, 3});
        Dense y = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {4, 1});
        mv_apply(Device::HIP, TransposeType::NO_TRANSPOSE, 1.0, A, x, 0.0, y);
        EXPECT((fetch<double>(y, 4) == std::vector<double>{2, 3, 5, -1}));
        Dense u = make(NumericType::FP32, 4, 1, std::vector<float>{1, 1, 1, 1});
        Dense t = make(NumericType::FP32, 2, 1, std::vector<float>{10, 10});
        mv_apply(Device::HIP, TransposeType::TRANSPOSE, 2.0, A, u, 1.0, t);  // 2 A^T 1 + t = (16, 12)
        EXPECT((fetch<float>(t, 2) == std::vector<float>{16, 12}));
        EXPECT(throws<std::runtime_error>([] {
            dispatch::Sparse::get().lookup_mv(Device::HIP, NumericType::FP32_16, NumericType::FP32,
                                              NumericType::FP32, IntType::INT32);
        }));
        EXPECT(throws<std::runtime_error>([] {
            dispatch::Lsqr::get().lookup_solve_sparse(Device::HIP, NumericType::FP32_BF16, NumericType::FP32,
                                                      NumericType::FP32, NumericType::FP32);
        }));
        Sparse bad = make_csr(NumericType::FP64, 2, 2, {0, 1, 2}, {0, 5}, {1, 1});
        Dense v = make(NumericType::FP64, 2, 1, std::vector<double>{1, 1});
        Dense w = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {2, 1});
        EXPECT(throws<std::invalid_argument>(
            [&] { mv_apply(Device::HIP, TransposeType::TRANSPOSE, 1.0, bad, v, 0.0, w); }));
    });
    run("csr_sketches_and_sketch_preconditioned_lsqr", [&] {
        // LsqrSketchPrecondTest with the CSR sketch (test_lsqr.cpp:461-523)
        const std::size_t m = 40, n = 5, k = 30;
        std::mt19937 gen(42);
        std::uniform_real_distribution<double> dist(-1.0, 1.0);
        std::vector<double> a(m * n), xt(n), bb(m, 0.0);
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = 0; i < m; ++i) a[i + j * m] = dist(gen);
        for (auto& v : xt) v = dist(gen);
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = 0; i < m; ++i) bb[i] += a[i + j * m] * xt[j];
        Dense A = make(NumericType::FP64, m, n, a);
        Dense b = make(NumericType::FP64, m, 1, bb);
        Sparse S = Sparse::create_csr(ctx, NumericType::FP64, IntType::INT32, {k, m}, k);
        generate_count_sketch(S);
        std::vector<int> rp(k + 1), ci(k);
        std::vector<double> sv(k);
        S.copy_to_host(rp.data(), ci.data(), sv.data());
        bool ok = true;
        for (std::size_t i = 0; i <= k; ++i) ok = ok && rp[i] == static_cast<int>(i);
        for (std::size_t i = 0; i < k; ++i) ok = ok && ci[i] >= 0 && ci[i] < static_cast<int>(m) && std::fabs(sv[i]) == 1.0;
        EXPECT(ok);
        Dense sk = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {k, n});
        apply_count_sketch(S, A, sk);
        auto skh = fetch<double>(sk, k * n);
        double maxerr = 0;
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = 0; i < k; ++i)
                maxerr = std::max(maxerr, std::fabs(skh[i + j * k] - sv[i] * a[ci[i] + j * m]));
        EXPECT(maxerr == 0.0);
        // the Gaussian CSR form: same shape, N(0,1) values, seeded
        Sparse G1 = Sparse::create_csr(ctx, NumericType::FP32, IntType::INT64, {k, m}, k);
        Sparse G2 = Sparse::create_csr(ctx, NumericType::FP32, IntType::INT64, {k, m}, k);
        create_gaussian_sketch(G1, 5);
        create_gaussian_sketch(G2, 5);
        std::vector<float> g1(k), g2(k);
        G1.copy_to_host(nullptr, nullptr, g1.data());
        G2.copy_to_host(nullptr, nullptr, g2.data());
        EXPECT(g1 == g2 && g1[0] != g1[1]);
        Sparse wrong = Sparse::create_csr(ctx, NumericType::FP32, IntType::INT32, {k, m}, k + 1);
        EXPECT(throws<std::invalid_argument>([&] { create_gaussian_sketch(wrong); }));
        // preconditioner from a full-rank sketch: the CountSketch of the reference test (k=30 rows)
        CountSketch H = create_count_sketch(k, m);
        Dense hs = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {k, n});
        apply_count_sketch(H, A, hs);
        generate_sketch_qr_precond(H, hs);
        Dense P = Dense::wrap_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, n}, k, hs.data());
        std::vector<int> rpa(m + 1);
        std::vector<int> cia;
        std::vector<double> va;
        for (std::size_t i = 0; i < m; ++i) {
            rpa[i] = static_cast<int>(cia.size());
            for (std::size_t j = 0; j < n; ++j) {
                cia.push_back(static_cast<int>(j));
                va.push_back(a[i + j * m]);
            }
        }
        rpa[m] = static_cast<int>(cia.size());
        Sparse As = make_csr(NumericType::FP64, m, n, rpa, cia, va);
        Dense x = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {n, 1});
        Workspace ws = create_lsqr_workspace(ctx, {m, n}, NumericType::FP64);
        add_precond(P, ws);
        int iters = 0;
        lsqr_solve(ctx, As, b, x, ws, 1e-10, 100, 1e-10, &iters);
        auto sol = fetch<double>(x, n);
        for (std::size_t i = 0; i < n; ++i) EXPECT(std::fabs(sol[i] - xt[i]) < 1e-6);
        std::printf("    sparse sketch-preconditioned iterations=%d\n", iters);
        // explicit duplicate must share A's pattern
        Sparse other = make_csr(NumericType::FP64, 4, 2, {0, 1, 2, 4, 6}, {0, 1, 0, 1, 1, 0}, {1, 1, 1, 1, 1, -1});
        Sparse A16 = make_csr(NumericType::FP16, 4, 2, {0, 1, 2, 4, 6}, {0, 1, 0, 1, 0, 1}, {1, 1, 1, 1, 1, -1});
        Sparse A64 = make_csr(NumericType::FP64, 4, 2, {0, 1, 2, 4, 6}, {0, 1, 0, 1, 0, 1}, {1, 1, 1, 1, 1, -1});
        Dense b4 = make(NumericType::FP64, 4, 1, std::vector<double>{1, 1, 2, 0});
        Dense x4 = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {2, 1});
        Workspace ws4 = create_lsqr_workspace(ctx, {4, 2}, NumericType::FP64);
        EXPECT(throws<std::invalid_argument>(
            [&] { lsqr_solve(ctx, A16, b4, x4, ws4, 1e-10, 20, 1e-10, nullptr, nullptr, 2, &other); }));
        lsqr_solve(ctx, A16, b4, x4, ws4, 1e-12, 20, 1e-12, nullptr, nullptr, 3, &A64);
        auto s4 = fetch<double>(x4, 2);
        EXPECT(std::fabs(s4[0] - 1) < 1e-10 && std::fabs(s4[1] - 1) < 1e-10);
    });
    std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "OK", g_fail);
    return g_fail ? 1 : 0;
}
