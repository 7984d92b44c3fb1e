// Command-line LSQR runner over the C++ API: reads A, x_true and b from
// MatrixMarket files, solves min |A x - b| with the requested storage and
// precisions (optionally with the CountSketch + QR right preconditioner),
// and reports iterations, solve time, final residual and solution error in
// the same "==>" lines as the reference's runner
// (benchmarks/solvers/cuda/lsqr_precision_runner.cu:457-718, options
// :264-420).
//
// Differences: the device is the GPU only (--device cpu is rejected: the
// CPU path of this build is the oracle used by the tests); the
// preconditioner's sketch is the hashed CountSketch with the reference's
// fixed seeds (the reference runner's sketch is time-seeded) and, for
// native 16-bit storage, is taken from an FP32 copy of A; the final
// residual is computed on the host in FP64 from the file's values.
// --check-inputs reads and validates the files and stops before any GPU
// work; --two-pass turns off the single-pass kernels.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "tomos/api.hpp"
#include "tomos/io.hpp"
#include "tomos_hip.h"

using namespace tomos;

namespace {

struct Args {
    std::string matrix, x_true, rhs, history;
    std::string device = "gpu";
    std::string mv_type = "fp32";
    bool mv_type_set = false;
    std::string value_type = "fp32", refine_type = "fp32", precond_type;
    double tol = 1e-8, max_relres = 1e-8;
    int max_iters = 200, restarts = 1;
    bool precond = false, normalize = false, explicit_dup = false, implicit_dup = false;
    bool check_inputs = false, two_pass = false;
};

void usage(const char* prog) {
    std::fprintf(stderr,
                 "Usage: %s --matrix A.mtx --x-true x.mtx --rhs b.mtx\n"
                 "          [--device gpu] [--mv-type fp64|fp32|fp16|bf16|fp32_16|fp32_bf16] [--value-type fp64|fp32]\n"
                 "          [--refine-type fp64|fp32] [--precond-type fp64|fp32]\n"
                 "          [--tol TOL] [--max-iters N] [--max-relres R] [--restarts N]\n"
                 "          [--residual-history PATH] [--precond] [--normalize-columns]\n"
                 "          [--implicit-duplicates] [--explicit-duplicates] [--check-inputs] [--two-pass]\n",
                 prog);
}

std::string lower(std::string s) {
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

NumericType vector_type(const std::string& raw, const char* what) {
    const std::string s = lower(raw);
    if (s == "fp64" || s == "double") return NumericType::FP64;
    if (s == "fp32" || s == "float" || s == "single") return NumericType::FP32;
    throw std::invalid_argument(std::string(what) + " '" + raw + "': this build supports fp64 and fp32");
}

NumericType storage_type(const std::string& raw) {
    const std::string s = lower(raw);
    if (s == "fp64" || s == "double") return NumericType::FP64;
    if (s == "fp32" || s == "float" || s == "single") return NumericType::FP32;
    if (s == "fp32_16") return NumericType::FP32_16;
    if (s == "fp32_bf16") return NumericType::FP32_BF16;
    if (s == "fp16" || s == "half") return NumericType::FP16;
    if (s == "bf16") return NumericType::BF16;
    throw std::invalid_argument("unknown --mv-type '" + raw + "'");
}

// FP32 -> IEEE binary16 bits, round to nearest even (subnormals, overflow
// to infinity, NaN kept quiet).
uint16_t to_half_bits(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7FFFFFFFu;
    if (ax >= 0x7F800000u) return static_cast<uint16_t>(sign | (ax > 0x7F800000u ? 0x7E00u : 0x7C00u));
    if (ax >= 0x477FF000u) return static_cast<uint16_t>(sign | 0x7C00u);  // rounds past 65504
    if (ax <= 0x33000000u) return static_cast<uint16_t>(sign);            // <= 2^-25: to zero (even)
    const int e = static_cast<int>(ax >> 23);
    const uint32_t mant = (ax & 0x7FFFFFu) | 0x800000u;
    uint32_t h, rem, tie;
    if (e < 113) {  // half subnormal: mant * 2^(e-126) units of 2^-24
        const int shift = 126 - e;
        h = mant >> shift;
        rem = mant & ((1u << shift) - 1u);
        tie = 1u << (shift - 1);
    } else {
        h = (static_cast<uint32_t>(e - 112) << 10) | ((mant >> 13) & 0x3FFu);
        rem = mant & 0x1FFFu;
        tie = 0x1000u;
    }
    if (rem > tie || (rem == tie && (h & 1u))) ++h;
    return static_cast<uint16_t>(sign | h);
}

// FP32 -> bf16 bits, round to nearest even (finite inputs; NaN kept NaN).
uint16_t to_bf16_bits(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    if ((x & 0x7FFFFFFFu) > 0x7F800000u) return static_cast<uint16_t>((x >> 16) | 0x40u);
    return static_cast<uint16_t>((static_cast<uint64_t>(x) + 0x7FFFu + ((x >> 16) & 1u)) >> 16);
}

Args parse(int argc, char** argv) {
    Args a;
    bool have_a = false, have_x = false, have_b = false;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        auto next = [&](const char* name) -> std::string {
            if (i + 1 >= argc) throw std::invalid_argument(std::string(name) + " needs a value");
            return argv[++i];
        };
        if (arg == "--matrix") { a.matrix = next("--matrix"); have_a = true; }
        else if (arg == "--x-true") { a.x_true = next("--x-true"); have_x = true; }
        else if (arg == "--rhs") { a.rhs = next("--rhs"); have_b = true; }
        else if (arg == "--device") a.device = lower(next("--device"));
        else if (arg == "--mv-type") { a.mv_type = next("--mv-type"); a.mv_type_set = true; }
        else if (arg == "--value-type") a.value_type = next("--value-type");
        else if (arg == "--refine-type") a.refine_type = next("--refine-type");
        else if (arg == "--precond-type") a.precond_type = next("--precond-type");
        else if (arg == "--tol") a.tol = std::stod(next("--tol"));
        else if (arg == "--max-iters") a.max_iters = std::stoi(next("--max-iters"));
        else if (arg == "--max-relres") a.max_relres = std::stod(next("--max-relres"));
        else if (arg == "--restarts") a.restarts = std::stoi(next("--restarts"));
        else if (arg == "--residual-history") a.history = next("--residual-history");
        else if (arg == "--precond") a.precond = true;
        else if (arg == "--normalize-columns") a.normalize = true;
        else if (arg == "--explicit-duplicates") a.explicit_dup = true;
        else if (arg == "--implicit-duplicates") a.implicit_dup = true;
        else if (arg == "--check-inputs") a.check_inputs = true;
        else if (arg == "--two-pass") a.two_pass = true;
        else if (arg == "-h" || arg == "--help") { usage(argv[0]); std::exit(0); }
        else throw std::invalid_argument("unrecognized argument '" + arg + "'");
    }
    if (!have_a || !have_x || !have_b) throw std::invalid_argument("--matrix, --x-true and --rhs are required");
    if (a.explicit_dup && a.implicit_dup)
        throw std::invalid_argument("--explicit-duplicates and --implicit-duplicates are mutually exclusive");
    if (a.precond_type.empty()) a.precond_type = a.value_type;
    return a;
}

std::vector<double> read_vector(const std::string& path, int64_t len, const char* what) {
    io::MatrixMarket v = io::read_matrix_market(path);
    if (!((v.cols == 1 && v.rows == len) || (v.rows == 1 && v.cols == len)))
        throw std::runtime_error(std::string(what) + " '" + path + "' is " + std::to_string(v.rows) + " x " +
                                 std::to_string(v.cols) + ", expected a vector of length " + std::to_string(len));
    return std::move(v.values);
}

double rel_l2(const std::vector<double>& a, const std::vector<double>& b) {
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return den > 0.0 ? std::sqrt(num / den) : std::sqrt(num);
}

template <class T>
Dense upload(Context& ctx, NumericType t, int64_t m, int64_t n, const std::vector<double>& v) {
    std::vector<T> h(v.begin(), v.end());
    Dense d = Dense::create_strided(ctx, Layout::COL_MAJOR, t, {static_cast<size_t>(m), static_cast<size_t>(n)});
    d.copy_from_host(h.data(), h.size() * sizeof(T));
    return d;
}

Dense upload_as(Context& ctx, NumericType t, int64_t m, int64_t n, const std::vector<double>& v) {
    return t == NumericType::FP64 ? upload<double>(ctx, t, m, n, v) : upload<float>(ctx, t, m, n, v);
}

std::vector<double> download(const Dense& d, int64_t count) {
    std::vector<double> out(static_cast<size_t>(count));
    if (d.get_value_type() == NumericType::FP64) {
        d.copy_to_host(out.data(), out.size() * sizeof(double));
    } else {
        std::vector<float> h(out.size());
        d.copy_to_host(h.data(), h.size() * sizeof(float));
        std::copy(h.begin(), h.end(), out.begin());
    }
    return out;
}

int run(const Args& args) {
    std::printf("==> Reading %s / %s / %s\n", args.matrix.c_str(), args.x_true.c_str(), args.rhs.c_str());
    std::fflush(stdout);
    io::MatrixMarket A = io::read_matrix_market(args.matrix);
    const int64_t m = A.rows, n = A.cols;
    if (m < 1 || n < 1) throw std::runtime_error("A is empty");
    std::vector<double> xt = read_vector(args.x_true, n, "x_true");
    std::vector<double> b = read_vector(args.rhs, m, "rhs");
    std::printf("==> A: %lld x %lld (%s, %lld entries)\n", static_cast<long long>(m), static_cast<long long>(n),
                A.coordinate ? "coordinate" : "array", static_cast<long long>(A.entries));
    const NumericType value = vector_type(args.value_type, "--value-type");
    const NumericType refine = vector_type(args.refine_type, "--refine-type");
    const NumericType pre_t = vector_type(args.precond_type, "--precond-type");
    NumericType store = storage_type(args.mv_type);
    NumericType compute = store == NumericType::FP32_16 ? NumericType::FP16
                          : store == NumericType::FP32_BF16 ? NumericType::BF16 : store;
    if (args.implicit_dup) {
        // A stored once packed; --mv-type then names the per-iteration tier
        if (args.mv_type_set && store != NumericType::FP32_16 && store != NumericType::FP32_BF16) {
            const std::string s = lower(args.mv_type);
            if (s != "fp32" && s != "float")
                throw std::invalid_argument("--implicit-duplicates stores A as fp32_16 or fp32_bf16; --mv-type " +
                                            args.mv_type + " does not fit");
        }
        store = store == NumericType::FP32_BF16 ? NumericType::FP32_BF16 : NumericType::FP32_16;
        compute = store == NumericType::FP32_16 ? NumericType::FP16 : NumericType::BF16;
    }
    if (args.device != "gpu" && args.device != "hip" && args.device != "cuda")
        throw std::invalid_argument("--device " + args.device +
                                    ": this build runs LSQR on the GPU (the CPU path is the test oracle)");
    std::printf("==> mv_storage=%s mv_compute=%s value_type=%s refine_type=%s precond=%s precond_type=%s "
                "normalize_columns=%s explicit_duplicates=%s implicit_duplicates=%s single_pass=%s\n",
                numeric_type_name(store), numeric_type_name(compute), numeric_type_name(value),
                numeric_type_name(refine), args.precond ? "yes" : "no", numeric_type_name(pre_t),
                args.normalize ? "yes" : "no", args.explicit_dup ? "yes" : "no", args.implicit_dup ? "yes" : "no",
                args.two_pass ? "no" : "yes");
    std::printf("==> max_iters=%d (per restart) restarts=%d (ceiling: %d iterations)\n", args.max_iters,
                args.restarts, args.max_iters * args.restarts);
    if (args.check_inputs) {
        // position-weighted sums pin the parsed values and their layout
        double sa = 0.0, wa = 0.0, sx = 0.0, sb = 0.0;
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) {
                const double v = A.values[i + j * m];
                sa += v;
                wa += v * static_cast<double>((i + 1) + 7 * (j + 1));
            }
        for (double v : xt) sx += v;
        for (double v : b) sb += v;
        std::printf("==> checksums: sum_A=%.17g weighted_A=%.17g sum_x=%.17g sum_b=%.17g\n", sa, wa, sx, sb);
        std::printf("==> inputs valid\n");
        return 0;
    }

    // column equilibration in FP64 (A' = A D^-1, solve A' z = b, x = D^-1 z)
    std::vector<double> Aw = A.values, colnorm(static_cast<size_t>(n), 1.0);
    if (args.normalize) {
        for (int64_t j = 0; j < n; ++j) {
            double s = 0.0;
            for (int64_t i = 0; i < m; ++i) s += Aw[i + j * m] * Aw[i + j * m];
            colnorm[j] = s > 0.0 ? std::sqrt(s) : 1.0;
            for (int64_t i = 0; i < m; ++i) Aw[i + j * m] /= colnorm[j];
        }
        std::printf("==> Column-normalized A (unit-norm columns; un-scaled after solving)\n");
    }

    Context ctx(Device::HIP);
    const bool native16 = store == NumericType::FP16 || store == NumericType::BF16;
    Dense Ad = [&]() {
        if (store == NumericType::FP64 || store == NumericType::FP32) return upload_as(ctx, store, m, n, Aw);
        if (native16) {
            std::vector<uint16_t> h(Aw.size());
            for (size_t k = 0; k < h.size(); ++k)
                h[k] = store == NumericType::FP16 ? to_half_bits(static_cast<float>(Aw[k]))
                                                  : to_bf16_bits(static_cast<float>(Aw[k]));
            Dense d = Dense::create_strided(ctx, Layout::COL_MAJOR, store,
                                            {static_cast<size_t>(m), static_cast<size_t>(n)});
            d.copy_from_host(h.data(), h.size() * sizeof(uint16_t));
            return d;
        }
        Dense a32 = upload<float>(ctx, NumericType::FP32, m, n, Aw);
        Dense p = Dense::create_strided(ctx, Layout::COL_MAJOR, store, {static_cast<size_t>(m), static_cast<size_t>(n)});
        if (store == NumericType::FP32_16) split_fp32_16(a32, p);
        else split_fp32_bf16(a32, p);
        return p;
    }();
    Dense bd = upload_as(ctx, refine, m, 1, b);
    Dense xd = upload_as(ctx, refine, n, 1, std::vector<double>(static_cast<size_t>(n), 0.0));

    Workspace ws = args.precond ? create_lsqr_workspace(ctx, {static_cast<size_t>(m), static_cast<size_t>(n)}, value, pre_t)
                                : create_lsqr_workspace(ctx, {static_cast<size_t>(m), static_cast<size_t>(n)}, value);
    std::vector<Dense> keep;
    keep.reserve(2);  // a_refine points into it
    // explicit duplicate at the refine precision: the restart residuals read it
    const bool dup = args.explicit_dup && store != refine;
    if (args.explicit_dup && !dup) std::printf("==> --explicit-duplicates has no effect here (mv_type == refine_type)\n");
    if (dup) {
        std::printf("==> Building explicit %s-precision duplicate of A (%lld x %lld)...\n", numeric_type_name(refine),
                    static_cast<long long>(m), static_cast<long long>(n));
        keep.push_back(upload_as(ctx, refine, m, n, Aw));
    }
    Dense* a_refine = dup ? &keep.back() : nullptr;
    if (args.precond) {
        const size_t k = 2 * static_cast<size_t>(n);
        std::printf("==> Building count-sketch + QR preconditioner (k=%zu = 2n rows, R^-1 at %s)\n", k,
                    numeric_type_name(pre_t));
        CountSketch S = create_count_sketch(k, static_cast<size_t>(m));
        generate_count_sketch(S);
        Dense sk = Dense::create_strided(ctx, Layout::COL_MAJOR, NumericType::FP64, {k, static_cast<size_t>(n)});
        if (native16) {  // the sketch reads an FP32 copy of A
            Dense a32 = upload<float>(ctx, NumericType::FP32, m, n, Aw);
            apply_count_sketch(S, a32, sk);
        } else {
            apply_count_sketch(S, Ad, sk);
        }
        generate_sketch_qr_precond(S, sk);
        std::vector<double> full = download(sk, static_cast<int64_t>(k) * n);
        std::vector<double> rinv(static_cast<size_t>(n * n));
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i) rinv[i + j * n] = full[i + j * static_cast<int64_t>(k)];
        keep.push_back(upload_as(ctx, pre_t, n, n, rinv));
        add_precond(keep.back(), ws);
    }
    ws.use_implicit_duplicate_ = args.implicit_dup;
    if (args.implicit_dup) {
        ws.mv_compute_type_ = compute;
        ws.mv_compute_type_set_ = true;
    }
    ws.single_pass_ = !args.two_pass;

    int iterations = 0;
    std::vector<double> history;
    const bool want_history = !args.history.empty();
    tomos_device_synchronize();
    const auto t0 = std::chrono::steady_clock::now();
    lsqr_solve(ctx, Ad, bd, xd, ws, args.tol, args.max_iters, args.max_relres, &iterations,
               want_history ? &history : nullptr, args.restarts, a_refine);
    tomos_device_synchronize();
    const auto t1 = std::chrono::steady_clock::now();
    const double solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();

    if (want_history) {
        FILE* hf = std::fopen(args.history.c_str(), "w");
        if (!hf) throw std::runtime_error("could not open '" + args.history + "' for writing");
        for (size_t i = 0; i < history.size(); ++i) std::fprintf(hf, "%zu %.16e\n", i + 1, history[i]);
        std::fclose(hf);
        std::printf("==> residual history (%zu entries) written to: %s\n", history.size(), args.history.c_str());
    }

    std::vector<double> x = download(xd, n);
    for (int64_t j = 0; j < n; ++j) x[j] /= colnorm[j];
    std::vector<double> r(static_cast<size_t>(m), 0.0);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) r[i] += A.values[i + j * m] * x[j];
    const double residual = rel_l2(r, b);
    const double sol_error = rel_l2(x, xt);
    std::printf("==> iterations:        %d\n", iterations);
    std::printf("==> solve time:        %.3f ms\n", solve_ms);
    std::printf("==> final residual:    %.6e  (||A*x-b|| / ||b||)\n", residual);
    std::printf("==> solution error:    %.6e  (||x-x_true|| / ||x_true||)\n", sol_error);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(parse(argc, argv));
    } catch (const std::invalid_argument& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        usage(argv[0]);
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
}
