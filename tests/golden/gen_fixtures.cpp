// Writes tests/golden/lsqr_known_answers.json: the known-answer systems of
// the reference's LSQR tests (test/core/solver/test_lsqr.cpp) as data.
// The random 40x5 system of LsqrSketchPrecondTest (:461-523) is rebuilt
// with the same standard-library engine and distribution it names
// (std::mt19937(42), std::uniform_real_distribution<double>(-1, 1), A filled
// column by column, then x_true).  Build: g++ -O2 -o gen gen_fixtures.cpp
#include <cstdio>
#include <random>
#include <string>
#include <vector>

static void emit_vec(FILE* f, const std::vector<double>& v) {
    std::fputc('[', f);
    for (size_t i = 0; i < v.size(); ++i) std::fprintf(f, "%s%.17g", i ? ", " : "", v[i]);
    std::fputc(']', f);
}

struct Sys {
    std::string name, source;
    size_t m, n;
    std::vector<double> a_rows;  // row-major m x n
    std::vector<double> b, x_expect;
    double x_tol, relres_max, tol;
    int max_iters;
};

int main(int argc, char** argv) {
    const char* out = argc > 1 ? argv[1] : "lsqr_known_answers.json";
    std::vector<Sys> sys;
    sys.push_back({"identity3", "test_lsqr.cpp:138-152", 3, 3, {1, 0, 0, 0, 1, 0, 0, 0, 1},
                   {1, 2, 3}, {1, 2, 3}, 1e-8, 1e-8, 1e-10, 50});
    sys.push_back({"diag124", "test_lsqr.cpp:154-168", 3, 3, {1, 0, 0, 0, 2, 0, 0, 0, 4},
                   {1, 4, 8}, {1, 2, 2}, 1e-8, 1e-8, 1e-10, 50});
    sys.push_back({"overdetermined_consistent", "test_lsqr.cpp:171-185", 4, 2,
                   {1, 0, 0, 1, 1, 1, 1, -1}, {1, 1, 2, 0}, {1, 1}, 1e-7, 1e-8, 1e-10, 100});
    sys.push_back({"overdetermined_inconsistent", "test_lsqr.cpp:187-200", 4, 2,
                   {1, 1, 1, 2, 1, 3, 1, 4}, {6, 5, 7, 10}, {3.5, 1.4}, 1e-6, 1.0, 1e-10, 200});
    {
        Sys s{"well_conditioned_6x3", "test_lsqr.cpp:203-234", 6, 3,
              {4, 1, 0, 0, 3, 1, 1, 0, 5, 2, 1, 1, 0, 2, 3, 1, 1, 2}, {}, {1, -1, 2},
              1e-7, 1e-8, 1e-10, 200};
        s.b.assign(6, 0.0);
        for (size_t i = 0; i < 6; ++i)
            for (size_t j = 0; j < 3; ++j) s.b[i] += s.a_rows[i * 3 + j] * s.x_expect[j];
        sys.push_back(s);
    }
    sys.push_back({"fp32_diag123", "test_lsqr.cpp:348-374", 3, 3, {1, 0, 0, 0, 2, 0, 0, 0, 3},
                   {1, 4, 9}, {1, 2, 3}, 1e-5, 1e-5, 1e-5, 50});

    FILE* f = std::fopen(out, "w");
    std::fprintf(f, "{\n  \"systems\": [\n");
    for (size_t k = 0; k < sys.size(); ++k) {
        const Sys& s = sys[k];
        std::fprintf(f, "    {\"name\": \"%s\", \"source\": \"%s\", \"m\": %zu, \"n\": %zu, \"a_rows\": ",
                     s.name.c_str(), s.source.c_str(), s.m, s.n);
        emit_vec(f, s.a_rows);
        std::fprintf(f, ", \"b\": ");
        emit_vec(f, s.b);
        std::fprintf(f, ", \"x_expect\": ");
        emit_vec(f, s.x_expect);
        std::fprintf(f, ", \"x_tol\": %g, \"relres_max\": %g, \"tol\": %g, \"max_iters\": %d}%s\n",
                     s.x_tol, s.relres_max, s.tol, s.max_iters, k + 1 < sys.size() ? "," : "");
    }
    // LsqrSketchPrecondTest: m=40, n=5, k=30
    const size_t m = 40, n = 5;
    std::mt19937 gen(42);
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    std::vector<double> a_colmajor(m * n);
    for (size_t j = 0; j < n; ++j)
        for (size_t i = 0; i < m; ++i) a_colmajor[i + j * m] = dist(gen);
    std::vector<double> x_true(n);
    for (auto& v : x_true) v = dist(gen);
    std::fprintf(f, "  ],\n  \"sketch_precond\": {\"source\": \"test_lsqr.cpp:461-523\", \"m\": 40, \"n\": 5, \"k\": 30, \"a_colmajor\": ");
    emit_vec(f, a_colmajor);
    std::fprintf(f, ", \"x_true\": ");
    emit_vec(f, x_true);
    std::fprintf(f, ", \"x_tol\": 1e-6, \"relres_max\": 1e-8, \"tol\": 1e-10, \"max_iters\": 100}\n}\n");
    std::fclose(f);
    return 0;
}
