// MatrixMarket reader/writer (format: NIST MatrixMarket exchange format,
// "array" and "coordinate" real/integer matrices).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "tomos/io.hpp"

namespace tomos {
namespace io {

namespace {

std::string lower(std::string s) {
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

[[noreturn]] void fail(const std::string& path, const std::string& why) {
    throw std::runtime_error("'" + path + "': " + why);
}

// Next line that is neither empty nor a '%' comment.
bool next_data_line(std::istream& in, std::string& line) {
    while (std::getline(in, line)) {
        size_t p = line.find_first_not_of(" \t\r");
        if (p == std::string::npos || line[p] == '%') continue;
        return true;
    }
    return false;
}

double parse_number(const std::string& path, const std::string& tok) {
    errno = 0;
    char* end = nullptr;
    const double v = std::strtod(tok.c_str(), &end);
    if (end == tok.c_str() || *end != '\0' || errno == ERANGE)
        fail(path, "bad number '" + tok + "'");
    return v;
}

int64_t parse_index(const std::string& path, const std::string& tok) {
    char* end = nullptr;
    const long long v = std::strtoll(tok.c_str(), &end, 10);
    if (end == tok.c_str() || *end != '\0') fail(path, "bad integer '" + tok + "'");
    return static_cast<int64_t>(v);
}

}  // namespace

MatrixMarket read_matrix_market(const std::string& path) {
    std::ifstream in(path);
    if (!in) fail(path, "could not open");
    std::string banner;
    if (!std::getline(in, banner)) fail(path, "empty file");
    std::istringstream bs(banner);
    std::string tag, object, format, field, symmetry;
    bs >> tag >> object >> format >> field >> symmetry;
    if (tag != "%%MatrixMarket") fail(path, "not a MatrixMarket file (no %%MatrixMarket banner)");
    object = lower(object);
    format = lower(format);
    field = lower(field);
    symmetry = lower(symmetry);
    if (object != "matrix") fail(path, "object '" + object + "' is not supported (matrix only)");
    if (format != "array" && format != "coordinate") fail(path, "format '" + format + "' is not supported");
    if (field != "real" && field != "integer" && field != "double")
        fail(path, "field '" + field + "' is not supported (real or integer)");
    if (symmetry != "general" && symmetry != "symmetric")
        fail(path, "symmetry '" + symmetry + "' is not supported (general or symmetric)");
    const bool sym = symmetry == "symmetric";

    MatrixMarket mm;
    mm.coordinate = format == "coordinate";
    std::string line;
    if (!next_data_line(in, line)) fail(path, "missing size line");
    std::istringstream ss(line);
    std::string t1, t2, t3;
    ss >> t1 >> t2;
    mm.rows = parse_index(path, t1);
    mm.cols = parse_index(path, t2);
    int64_t nnz = 0;
    if (mm.coordinate) {
        ss >> t3;
        nnz = parse_index(path, t3);
    }
    if (mm.rows < 0 || mm.cols < 0 || nnz < 0) fail(path, "negative size");
    if (sym && mm.rows != mm.cols) fail(path, "symmetric matrix is not square");
    mm.values.assign(static_cast<size_t>(mm.rows) * static_cast<size_t>(mm.cols), 0.0);

    if (!mm.coordinate) {
        // column-major; symmetric arrays list the lower triangle only
        int64_t i = 0, j = 0;
        const int64_t expect = sym ? mm.rows * (mm.rows + 1) / 2 : mm.rows * mm.cols;
        for (int64_t k = 0; k < expect; ++k) {
            if (!next_data_line(in, line)) fail(path, "truncated: " + std::to_string(k) + " of " +
                                                          std::to_string(expect) + " values");
            std::istringstream ls(line);
            std::string tok;
            ls >> tok;
            const double v = parse_number(path, tok);
            mm.values[static_cast<size_t>(i + j * mm.rows)] = v;
            if (sym) mm.values[static_cast<size_t>(j + i * mm.rows)] = v;
            if (++i == mm.rows) {
                ++j;
                i = sym ? j : 0;
            }
        }
        mm.entries = expect;
    } else {
        for (int64_t k = 0; k < nnz; ++k) {
            if (!next_data_line(in, line)) fail(path, "truncated: " + std::to_string(k) + " of " +
                                                          std::to_string(nnz) + " entries");
            std::istringstream ls(line);
            std::string ti, tj, tv;
            ls >> ti >> tj >> tv;
            const int64_t i = parse_index(path, ti) - 1, j = parse_index(path, tj) - 1;
            if (i < 0 || i >= mm.rows || j < 0 || j >= mm.cols)
                fail(path, "entry " + std::to_string(k + 1) + " out of range");
            const double v = parse_number(path, tv);
            mm.values[static_cast<size_t>(i + j * mm.rows)] += v;
            if (sym && i != j) mm.values[static_cast<size_t>(j + i * mm.rows)] += v;
        }
        mm.entries = nnz;
    }
    return mm;
}

void write_matrix_market_array(const std::string& path, int64_t rows, int64_t cols, const double* values) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) fail(path, "could not open for writing");
    std::fprintf(f, "%%%%MatrixMarket matrix array real general\n%lld %lld\n", static_cast<long long>(rows),
                 static_cast<long long>(cols));
    const size_t count = static_cast<size_t>(rows) * static_cast<size_t>(cols);
    for (size_t k = 0; k < count; ++k) std::fprintf(f, "%.16e\n", values[k]);
    if (std::fclose(f) != 0) fail(path, "write failed");
}

}  // namespace io
}  // namespace tomos
