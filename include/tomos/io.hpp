// MatrixMarket input/output for the LSQR runner (the reference reads its
// test problems through src/io/mmio.c, see
// benchmarks/solvers/cuda/lsqr_precision_runner.cu:101-125).  Dense
// matrices come back column-major in FP64 whatever the file's layout.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace tomos {
namespace io {

struct MatrixMarket {
    int64_t rows = 0, cols = 0;
    std::vector<double> values;  // rows * cols, column-major
    bool coordinate = false;     // file was a (sparse) coordinate file
    int64_t entries = 0;         // nonzeros listed (coordinate) or rows * cols (array)
};

// Reads "%%MatrixMarket matrix array|coordinate real|integer general|symmetric"
// (array: column-major values; coordinate: 1-based "i j v" triplets, summed
// when repeated, mirrored for symmetric).  Throws std::runtime_error with
// the path and the reason on any malformed or unsupported file.
MatrixMarket read_matrix_market(const std::string& path);

// Writes an array file: banner, "rows cols", then one value per line in
// column-major order with 17 significant digits.
void write_matrix_market_array(const std::string& path, int64_t rows, int64_t cols, const double* values);

}  // namespace io
}  // namespace tomos
