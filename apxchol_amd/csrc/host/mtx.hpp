// MatrixMarket files at the boundary, and the reading of an input matrix as
// an assembled Laplacian/SDDM operator or as a graph adjacency matrix -- the
// reference's CLI input layer (src/mtx_input.{h,cpp}; src/main.cpp:14-98, whose
// reader is the fast_matrix_market library). Host code only.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "factor.hpp"

namespace apxchol_amd {

// An owned CSC matrix (row indices ascending within a column, no duplicates).
struct csc_matrix {
    std::int64_t rows = 0, cols = 0;
    std::vector<std::int64_t> colptr;
    std::vector<std::int32_t> rowidx;
    std::vector<double> vals;
    csc_view view() const { return {cols, colptr.data(), rowidx.data(), vals.data()}; }
};

struct mtx_header {
    bool coordinate = true;   // false: `array` (dense, column-major)
    bool pattern = false;     // `pattern` field: every stored entry is 1
    bool symmetric = false;   // `symmetric`: one triangle stored, mirrored on read
};

// Reads a `matrix coordinate` file (real, integer or pattern; general or
// symmetric). Symmetric files get their off-diagonal entries mirrored;
// duplicate entries are summed. Throws std::runtime_error on a malformed file.
csc_matrix read_mtx(const std::string& path, mtx_header* hdr = nullptr);
// Reads a dense vector: `matrix array` with one column, or a coordinate n x 1.
std::vector<double> read_mtx_vector(const std::string& path);
// Writes a coordinate file (general, or symmetric: the lower triangle) and a
// dense `array` vector, both with 17 significant digits (round-trip exact).
void write_mtx(const std::string& path, const csc_matrix& A, bool symmetric);
void write_mtx_vector(const std::string& path, const double* x, std::int64_t n);

// How to read the matrix (mtx_input.h:36-40).
enum class input_kind : int { automatic = 0, laplacian = 1, adjacency = 2 };

// One pass of structural facts (mtx_input.h:46-73). Column sums stand in for
// row sums; the sum slack is 1e-10 * max(|diagonal|, 1).
struct input_scan {
    std::int64_t n = 0;
    std::int64_t offdiag_nnz = 0, offdiag_pos = 0, offdiag_neg = 0;
    std::int64_t upper_nnz = 0, lower_nnz = 0;
    std::int64_t diag_pos = 0;
    std::int64_t deficient_rows = 0;
    double worst_deficit = 0.0;
    std::int64_t worst_row = -1;
    std::int64_t excess_rows = 0;
};
input_scan scan_input(const csc_view& A);

// Resolves `requested` against the scan (mtx_input.cpp:62-147): automatic
// picks laplacian when every off-diagonal is <= 0 (or there are none),
// adjacency when every one is >= 0, and for mixed signs laplacian only if
// every row has a positive diagonal; otherwise, and for a one-sided matrix
// whatever was requested, throws std::runtime_error. `reason` says why.
input_kind resolve_input_kind(input_kind requested, const input_scan& s, bool pattern_field, std::string& reason);

// L = D - A from a weighted adjacency matrix (mtx_input.cpp:149-175): edge
// weights |a_ij|, self-loops dropped, D the weighted degree, every vertex
// keeps an explicit diagonal entry (0 for an isolated one).
csc_matrix adjacency_to_laplacian(const csc_view& A);

// One line naming the reading and the operator's class, and the diagnosis of
// a solve that did not converge (empty when nothing suspicious was found).
std::string describe_input(input_kind kind, const input_scan& s, const std::string& reason);
std::string diagnose_failure(input_kind kind, const input_scan& s);

}  // namespace apxchol_amd
