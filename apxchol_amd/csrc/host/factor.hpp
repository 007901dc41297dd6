// Host-side types of the apxchol MI355X solve path: the input operator view,
// the approximate-Cholesky factorization, and the options that drive it.
#pragma once
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace apxchol_amd {

using node_t = std::uint32_t;   // vertex / row ids
using off_t64 = std::uint64_t;  // cumulative offsets into nnz-sized arrays

// A symmetric sparse operator in compressed-sparse-column form with BOTH
// triangles stored (what Eigen::SparseMatrix<double> holds for the reference's
// apxchol::solve / cpu_solver inputs, include/apxchol/solver/solve.h:136-138).
// Row indices need not be sorted within a column.
struct csc_view {
    std::int64_t n = 0;
    const std::int64_t* colptr = nullptr;  // n+1
    const std::int32_t* rowidx = nullptr;  // colptr[n]
    const double* vals = nullptr;          // colptr[n]
};

// Selection knobs of the partitioner (solver/factor_options.h:22-54).
struct partition_options {
    double degree_quantile = 0.2;
    double degree_multiplier = 2.0;
    bool degree_tiebreak = true;
};

// solver/factor_options.h:56-156, the subset the block_greedy + tree path reads.
struct factor_options {
    unsigned seed = 42;
    partition_options partition;
    double min_is_fraction = 0.05;
    std::size_t omp_threshold = 2000;
    // SIZE_MAX = defer to the partitioner (block_greedy declares 500).
    std::size_t parallel_residual_threshold = SIZE_MAX;
    // Host threads of the factorization (0: APXCHOL_AMD_FACTOR_THREADS, else
    // the CPUs this process may use, capped by OMP_NUM_THREADS and 16) and the
    // smallest round run in parallel (0: APXCHOL_AMD_FACTOR_PAR_MIN, else 2048). The factor does not depend on
    // either.
    int threads = 0;
    std::size_t parallel_min = 0;
};

// Per-round statistics (solver/factorization.h:45-56).
struct round_stats {
    std::size_t active = 0;
    std::size_t is_size = 0;
    double avg_deg = 0.0;
};

// The approximate Cholesky factor P^T L L^T P ~= A (solver/factorization.h:22-56):
// L lower triangular in CSC, diagonal FIRST in every column, rows ascending;
// perm[v] = elimination position of original vertex v.
struct factorization {
    node_t n = 0;
    std::vector<off_t64> outer;   // n+1
    std::vector<node_t> inner;    // nnz
    std::vector<float> vals;      // nnz (fp32, as factor_value_t)
    std::vector<node_t> perm;     // n
    bool sddm = false;
    std::int64_t lumped_offdiag = 0;
    std::vector<round_stats> rounds;
    int threads = 1;  // host threads the factorization ran on (the factor does not depend on it)
    off_t64 nnz() const { return outer.empty() ? 0 : outer.back(); }
};

// Operator contract + M-matrix lumping (operator_class.h:1-212): throws
// std::invalid_argument on a non-square / asymmetric / non-positive-diagonal
// operator; returns the matrix the preconditioner is built from (positive
// off-diagonal pairs moved onto the diagonal) in `lumped` when any exist.
struct operator_check {
    bool needs_lumping = false;
    std::int64_t lumped_entries = 0;
};
operator_check check_operator(const csc_view& A);

// Factorize (src/factorization.cpp:125-175 with vec_pool storage and the
// block_greedy partitioner): the factor is a pure function of (A, seed), the
// reference's one-thread factor, whatever opts.threads is.
factorization factorize(const csc_view& A, const factor_options& opts = {});

}  // namespace apxchol_amd
