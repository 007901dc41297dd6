// Host preparation of the factor for the GPU sweeps and of the PCG operator.
#pragma once
#include "factor.hpp"

#include <cstdint>
#include <vector>

namespace apxchol_amd {

// Default relative threshold of the compacting factor drop
// (solver/sptrsv/factor_drop.h:70).
inline constexpr double kFactorDropRelDefault = 1e-4;

// IEEE binary16 narrowing, round-to-nearest-even with IEEE subnormals,
// overflow to infinity and NaN kept (lowprec.h:127-175 states the contract).
std::uint16_t fp16_round(float x);
float fp16_widen(std::uint16_t h);

// What the compacting drop did (solver/sptrsv/factor_drop.h:101-108).
struct drop_stats {
    double rel = 0.0;
    std::uint64_t nnz_factor = 0, nnz_stored = 0, dropped = 0;
};

// The factor's L11 block (m = n for SDDM, n-1 for a Laplacian: the last row and
// column are dropped, solver/preconditioner.h:228-231) after the drop, as the
// two sweeps read it. Row lists exclude the diagonal, which sits in `diag`.
//   forward sweep (L y = x):   row i of L,   entries (j < i, L_ij), j ascending
//   back sweep    (L^T z = y): column j of L, entries (i > j, L_ij), i descending
// fp16 storage (per-column scaled, lowprec.h / cuda_host.h:16-34): values are
// fp16(L_ij / s_j), diag[j] = fp32(L_jj / s_j) + the column's rounding
// residual, inv_scale[j] = fp32(1 / s_j); fp32 storage: values L_ij, diag
// L_jj, inv_scale empty.
struct trsv_factor {
    node_t m = 0;
    bool fp16 = false;
    drop_stats drop;
    std::vector<std::uint32_t> f_ptr, b_ptr;   // m+1
    std::vector<std::int32_t> f_idx, b_idx;    // off-diagonal count
    std::vector<float> f_val, b_val;           // fp32 storage
    std::vector<std::uint16_t> f_h, b_h;       // fp16 storage
    std::vector<float> diag;                   // m
    std::vector<float> inv_scale;              // m (fp16 only)
    // Level boundaries in factor order (round-as-level, omp.h:870-945):
    // level l = rows [level_ptr[l], level_ptr[l+1]); elimination rounds first,
    // then one level per residual-peel column.
    std::vector<std::uint32_t> level_ptr;
    std::uint64_t offdiag() const { return f_idx.size(); }
};

// rel < 0 reads APXCHOL_FACTOR_DROP like the reference (unset = 1e-4, <= 0 off).
trsv_factor prepare_trsv(const factorization& F, bool fp16, double rel = -1.0);

// Renumbers the rows of every level by the original vertex order (a level is
// an independent set of the elimination, so any order inside it is the same
// factor): position q moves to Q[q], levels keep their ranges. Rewrites T's
// row / column lists (entries re-sorted: forward rows ascending, back
// columns descending), diag and inv_scale, and returns Q (m entries). The
// vertex at factor position q is the v with perm[v] == q.
// With by_depth the levels are first replaced by the topological depth of
// the forward sweep's dependency graph (level of a row = 1 + the deepest
// level it reads; omp.h:865-1038 offers the same choice), which merges rounds
// and peel columns that do not depend on each other; the back sweep walks the
// same levels in reverse.
std::vector<std::uint32_t> relabel_within_levels(trsv_factor& T, const std::vector<node_t>& perm,
                                                 bool by_depth = false);

// P A P^T as a full-symmetric CSR with rows' columns ascending
// (solver/pcg_cuda.h:304-404, solve.cpp:320-476); values stored fp32 when every
// one round-trips fp32 exactly.
struct permuted_operator {
    std::int64_t n = 0;
    std::vector<std::int64_t> ptr;
    std::vector<std::int32_t> idx;
    std::vector<double> val;
    std::vector<float> val32;
    bool fp32_exact = false;
};
permuted_operator build_permuted_operator(const csc_view& A, const std::vector<node_t>& perm);

// Sliced-ELL form of the operator (device.hpp's sell_desc): rows with more
// than long_len entries are listed in longrows; the others are packed 64 to a
// slice in the order `rows` gives (a permutation of the n rows; with sigma > 1
// stably sorted by decreasing length inside windows of sigma of them) and a
// row map records it.
struct sell_layout {
    std::vector<std::uint32_t> soff;
    std::vector<std::int32_t> idx;
    std::vector<double> val;
    std::vector<float> val32;
    std::vector<std::int32_t> rowmap;
    std::vector<std::int32_t> longrows;
    std::uint64_t slots = 0;              // padded entries of the slices
};
// Padded entries of the unsorted slicing of `rows` (empty: 0..n-1) without
// the rows longer than long_len.
std::uint64_t sell_slots(const permuted_operator& op, std::int64_t long_len, const std::vector<std::int32_t>& rows);
sell_layout build_sell(const permuted_operator& op, std::int64_t long_len, std::int64_t sigma,
                       const std::vector<std::int32_t>& rows);

}  // namespace apxchol_amd
