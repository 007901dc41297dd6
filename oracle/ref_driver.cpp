// TEST INFRASTRUCTURE ONLY -- never linked into, loaded by, or called from the
// product library. This is a thin extern "C" driver over the reference's own
// Eigen-free headers, compiled in place from the reference tree by
// oracle/Makefile into oracle/_ref/libapxref.so. It exists so the oracle and the
// product's host code can be checked against the reference's real code:
//
//   * graph<vec_pool_incidence> (graph/graph.h, graph/incidence_list.h),
//     block_greedy_partitioner (solver/partition/block_greedy.h),
//     baumann_kyng_partitioner (solver/partition/baumann_kyng.h), selection
//     (solver/partitioner.h), prune_and_degrees / is_degree_threshold /
//     parallel_degree_quantile (solver/partitioner_helpers.h) and
//     tree_elimination::sample_clique (solver/elimination/elimination.h) are the
//     reference's; the round loop that orchestrates them is restated below from
//     solver/factorization_impl.h:667-1338 and src/factorization.cpp:25-107,
//     because those two files include Eigen (absent from this image) and cannot
//     be compiled here;
//   * omp_sptrsv (solver/sptrsv/omp.h) -- the reference's CPU SpTRSV: compacting
//     drop, fp16 narrowing, CSC->CSR transpose, round-as-level level sets and
//     the level-set sweeps -- is the reference's, called as is;
//   * cuda_host.h's GPU host prep (L11 int32 copy, drop, fp16 narrowing) is the
//     reference's, called as is.
//
// Built single-threaded (no -fopenmp), with -DAPXCHOL_POOL_FP32 (the reference's
// CMake default, CMakeLists.txt:79) and -ffp-contract=off (ISO C++ arithmetic,
// what the reference's g++ -std=c++23 build does).
#include "apxchol/graph/graph.h"
#include "apxchol/solver/partitioner.h"
#include "apxchol/solver/partitioner_helpers.h"
#include "apxchol/solver/partition/block_greedy.h"
#include "apxchol/solver/partition/baumann_kyng.h"
#include "apxchol/solver/elimination/elimination.h"
#include "apxchol/solver/sptrsv/omp.h"
#include "apxchol/solver/sptrsv/cuda_host.h"
#include "apxchol/lowprec.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <span>
#include <vector>

using namespace apxchol;

namespace {

using ref_graph = graph<vec_pool_incidence>;

struct ref_col {
    node_index vertex = 0;
    float diag = 0.0f;
    std::vector<std::pair<node_index, float>> entries;
};

struct ref_ws {
    std::vector<weighted_neighbor> nbrs;
    std::vector<deferred_edge> ebuf;
    std::vector<std::pair<node_index, double>> exbuf;
    std::vector<node_index> first;
    std::vector<node_index> touched;
};

// Restated from solver/factorization_impl.h:164-234 (process_vertex, the
// dedup_inline branch taken by every elimination path).
void drv_process_vertex(const tree_elimination& elim, ref_graph& G, node_index v,
                        std::uint64_t run_seed, ref_col& col, ref_ws& ws) {
    static constexpr node_index npos = node_index(-1);
    auto& nbrs = ws.nbrs;
    nbrs.clear();
    ws.ebuf.clear();
    ws.exbuf.clear();
    double edge_deg = 0.0;
    if (ws.first.size() < size_t(G.n())) ws.first.assign(G.n(), npos);
    ws.touched.clear();
    for (auto [u, w] : G.neighbors(v)) {
        if (!G.is_active(u)) continue;
        if (ws.first[u] == npos) {
            ws.first[u] = static_cast<node_index>(nbrs.size());
            ws.touched.push_back(u);
            nbrs.push_back({u, w});
        } else {
            nbrs[ws.first[u]].weight += w;
        }
        edge_deg += w;
    }
    for (auto t : ws.touched) ws.first[t] = npos;
    col.entries.clear();
    if (nbrs.empty()) {
        const double d = G.excess(v);
        col.vertex = v;
        col.diag = static_cast<float>(d > 0.0 ? std::sqrt(d) : 1.0);
        return;
    }
    double total_deg = edge_deg + G.excess(v);
    if (total_deg <= 0.0) total_deg = 1.0;
    const double sqrt_deg = std::sqrt(total_deg);
    col.vertex = v;
    col.diag = static_cast<float>(sqrt_deg);
    for (const auto& [u, w] : nbrs)
        col.entries.emplace_back(u, static_cast<float>(w / sqrt_deg));
    const double ev = G.excess(v);
    if (ev > 0.0)
        for (const auto& [u, w] : nbrs) ws.exbuf.emplace_back(u, w * ev / total_deg);
    const std::uint64_t vseed = run_seed ^ ((std::uint64_t(v) + 1) * 0x9E3779B97F4A7C15ULL);
    elim.sample_clique(std::span<weighted_neighbor>(nbrs), total_deg, vseed, edge_emitter(ws.ebuf));
}

// Serial (T=1) elimination of one vertex and application of its updates:
// factorization_impl.h:458-475 (serial round path) and 568-587 (peel_one).
void drv_eliminate(const tree_elimination& elim, ref_graph& G, node_index v, std::uint64_t seed,
                   std::vector<ref_col>& cols, ref_ws& ws) {
    ref_col c;
    drv_process_vertex(elim, G, v, seed, c, ws);
    cols.push_back(std::move(c));
    for (auto [a, b, w] : ws.ebuf) G.add_edge(a, b, w);
    for (auto [u, d] : ws.exbuf) G.excess(u) += d;
    G.deactivate(v);
}

// factorization_impl.h:57-86
bool drv_selection_should_handoff(size_t selected, size_t candidates, size_t active, double min_fraction,
                                  size_t handoff, size_t omp_threshold) {
    if (selected == 0) return true;
    if (omp_threshold != 0 && selected >= omp_threshold) return false;
    if (handoff != std::numeric_limits<size_t>::max()) {
        const size_t protected_tail = handoff > std::numeric_limits<size_t>::max() / 2
                                          ? std::numeric_limits<size_t>::max() : 2 * handoff;
        if (active <= protected_tail) return false;
    }
    return selected < candidates * min_fraction;
}

// factorization_impl.h:101-112
double drv_adaptive_yield(double base, size_t active, double avg_deg, size_t handoff) {
    if (!(base > 0.0) || handoff == std::numeric_limits<size_t>::max() ||
        handoff > std::numeric_limits<size_t>::max() / 4 || active <= 4 * handoff ||
        avg_deg < static_cast<double>(handoff))
        return base;
    return std::max(base, std::min(0.15, 3.0 * base));
}

struct ref_factor {
    node_index n = 0;
    bool sddm = false;
    std::vector<edge_index> outer;
    std::vector<node_index> inner;
    std::vector<float> vals;
    std::vector<node_index> perm;
    std::vector<std::uint64_t> round_is;
    std::vector<std::uint64_t> round_active;
};

}  // namespace

extern "C" {

// The reference factorization at one thread, default factor_options except the
// seed: block_greedy (quantile 0.2, (degree,index) tie-break), yield handoff
// at min_is_fraction 0.05, residual coalesce + baumann_kyng down to 500, natural
// peel, tree_elimination. Input: FULL symmetric CSC (both triangles, rows
// sorted within each column) of a Laplacian/SDDM M-matrix.
void* ref_factorize(std::uint32_t n, const std::int64_t* colptr, const std::int32_t* rowidx,
                    const double* vals, unsigned seed) {
    // make_graph (graph/conversions.h:45-216, serial branch + excess phase).
    ref_graph G(n);
    std::vector<double> diag(n, 0.0);
    for (std::uint32_t k = 0; k < n; ++k)
        for (std::int64_t p = colptr[k]; p < colptr[k + 1]; ++p) {
            const std::uint32_t r = static_cast<std::uint32_t>(rowidx[p]);
            if (r > k) G.add_edge(r, k, -vals[p]);
            else if (r == k) diag[k] = vals[p];
        }
    for (std::uint32_t k = 0; k < n; ++k) {
        double wdeg = 0.0;
        for (std::int64_t p = colptr[k]; p < colptr[k + 1]; ++p)
            if (static_cast<std::uint32_t>(rowidx[p]) != k) wdeg -= vals[p];
        double excess = diag[k] - wdeg;
        excess = excess > diag[k] * 1e-12 ? excess : 0.0;
        G.excess(k) = excess;
    }

    // factorize_impl (factorization_impl.h:667-1338) at T=1.
    factor_options opts;
    opts.seed = seed;
    auto* F = new ref_factor;
    F->n = n;
    if (n == 0) return F;
    for (node_index v = 0; v < n; ++v)
        if (G.excess(v) > 0.0) { F->sddm = true; break; }
    const tree_elimination elim{.exact_clique_max_degree = opts.exact_clique_max_degree};
    ref_ws ws;
    std::vector<ref_col> cols;
    cols.reserve(n);
    const size_t residual_thresh = block_greedy_partitioner::residual_handoff_threshold;
    std::vector<node_index> active(n);
    std::iota(active.begin(), active.end(), node_index{0});

    selection sel;
    std::vector<node_index> pre_degrees, pre_scratch, pre_eligible, live_degrees;
    std::vector<std::array<size_t, 256>> pre_hist;
    std::vector<size_t> pre_offsets;
    double last_avg_degree = 0.0;
    size_t last_candidate_count = 0;
    block_greedy_partitioner bg;

    auto record_round = [&](size_t act, size_t is) {
        F->round_active.push_back(act);
        F->round_is.push_back(is);
    };

    while (active.size() > 1) {
        sel.reset(G.n());
        partition_context pctx{.options = opts.partition, .seed = opts.seed,
                               .omp_threshold = opts.omp_threshold, .cp = nullptr, .degrees = {}};
        const std::span<const node_index> act(active);
        const double avg_deg = prune_and_degrees(G, act, pre_degrees, opts.omp_threshold);
        const double q = opts.partition.degree_quantile;
        const double thr = q > 0.0 && q < 1.0 && act.size() > opts.omp_threshold
            ? static_cast<double>(parallel_degree_quantile(pre_degrees, act.size(), q, pre_hist))
            : is_degree_threshold(pre_degrees, act.size(), avg_deg, opts.partition, pre_scratch);
        if (live_degrees.size() < size_t(G.n())) live_degrees.resize(G.n());
        pre_eligible.clear();
        for (size_t i = 0; i < act.size(); ++i) {
            live_degrees[act[i]] = pre_degrees[i];
            if (pre_degrees[i] <= thr) pre_eligible.push_back(act[i]);
        }
        last_candidate_count = pre_eligible.size();
        last_avg_degree = avg_deg;
        pctx.degrees = live_degrees;
        bg.find_partition(G, std::span<const node_index>(pre_eligible), pctx, sel);
        const partition_result& part = sel.finalize();

        const double min_yield = drv_adaptive_yield(opts.min_is_fraction, active.size(), last_avg_degree,
                                                    residual_thresh);
        if (drv_selection_should_handoff(part.num_regions(), last_candidate_count, active.size(), min_yield,
                                         residual_thresh, opts.omp_threshold))
            break;
        record_round(active.size(), part.num_regions());
        for (node_index v : part.data) drv_eliminate(elim, G, v, opts.seed, cols, ws);
        std::erase_if(active, [&](node_index v) { return !G.is_active(v); });
    }

    // Residual coalesce (factorization_impl.h:1133-1160) -- the reference's own
    // estimator and rebuild.
    if (active.size() > residual_thresh && active.size() >= 1024) {
        const double ratio = detail::residual_coalescer<vec_pool_incidence>::estimate(G, active);
        if (ratio >= 4.0) (void)detail::residual_coalescer<vec_pool_incidence>::rebuild(G, active);
    }

    // BK residual loop (factorization_impl.h:1161-1263).
    if (active.size() > residual_thresh) {
        baumann_kyng_partitioner bk;
        if (last_avg_degree > 0.0) bk.est_avg_degree = last_avg_degree;
        size_t empty = 0;
        while (active.size() > residual_thresh) {
            sel.reset(G.n());
            partition_context pctx{.options = opts.partition, .seed = opts.seed,
                                   .omp_threshold = opts.omp_threshold, .cp = nullptr, .degrees = {}};
            bk.find_partition(G, std::span<const node_index>(active), pctx, sel);
            const partition_result& part = sel.finalize();
            if (part.num_regions() == 0) {
                if (++empty >= 64) break;
                continue;
            }
            empty = 0;
            record_round(active.size(), part.num_regions());
            for (node_index v : part.data) drv_eliminate(elim, G, v, opts.seed, cols, ws);
            std::erase_if(active, [&](node_index v) { return !G.is_active(v); });
        }
    }
    for (node_index v : active) drv_eliminate(elim, G, v, opts.seed, cols, ws);

    // build_csc / assemble_csc (src/factorization.cpp:25-107).
    F->perm.assign(n, 0);
    for (node_index i = 0; i < n; ++i) F->perm[cols[i].vertex] = i;
    std::vector<node_index> cnt(n, 1);
    for (const auto& c : cols) cnt[F->perm[c.vertex]] += static_cast<node_index>(c.entries.size());
    F->outer.assign(size_t(n) + 1, 0);
    for (node_index c = 0; c < n; ++c) F->outer[c + 1] = F->outer[c] + cnt[c];
    F->inner.resize(F->outer[n]);
    F->vals.resize(F->outer[n]);
    std::vector<std::pair<node_index, float>> ce;
    for (const auto& c : cols) {
        const node_index pc = F->perm[c.vertex];
        ce.clear();
        for (const auto& [nb, val] : c.entries) ce.emplace_back(F->perm[nb], -val);
        std::sort(ce.begin(), ce.end());
        edge_index pos = F->outer[pc];
        F->inner[pos] = pc;
        F->vals[pos] = c.diag;
        ++pos;
        for (auto [r, val] : ce) { F->inner[pos] = r; F->vals[pos] = val; ++pos; }
    }
    return F;
}

void ref_factor_sizes(void* h, std::uint64_t* out /* n, nnz, nrounds, sddm */) {
    auto* F = static_cast<ref_factor*>(h);
    out[0] = F->n;
    out[1] = F->outer.empty() ? 0 : F->outer.back();
    out[2] = F->round_is.size();
    out[3] = F->sddm ? 1 : 0;
}

void ref_factor_export(void* h, std::uint32_t* outer, std::uint32_t* inner, float* vals, std::uint32_t* perm,
                       std::uint64_t* round_is, std::uint64_t* round_active) {
    auto* F = static_cast<ref_factor*>(h);
    std::copy(F->outer.begin(), F->outer.end(), outer);
    std::copy(F->inner.begin(), F->inner.end(), inner);
    std::copy(F->vals.begin(), F->vals.end(), vals);
    std::copy(F->perm.begin(), F->perm.end(), perm);
    std::copy(F->round_is.begin(), F->round_is.end(), round_is);
    std::copy(F->round_active.begin(), F->round_active.end(), round_active);
}

void ref_factor_free(void* h) { delete static_cast<ref_factor*>(h); }

// tree_elimination::sample_clique as is. Writes up to d-1 edges (u, v, w).
std::int64_t ref_sample_clique(std::int64_t d, const std::uint32_t* verts, const double* weights, double deg,
                               std::uint64_t seed, std::uint32_t* out_u, std::uint32_t* out_v, double* out_w) {
    std::vector<weighted_neighbor> nb(static_cast<size_t>(d));
    for (std::int64_t i = 0; i < d; ++i) nb[i] = {verts[i], weights[i]};
    std::vector<deferred_edge> buf;
    tree_elimination{}.sample_clique(std::span<weighted_neighbor>(nb), deg, seed, edge_emitter(buf));
    for (size_t i = 0; i < buf.size(); ++i) { out_u[i] = buf[i].u; out_v[i] = buf[i].v; out_w[i] = buf[i].w; }
    return static_cast<std::int64_t>(buf.size());
}

std::uint16_t ref_fp16_round(float x) { return fp16_t(x).bits; }
float ref_fp16_widen(std::uint16_t h) { return fp16_t::widen_bits(h); }

// ---- omp_sptrsv (solver/sptrsv/omp.h) as is ------------------------------
struct ref_trsv {
    sparse_csc L;
    omp_sptrsv s;
};

// Factor CSC (diagonal first per column, rows ascending) of order n; m = n
// (SDDM) or n-1 (Laplacian). bounds: cumulative round IS sizes (nb entries
// starting at 0), clamped to m, as apx_cholesky::install_factor hands them
// over (preconditioner.h:238-249); nb = 0 means "no rounds" (topological
// levels). Storage (fp32/fp16) and the drop threshold come from the
// APXCHOL_SPTRSV_FP16 / APXCHOL_FACTOR_DROP environment, as in the reference.
void* ref_trsv_setup(std::uint32_t n, const std::uint32_t* outer, const std::uint32_t* inner, const float* vals,
                     std::uint32_t m, const std::uint32_t* bounds, std::int64_t nb) {
    auto* T = new ref_trsv;
    T->L.resize(n, n);
    std::copy(outer, outer + n + 1, T->L.outer_.begin());
    T->L.resizeNonZeros(outer[n]);
    std::copy(inner, inner + outer[n], T->L.inner_.begin());
    std::copy(vals, vals + outer[n], T->L.vals_.begin());
    if (nb > 0) T->s.set_round_bounds(std::vector<node_index>(bounds, bounds + nb));
    T->s.setup(T->L, m);
    return T;
}

void ref_trsv_free(void* h) { delete static_cast<ref_trsv*>(h); }

// out: stored_nnz, fp16 flag, fwd level count, bck level count, dropped
void ref_trsv_info(void* h, std::uint64_t* out) {
    auto* T = static_cast<ref_trsv*>(h);
    out[0] = T->s.stored_nnz();
    out[1] = T->s.fp16() ? 1 : 0;
    out[2] = static_cast<std::uint64_t>(T->s.num_fwd_levels());
    out[3] = static_cast<std::uint64_t>(T->s.num_bck_levels());
    out[4] = T->s.drop_stats().dropped;
}

// CSR (forward) and CSC (back) arrays. vals32 (nnz floats) are filled on the
// fp32 storage, vals16 (nnz uint16 bit patterns) on the fp16 storage; diag /
// inv_scale / scale (m floats) only on fp16.
void ref_trsv_export(void* h, std::uint32_t* csr_ptr, std::uint32_t* csr_idx, float* csr_v32,
                     std::uint16_t* csr_v16, std::uint32_t* csc_ptr, std::uint32_t* csc_idx, float* csc_v32,
                     std::uint16_t* csc_v16, float* diag, float* scale) {
    auto* T = static_cast<ref_trsv*>(h);
    const auto& s = T->s;
    std::copy(s.csr_row_ptr().begin(), s.csr_row_ptr().end(), csr_ptr);
    std::copy(s.csr_col_idx().begin(), s.csr_col_idx().end(), csr_idx);
    std::copy(s.csc_col_ptr().begin(), s.csc_col_ptr().end(), csc_ptr);
    std::copy(s.csc_row_idx().begin(), s.csc_row_idx().end(), csc_idx);
    if (s.fp16()) {
        for (size_t i = 0; i < s.csr_vals16().size(); ++i) csr_v16[i] = s.csr_vals16()[i].bits;
        for (size_t i = 0; i < s.csc_vals16().size(); ++i) csc_v16[i] = s.csc_vals16()[i].bits;
        std::copy(s.fp16_diag().begin(), s.fp16_diag().end(), diag);
        std::copy(s.col_scales().begin(), s.col_scales().end(), scale);
    } else {
        std::copy(s.csr_vals().begin(), s.csr_vals().end(), csr_v32);
        std::copy(s.csc_vals().begin(), s.csc_vals().end(), csc_v32);
    }
}

// Level sizes and per-level off-diagonal work (omp_sptrsv::level_stats).
void ref_trsv_level_stats(void* h, int fwd, std::int64_t* sizes, std::int64_t* work) {
    auto* T = static_cast<ref_trsv*>(h);
    std::vector<int> sz;
    std::vector<long long> wk;
    T->s.level_stats(fwd != 0, sz, wk);
    for (size_t l = 0; l < sz.size(); ++l) { sizes[l] = sz[l]; work[l] = wk[l]; }
}

void ref_trsv_forward(void* h, const double* x, double* y) { static_cast<ref_trsv*>(h)->s.forward_solve(x, y); }
void ref_trsv_back(void* h, const double* x, double* y) { static_cast<ref_trsv*>(h)->s.transpose_solve(x, y); }

// ---- cuda_host.h GPU host prep as is -------------------------------------
// L11 (int32 CSC), the drop with the GPU keep predicate, then (fp16) the
// per-column-scaled narrowing. Returns the stored nnz; arrays sized by the
// caller from a first call with null outputs.
std::int64_t ref_gpu_prep(std::uint32_t n, const std::uint32_t* outer, const std::uint32_t* inner, const float* vals,
                          std::int64_t m, int fp16, double rel, std::int32_t* out_ptr, std::int32_t* out_idx,
                          float* out_v32, std::uint16_t* out_v16, float* out_diag, float* out_inv_scale) {
    sparse_csc L;
    L.resize(n, n);
    std::copy(outer, outer + n + 1, L.outer_.begin());
    L.resizeNonZeros(outer[n]);
    std::copy(inner, inner + outer[n], L.inner_.begin());
    std::copy(vals, vals + outer[n], L.vals_.begin());
    auto A = cuda_host::build_L11_csc_int<float>(L, m);
    const auto s = cuda_host::column_scales(A);
    (void)cuda_host::apply_factor_drop(A, s, rel, fp16 != 0);
    if (!out_ptr) return A.nnz;
    std::copy(A.ptr.begin(), A.ptr.end(), out_ptr);
    std::copy(A.idx.get(), A.idx.get() + A.nnz, out_idx);
    if (fp16) {
        auto h16 = cuda_host::narrow_fp16_scaled(A, s);
        std::copy(h16.vals.get(), h16.vals.get() + A.nnz, out_v16);
        std::copy(h16.diag.begin(), h16.diag.end(), out_diag);
        std::copy(h16.inv_scale.begin(), h16.inv_scale.end(), out_inv_scale);
    } else {
        std::copy(A.vals.get(), A.vals.get() + A.nnz, out_v32);
    }
    return A.nnz;
}

}  // extern "C"
