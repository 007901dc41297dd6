// Host preparation of the factor for the GPU sweeps (the reference's
// omp_sptrsv::setup_impl, solver/sptrsv/omp.h:505-861, and cuda_sptrsv::setup,
// solver/sptrsv/cuda.h:233-459 / cuda_host.h:62-232, restated) and of the
// permuted PCG operator (solver/pcg_cuda.h:304-404).
#include "trsv_prep.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace apxchol_amd {

std::uint16_t fp16_round(float x) {
    const _Float16 h = static_cast<_Float16>(x);   // IEEE RNE (compiler-rt / F16C)
    std::uint16_t b;
    std::memcpy(&b, &h, sizeof b);
    return b;
}

float fp16_widen(std::uint16_t b) {
    _Float16 h;
    std::memcpy(&h, &b, sizeof h);
    return static_cast<float>(h);
}

namespace {

bool fp16_zero_or_subnormal(std::uint16_t b) { return (b & 0x7c00u) == 0; }

// What the fp16 storage keeps for an off-diagonal v of a column with scale s:
// fp16(v / s), its subnormals flushed to signed zero (omp.h:353-366).
std::uint16_t narrow16(float v, float s) {
    const std::uint16_t b = fp16_round(v / s);
    return fp16_zero_or_subnormal(b) ? static_cast<std::uint16_t>(b & 0x8000u) : b;
}

double drop_rel_from_env() {
    const char* e = std::getenv("APXCHOL_FACTOR_DROP");
    if (!e || !*e) return kFactorDropRelDefault;
    const double r = std::atof(e);
    return r > 0.0 ? r : 0.0;
}

}  // namespace

trsv_factor prepare_trsv(const factorization& F, bool fp16, double rel) {
    if (rel < 0.0) rel = drop_rel_from_env();
    trsv_factor T;
    const node_t n = F.n;
    const node_t m = F.sddm ? n : (n > 0 ? n - 1 : 0);
    T.m = m;
    T.fp16 = fp16;
    if (F.nnz() > std::numeric_limits<std::int32_t>::max())
        throw std::runtime_error("apxchol_amd: factor nnz exceeds the 32-bit index range of the device layout");

    // L11: columns 0..m-1 without row n-1 on the Laplacian path. Columns keep
    // the diagonal first, rows ascending.
    std::vector<std::uint64_t> c_ptr(static_cast<std::size_t>(m) + 1, 0);
    std::vector<node_t> c_row;
    std::vector<float> c_val;
    c_row.reserve(F.nnz());
    c_val.reserve(F.nnz());
    for (node_t j = 0; j < m; ++j) {
        for (off_t64 p = F.outer[j]; p < F.outer[j + 1]; ++p) {
            if (F.inner[p] >= m) continue;
            c_row.push_back(F.inner[p]);
            c_val.push_back(F.vals[p]);
        }
        c_ptr[j + 1] = c_row.size();
    }
    T.drop.rel = rel;
    T.drop.nnz_factor = c_row.size();

    // Column scale s_j = max |off-diagonal| (1 if none), before the drop
    // (factor_drop.h:83-89); the fp16 storage falls back to s_j = 1 where
    // 1/s_j or L_jj/s_j would overflow fp32 (omp.h:601-615).
    std::vector<float> scale(m, 1.0f);
    if (rel > 0.0 || fp16) {
        for (node_t j = 0; j < m; ++j) {
            double mx = 0.0;
            for (std::uint64_t p = c_ptr[j] + 1; p < c_ptr[j + 1]; ++p)
                mx = std::max(mx, std::fabs(static_cast<double>(c_val[p])));
            scale[j] = mx > 0.0 ? static_cast<float>(mx) : 1.0f;
            if (fp16) {
                const float s = scale[j];
                const float d = c_val[c_ptr[j]];
                if (!std::isfinite(1.0f / s) || !std::isfinite(d / s)) scale[j] = 1.0f;
            }
        }
    }

    // Compacting drop (factor_drop.h:131-211): keep an off-diagonal iff
    // |v| >= rel * s_j and the storage keeps it nonzero; spread each column's
    // dropped mass over its kept off-diagonals in proportion to |v|.
    auto keep = [&](float v, float s) {
        if (!(std::fabs(static_cast<double>(v)) >= rel * static_cast<double>(s))) return false;
        if (fp16) return !fp16_zero_or_subnormal(fp16_round(v / s));
        return v != 0.0f;
    };
    if (rel > 0.0) {
        std::vector<std::uint64_t> d_ptr(static_cast<std::size_t>(m) + 1, 0);
        std::vector<node_t> d_row;
        std::vector<float> d_val;
        d_row.reserve(c_row.size());
        d_val.reserve(c_row.size());
        for (node_t j = 0; j < m; ++j) {
            const float s = scale[j];
            const std::uint64_t first = d_row.size();
            double dropped_sum = 0.0, kept_abs = 0.0;
            for (std::uint64_t p = c_ptr[j]; p < c_ptr[j + 1]; ++p) {
                const float v = c_val[p];
                if (c_row[p] != j) {
                    if (!keep(v, s)) { dropped_sum += static_cast<double>(v); continue; }
                    kept_abs += std::fabs(static_cast<double>(v));
                }
                d_row.push_back(c_row[p]);
                d_val.push_back(v);
            }
            if (dropped_sum != 0.0 && kept_abs > 0.0) {
                const double per_abs = dropped_sum / kept_abs;
                for (std::uint64_t q = first; q < d_row.size(); ++q) {
                    if (d_row[q] == j) continue;
                    const double v = static_cast<double>(d_val[q]);
                    d_val[q] = static_cast<float>(v + per_abs * std::fabs(v));
                }
            }
            d_ptr[j + 1] = d_row.size();
        }
        if (d_row.size() != c_row.size()) {
            c_ptr.swap(d_ptr);
            c_row.swap(d_row);
            c_val.swap(d_val);
        }
    }
    T.drop.nnz_stored = c_row.size();
    T.drop.dropped = T.drop.nnz_factor - T.drop.nnz_stored;

    // Diagonal and stored values.
    T.diag.resize(m);
    std::vector<std::uint16_t> c_h;
    if (fp16) {
        c_h.resize(c_row.size());
        T.inv_scale.resize(m);
        for (node_t j = 0; j < m; ++j) {
            const float s = scale[j];
            T.inv_scale[j] = 1.0f / s;
            const float d = c_val[c_ptr[j]] / s;
            double resid = 0.0;
            for (std::uint64_t p = c_ptr[j]; p < c_ptr[j + 1]; ++p) {
                const std::uint16_t h = narrow16(c_val[p], s);
                c_h[p] = h;
                if (c_row[p] == j) continue;
                resid += static_cast<double>(c_val[p]) / static_cast<double>(s) - static_cast<double>(fp16_widen(h));
            }
            T.diag[j] = static_cast<float>(static_cast<double>(d) + resid);
            const double r = static_cast<double>(T.inv_scale[j]);
            if (!(std::isfinite(T.diag[j]) && T.diag[j] != 0.0f && std::isfinite(r * r)))
                throw std::runtime_error("apxchol_amd: fp16 factor storage cannot represent column " + std::to_string(j));
        }
    } else {
        for (node_t j = 0; j < m; ++j) T.diag[j] = c_val[c_ptr[j]];
    }

    // Back sweep: the columns without their diagonal, rows DESCENDING -- i.e.
    // ascending sweep position (the back sweep walks rows m-1 .. 0), as the
    // forward rows are: in both directions a row's entries run from its
    // earliest dependency to its latest.
    const std::uint64_t off = c_row.size() - m;
    T.b_ptr.assign(static_cast<std::size_t>(m) + 1, 0);
    T.b_idx.resize(off);
    if (fp16) T.b_h.resize(off); else T.b_val.resize(off);
    {
        std::uint64_t o = 0;
        for (node_t j = 0; j < m; ++j) {
            for (std::uint64_t p = c_ptr[j + 1]; p-- > c_ptr[j] + 1; ++o) {
                T.b_idx[o] = static_cast<std::int32_t>(c_row[p]);
                if (fp16) T.b_h[o] = c_h[p]; else T.b_val[o] = c_val[p];
            }
            T.b_ptr[j + 1] = static_cast<std::uint32_t>(o);
        }
    }
    // Forward sweep: rows of L (counting-sort transpose; column order gives
    // ascending columns within each row).
    T.f_ptr.assign(static_cast<std::size_t>(m) + 1, 0);
    for (std::uint64_t o = 0; o < off; ++o) ++T.f_ptr[T.b_idx[o] + 1];
    for (node_t i = 0; i < m; ++i) T.f_ptr[i + 1] += T.f_ptr[i];
    T.f_idx.resize(off);
    if (fp16) T.f_h.resize(off); else T.f_val.resize(off);
    {
        std::vector<std::uint32_t> cur(T.f_ptr.begin(), T.f_ptr.end() - 1);
        for (node_t j = 0; j < m; ++j)
            for (std::uint32_t o = T.b_ptr[j]; o < T.b_ptr[j + 1]; ++o) {
                const std::uint32_t i = static_cast<std::uint32_t>(T.b_idx[o]);
                const std::uint32_t q = cur[i]++;
                T.f_idx[q] = static_cast<std::int32_t>(j);
                if (fp16) T.f_h[q] = T.b_h[o]; else T.f_val[q] = T.b_val[o];
            }
    }

    // Round-as-level boundaries (omp.h:890-945): cumulative IS sizes clamped
    // to m, then one level per residual-peel column.
    T.level_ptr.assign(1, 0);
    std::uint64_t acc = 0;
    for (const auto& r : F.rounds) {
        acc = std::min<std::uint64_t>(acc + r.is_size, m);
        if (acc > T.level_ptr.back()) T.level_ptr.push_back(static_cast<std::uint32_t>(acc));
    }
    for (std::uint64_t c = T.level_ptr.back(); c < m; ++c) T.level_ptr.push_back(static_cast<std::uint32_t>(c + 1));
    return T;
}

std::vector<std::uint32_t> relabel_within_levels(trsv_factor& T, const std::vector<node_t>& perm, bool by_depth) {
    const std::uint32_t m = T.m;
    std::vector<std::uint32_t> vert(perm.size());
    for (std::size_t v = 0; v < perm.size(); ++v) vert[perm[v]] = static_cast<std::uint32_t>(v);
    std::vector<std::uint32_t> Q(m), order;
    if (by_depth) {
        // Depth of every row in factor order (its dependencies precede it).
        std::vector<std::uint32_t> depth(m, 0);
        std::uint32_t levels = m > 0 ? 1 : 0;
        for (std::uint32_t r = 0; r < m; ++r) {
            std::uint32_t dr = 0;
            for (std::uint32_t p = T.f_ptr[r]; p < T.f_ptr[r + 1]; ++p) {
                const std::uint32_t j = static_cast<std::uint32_t>(T.f_idx[p]);
                if (j >= r) throw std::runtime_error("apxchol_amd: forward dependency not before its row");
                dr = std::max(dr, depth[j] + 1);
            }
            depth[r] = dr;
            levels = std::max(levels, dr + 1);
        }
        // Rows grouped by depth (stable: factor order inside a group).
        std::vector<std::uint32_t> lp(static_cast<std::size_t>(levels) + 1, 0);
        for (std::uint32_t r = 0; r < m; ++r) ++lp[depth[r] + 1];
        for (std::uint32_t l = 0; l < levels; ++l) lp[l + 1] += lp[l];
        std::vector<std::uint32_t> at(lp.begin(), lp.end() - 1), byd(m);
        for (std::uint32_t r = 0; r < m; ++r) byd[at[depth[r]]++] = r;
        for (std::uint32_t l = 0; l < levels; ++l) {
            const std::uint32_t a = lp[l], b = lp[l + 1];
            std::sort(byd.begin() + a, byd.begin() + b, [&](std::uint32_t x, std::uint32_t y) { return vert[x] < vert[y]; });
            for (std::uint32_t k = a; k < b; ++k) Q[byd[k]] = k;
        }
        T.level_ptr.swap(lp);
    } else {
        for (std::size_t l = 0; l + 1 < T.level_ptr.size(); ++l) {
            const std::uint32_t a = T.level_ptr[l], b = T.level_ptr[l + 1];
            order.resize(b - a);
            for (std::uint32_t q = a; q < b; ++q) order[q - a] = q;
            std::sort(order.begin(), order.end(), [&](std::uint32_t x, std::uint32_t y) { return vert[x] < vert[y]; });
            for (std::uint32_t k = 0; k < b - a; ++k) Q[order[k]] = a + k;
        }
    }
    auto relist = [&](std::vector<std::uint32_t>& ptr, std::vector<std::int32_t>& idx, std::vector<float>& val,
                      std::vector<std::uint16_t>& h, bool descending) {
        std::vector<std::uint32_t> nptr(static_cast<std::size_t>(m) + 1, 0);
        for (std::uint32_t r = 0; r < m; ++r) nptr[Q[r] + 1] = ptr[r + 1] - ptr[r];
        for (std::uint32_t r = 0; r < m; ++r) nptr[r + 1] += nptr[r];
        std::vector<std::int32_t> nidx(idx.size());
        std::vector<float> nval(val.size());
        std::vector<std::uint16_t> nh(h.size());
        std::vector<std::pair<std::int32_t, std::uint32_t>> ent;   // (new index, old slot)
        for (std::uint32_t r = 0; r < m; ++r) {
            ent.clear();
            for (std::uint32_t p = ptr[r]; p < ptr[r + 1]; ++p)
                ent.emplace_back(static_cast<std::int32_t>(Q[static_cast<std::uint32_t>(idx[p])]), p);
            if (descending) std::sort(ent.begin(), ent.end(), [](auto& x, auto& y) { return x.first > y.first; });
            else std::sort(ent.begin(), ent.end(), [](auto& x, auto& y) { return x.first < y.first; });
            std::uint32_t o = nptr[Q[r]];
            for (const auto& [j, p] : ent) {
                nidx[o] = j;
                if (!val.empty()) nval[o] = val[p];
                if (!h.empty()) nh[o] = h[p];
                ++o;
            }
        }
        ptr.swap(nptr);
        idx.swap(nidx);
        val.swap(nval);
        h.swap(nh);
    };
    relist(T.f_ptr, T.f_idx, T.f_val, T.f_h, false);
    relist(T.b_ptr, T.b_idx, T.b_val, T.b_h, true);
    auto move_rows = [&](std::vector<float>& a) {
        if (a.empty()) return;
        std::vector<float> b(a.size());
        for (std::uint32_t r = 0; r < m; ++r) b[Q[r]] = a[r];
        a.swap(b);
    };
    move_rows(T.diag);
    move_rows(T.inv_scale);
    return Q;
}

permuted_operator build_permuted_operator(const csc_view& A, const std::vector<node_t>& perm) {
    permuted_operator op;
    const std::int64_t n = A.n;
    op.n = n;
    op.ptr.assign(static_cast<std::size_t>(n) + 1, 0);
    for (std::int64_t k = 0; k < n; ++k) op.ptr[perm[k] + 1] = A.colptr[k + 1] - A.colptr[k];
    for (std::int64_t i = 0; i < n; ++i) op.ptr[i + 1] += op.ptr[i];
    const std::int64_t nnz = op.ptr[n];
    op.idx.resize(nnz);
    op.val.resize(nnz);
    std::vector<std::pair<std::int32_t, double>> row;
    bool exact = true;
    for (std::int64_t k = 0; k < n; ++k) {
        row.clear();
        for (std::int64_t p = A.colptr[k]; p < A.colptr[k + 1]; ++p) {
            row.emplace_back(static_cast<std::int32_t>(perm[A.rowidx[p]]), A.vals[p]);
            exact = exact && static_cast<double>(static_cast<float>(A.vals[p])) == A.vals[p];
        }
        std::sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        std::int64_t o = op.ptr[perm[k]];
        for (const auto& [c, v] : row) { op.idx[o] = c; op.val[o] = v; ++o; }
    }
    op.fp32_exact = exact;
    if (exact) {
        op.val32.resize(nnz);
        for (std::int64_t p = 0; p < nnz; ++p) op.val32[p] = static_cast<float>(op.val[p]);
        std::vector<double>().swap(op.val);
    }
    return op;
}

std::uint64_t sell_slots(const permuted_operator& op, std::int64_t long_len, const std::vector<std::int32_t>& rows) {
    const std::int64_t n = op.n;
    std::uint64_t slots = 0;
    std::int64_t in_slice = 0, w = 0;
    for (std::int64_t i = 0; i < n; ++i) {
        const std::int64_t r = rows.empty() ? i : rows[i];
        const std::int64_t len = op.ptr[r + 1] - op.ptr[r];
        if (len > long_len) continue;
        w = std::max(w, len);
        if (++in_slice == 64) { slots += static_cast<std::uint64_t>(w) * 64; in_slice = 0; w = 0; }
    }
    return slots + static_cast<std::uint64_t>(w) * 64;
}

sell_layout build_sell(const permuted_operator& op, std::int64_t long_len, std::int64_t sigma,
                       const std::vector<std::int32_t>& rows) {
    sell_layout s;
    const std::int64_t n = op.n;
    if (static_cast<std::int64_t>(rows.size()) != n) throw std::invalid_argument("apxchol_amd: sliced ELL row order");
    auto len = [&](std::int64_t r) { return op.ptr[r + 1] - op.ptr[r]; };
    std::vector<std::int32_t> order;   // short rows in slice order
    order.reserve(static_cast<std::size_t>(n));
    const std::int64_t win = std::max<std::int64_t>(sigma, 1);
    for (std::int64_t w0 = 0; w0 < n; w0 += win) {
        const std::size_t first = order.size();
        for (std::int64_t i = w0; i < std::min(n, w0 + win); ++i) {
            const std::int64_t r = rows[i];
            if (r < 0 || r >= n) throw std::invalid_argument("apxchol_amd: sliced ELL row order");
            if (len(r) > long_len) {
                s.longrows.push_back(static_cast<std::int32_t>(r));
            } else {
                order.push_back(static_cast<std::int32_t>(r));
            }
        }
        if (sigma > 1)
            std::stable_sort(order.begin() + static_cast<std::ptrdiff_t>(first), order.end(),
                             [&](std::int32_t a, std::int32_t b) { return len(a) > len(b); });
    }
    const std::int64_t nrows = static_cast<std::int64_t>(order.size());
    const std::int64_t nslices = (nrows + 63) / 64;
    s.soff.assign(static_cast<std::size_t>(nslices) + 1, 0);
    for (std::int64_t q = 0; q < nslices; ++q) {
        std::int64_t w = 0;
        for (std::int64_t l = q * 64; l < std::min(nrows, q * 64 + 64); ++l) w = std::max(w, len(order[l]));
        const std::uint64_t next = static_cast<std::uint64_t>(s.soff[q]) + static_cast<std::uint64_t>(w);
        if (next > UINT32_MAX / 64) throw std::runtime_error("apxchol_amd: sliced ELL too large");
        s.soff[q + 1] = static_cast<std::uint32_t>(next);
    }
    s.slots = static_cast<std::uint64_t>(s.soff[nslices]) * 64;
    const bool f32 = op.fp32_exact;
    s.idx.assign(s.slots, 0);
    if (f32) s.val32.assign(s.slots, 0.0f); else s.val.assign(s.slots, 0.0);
    s.rowmap.assign(static_cast<std::size_t>(nslices) * 64, -1);
    for (std::int64_t q = 0; q < nslices; ++q) {
        const std::int64_t w = s.soff[q + 1] - s.soff[q];
        for (std::int64_t l = 0; l < 64; ++l) {
            const std::int64_t pos = q * 64 + l;
            if (pos >= nrows) continue;
            const std::int64_t r = order[pos];
            s.rowmap[pos] = static_cast<std::int32_t>(r);
            const std::int64_t b = op.ptr[r], e = op.ptr[r + 1];
            for (std::int64_t k = 0; k < w; ++k) {
                const std::size_t at = (static_cast<std::size_t>(s.soff[q]) + static_cast<std::size_t>(k)) * 64 +
                                       static_cast<std::size_t>(l);
                if (b + k < e) {
                    s.idx[at] = op.idx[b + k];
                    if (f32) s.val32[at] = op.val32[b + k]; else s.val[at] = op.val[b + k];
                } else {
                    s.idx[at] = static_cast<std::int32_t>(r);   // padding: value 0
                }
            }
        }
    }
    return s;
}

}  // namespace apxchol_amd
