// ORACLE -- TEST INFRASTRUCTURE ONLY. A CPU restatement of the reference's
// solve path at one thread (what the reference computes with OMP_NUM_THREADS=1),
// used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
// checker and the baseline -- never by the product library, which has no CPU
// fallback. Built by oracle/Makefile into oracle/liboracle.so.
//
// Restated functions and the reference lines they follow:
//   L11 copy ................. include/apxchol/solver/sptrsv/omp.h:535-584
//   column scale ............. include/apxchol/solver/sptrsv/factor_drop.h:83-89
//   fp16 scale fallback ...... omp.h:601-621
//   compacting drop .......... factor_drop.h:131-211 (keep predicate omp.h:383-390)
//   fp16 narrowing / diag .... omp.h:353-366, 420-428, 666-679, 707-750
//   CSC->CSR transpose ....... include/apxchol/solver/sptrsv/transpose.h:117-244 (serial order)
//   round-as-level levels .... omp.h:890-945 (topological: 946-997)
//   level-set sweeps ......... omp.h:1094-1119, 1273-1325, 1343-1445 (T=1: thin levels
//                              dot_thin, fat levels the sequential / fp16 SIMD-order sum)
//   apply (center-k) ......... include/apxchol/solver/preconditioner.h:344-529, env_knobs.h:69-125
//   deterministic reductions . preconditioner.h:44-127 (T=1)
//   operator build ........... src/solve.cpp:320-444 (row-major full symmetric, fp32 iff exact)
//   PCG ...................... src/solve.cpp:93-269, 478-654
//
// Parity status of each piece is recorded in the tests that check it:
// tests/test_cpu_parity.py (against the reference's own headers compiled into
// oracle/_ref, its fixture files and known answers, and closed forms) and
// tests/test_cpu_reference_cells.py (iteration counts the reference published).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

using u32 = std::uint32_t;
using u64 = std::uint64_t;

constexpr u32 kThinLevel = 1024;   // omp.h:32 kSpTRSVOMPThreshold

std::uint16_t half_round(float x) {
    const _Float16 h = static_cast<_Float16>(x);
    std::uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}
float half_widen(std::uint16_t b) {
    _Float16 h;
    std::memcpy(&h, &b, 2);
    return static_cast<float>(h);
}
bool half_zero_or_sub(std::uint16_t b) { return (b & 0x7c00u) == 0; }

struct oracle {
    u32 n = 0, m = 0;
    bool sddm = false, fp16 = false;
    // SpTRSV storage (omp_sptrsv): CSR of L11 (diagonal LAST per row) and CSC
    // of L11 (diagonal FIRST per column).
    std::vector<u64> csr_ptr, csc_ptr;
    std::vector<u32> csr_idx, csc_idx;
    std::vector<float> csr_v, csc_v;
    std::vector<std::uint16_t> csr_h, csc_h;
    std::vector<float> diag16, inv_scale16;   // fp16 storage only
    std::vector<std::vector<u32>> fwd_levels, bck_levels;
    u64 nnz_factor = 0, nnz_stored = 0;
    std::vector<u32> perm;
    // PCG operator (original order).
    std::vector<u64> op_ptr;
    std::vector<u32> op_idx;
    std::vector<double> op_v;
    std::vector<float> op_f;
    bool op32 = false;
    int center_k = 10;
    bool center_mode = true;
    u64 apply_count = 0;
    std::vector<double> s1, s2;
};

// sum over q in [p, end) of vals[q] * y[idx[q]], 4-way (omp.h:1254-1266).
template <class Get>
double dot_thin(Get get, const u32* idx, u64 p, u64 end, const double* y) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (; p + 4 <= end; p += 4) {
        s0 += get(p + 0) * y[idx[p + 0]];
        s1 += get(p + 1) * y[idx[p + 1]];
        s2 += get(p + 2) * y[idx[p + 2]];
        s3 += get(p + 3) * y[idx[p + 3]];
    }
    double sum = (s0 + s1) + (s2 + s3);
    for (; p < end; ++p) sum += get(p) * y[idx[p]];
    return sum;
}

// Fat-level sum of the fp16 storage (omp.h:1290-1325 summation order).
template <class Get>
double dot_fat16(Get get, const u32* idx, u64 p, u64 end, const double* y) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (; p + 8 <= end; p += 8) {
        s0 += get(p + 0) * y[idx[p + 0]];
        s1 += get(p + 1) * y[idx[p + 1]];
        s2 += get(p + 2) * y[idx[p + 2]];
        s3 += get(p + 3) * y[idx[p + 3]];
        s0 += get(p + 4) * y[idx[p + 4]];
        s1 += get(p + 5) * y[idx[p + 5]];
        s2 += get(p + 6) * y[idx[p + 6]];
        s3 += get(p + 7) * y[idx[p + 7]];
    }
    if (p + 4 <= end) {
        s0 += get(p + 0) * y[idx[p + 0]];
        s1 += get(p + 1) * y[idx[p + 1]];
        s2 += get(p + 2) * y[idx[p + 2]];
        s3 += get(p + 3) * y[idx[p + 3]];
        p += 4;
    }
    double sum = (s0 + s1) + (s2 + s3);
    for (; p < end; ++p) sum += get(p) * y[idx[p]];
    return sum;
}

// One row (forward: CSR row v) or column (back: CSC column v) of a sweep,
// omp.h:1343-1385 (solve_row): the thin-level 4-way sum, the fat-level plain
// sum (fp32) or SIMD-order sum (fp16), then y_out[v] = (rhs - sum) / diag.
void solve_row(const oracle& o, bool fwd, bool thin, u32 v, const double* x_in, double* y_out) {
    const auto& ptr = fwd ? o.csr_ptr : o.csc_ptr;
    const auto& idx = fwd ? o.csr_idx : o.csc_idx;
    const auto& v32 = fwd ? o.csr_v : o.csc_v;
    const auto& v16 = fwd ? o.csr_h : o.csc_h;
    auto get32 = [&](u64 p) { return static_cast<double>(v32[p]); };
    auto get16 = [&](u64 p) { return static_cast<double>(half_widen(v16[p])); };
    const u64 p0 = fwd ? ptr[v] : ptr[v] + 1;
    const u64 p1 = fwd ? ptr[v + 1] - 1 : ptr[v + 1];
    double sum;
    if (o.fp16) {
        sum = thin ? dot_thin(get16, idx.data(), p0, p1, y_out) : dot_fat16(get16, idx.data(), p0, p1, y_out);
    } else if (thin) {
        sum = dot_thin(get32, idx.data(), p0, p1, y_out);
    } else {
        sum = 0.0;
        for (u64 p = p0; p < p1; ++p) sum += get32(p) * y_out[idx[p]];
    }
    double rhs = x_in[v];
    if (!fwd && o.fp16) {
        const double r = static_cast<double>(o.inv_scale16[v]);
        rhs = x_in[v] * (r * r);
    }
    const double dg = o.fp16 ? static_cast<double>(o.diag16[v])
                             : static_cast<double>(fwd ? v32[ptr[v + 1] - 1] : v32[ptr[v]]);
    y_out[v] = (rhs - sum) / dg;
}

void sweep(const oracle& o, bool fwd, const double* x_in, double* y_out) {
    for (const auto& level : fwd ? o.fwd_levels : o.bck_levels) {
        const bool thin = level.size() <= kThinLevel;
        for (const u32 v : level) solve_row(o, fwd, thin, v, x_in, y_out);
    }
}

// T=1 deterministic sum (preconditioner.h:108-125).
double det_sum(const double* v, u64 n) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    u64 i = 0;
    for (; i + 4 <= n; i += 4) { s0 += v[i]; s1 += v[i + 1]; s2 += v[i + 2]; s3 += v[i + 3]; }
    double s = (s0 + s1) + (s2 + s3);
    for (; i < n; ++i) s += v[i];
    return s;
}

bool take_center(oracle& o) {
    if (!o.center_mode) return true;
    return (++o.apply_count % static_cast<u64>(o.center_k)) == 0;
}

// apply_core (preconditioner.h:403-529) at T=1.
void apply_core(oracle& o, const double* b, double b_sum, bool center, double* z, double* rz_out) {
    const u32 n = o.n;
    const u32* P = o.perm.data();
    auto gather_out = [&](const double* src, double shift) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        u64 v = 0;
        for (; v + 4 <= n; v += 4) {
            const double z0 = src[P[v + 0]] - shift;
            const double z1 = src[P[v + 1]] - shift;
            const double z2 = src[P[v + 2]] - shift;
            const double z3 = src[P[v + 3]] - shift;
            z[v + 0] = z0; z[v + 1] = z1; z[v + 2] = z2; z[v + 3] = z3;
            s0 += b[v + 0] * z0; s1 += b[v + 1] * z1; s2 += b[v + 2] * z2; s3 += b[v + 3] * z3;
        }
        double s = (s0 + s1) + (s2 + s3);
        for (; v < n; ++v) {
            const double zv = src[P[v]] - shift;
            z[v] = zv;
            s += b[v] * zv;
        }
        if (rz_out) *rz_out = s;
    };
    o.s1.assign(n, 0.0);
    o.s2.assign(n, 0.0);
    if (o.sddm) {
        for (u32 v = 0; v < n; ++v) z[P[v]] = b[v];
        sweep(o, true, z, o.s1.data());
        sweep(o, false, o.s1.data(), o.s2.data());
        gather_out(o.s2.data(), 0.0);
    } else {
        if (center) {
            const double b_mean = b_sum / static_cast<double>(n);
            for (u32 v = 0; v < n; ++v) z[P[v]] = b[v] - b_mean;
        } else {
            for (u32 v = 0; v < n; ++v) z[P[v]] = b[v];
        }
        sweep(o, true, z, o.s1.data());
        sweep(o, false, o.s1.data(), o.s2.data());
        o.s2[o.m] = 0.0;
        const double z_mean = center ? det_sum(o.s2.data(), n) / static_cast<double>(n) : 0.0;
        gather_out(o.s2.data(), z_mean);
    }
}

// parallel_spmv_csr at T=1 (solve.cpp:122-161): y = A x, returns x.y.
double spmv(const oracle& o, const double* x, double* y) {
    double xy = 0.0;
    for (u32 i = 0; i < o.n; ++i) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        u64 k = o.op_ptr[i];
        const u64 e = o.op_ptr[i + 1];
        auto val = [&](u64 q) { return o.op32 ? static_cast<double>(o.op_f[q]) : o.op_v[q]; };
        for (; k + 4 <= e; k += 4) {
            s0 += val(k + 0) * x[o.op_idx[k + 0]];
            s1 += val(k + 1) * x[o.op_idx[k + 1]];
            s2 += val(k + 2) * x[o.op_idx[k + 2]];
            s3 += val(k + 3) * x[o.op_idx[k + 3]];
        }
        double sum = (s0 + s1) + (s2 + s3);
        for (; k < e; ++k) sum += val(k) * x[o.op_idx[k]];
        y[i] = sum;
        xy += x[i] * sum;
    }
    return xy;
}

void build_levels(oracle& o, const u64* round_is, u64 nrounds) {
    const u32 m = o.m;
    o.fwd_levels.clear();
    o.bck_levels.clear();
    // apx_cholesky::install_factor bounds (preconditioner.h:238-249).
    std::vector<u32> bounds;
    if (nrounds > 0) {
        bounds.push_back(0);
        u32 acc = 0;
        for (u64 r = 0; r < nrounds; ++r) {
            acc += static_cast<u32>(round_is[r]);
            if (acc > m) acc = m;
            bounds.push_back(acc);
        }
    }
    if (!bounds.empty() && bounds.back() <= m) {
        std::vector<int> rof(m);
        const u32 R = static_cast<u32>(bounds.size()) - 1;
        const u32 last = bounds.back();
        u32 r = 0;
        for (u32 c = 0; c < m; ++c) {
            if (c >= last) { rof[c] = static_cast<int>(R + (c - last)); continue; }
            while (r + 1 < bounds.size() && c >= bounds[r + 1]) ++r;
            rof[c] = static_cast<int>(r);
        }
        int maxd = 0;
        for (u32 c = 0; c < m; ++c) maxd = std::max(maxd, rof[c]);
        if (m == 0) return;
        o.fwd_levels.assign(maxd + 1, {});
        o.bck_levels.assign(maxd + 1, {});
        for (u32 c = 0; c < m; ++c) {
            o.fwd_levels[rof[c]].push_back(c);
            o.bck_levels[maxd - rof[c]].push_back(c);
        }
        return;
    }
    // Topological depth levels (omp.h:946-997).
    std::vector<int> depth(m, 0);
    int maxd = 0;
    for (u32 i = 0; i < m; ++i) {
        int d = 0;
        for (u64 p = o.csr_ptr[i]; p < o.csr_ptr[i + 1] - 1; ++p) d = std::max(d, depth[o.csr_idx[p]] + 1);
        depth[i] = d;
        maxd = std::max(maxd, d);
    }
    o.fwd_levels.assign(m ? maxd + 1 : 0, {});
    for (u32 i = 0; i < m; ++i) o.fwd_levels[depth[i]].push_back(i);
    std::fill(depth.begin(), depth.end(), 0);
    maxd = 0;
    for (u32 j = m; j-- > 0;) {
        int d = 0;
        for (u64 p = o.csc_ptr[j]; p < o.csc_ptr[j + 1]; ++p)
            if (o.csc_idx[p] > j) d = std::max(d, depth[o.csc_idx[p]] + 1);
        depth[j] = d;
        maxd = std::max(maxd, d);
    }
    o.bck_levels.assign(m ? maxd + 1 : 0, {});
    for (u32 j = 0; j < m; ++j) o.bck_levels[depth[j]].push_back(j);
}

}  // namespace

extern "C" {

// Build the oracle from the operator A (full symmetric CSC, both triangles)
// and a factor (CSC of L, diagonal first, perm, per-round IS sizes).
// fp16: per-column-scaled fp16 storage; rel: drop threshold (<= 0 off);
// use_rounds: round-as-level (1) or topological levels (0).
void* or_create(u32 n, const std::int64_t* colptr, const std::int32_t* rowidx, const double* avals, const u64* outer,
                const u32* inner, const float* fvals, const u32* perm, const u64* round_is, u64 nrounds, int sddm,
                int fp16, double rel, int use_rounds, int center_k) {
    auto* o = new oracle;
    o->n = n;
    o->sddm = sddm != 0;
    o->fp16 = fp16 != 0;
    o->m = o->sddm ? n : (n ? n - 1 : 0);
    o->center_k = center_k >= 1 ? center_k : 10;
    o->perm.assign(perm, perm + n);
    const u32 m = o->m;

    // L11 (omp.h:535-584).
    std::vector<u64> c_ptr(m + 1, 0);
    std::vector<u32> c_idx;
    std::vector<float> c_val;
    for (u32 j = 0; j < m; ++j) {
        for (u64 p = outer[j]; p < outer[j + 1]; ++p) {
            if (!o->sddm && inner[p] == n - 1) continue;
            c_idx.push_back(inner[p]);
            c_val.push_back(fvals[p]);
        }
        c_ptr[j + 1] = c_idx.size();
    }
    o->nnz_factor = c_idx.size();
    // Column scales (factor_drop.h:83-89) + fp16 fallback (omp.h:601-615).
    std::vector<float> scale(m, 1.0f);
    if (o->fp16 || rel > 0.0) {
        for (u32 j = 0; j < m; ++j) {
            double mx = 0.0;
            for (u64 p = c_ptr[j] + 1; p < c_ptr[j + 1]; ++p) mx = std::max(mx, std::fabs(static_cast<double>(c_val[p])));
            scale[j] = mx > 0.0 ? static_cast<float>(mx) : 1.0f;
        }
        if (o->fp16)
            for (u32 j = 0; j < m; ++j) {
                const float s = scale[j];
                const float d = c_val[c_ptr[j]];
                if (!std::isfinite(1.0f / s) || !std::isfinite(d / s)) scale[j] = 1.0f;
            }
    }
    // Drop (factor_drop.h:131-211).
    auto keep = [&](float v, float s) {
        if (!(std::fabs(static_cast<double>(v)) >= rel * static_cast<double>(s))) return false;
        if (o->fp16) {
            const std::uint16_t h = half_round(static_cast<float>(v) / s);
            return !((h & 0x7fffu) == 0 || ((h & 0x7c00u) == 0 && (h & 0x03ffu) != 0));
        }
        return v != 0.0f;
    };
    if (rel > 0.0) {
        std::vector<u64> d_ptr(m + 1, 0);
        std::vector<u32> d_idx;
        std::vector<float> d_val;
        for (u32 j = 0; j < m; ++j) {
            const float s = scale[j];
            const u64 first = d_idx.size();
            double dropped = 0.0, kept_abs = 0.0;
            for (u64 p = c_ptr[j]; p < c_ptr[j + 1]; ++p) {
                const float v = c_val[p];
                if (c_idx[p] != j) {
                    if (!keep(v, s)) { dropped += static_cast<double>(v); continue; }
                    kept_abs += std::fabs(static_cast<double>(v));
                }
                d_idx.push_back(c_idx[p]);
                d_val.push_back(v);
            }
            if (dropped != 0.0 && kept_abs > 0.0) {
                const double per = dropped / kept_abs;
                for (u64 q = first; q < d_idx.size(); ++q) {
                    if (d_idx[q] == j) continue;
                    const double v = static_cast<double>(d_val[q]);
                    d_val[q] = static_cast<float>(v + per * std::fabs(v));
                }
            }
            d_ptr[j + 1] = d_idx.size();
        }
        c_ptr.swap(d_ptr);
        c_idx.swap(d_idx);
        c_val.swap(d_val);
    }
    o->nnz_stored = c_idx.size();
    const u64 nnz = c_idx.size();
    // CSC copy (+ fp16 narrowing, diag with the rounding residual).
    o->csc_ptr = c_ptr;
    o->csc_idx = c_idx;
    if (o->fp16) {
        o->csc_h.resize(nnz);
        o->diag16.resize(m);
        o->inv_scale16.resize(m);
        for (u32 j = 0; j < m; ++j) {
            const float s = scale[j];
            o->inv_scale16[j] = 1.0f / s;
            o->diag16[j] = static_cast<float>(static_cast<double>(static_cast<float>(c_val[c_ptr[j]]) / s));
            double resid = 0.0;
            for (u64 p = c_ptr[j]; p < c_ptr[j + 1]; ++p) {
                std::uint16_t h = half_round(c_val[p] / s);
                if ((h & 0x7c00u) == 0 && (h & 0x03ffu) != 0) h = static_cast<std::uint16_t>(h & 0x8000u);
                o->csc_h[p] = h;
                if (c_idx[p] != j) {
                    const double x = static_cast<double>(c_val[p]) / static_cast<double>(s);
                    resid += x - static_cast<double>(half_widen(h));
                }
            }
            o->diag16[j] = static_cast<float>(static_cast<double>(o->diag16[j]) + resid);
        }
    } else {
        o->csc_v = c_val;
    }
    // Transpose (serial column-order scatter, transpose.h:66).
    o->csr_ptr.assign(m + 1, 0);
    for (u64 p = 0; p < nnz; ++p) ++o->csr_ptr[c_idx[p] + 1];
    for (u32 i = 0; i < m; ++i) o->csr_ptr[i + 1] += o->csr_ptr[i];
    o->csr_idx.resize(nnz);
    if (o->fp16) o->csr_h.resize(nnz); else o->csr_v.resize(nnz);
    std::vector<u64> cur(o->csr_ptr.begin(), o->csr_ptr.end() - 1);
    for (u32 j = 0; j < m; ++j)
        for (u64 p = c_ptr[j]; p < c_ptr[j + 1]; ++p) {
            const u64 q = cur[c_idx[p]]++;
            o->csr_idx[q] = j;
            if (o->fp16) o->csr_h[q] = o->csc_h[p]; else o->csr_v[q] = c_val[p];
        }
    build_levels(*o, round_is, use_rounds ? nrounds : 0);

    // PCG operator: row-major full symmetric copy of the lower triangle
    // (solve.cpp:330-421), rows sorted; fp32 iff every value round-trips.
    o->op_ptr.assign(n + 1, 0);
    for (u32 k = 0; k < n; ++k)
        for (std::int64_t p = colptr[k]; p < colptr[k + 1]; ++p) {
            const u32 r = static_cast<u32>(rowidx[p]);
            if (r < k) continue;
            ++o->op_ptr[r + 1];
            if (r != k) ++o->op_ptr[k + 1];
        }
    for (u32 i = 0; i < n; ++i) o->op_ptr[i + 1] += o->op_ptr[i];
    o->op_idx.resize(o->op_ptr[n]);
    o->op_v.resize(o->op_ptr[n]);
    std::vector<u64> pos(o->op_ptr.begin(), o->op_ptr.end() - 1);
    for (u32 k = 0; k < n; ++k)
        for (std::int64_t p = colptr[k]; p < colptr[k + 1]; ++p) {
            const u32 r = static_cast<u32>(rowidx[p]);
            if (r < k) continue;
            u64 s = pos[r]++;
            o->op_idx[s] = k;
            o->op_v[s] = avals[p];
            if (r != k) {
                s = pos[k]++;
                o->op_idx[s] = r;
                o->op_v[s] = avals[p];
            }
        }
    std::vector<std::pair<u32, double>> kv;
    for (u32 i = 0; i < n; ++i) {
        kv.clear();
        for (u64 p = o->op_ptr[i]; p < o->op_ptr[i + 1]; ++p) kv.emplace_back(o->op_idx[p], o->op_v[p]);
        std::sort(kv.begin(), kv.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (u64 p = o->op_ptr[i]; p < o->op_ptr[i + 1]; ++p) {
            o->op_idx[p] = kv[p - o->op_ptr[i]].first;
            o->op_v[p] = kv[p - o->op_ptr[i]].second;
        }
    }
    bool exact = true;
    for (std::int64_t p = 0; p < colptr[n]; ++p)
        exact = exact && static_cast<double>(static_cast<float>(avals[p])) == avals[p];
    o->op32 = exact;
    if (exact) {
        o->op_f.resize(o->op_v.size());
        for (size_t p = 0; p < o->op_v.size(); ++p) o->op_f[p] = static_cast<float>(o->op_v[p]);
    }
    return o;
}

void or_destroy(void* h) { delete static_cast<oracle*>(h); }

// info: m, nnz_factor, nnz_stored, fwd levels, bck levels, op32
void or_info(void* h, u64* out) {
    auto* o = static_cast<oracle*>(h);
    out[0] = o->m;
    out[1] = o->nnz_factor;
    out[2] = o->nnz_stored;
    out[3] = o->fwd_levels.size();
    out[4] = o->bck_levels.size();
    out[5] = o->op32 ? 1 : 0;
}

void or_export(void* h, u64* csr_ptr, u32* csr_idx, float* csr_v, std::uint16_t* csr_h, u64* csc_ptr, u32* csc_idx,
               float* csc_v, std::uint16_t* csc_h, float* diag16, float* inv_scale16) {
    auto* o = static_cast<oracle*>(h);
    auto cp = [](auto* dst, const auto& v) { if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0])); };
    cp(csr_ptr, o->csr_ptr); cp(csr_idx, o->csr_idx); cp(csr_v, o->csr_v); cp(csr_h, o->csr_h);
    cp(csc_ptr, o->csc_ptr); cp(csc_idx, o->csc_idx); cp(csc_v, o->csc_v); cp(csc_h, o->csc_h);
    cp(diag16, o->diag16); cp(inv_scale16, o->inv_scale16);
}

void or_level_sizes(void* h, int fwd, std::int64_t* sizes) {
    auto* o = static_cast<oracle*>(h);
    const auto& lv = fwd ? o->fwd_levels : o->bck_levels;
    for (size_t l = 0; l < lv.size(); ++l) sizes[l] = static_cast<std::int64_t>(lv[l].size());
}

void or_forward(void* h, const double* x, double* y) {
    auto* o = static_cast<oracle*>(h);
    sweep(*o, true, x, y);
}
void or_back(void* h, const double* x, double* y) {
    auto* o = static_cast<oracle*>(h);
    sweep(*o, false, x, y);
}

// cpu_solver::apply(r) = _solve_impl (preconditioner.h:344-353).
void or_apply(void* h, const double* r, double* z) {
    auto* o = static_cast<oracle*>(h);
    const bool center = !o->sddm && take_center(*o);
    const double b_sum = center ? det_sum(r, o->n) : 0.0;
    apply_core(*o, r, b_sum, center, z, nullptr);
}
void or_reset_apply_count(void* h) { static_cast<oracle*>(h)->apply_count = 0; }

double or_spmv(void* h, const double* x, double* y) { return spmv(*static_cast<oracle*>(h), x, y); }

// cpu_solver::solve_impl (solve.cpp:478-654) at T=1. out: iterations, residual.
void or_solve(void* h, const double* b, double* x, const double* x0, double tol, int max_iter, int stag_window,
              double* out) {
    auto* o = static_cast<oracle*>(h);
    const u32 n = o->n;
    std::vector<double> r(n), z(n), p(n), Ap(n);
    const bool laplacian = !o->sddm;
    auto center_x = [&]() {
        const double mean = det_sum(x, n) / static_cast<double>(n);
        for (u32 i = 0; i < n; ++i) x[i] -= mean;
    };
    auto init_residual = [&](const double* ax0, double& rr, double& rs) {
        double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0, s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        u64 i = 0;
        for (; i + 4 <= n; i += 4) {
            double rv[4];
            for (int k = 0; k < 4; ++k) rv[k] = ax0 ? b[i + k] - ax0[i + k] : b[i + k];
            for (int k = 0; k < 4; ++k) r[i + k] = rv[k];
            q0 += rv[0] * rv[0]; q1 += rv[1] * rv[1]; q2 += rv[2] * rv[2]; q3 += rv[3] * rv[3];
            s0 += rv[0]; s1 += rv[1]; s2 += rv[2]; s3 += rv[3];
        }
        double q = (q0 + q1) + (q2 + q3), s = (s0 + s1) + (s2 + s3);
        for (; i < n; ++i) {
            const double ri = ax0 ? b[i] - ax0[i] : b[i];
            r[i] = ri;
            q += ri * ri;
            s += ri;
        }
        rr = q;
        rs = s;
    };
    double bnorm, rr, rs;
    out[0] = 0;
    out[1] = 0.0;
    bool x0_nonzero = false;
    if (x0) for (u32 i = 0; i < n && !x0_nonzero; ++i) x0_nonzero = x0[i] != 0.0;
    if (x0 && x0_nonzero) {
        double bb = 0.0;   // Eigen's b.norm(); any order is fine at the tolerance it feeds
        for (u32 i = 0; i < n; ++i) bb += b[i] * b[i];
        bnorm = std::sqrt(bb);
        if (bnorm == 0.0) { std::fill(x, x + n, 0.0); return; }
        std::copy(x0, x0 + n, x);
        spmv(*o, x, Ap.data());
        init_residual(Ap.data(), rr, rs);
        out[0] = 0;
        out[1] = std::sqrt(rr) / bnorm;
        if (out[1] < tol) { if (laplacian) center_x(); return; }
    } else {
        init_residual(nullptr, rr, rs);
        bnorm = std::sqrt(rr);
        if (bnorm == 0.0) { std::fill(x, x + n, 0.0); return; }
        std::fill(x, x + n, 0.0);
        out[0] = 0;
        out[1] = 1.0;
    }
    o->apply_count = 0;
    auto apply_fused = [&](double r_sum) {
        const bool center = !o->sddm && take_center(*o);
        double rz = 0.0;
        apply_core(*o, r.data(), r_sum, center, z.data(), &rz);
        return rz;
    };
    double rz = apply_fused(rs);
    p = z;
    double prev = 1.0;
    for (int i = 0; i < max_iter; ++i) {
        const double pAp = spmv(*o, p.data(), Ap.data());
        if (pAp <= 0.0) break;
        const double alpha = rz / pAp;
        double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0, s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        u64 k = 0;
        for (; k + 4 <= n; k += 4) {
            x[k + 0] += alpha * p[k + 0];
            x[k + 1] += alpha * p[k + 1];
            x[k + 2] += alpha * p[k + 2];
            x[k + 3] += alpha * p[k + 3];
            const double r0 = r[k + 0] - alpha * Ap[k + 0];
            const double r1 = r[k + 1] - alpha * Ap[k + 1];
            const double r2 = r[k + 2] - alpha * Ap[k + 2];
            const double r3 = r[k + 3] - alpha * Ap[k + 3];
            r[k + 0] = r0; r[k + 1] = r1; r[k + 2] = r2; r[k + 3] = r3;
            q0 += r0 * r0; q1 += r1 * r1; q2 += r2 * r2; q3 += r3 * r3;
            s0 += r0; s1 += r1; s2 += r2; s3 += r3;
        }
        double q = (q0 + q1) + (q2 + q3), s = (s0 + s1) + (s2 + s3);
        for (; k < n; ++k) {
            x[k] += alpha * p[k];
            const double ri = r[k] - alpha * Ap[k];
            r[k] = ri;
            q += ri * ri;
            s += ri;
        }
        rr = q;
        rs = s;
        const double rn = std::sqrt(rr) / bnorm;
        out[0] = i + 1;
        out[1] = rn;
        if (rn < tol) break;
        if (stag_window > 0 && (i + 1) % stag_window == 0) {
            if (rn > prev * 0.5) break;
            prev = rn;
        }
        const double rz_new = apply_fused(rs);
        const double beta = rz_new / rz;
        for (u32 j = 0; j < n; ++j) p[j] = z[j] + beta * p[j];
        rz = rz_new;
    }
    if (laplacian) center_x();
}

}  // extern "C"

// ---- the OpenMP CPU path (timed CPU baseline, not the checker) ---------------
//
// The reference's multi-threaded CPU solve: the level-set sweeps of
// omp.h:1094-1119 (one parallel region per sweep; a thin level -- at most 1024
// rows -- runs on one thread, a fat level is split statically over the team,
// the implicit barrier carrying the level dependency) with solve_row's per-row
// arithmetic, so a sweep gives the one-thread bits; and the fused passes of
// src/solve.cpp:122-269 and preconditioner.h:403-529 over static_chunk
// partitions with per-thread partials summed in thread order
// (preconditioner.h:44-127), so a solve is bit-identical run to run at a fixed
// thread count. Each pass runs its team only above 2000 elements
// (fused_omp_min, preconditioner.h:61).
#include <omp.h>

namespace {

constexpr u64 kFusedOmpMin = 2000;
constexpr int kPartStride = 8;

struct team {
    int threads = 1;
    std::vector<double> part;
    explicit team(int t) : threads(t < 1 ? 1 : t), part(static_cast<size_t>(threads) * kPartStride, 0.0) {}
    void ids(int& tid, int& nt) const { tid = omp_get_thread_num(); nt = omp_get_num_threads(); }
    double reduce(int nt, int slot = 0) const {
        double s = 0.0;
        for (int t = 0; t < nt; ++t) s += part[static_cast<size_t>(t) * kPartStride + slot];
        return s;
    }
};

void static_chunk(u64 n, int tid, int nt, u64& lo, u64& hi) {
    const u64 base = n / static_cast<u64>(nt), rem = n % static_cast<u64>(nt);
    lo = static_cast<u64>(tid) * base + std::min<u64>(static_cast<u64>(tid), rem);
    hi = lo + base + (static_cast<u64>(tid) < rem ? 1 : 0);
}

void sweep_omp(const oracle& o, team& tm, bool fwd, const double* x_in, double* y_out) {
    const auto& levels = fwd ? o.fwd_levels : o.bck_levels;
    const size_t L = levels.size();
#pragma omp parallel num_threads(tm.threads)
    for (size_t l = 0; l < L; ++l) {
        const auto& level = levels[l];
        const long long sz = static_cast<long long>(level.size());
        if (level.size() <= kThinLevel) {
#pragma omp single
            for (long long k = 0; k < sz; ++k) solve_row(o, fwd, true, level[k], x_in, y_out);
        } else {
#pragma omp for schedule(static)
            for (long long k = 0; k < sz; ++k) solve_row(o, fwd, false, level[k], x_in, y_out);
        }
    }
}

double det_sum_omp(team& tm, const double* v, u64 n) {
    int used = 1;
#pragma omp parallel num_threads(tm.threads) if (n > kFusedOmpMin)
    {
        int tid, nt;
        tm.ids(tid, nt);
        if (tid == 0) used = nt;
        u64 lo, hi;
        static_chunk(n, tid, nt, lo, hi);
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        u64 i = lo;
        for (; i + 4 <= hi; i += 4) { s0 += v[i]; s1 += v[i + 1]; s2 += v[i + 2]; s3 += v[i + 3]; }
        double s = (s0 + s1) + (s2 + s3);
        for (; i < hi; ++i) s += v[i];
        tm.part[static_cast<size_t>(tid) * kPartStride] = s;
    }
    return tm.reduce(used);
}

void apply_core_omp(oracle& o, team& tm, const double* b, double b_sum, bool center, double* z, double* rz_out) {
    const u64 n = o.n;
    const u32* P = o.perm.data();
    const bool par = n > kFusedOmpMin;
    auto gather_out = [&](const double* src, double shift) {
        int used = 1;
#pragma omp parallel num_threads(tm.threads) if (par)
        {
            int tid, nt;
            tm.ids(tid, nt);
            if (tid == 0) used = nt;
            u64 lo, hi;
            static_chunk(n, tid, nt, lo, hi);
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            u64 v = lo;
            for (; v + 4 <= hi; v += 4) {
                const double z0 = src[P[v + 0]] - shift;
                const double z1 = src[P[v + 1]] - shift;
                const double z2 = src[P[v + 2]] - shift;
                const double z3 = src[P[v + 3]] - shift;
                z[v + 0] = z0; z[v + 1] = z1; z[v + 2] = z2; z[v + 3] = z3;
                s0 += b[v + 0] * z0; s1 += b[v + 1] * z1; s2 += b[v + 2] * z2; s3 += b[v + 3] * z3;
            }
            double s = (s0 + s1) + (s2 + s3);
            for (; v < hi; ++v) {
                const double zv = src[P[v]] - shift;
                z[v] = zv;
                s += b[v] * zv;
            }
            tm.part[static_cast<size_t>(tid) * kPartStride] = s;
        }
        if (rz_out) *rz_out = tm.reduce(used);
    };
    // Scatter, with the input centring on a centring Laplacian application
    // (b - 0.0 is b exactly, so one loop serves the plain scatter too).
    const double b_mean = (!o.sddm && center) ? b_sum / static_cast<double>(n) : 0.0;
#pragma omp parallel for schedule(static) num_threads(tm.threads) if (par)
    for (long long v = 0; v < static_cast<long long>(n); ++v) z[P[v]] = b[v] - b_mean;
    sweep_omp(o, tm, true, z, o.s1.data());
    sweep_omp(o, tm, false, o.s1.data(), o.s2.data());
    if (o.sddm) {
        gather_out(o.s2.data(), 0.0);
    } else {
        o.s2[o.m] = 0.0;
        const double z_mean = center ? det_sum_omp(tm, o.s2.data(), n) / static_cast<double>(n) : 0.0;
        gather_out(o.s2.data(), z_mean);
    }
}

double spmv_omp(const oracle& o, team& tm, const double* x, double* y) {
    const u64 n = o.n;
    int used = 1;
#pragma omp parallel num_threads(tm.threads) if (n > kFusedOmpMin)
    {
        int tid, nt;
        tm.ids(tid, nt);
        if (tid == 0) used = nt;
        u64 lo, hi;
        static_chunk(n, tid, nt, lo, hi);
        auto val = [&](u64 q) { return o.op32 ? static_cast<double>(o.op_f[q]) : o.op_v[q]; };
        double xy = 0.0;
        for (u64 i = lo; i < hi; ++i) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            u64 k = o.op_ptr[i];
            const u64 e = o.op_ptr[i + 1];
            for (; k + 4 <= e; k += 4) {
                s0 += val(k + 0) * x[o.op_idx[k + 0]];
                s1 += val(k + 1) * x[o.op_idx[k + 1]];
                s2 += val(k + 2) * x[o.op_idx[k + 2]];
                s3 += val(k + 3) * x[o.op_idx[k + 3]];
            }
            double sum = (s0 + s1) + (s2 + s3);
            for (; k < e; ++k) sum += val(k) * x[o.op_idx[k]];
            y[i] = sum;
            xy += x[i] * sum;
        }
        tm.part[static_cast<size_t>(tid) * kPartStride] = xy;
    }
    return tm.reduce(used);
}

// r = b - A x0 (ax0 null: r = b) or x += alpha p; r -= alpha Ap (update):
// r.r and sum r of the new r, per-thread 4-way accumulators.
void fused_r_omp(team& tm, u64 n, double* r, const double* b, const double* ax0, double* x, const double* p,
                 const double* ap, double alpha, bool update, double& rr, double& rs) {
    int used = 1;
#pragma omp parallel num_threads(tm.threads) if (n > kFusedOmpMin)
    {
        int tid, nt;
        tm.ids(tid, nt);
        if (tid == 0) used = nt;
        u64 lo, hi;
        static_chunk(n, tid, nt, lo, hi);
        double q[4] = {0.0, 0.0, 0.0, 0.0}, sm[4] = {0.0, 0.0, 0.0, 0.0};
        auto step = [&](u64 i) {
            double ri;
            if (update) {
                x[i] += alpha * p[i];
                ri = r[i] - alpha * ap[i];
            } else {
                ri = ax0 ? b[i] - ax0[i] : b[i];
            }
            r[i] = ri;
            return ri;
        };
        u64 i = lo;
        for (; i + 4 <= hi; i += 4)
            for (int k = 0; k < 4; ++k) {
                const double ri = step(i + k);
                q[k] += ri * ri;
                sm[k] += ri;
            }
        double qq = (q[0] + q[1]) + (q[2] + q[3]), ss = (sm[0] + sm[1]) + (sm[2] + sm[3]);
        for (; i < hi; ++i) {
            const double ri = step(i);
            qq += ri * ri;
            ss += ri;
        }
        tm.part[static_cast<size_t>(tid) * kPartStride + 0] = qq;
        tm.part[static_cast<size_t>(tid) * kPartStride + 1] = ss;
    }
    rr = tm.reduce(used, 0);
    rs = tm.reduce(used, 1);
}

}  // namespace

extern "C" {

// The OpenMP CPU path's forward / back sweep and apply at `threads` threads.
void or_forward_omp(void* h, const double* x, double* y, int threads) {
    team tm(threads);
    sweep_omp(*static_cast<oracle*>(h), tm, true, x, y);
}
void or_back_omp(void* h, const double* x, double* y, int threads) {
    team tm(threads);
    sweep_omp(*static_cast<oracle*>(h), tm, false, x, y);
}
void or_apply_omp(void* h, const double* r, double* z, int threads) {
    auto* o = static_cast<oracle*>(h);
    team tm(threads);
    o->s1.assign(o->n, 0.0);
    o->s2.assign(o->n, 0.0);
    const bool center = !o->sddm && take_center(*o);
    const double b_sum = center ? det_sum_omp(tm, r, o->n) : 0.0;
    apply_core_omp(*o, tm, r, b_sum, center, z, nullptr);
}

// cpu_solver::solve_impl (solve.cpp:478-654) on `threads` threads, x0 = 0.
// out: iterations, residual.
void or_solve_omp(void* h, const double* b, double* x, double tol, int max_iter, int stag_window, int threads,
                  double* out) {
    auto* o = static_cast<oracle*>(h);
    const u64 n = o->n;
    team tm(threads);
    std::vector<double> r(n), z(n), p(n), Ap(n);
    o->s1.assign(n, 0.0);
    o->s2.assign(n, 0.0);
    out[0] = 0;
    out[1] = 0.0;
    double rr, rs;
    fused_r_omp(tm, n, r.data(), b, nullptr, nullptr, nullptr, nullptr, 0.0, false, rr, rs);
    const double bnorm = std::sqrt(rr);
#pragma omp parallel for schedule(static) num_threads(tm.threads) if (n > kFusedOmpMin)
    for (long long i = 0; i < static_cast<long long>(n); ++i) x[i] = 0.0;
    if (bnorm == 0.0) return;
    out[1] = 1.0;
    o->apply_count = 0;
    auto apply_fused = [&](double r_sum) {
        const bool center = !o->sddm && take_center(*o);
        double rz = 0.0;
        apply_core_omp(*o, tm, r.data(), r_sum, center, z.data(), &rz);
        return rz;
    };
    double rz = apply_fused(rs);
#pragma omp parallel for schedule(static) num_threads(tm.threads) if (n > kFusedOmpMin)
    for (long long i = 0; i < static_cast<long long>(n); ++i) p[i] = z[i];
    double prev = 1.0;
    for (int i = 0; i < max_iter; ++i) {
        const double pAp = spmv_omp(*o, tm, p.data(), Ap.data());
        if (pAp <= 0.0) break;
        const double alpha = rz / pAp;
        fused_r_omp(tm, n, r.data(), nullptr, nullptr, x, p.data(), Ap.data(), alpha, true, rr, rs);
        const double rn = std::sqrt(rr) / bnorm;
        out[0] = i + 1;
        out[1] = rn;
        if (rn < tol) break;
        if (stag_window > 0 && (i + 1) % stag_window == 0) {
            if (rn > prev * 0.5) break;
            prev = rn;
        }
        const double rz_new = apply_fused(rs);
        const double beta = rz_new / rz;
#pragma omp parallel for schedule(static) num_threads(tm.threads) if (n > kFusedOmpMin)
        for (long long j = 0; j < static_cast<long long>(n); ++j) p[j] = z[j] + beta * p[j];
        rz = rz_new;
    }
    if (!o->sddm) {
        const double mean = det_sum_omp(tm, x, n) / static_cast<double>(n);
#pragma omp parallel for schedule(static) num_threads(tm.threads) if (n > kFusedOmpMin)
        for (long long i = 0; i < static_cast<long long>(n); ++i) x[i] -= mean;
    }
}

}  // extern "C"
