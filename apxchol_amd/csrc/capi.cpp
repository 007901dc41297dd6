// extern "C" boundary of the MI355X solve path (include/apxchol_amd.h). No
// exception crosses it: every failure becomes an apx_status plus a
// thread-local message.
#include "../../include/apxchol_amd.h"

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "host/factor.hpp"
#include "host/gpu_solver.hpp"
#include "host/group.hpp"
#include "host/mtx.hpp"
#include "host/server.hpp"
#include "host/trsv_prep.hpp"
#include "kernels/launch.hpp"

using namespace apxchol_amd;

struct apx_factor_s {
    factorization F;
};
struct apx_solver_s {
    std::shared_ptr<gpu_solver> s;   // shared with a server that serves it
};
struct apx_server_s {
    std::unique_ptr<solve_server> s;
};
struct apx_group_s {
    std::unique_ptr<solver_group> g;
};
struct apx_matrix_s {
    csc_matrix A;
};

namespace {

thread_local std::string g_err;

template <class Fn>
apx_status guarded(Fn&& fn) {
    try {
        fn();
        return APX_OK;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return APX_ERR_INVALID;
    } catch (const std::bad_alloc& e) {
        g_err = std::string("out of memory: ") + e.what();
        return APX_ERR_MEMORY;
    } catch (const std::exception& e) {
        g_err = e.what();
        return APX_ERR_RUNTIME;
    } catch (...) {
        g_err = "unknown error";
        return APX_ERR_RUNTIME;
    }
}

csc_view make_view(int64_t n, const int64_t* colptr, const int32_t* rowidx, const double* vals) {
    if (n < 0) throw std::invalid_argument("n must be >= 0");
    if (n > 0 && (!colptr || !rowidx || !vals)) throw std::invalid_argument("null matrix array");
    csc_view v;
    v.n = n;
    v.colptr = colptr;
    v.rowidx = rowidx;
    v.vals = vals;
    return v;
}

solve_options to_opts(const apx_options* o) {
    solve_options so;
    apx_options d;
    apx_options_default(&d);
    if (!o) o = &d;
    so.tol = o->tol;
    so.max_iter = o->max_iter;
    so.stagnation_window = o->stagnation_window;
    so.factor_opts.seed = o->seed;
    if (o->precision != 0 && o->precision != 1) throw std::invalid_argument("precision must be 0 (fp64) or 1 (fp32)");
    if (o->schedule < 0 || o->schedule > 2)
        throw std::invalid_argument("schedule must be 0 (dataflow), 1 (levels) or 2 (hybrid)");
    so.precision = o->precision ? sweep_precision::fp32 : sweep_precision::fp64;
    so.fp16_storage = o->fp16_storage != 0;
    so.schedule = static_cast<sweep_schedule>(o->schedule);
    so.tail_rows = o->tail_rows;
    so.drop_rel = o->drop_rel;
    so.keep_factor_values = o->keep_factor != 0;
    so.iters_per_check = o->iters_per_check > 0 ? o->iters_per_check : 2;
    so.sweep_chunk = o->sweep_chunk;
    so.sweep_blocks = o->sweep_blocks;
    so.graph = o->graph != 0;
    so.relabel = o->relabel != 0;
    so.lane_batches = o->lane_batches;
    so.depth_levels = o->depth_levels != 0;
    so.tail_batches = o->tail_batches;
    so.tail_cut = o->tail_cut;
    return so;
}

void fill_info(const solve_result& r, apx_solve_info* info) {
    if (!info) return;
    info->iterations = r.iterations;
    info->residual = r.residual;
    info->lumped_offdiag = r.lumped_offdiag;
    info->solve_vram_mb = r.solve_vram_mb;
    info->status = r.status;
    info->factorize_ms = r.timings.factorize_ms;
    info->sptrsv_setup_ms = r.timings.sptrsv_setup_ms;
    info->pcg_setup_ms = r.timings.pcg_setup_ms;
    info->pcg_ms = r.timings.pcg_ms;
}

void need(const void* p, const char* what) {
    if (!p) throw std::invalid_argument(std::string("null ") + what);
}

}  // namespace

extern "C" {

void apx_options_default(apx_options* o) {
    if (!o) return;
    o->tol = 1e-8;
    o->max_iter = 200;
    o->stagnation_window = 50;
    o->seed = 42;
    o->precision = 0;
    o->fp16_storage = 0;
    o->schedule = 2;
    o->drop_rel = -1.0;
    o->keep_factor = 0;
    o->iters_per_check = 2;
    o->sweep_chunk = 0;
    o->sweep_blocks = 0;
    o->tail_rows = 0;
    o->graph = 1;
    o->relabel = 1;
    o->lane_batches = 0;
    o->depth_levels = 1;
    o->tail_batches = 0;
    o->tail_cut = 0;
}

const char* apx_last_error(void) { return g_err.c_str(); }

uint16_t apx_fp16_round(float x) { return fp16_round(x); }
float apx_fp16_widen(uint16_t h) { return fp16_widen(h); }

apx_status apx_factorize(int64_t n, const int64_t* colptr, const int32_t* rowidx, const double* vals, uint32_t seed,
                         apx_factor_t* out) {
    return guarded([&] {
        need(out, "out");
        factor_options fo;
        fo.seed = seed;
        auto* f = new apx_factor_s{factorize(make_view(n, colptr, rowidx, vals), fo)};
        *out = f;
    });
}

apx_status apx_factor_info(apx_factor_t f, int64_t* out) {
    return guarded([&] {
        need(f, "factor");
        need(out, "out");
        out[0] = f->F.n;
        out[1] = static_cast<int64_t>(f->F.nnz());
        out[2] = static_cast<int64_t>(f->F.rounds.size());
        out[3] = f->F.sddm ? 1 : 0;
        out[4] = f->F.lumped_offdiag;
        out[5] = f->F.threads;
        out[6] = 0;
    });
}

apx_status apx_factor_export(apx_factor_t f, uint64_t* outer, uint32_t* inner, float* vals, uint32_t* perm,
                             uint64_t* round_is) {
    return guarded([&] {
        need(f, "factor");
        const auto& F = f->F;
        if (outer) std::memcpy(outer, F.outer.data(), F.outer.size() * sizeof(uint64_t));
        if (inner) std::memcpy(inner, F.inner.data(), F.inner.size() * sizeof(uint32_t));
        if (vals) std::memcpy(vals, F.vals.data(), F.vals.size() * sizeof(float));
        if (perm) std::memcpy(perm, F.perm.data(), F.perm.size() * sizeof(uint32_t));
        if (round_is)
            for (std::size_t r = 0; r < F.rounds.size(); ++r) round_is[r] = F.rounds[r].is_size;
    });
}

void apx_factor_destroy(apx_factor_t f) { delete f; }

apx_status apx_trsv_sizes(apx_factor_t f, int32_t fp16, double drop_rel, int64_t* sizes) {
    return guarded([&] {
        need(f, "factor");
        need(sizes, "sizes");
        const trsv_factor T = prepare_trsv(f->F, fp16 != 0, drop_rel);
        sizes[0] = T.m;
        sizes[1] = static_cast<int64_t>(T.offdiag());
        sizes[2] = static_cast<int64_t>(T.level_ptr.size()) - 1;
        sizes[3] = static_cast<int64_t>(T.drop.nnz_factor);
        sizes[4] = static_cast<int64_t>(T.drop.nnz_stored);
        sizes[5] = T.fp16 ? 1 : 0;
    });
}

apx_status apx_trsv_export(apx_factor_t f, int32_t fp16, double drop_rel, uint32_t* f_ptr, int32_t* f_idx,
                           float* f_val, uint16_t* f_h, uint32_t* b_ptr, int32_t* b_idx, float* b_val, uint16_t* b_h,
                           float* diag, float* inv_scale, uint32_t* level_ptr) {
    return guarded([&] {
        need(f, "factor");
        const trsv_factor T = prepare_trsv(f->F, fp16 != 0, drop_rel);
        auto cp = [](auto* dst, const auto& v) {
            if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
        };
        cp(f_ptr, T.f_ptr); cp(f_idx, T.f_idx); cp(f_val, T.f_val); cp(f_h, T.f_h);
        cp(b_ptr, T.b_ptr); cp(b_idx, T.b_idx); cp(b_val, T.b_val); cp(b_h, T.b_h);
        cp(diag, T.diag); cp(inv_scale, T.inv_scale); cp(level_ptr, T.level_ptr);
    });
}

apx_status apx_solver_create(int64_t n, const int64_t* colptr, const int32_t* rowidx, const double* vals,
                             const apx_options* opts, apx_solver_t* out) {
    return guarded([&] {
        need(out, "out");
        auto* s = new apx_solver_s;
        try {
            s->s = std::make_shared<gpu_solver>(make_view(n, colptr, rowidx, vals), to_opts(opts));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

apx_status apx_solver_create_with_factor(int64_t n, const int64_t* colptr, const int32_t* rowidx, const double* vals,
                                         apx_factor_t f, const apx_options* opts, apx_solver_t* out) {
    return guarded([&] {
        need(out, "out");
        need(f, "factor");
        auto* s = new apx_solver_s;
        try {
            s->s = std::make_shared<gpu_solver>(make_view(n, colptr, rowidx, vals), f->F, to_opts(opts));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

void apx_solver_destroy(apx_solver_t s) { delete s; }

apx_status apx_solver_info_get(apx_solver_t s, apx_solver_info* info) {
    return guarded([&] {
        need(s, "solver");
        need(info, "info");
        const solver_stats& st = s->s->stats();
        info->n = st.n; info->m = st.m; info->sddm = st.sddm ? 1 : 0;
        info->factor_nnz = st.factor_nnz; info->stored_nnz = st.stored_nnz; info->offdiag = st.offdiag;
        info->dropped = st.dropped; info->rounds = st.rounds; info->levels = st.levels;
        info->fwd_batches = st.fwd_batches; info->bck_batches = st.bck_batches; info->op_nnz = st.op_nnz;
        info->op_fp32 = st.op_fp32 ? 1 : 0; info->spmv_lanes = st.spmv_lanes; info->dataflow_grid = st.dataflow_grid;
        info->fp16_storage = s->s->trsv().fp16 ? 1 : 0;
        info->device_bytes = st.device_bytes;
        info->spmv_format = st.spmv_format;
        info->spmv_sell_slots = st.spmv_sell_slots;
        info->split_couplings[0] = st.split_couplings[0];
        info->split_couplings[1] = st.split_couplings[1];
        for (int k = 0; k < 2; ++k) {
            info->level_bytes[k] = st.level_bytes[k];
            info->tail_bytes[k] = st.tail_bytes[k];
            info->level_launches[k] = st.level_launches[k];
            info->tail_runs[k] = st.tail_runs[k];
            info->lane_levels[k] = st.lane_levels[k];
        }
    });
}

apx_status apx_solver_solve(apx_solver_t s, const double* d_b, double* d_x, const double* d_x0, double tol,
                            int32_t max_iter, apx_solve_info* info, void* stream) {
    return guarded([&] {
        need(s, "solver");
        if (s->s->rows() > 0) { need(d_b, "b"); need(d_x, "x"); }
        fill_info(s->s->solve_device(d_b, d_x, tol, max_iter, d_x0, static_cast<hipStream_t>(stream)), info);
    });
}

apx_status apx_group_create(const apx_solver_t* solvers, int32_t count, apx_group_t* out) {
    return guarded([&] {
        need(out, "out");
        *out = nullptr;
        if (count <= 0) throw std::invalid_argument("a group needs at least one solver");
        need(solvers, "solvers");
        std::vector<std::shared_ptr<const gpu_solver>> members;
        for (int32_t i = 0; i < count; ++i) {
            need(solvers[i], "solver");
            members.push_back(solvers[i]->s);
        }
        auto g = std::make_unique<apx_group_s>();
        g->g = std::make_unique<solver_group>(std::move(members));
        *out = g.release();
    });
}

void apx_group_destroy(apx_group_t g) { delete g; }

apx_status apx_group_info(apx_group_t g, int32_t* members, int32_t* launches_per_iteration) {
    return guarded([&] {
        need(g, "group");
        if (members) *members = g->g->size();
        if (launches_per_iteration) *launches_per_iteration = g->g->launches_per_iteration();
    });
}

apx_status apx_group_solve(apx_group_t g, const double* const* d_b, double* const* d_x, double tol, int32_t max_iter,
                           apx_solve_info* infos, void* stream) {
    return guarded([&] {
        need(g, "group");
        need(d_b, "b");
        need(d_x, "x");
        const int N = g->g->size();
        std::vector<const double*> b(d_b, d_b + N);
        std::vector<double*> x(d_x, d_x + N);
        const auto r = g->g->solve(b, x, tol, max_iter, static_cast<hipStream_t>(stream));
        if (infos)
            for (int i = 0; i < N; ++i) fill_info(r[i], infos + i);
    });
}

apx_status apx_solver_solve_host(apx_solver_t s, const double* b, double* x, const double* x0, double tol,
                                 int32_t max_iter, apx_solve_info* info) {
    return guarded([&] {
        need(s, "solver");
        const std::size_t n = static_cast<std::size_t>(s->s->rows());
        if (n > 0) { need(b, "b"); need(x, "x"); }
        std::vector<double> bv(b, b + n), x0v;
        if (x0) x0v.assign(x0, x0 + n);
        solve_result r = s->s->solve(bv, tol, max_iter, x0 ? &x0v : nullptr);
        if (n > 0) std::memcpy(x, r.x.data(), n * sizeof(double));
        fill_info(r, info);
    });
}

apx_status apx_solver_solve_batch(apx_solver_t s, int32_t k, const double* d_B, double* d_X, double tol,
                                  int32_t max_iter, apx_solve_info* infos, void* stream) {
    return guarded([&] {
        need(s, "solver");
        if (k < 0) throw std::invalid_argument("k must be >= 0");
        if (k > 0 && s->s->rows() > 0) { need(d_B, "B"); need(d_X, "X"); }
        const auto r = s->s->solve_batch_device(k, d_B, d_X, tol, max_iter, static_cast<hipStream_t>(stream));
        if (infos)
            for (int32_t c = 0; c < k; ++c) fill_info(r[c], infos + c);
    });
}

apx_status apx_solver_solve_batch_host(apx_solver_t s, int32_t k, const double* B, double* X, double tol,
                                       int32_t max_iter, apx_solve_info* infos) {
    return guarded([&] {
        need(s, "solver");
        if (k < 0) throw std::invalid_argument("k must be >= 0");
        const std::size_t total = static_cast<std::size_t>(k) * static_cast<std::size_t>(s->s->rows());
        if (total > 0) { need(B, "B"); need(X, "X"); }
        std::vector<double> bv(B, B + total), xv;
        const auto r = s->s->solve_batch(k, bv, xv, tol, max_iter);
        if (total > 0) std::memcpy(X, xv.data(), total * sizeof(double));
        if (infos)
            for (int32_t c = 0; c < k; ++c) fill_info(r[c], infos + c);
    });
}

apx_status apx_solver_apply_batch(apx_solver_t s, int32_t k, const double* d_R, double* d_Z, void* stream) {
    return guarded([&] {
        need(s, "solver");
        if (k < 0) throw std::invalid_argument("k must be >= 0");
        if (k > 0 && s->s->rows() > 0) { need(d_R, "R"); need(d_Z, "Z"); }
        s->s->apply_batch_device(k, d_R, d_Z, static_cast<hipStream_t>(stream));
    });
}

apx_status apx_solver_autotune(apx_solver_t s, int32_t max_block_width, apx_autotune_result* out, void* stream) {
    return guarded([&] {
        need(s, "solver");
        const auto r = s->s->autotune(max_block_width > 0 ? max_block_width : 64, static_cast<hipStream_t>(stream));
        if (out) {
            out->spmv_lanes = r.spmv_lanes;
            out->block_width = r.block_width;
            for (int k = 0; k < 7; ++k) out->spmv_us[k] = r.spmv_us[k];
            for (int k = 0; k < 4; ++k) out->sweep_pair_us[k] = r.sweep_pair_us[k];
            out->spmv_sell = r.spmv_sell;
            out->spmv_sell_us = r.spmv_sell_us;
        }
    });
}

apx_status apx_server_create(uint64_t memory_budget, int32_t slots, int32_t iters_per_check, apx_server_t* out) {
    return guarded([&] {
        need(out, "out");
        server_options o;
        o.memory_budget = memory_budget;
        o.slots = slots >= 0 ? slots : 32;
        o.iters_per_check = iters_per_check > 0 ? iters_per_check : 2;
        auto h = std::make_unique<apx_server_s>();
        h->s = std::make_unique<solve_server>(o);
        *out = h.release();
    });
}

void apx_server_destroy(apx_server_t srv) { delete srv; }

apx_status apx_server_set_narrow_block(apx_server_t srv, int32_t slots) {
    return guarded([&] {
        need(srv, "server");
        srv->s->set_small_slots(slots);
    });
}

apx_status apx_server_set_single_lane(apx_server_t srv, int32_t on) {
    return guarded([&] {
        need(srv, "server");
        srv->s->set_single_lane(on != 0);
    });
}

apx_status apx_server_add(apx_server_t srv, apx_solver_t solver, int32_t* factor_id) {
    return guarded([&] {
        need(srv, "server");
        need(solver, "solver");
        const int id = srv->s->add(solver->s);
        if (factor_id) *factor_id = id;
    });
}

apx_status apx_server_submit(apx_server_t srv, int32_t factor, const double* b, double tol, int32_t max_iter,
                             int64_t* request_id) {
    return guarded([&] {
        need(srv, "server");
        need(b, "b");
        const int64_t id = srv->s->submit(factor, b, tol, max_iter);
        if (request_id) *request_id = id;
    });
}

apx_status apx_server_wait(apx_server_t srv, int64_t request_id, double* x, apx_solve_info* info, double timeout_s,
                           int32_t* ready) {
    return guarded([&] {
        need(srv, "server");
        solve_result r;
        const bool ok = srv->s->wait(request_id, x, &r, timeout_s);
        if (ready) *ready = ok ? 1 : 0;
        if (ok) fill_info(r, info);
    });
}

apx_status apx_server_get_stats(apx_server_t srv, apx_server_stats* st) {
    return guarded([&] {
        need(srv, "server");
        need(st, "stats");
        const server_stats s = srv->s->stats();
        st->submitted = s.submitted;
        st->completed = s.completed;
        st->factors = s.factors;
        st->p50_ms = s.p50_ms;
        st->p99_ms = s.p99_ms;
        st->mean_ms = s.mean_ms;
        st->solves_per_s = s.solves_per_s;
        st->memory_used = s.memory_used;
        st->memory_budget = s.memory_budget;
        st->single_lane_completed = s.single_lane_completed;
    });
}

apx_status apx_solver_apply(apx_solver_t s, const double* d_r, double* d_z, void* stream) {
    return guarded([&] {
        need(s, "solver");
        if (s->s->rows() > 0) { need(d_r, "r"); need(d_z, "z"); }
        s->s->apply_device(d_r, d_z, static_cast<hipStream_t>(stream));
    });
}

apx_status apx_solver_apply_host(apx_solver_t s, const double* r, double* z) {
    return guarded([&] {
        need(s, "solver");
        const std::size_t n = static_cast<std::size_t>(s->s->rows());
        if (n > 0) { need(r, "r"); need(z, "z"); }
        const std::vector<double> zv = s->s->apply(std::vector<double>(r, r + n));
        if (n > 0) std::memcpy(z, zv.data(), n * sizeof(double));
    });
}

apx_status apx_solver_reset_apply_count(apx_solver_t s) {
    return guarded([&] {
        need(s, "solver");
        s->s->reset_apply_count();
    });
}

apx_status apx_solver_sweep(apx_solver_t s, int32_t backward, const double* d_rhs, double* d_out, void* stream) {
    return guarded([&] {
        need(s, "solver");
        need(d_rhs, "rhs");
        need(d_out, "out");
        s->s->sweep_device(backward != 0, d_rhs, d_out, static_cast<hipStream_t>(stream));
    });
}

apx_status apx_solver_solve_LLt(apx_solver_t s, const double* d_x, double* d_y, void* stream) {
    return guarded([&] {
        need(s, "solver");
        need(d_x, "x");
        need(d_y, "y");
        s->s->solve_LLt_device(d_x, d_y, static_cast<hipStream_t>(stream));
    });
}

apx_status apx_solver_spmv(apx_solver_t s, const double* d_x, double* d_y, void* stream) {
    return guarded([&] {
        need(s, "solver");
        need(d_x, "x");
        need(d_y, "y");
        s->s->spmv_device(d_x, d_y, static_cast<hipStream_t>(stream));
    });
}

apx_status apx_solver_spmv_batch(apx_solver_t s, int32_t k, const double* d_X, double* d_Y, void* stream) {
    return guarded([&] {
        need(s, "solver");
        need(d_X, "X");
        need(d_Y, "Y");
        s->s->spmv_batch_device(k, d_X, d_Y, static_cast<hipStream_t>(stream));
    });
}

apx_status apx_solver_set_timing(apx_solver_t s, int32_t on) {
    return guarded([&] {
        need(s, "solver");
        s->s->set_timing(on != 0);
    });
}

apx_status apx_solver_take_timing(apx_solver_t s, double* ms, int64_t* launches) {
    return guarded([&] {
        need(s, "solver");
        need(ms, "ms");
        need(launches, "launches");
        s->s->take_timing(ms, launches);
    });
}

apx_status apx_solver_export_factor(apx_solver_t s, uint64_t* outer, uint32_t* inner, float* vals, uint32_t* perm) {
    return guarded([&] {
        need(s, "solver");
        const auto& F = s->s->factor();
        if ((inner || vals) && F.nnz() > 0 && F.vals.empty())
            throw std::invalid_argument("factor values were released (create the solver with keep_factor = 1)");
        if (outer) std::memcpy(outer, F.outer.data(), F.outer.size() * sizeof(uint64_t));
        if (inner) std::memcpy(inner, F.inner.data(), F.inner.size() * sizeof(uint32_t));
        if (vals) std::memcpy(vals, F.vals.data(), F.vals.size() * sizeof(float));
        if (perm) std::memcpy(perm, F.perm.data(), F.perm.size() * sizeof(uint32_t));
    });
}

// ---- stand-alone PCG pieces ----

namespace {
int piece_n(int64_t n) {
    if (n < 0 || n > INT32_MAX) throw std::invalid_argument("n must be in [0, 2^31 - 1]");
    return static_cast<int>(n);
}
}  // namespace

int32_t apx_pcg_blocks(int64_t units) { return pcg_blocks(units); }

int32_t apx_pcg_spmv_lanes_for(double avg_nnz_per_row) {
    // pcg_cuda_kernels.h:64-69: the smallest power of two >= avg / sqrt(2), <= 32.
    const double target = avg_nnz_per_row / 1.4142135623730951;
    int32_t lanes = 1;
    while (lanes < 32 && static_cast<double>(lanes) < target) lanes <<= 1;
    return lanes;
}

apx_status apx_pcg_spmv_pAp(int32_t n, const int32_t* d_rowptr, const int32_t* d_colidx, const void* d_vals,
                            int32_t vals_fp32, const double* d_x, double* d_y, double* d_part, int32_t lanes,
                            void* stream) {
    return guarded([&] {
        piece_n(n);
        if (lanes < 1 || lanes > 64 || (lanes & (lanes - 1)) != 0)
            throw std::invalid_argument("lanes must be a power of two in [1, 64]");
        if (n == 0) return;
        need(d_rowptr, "rowptr"); need(d_colidx, "colidx"); need(d_vals, "vals");
        need(d_x, "x"); need(d_y, "y"); need(d_part, "part");
        launch_piece_spmv_pAp(n, reinterpret_cast<const unsigned*>(d_rowptr), d_colidx, d_vals, vals_fp32 != 0, lanes,
                              d_x, d_y, d_part, static_cast<hipStream_t>(stream));
    });
}

apx_status apx_pcg_spmv_pAp_noalloc(int32_t n, const int32_t* d_rowptr, const int32_t* d_colidx, const void* d_vals,
                                    int32_t vals_fp32, const double* d_x, double* d_y, double* d_pAp, int32_t lanes,
                                    void* stream) {
    return guarded([&] {
        piece_n(n);
        if (lanes < 1 || lanes > 64 || (lanes & (lanes - 1)) != 0)
            throw std::invalid_argument("lanes must be a power of two in [1, 64]");
        need(d_pAp, "pAp");
        if (n > 0) {
            need(d_rowptr, "rowptr"); need(d_colidx, "colidx"); need(d_vals, "vals");
            need(d_x, "x"); need(d_y, "y");
        }
        launch_piece_spmv_pAp_scratch(n, reinterpret_cast<const unsigned*>(d_rowptr), d_colidx, d_vals, vals_fp32 != 0,
                                      lanes, d_x, d_y, d_pAp, static_cast<hipStream_t>(stream));
    });
}

apx_status apx_pcg_update_xr(int64_t n, double* d_x, const double* d_p, double* d_r, const double* d_Ap, double alpha,
                             double* d_part, void* stream) {
    return guarded([&] {
        const int k = piece_n(n);
        if (k == 0) return;
        need(d_x, "x"); need(d_p, "p"); need(d_r, "r"); need(d_Ap, "Ap"); need(d_part, "part");
        launch_piece_update_xr(k, d_x, d_p, d_r, d_Ap, alpha, d_part, static_cast<hipStream_t>(stream));
    });
}

apx_status apx_pcg_dot(int64_t n, const double* d_a, const double* d_b, double* d_part, void* stream) {
    return guarded([&] {
        const int k = piece_n(n);
        if (k == 0) return;
        need(d_a, "a"); need(d_b, "b"); need(d_part, "part");
        launch_piece_dot(k, d_a, d_b, d_part, static_cast<hipStream_t>(stream));
    });
}

apx_status apx_pcg_update_p(int64_t n, double* d_p, const double* d_z, double beta, void* stream) {
    return guarded([&] {
        const int k = piece_n(n);
        if (k == 0) return;
        need(d_p, "p"); need(d_z, "z");
        launch_piece_update_p(k, d_p, d_z, beta, static_cast<hipStream_t>(stream));
    });
}

apx_status apx_pcg_reduce_partials(const double* d_part, int32_t blocks, double* d_out, void* stream) {
    return guarded([&] {
        if (blocks < 0 || blocks > APX_PCG_MAX_BLOCKS)
            throw std::invalid_argument("blocks must be in [0, APX_PCG_MAX_BLOCKS]");
        need(d_out, "out");
        if (blocks > 0) need(d_part, "part");
        launch_piece_reduce(d_part, blocks, d_out, static_cast<hipStream_t>(stream));
    });
}

// ---- MatrixMarket files and input interpretation ----

namespace {
input_scan to_scan(const apx_input_scan* s) {
    need(s, "scan");
    input_scan r;
    r.n = s->n; r.offdiag_nnz = s->offdiag_nnz; r.offdiag_pos = s->offdiag_pos; r.offdiag_neg = s->offdiag_neg;
    r.upper_nnz = s->upper_nnz; r.lower_nnz = s->lower_nnz; r.diag_pos = s->diag_pos;
    r.deficient_rows = s->deficient_rows; r.worst_deficit = s->worst_deficit; r.worst_row = s->worst_row;
    r.excess_rows = s->excess_rows;
    return r;
}
input_kind to_kind(int32_t k) {
    if (k < 0 || k > 2) throw std::invalid_argument("input kind must be 0 (automatic), 1 (laplacian) or 2 (adjacency)");
    return static_cast<input_kind>(k);
}
void copy_text(const std::string& t, char* out, int64_t len) {
    if (!out || len <= 0) return;
    const std::size_t k = std::min<std::size_t>(t.size(), static_cast<std::size_t>(len - 1));
    std::memcpy(out, t.data(), k);
    out[k] = '\0';
}
}  // namespace

apx_status apx_mtx_read(const char* path, apx_matrix_t* out, int32_t* pattern_field) {
    return guarded([&] {
        need(path, "path");
        need(out, "out");
        mtx_header h;
        auto m = std::make_unique<apx_matrix_s>();
        m->A = read_mtx(path, &h);
        if (pattern_field) *pattern_field = h.pattern ? 1 : 0;
        *out = m.release();
    });
}

apx_status apx_matrix_create(int64_t n, const int64_t* colptr, const int32_t* rowidx, const double* vals,
                             apx_matrix_t* out) {
    return guarded([&] {
        need(out, "out");
        const csc_view v = make_view(n, colptr, rowidx, vals);
        auto m = std::make_unique<apx_matrix_s>();
        m->A.rows = m->A.cols = n;
        m->A.colptr.assign(v.colptr, v.colptr + n + 1);
        if (m->A.colptr[0] != 0) throw std::invalid_argument("colptr[0] must be 0");
        for (int64_t j = 0; j < n; ++j)
            if (m->A.colptr[j + 1] < m->A.colptr[j]) throw std::invalid_argument("colptr must be non-decreasing");
        const int64_t nnz = m->A.colptr[n];
        m->A.rowidx.assign(v.rowidx, v.rowidx + nnz);
        m->A.vals.assign(v.vals, v.vals + nnz);
        for (const int32_t i : m->A.rowidx)
            if (i < 0 || i >= n) throw std::invalid_argument("row index out of range");
        *out = m.release();
    });
}

apx_status apx_matrix_size(apx_matrix_t m, int64_t* rows, int64_t* cols, int64_t* nnz) {
    return guarded([&] {
        need(m, "matrix");
        if (rows) *rows = m->A.rows;
        if (cols) *cols = m->A.cols;
        if (nnz) *nnz = static_cast<int64_t>(m->A.rowidx.size());
    });
}

apx_status apx_matrix_export(apx_matrix_t m, int64_t* colptr, int32_t* rowidx, double* vals) {
    return guarded([&] {
        need(m, "matrix");
        if (colptr) std::memcpy(colptr, m->A.colptr.data(), m->A.colptr.size() * sizeof(int64_t));
        if (rowidx) std::memcpy(rowidx, m->A.rowidx.data(), m->A.rowidx.size() * sizeof(int32_t));
        if (vals) std::memcpy(vals, m->A.vals.data(), m->A.vals.size() * sizeof(double));
    });
}

void apx_matrix_destroy(apx_matrix_t m) { delete m; }

apx_status apx_mtx_write(const char* path, apx_matrix_t m, int32_t symmetric) {
    return guarded([&] {
        need(path, "path");
        need(m, "matrix");
        write_mtx(path, m->A, symmetric != 0);
    });
}

apx_status apx_mtx_read_vector(const char* path, double* x, int64_t cap, int64_t* n) {
    return guarded([&] {
        need(path, "path");
        need(n, "n");
        const std::vector<double> v = read_mtx_vector(path);
        *n = static_cast<int64_t>(v.size());
        if (x && cap >= *n && !v.empty()) std::memcpy(x, v.data(), v.size() * sizeof(double));
    });
}

apx_status apx_mtx_write_vector(const char* path, const double* x, int64_t n) {
    return guarded([&] {
        need(path, "path");
        if (n < 0) throw std::invalid_argument("n must be >= 0");
        if (n > 0) need(x, "x");
        write_mtx_vector(path, x, n);
    });
}

apx_status apx_input_scan_get(apx_matrix_t m, apx_input_scan* s) {
    return guarded([&] {
        need(m, "matrix");
        need(s, "scan");
        if (m->A.rows != m->A.cols) throw std::invalid_argument("matrix must be square");
        const input_scan r = scan_input(m->A.view());
        s->n = r.n; s->offdiag_nnz = r.offdiag_nnz; s->offdiag_pos = r.offdiag_pos; s->offdiag_neg = r.offdiag_neg;
        s->upper_nnz = r.upper_nnz; s->lower_nnz = r.lower_nnz; s->diag_pos = r.diag_pos;
        s->deficient_rows = r.deficient_rows; s->worst_deficit = r.worst_deficit; s->worst_row = r.worst_row;
        s->excess_rows = r.excess_rows;
    });
}

apx_status apx_input_resolve(int32_t requested, const apx_input_scan* s, int32_t pattern_field, int32_t* kind,
                             char* reason, int64_t reason_len) {
    return guarded([&] {
        need(kind, "kind");
        std::string why;
        *kind = static_cast<int32_t>(resolve_input_kind(to_kind(requested), to_scan(s), pattern_field != 0, why));
        copy_text(why, reason, reason_len);
    });
}

apx_status apx_adjacency_to_laplacian(apx_matrix_t a, apx_matrix_t* out) {
    return guarded([&] {
        need(a, "matrix");
        need(out, "out");
        if (a->A.rows != a->A.cols) throw std::invalid_argument("matrix must be square");
        auto m = std::make_unique<apx_matrix_s>();
        m->A = adjacency_to_laplacian(a->A.view());
        *out = m.release();
    });
}

apx_status apx_input_describe(int32_t kind, const apx_input_scan* s, const char* reason, char* out, int64_t out_len) {
    return guarded([&] { copy_text(describe_input(to_kind(kind), to_scan(s), reason ? reason : ""), out, out_len); });
}

apx_status apx_input_diagnose(int32_t kind, const apx_input_scan* s, char* out, int64_t out_len) {
    return guarded([&] { copy_text(diagnose_failure(to_kind(kind), to_scan(s)), out, out_len); });
}

apx_status apx_solve(int64_t n, const int64_t* colptr, const int32_t* rowidx, const double* vals, const double* b,
                     double* x, const apx_options* opts, apx_solve_info* info) {
    return guarded([&] {
        const csc_view A = make_view(n, colptr, rowidx, vals);
        if (n > 0) { need(b, "b"); need(x, "x"); }
        const gpu_solver slv(A, to_opts(opts));
        solve_result r = slv.solve(std::vector<double>(b, b + n));
        if (n > 0) std::memcpy(x, r.x.data(), static_cast<std::size_t>(n) * sizeof(double));
        fill_info(r, info);
    });
}

}  // extern "C"
