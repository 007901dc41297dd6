// Launchers of the gfx950 kernels (one translation unit so the templates
// instantiate once).
#include "launch.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>

#include "pcg.hip"
#include "sweep.hip"
#include "block.hip"
#include "pieces.hip"
#include "group.hip"

namespace apxchol_amd {

static_assert(kWavesPerBlock == kLevelWavesPerBlock, "level launches: host and device waves per workgroup");

int vec_block_cap() {
    static const int cap = [] {
        const char* e = std::getenv("APXCHOL_AMD_VEC_BLOCKS");
        const int v = (e && *e) ? std::atoi(e) : 0;
        return (v >= 1 && v <= kMaxBlocks) ? v : kMaxBlocks;
    }();
    return cap;
}

namespace {

// Launch `kernel`, carrying the hook's timing events when it hands some out.
template <class K, class... A>
void launch_timed(const launch_hook& hook, int kind, K kernel, dim3 grid, dim3 block, unsigned lds, hipStream_t s,
                  A... args) {
    hipEvent_t ev[2];
    if (hook.events(kind, ev)) hipExtLaunchKernelGGL(kernel, grid, block, lds, s, ev[0], ev[1], 0, args...);
    else hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
}

void check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("apxchol_amd: launch of ") + what + " failed: " + hipGetErrorString(e));
}

void hip_check_launch(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("apxchol_amd: ") + what + " failed: " + hipGetErrorString(e));
}

template <class TI, class TO, bool F16>
int dataflow_grid_for() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sweep_dataflow<TI, TO, F16>, kBlock, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    return per_cu * cus;
}

// Workgroups of one level launch: one wave per batch, up to this many.
inline constexpr int kLevelMaxBlocks = kLevelLaunchMaxBlocks;

template <class TI, class TO, bool F16>
void sweep_t(const sweep_desc& d, const void* rhs, void* out, const pcg_state* st, int* err, const sweep_sched& sc,
             hipStream_t s, const launch_hook& hook) {
    const TI* in = static_cast<const TI*>(rhs);
    TO* o = static_cast<TO*>(out);
    auto level_launch = [&](int l) {
        const int b0 = sc.level_first[l], b1 = sc.level_first[l + 1];
        if (b1 <= b0) return;
        if (sc.level_diag && sc.level_diag[l]) {
            // No entries in the level: one load per row.
            const int p0 = sc.level_pos[l], p1 = sc.level_pos[l + 1];
            launch_timed(hook, 0, sweep_diag_level<TI, TO>, dim3(vec_blocks(p1 - p0)), dim3(kBlock), 0, s, d, p0, p1,
                         in, o, st);
            return;
        }
        int g = (b1 - b0 + kWavesPerBlock - 1) / kWavesPerBlock;
        const int lb = sc.level_lb ? sc.level_lb[l] : -1;
        if (lb >= 0) {
            if (static_cast<long long>(lb) + (b1 - b0) > sc.lane_batches)
                throw std::runtime_error("apxchol_amd: level outside the lane layout");
            // Big levels: several batches per wave (fewer waves to dispatch).
            const int nlb = b1 - b0;
            const int bpw = (sc.lane_bpw4 > 0 && nlb >= sc.lane_bpw4) ? 4 : (sc.lane_bpw2 > 0 && nlb >= sc.lane_bpw2) ? 2 : 1;
            const int gb = (nlb + kWavesPerBlock * bpw - 1) / (kWavesPerBlock * bpw);
            if (bpw == 4)
                launch_timed(hook, 0, sweep_lanes<TI, TO, F16, kBlock, 4>, dim3(gb), dim3(kBlock), 0, s, d, lb, nlb,
                             sc.lane_rec_dev, sc.lane_ent_dev, in, o, st);
            else if (bpw == 2)
                launch_timed(hook, 0, sweep_lanes<TI, TO, F16, kBlock, 2>, dim3(gb), dim3(kBlock), 0, s, d, lb, nlb,
                             sc.lane_rec_dev, sc.lane_ent_dev, in, o, st);
            else
                launch_timed(hook, 0, sweep_lanes<TI, TO, F16>, dim3(gb), dim3(kBlock), 0, s, d, lb, nlb,
                             sc.lane_rec_dev, sc.lane_ent_dev, in, o, st);
            return;
        }
        if (g > kLevelMaxBlocks) g = kLevelMaxBlocks;
        launch_timed(hook, 0, sweep_levels<TI, TO, F16>, dim3(g), dim3(kBlock), 0, s, d, b0, b1, in, o, st, err);
    };
    if (sc.kind == 0) {
        launch_timed(hook, 3, sweep_dataflow<TI, TO, F16>, dim3(sc.grid), dim3(kBlock), 0, s, d, in, o, st,
                     sc.grid * kWavesPerBlock, err);
        check_launch("sweep_dataflow");
    } else if (sc.kind == 1) {
        for (int l = 0; l < sc.nlevels; ++l) level_launch(l);
        check_launch("sweep_levels");
    } else {
        static const bool attr = [] {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&sweep_tail<TO>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsBytes);
            return true;
        }();
        (void)attr;
        // Tail run k's batches [bf, bl) and its coupling values.
        auto run_of = [&](int k, int& bf, int& bl, tail_val<TO>*& tv) {
            bf = sc.level_first[sc.seg_l0[k]];
            bl = sc.level_first[sc.seg_l1[k]];
            const int base = sc.seg_base[k];
            if (bl <= bf || base < 0 || base + (bl - bf) > sc.tail_batches ||
                static_cast<size_t>(sc.tail_batches) * kWave * sizeof(tail_val<TO>) > sc.tail_val_bytes)
                throw std::runtime_error("apxchol_amd: tail run outside the tail layout");
            tv = static_cast<tail_val<TO>*>(sc.tail_val) + static_cast<size_t>(base) * kWave;
        };
        const bool can_split = sc.seg_split && sc.side && sc.ev_fork && sc.ev_join;
        const auto split_at = [&](int k) { return can_split && k < sc.nseg && sc.seg_tail[k] ? sc.seg_split[k] : -1; };
        // One wave per workgroup (a batch each), as the lane levels.
        const auto couple = [&](hipStream_t on, int bf, int bl, int plo, int phi, int cont, tail_val<TO>* tv) {
            launch_timed(hook, 2, sweep_couple<TI, TO, F16, kWave>, dim3(bl - bf), dim3(kWave), 0, on, d, bf, bl, plo,
                         phi, cont, in, static_cast<const TO*>(o), tv, st, 1, static_cast<const batch_state*>(nullptr));
        };
        for (int k = 0; k < sc.nseg; ++k) {
            if (!sc.seg_tail[k]) {
                for (int l = sc.seg_l0[k]; l < sc.seg_l1[k]; ++l) level_launch(l);
                continue;
            }
            const int p0 = sc.seg_p0[k], p1 = sc.seg_p1[k];
            int bf, bl;
            tail_val<TO>* tv;
            run_of(k, bf, bl, tv);
            const size_t lds = tail_lds_bytes(p1 - p0, bl - bf, sizeof(TO));
            if (lds > static_cast<size_t>(kTailLdsBytes)) throw std::runtime_error("apxchol_amd: tail run exceeds LDS");
            const int ps = split_at(k);
            if (ps > 0) {
                // The entries before ps were summed on the side stream while
                // the run before this one walked; add the rest.
                hip_check_launch(hipStreamWaitEvent(s, sc.ev_join, 0), "hipStreamWaitEvent");
                couple(s, bf, bl, ps, p0, 1, tv);
            } else {
                couple(s, bf, bl, 0, p0, 0, tv);
            }
            const int pn = split_at(k + 1);
            if (pn > 0) {
                // Everything before this run is final here: fork the next
                // run's coupling over the entries before this run.
                if (pn != p0) throw std::runtime_error("apxchol_amd: coupling split not at the run before");
                int nbf, nbl;
                tail_val<TO>* ntv;
                run_of(k + 1, nbf, nbl, ntv);
                hip_check_launch(hipEventRecord(sc.ev_fork, s), "hipEventRecord");
                hip_check_launch(hipStreamWaitEvent(sc.side, sc.ev_fork, 0), "hipStreamWaitEvent");
                couple(sc.side, nbf, nbl, 0, pn, 0, ntv);
                hip_check_launch(hipEventRecord(sc.ev_join, sc.side), "hipEventRecord");
            }
            launch_timed(hook, 1, sweep_tail<TO>, dim3(1), dim3(kTailBlock), static_cast<unsigned>(lds), s, sc.tail,
                         sc.seg_base[k], bl - bf, sc.seg_phases[k], p1 - p0, d.backward ? d.m - 1 - p0 : p0,
                         static_cast<int>(d.backward), o, static_cast<const tail_val<TO>*>(tv), st, 1,
                         static_cast<const batch_state*>(nullptr));
        }
        check_launch("sweep_hybrid");
    }
}

}  // namespace

int sweep_dataflow_grid(vec_prec p, bool fp16) {
    if (p == vec_prec::fp64) return fp16 ? dataflow_grid_for<double, double, true>() : dataflow_grid_for<double, double, false>();
    return fp16 ? dataflow_grid_for<float, float, true>() : dataflow_grid_for<float, float, false>();
}

int sweep_tail_max_rows(vec_prec p) {
    return kTailLdsBytes / static_cast<int>(p == vec_prec::fp64 ? sizeof(double) : sizeof(float)) - 1;
}

std::size_t sweep_tail_lds(vec_prec p, int rows, int batches) {
    return tail_lds_bytes(rows, batches, p == vec_prec::fp64 ? sizeof(double) : sizeof(float));
}

bool sweep_tail_fits(vec_prec p, int rows, int batches, int levels) {
    return levels <= kTailWalkMaxLevels &&
           tail_lds_bytes(rows, batches, p == vec_prec::fp64 ? sizeof(double) : sizeof(float)) <=
               static_cast<std::size_t>(kTailLdsBytes);
}

void launch_sweep(const sweep_desc& d, vec_prec p, bool fp16, const void* rhs, void* out, const pcg_state* st,
                  int* err, const sweep_sched& sc, hipStream_t s, const launch_hook& hook) {
    if (d.nb == 0) return;
    if (p == vec_prec::fp64) {
        if (fp16) sweep_t<double, double, true>(d, rhs, out, st, err, sc, s, hook);
        else sweep_t<double, double, false>(d, rhs, out, st, err, sc, s, hook);
    } else if (!d.backward) {
        if (fp16) sweep_t<double, float, true>(d, rhs, out, st, err, sc, s, hook);
        else sweep_t<double, float, false>(d, rhs, out, st, err, sc, s, hook);
    } else {
        if (fp16) sweep_t<float, float, true>(d, rhs, out, st, err, sc, s, hook);
        else sweep_t<float, float, false>(d, rhs, out, st, err, sc, s, hook);
    }
}

void launch_fill_outputs(vec_prec p, int n, int m, void* y, void* z, const pcg_state* st, hipStream_t s) {
    const int g = vec_blocks(n);
    if (p == vec_prec::fp64) fill_outputs<double><<<g, kBlock, 0, s>>>(n, m, static_cast<double*>(y), static_cast<double*>(z), st);
    else fill_outputs<float><<<g, kBlock, 0, s>>>(n, m, static_cast<float*>(y), static_cast<float*>(z), st);
    check_launch("fill_outputs");
}

// Stage 2 of a PCG reduction over the partials of an nb-workgroup producer.
void finish(int op, int n, const double* part, int nb, pcg_state* st, hipStream_t s) {
    pcg_finish<<<1, kBlock, 0, s>>>(op, n, part, nb, st);
    check_launch("pcg_finish");
}

template <class VT>
static void spmv_t(int n, const unsigned* ptr, const int* idx, const VT* val, int lanes, const double* x, double* y,
                   pcg_state* st, double* part, bool want_dot, hipStream_t s, const launch_hook& hook) {
    const int g = vec_blocks(static_cast<long long>(n) * lanes);
    const int w = want_dot ? 1 : 0;
    auto go = [&](auto kernel) {
        launch_timed(hook, 4, kernel, dim3(g), dim3(kBlock), 0, s, n, ptr, idx, val, x, y, st, part, w);
    };
    switch (lanes) {
        case 1: go(spmv_pAp<VT, 1>); break;
        case 2: go(spmv_pAp<VT, 2>); break;
        case 4: go(spmv_pAp<VT, 4>); break;
        case 8: go(spmv_pAp<VT, 8>); break;
        case 16: go(spmv_pAp<VT, 16>); break;
        case 32: go(spmv_pAp<VT, 32>); break;
        default: go(spmv_pAp<VT, 64>); break;
    }
    check_launch("spmv_pAp");
    if (want_dot) finish(kFinPAp, n, part, g, st, s);
}

void launch_spmv(int n, const unsigned* ptr, const int* idx, const void* val, bool val32, int lanes, const double* x,
                 double* y, pcg_state* st, double* part, bool want_dot, hipStream_t s, const launch_hook& hook) {
    if (want_dot && !st) throw std::invalid_argument("apxchol_amd: SpMV dot product without a solve state");
    if (val32) spmv_t<float>(n, ptr, idx, static_cast<const float*>(val), lanes, x, y, st, part, want_dot, s, hook);
    else spmv_t<double>(n, ptr, idx, static_cast<const double*>(val), lanes, x, y, st, part, want_dot, s, hook);
}

void launch_spmv_sell(const sell_desc& a, bool val32, const double* x, double* y, pcg_state* st, double* part,
                      bool want_dot, hipStream_t s, const launch_hook& hook) {
    if (want_dot && !st) throw std::invalid_argument("apxchol_amd: SpMV dot product without a solve state");
    const long long items = static_cast<long long>(a.nslices) + a.nlong;
    if (items <= 0) return;
    long long g = (items + kWavesPerBlock - 1) / kWavesPerBlock;
    if (g > a.max_blocks) g = a.max_blocks;
    g = std::max<long long>(kXcds, g / kXcds * kXcds);   // whole bands (spmv_sell)
    const int w = want_dot ? 1 : 0;
    auto go = [&](auto kernel) {
        launch_timed(hook, 4, kernel, dim3(static_cast<unsigned>(g)), dim3(kBlock), 0, s, a, x, y, st, part, w);
    };
    if (!a.rowmap) throw std::invalid_argument("apxchol_amd: sliced ELL without a row map");
    if (val32) go(spmv_sell<float>);
    else go(spmv_sell<double>);
    check_launch("spmv_sell");
    if (want_dot) finish(kFinPAp, a.n, part, static_cast<int>(g), st, s);
}

void launch_scatter_init(int n, const unsigned* perm, const double* b, const double* x0, double* r, double* x,
                         hipStream_t s) {
    scatter_init<<<vec_blocks(n), kBlock, 0, s>>>(n, perm, b, x0, r, x);
    check_launch("scatter_init");
}

void launch_norm2_b(int n, const double* b, pcg_state* st, double* part, hipStream_t s) {
    const int g = vec_blocks(n);
    norm2_b<<<g, kBlock, 0, s>>>(n, b, part);
    check_launch("norm2_b");
    finish(kFinBnorm, n, part, g, st, s);
}

void launch_init_residual(int n, double* r, const double* ax0, pcg_state* st, double* part, hipStream_t s) {
    const int g = vec_blocks(n);
    init_residual<<<g, kBlock, 0, s>>>(n, r, ax0, part);
    check_launch("init_residual");
    finish(ax0 ? kFinInitX0 : kFinInit, n, part, g, st, s);
}

void launch_update_xr(vec_prec p, int n, int m, double* x, const double* pv, double* r, const double* ap, void* y,
                      void* z, pcg_state* st, double* part, hipStream_t s) {
    const int g = vec_blocks(n);
    if (p == vec_prec::fp64)
        update_xr<double><<<g, kBlock, 0, s>>>(n, m, x, pv, r, ap, static_cast<double*>(y), static_cast<double*>(z), st, part);
    else
        update_xr<float><<<g, kBlock, 0, s>>>(n, m, x, pv, r, ap, static_cast<float*>(y), static_cast<float*>(z), st, part);
    check_launch("update_xr");
    finish(kFinXR, n, part, g, st, s);
}

void launch_dot_center(vec_prec p, int n, const double* r, const void* out, pcg_state* st, double* part, bool first,
                       hipStream_t s) {
    const int g = vec_blocks(n);
    if (p == vec_prec::fp64) dot_center<double><<<g, kBlock, 0, s>>>(n, r, static_cast<const double*>(out), st, part);
    else dot_center<float><<<g, kBlock, 0, s>>>(n, r, static_cast<const float*>(out), st, part);
    check_launch("dot_center");
    finish(first ? kFinDotFirst : kFinDot, n, part, g, st, s);
}

void launch_update_p(vec_prec p, int n, double* pv, const void* out, const pcg_state* st, bool first, hipStream_t s) {
    const int g = vec_blocks(n);
    if (p == vec_prec::fp64) update_p<double><<<g, kBlock, 0, s>>>(n, pv, static_cast<const double*>(out), st, first ? 1 : 0);
    else update_p<float><<<g, kBlock, 0, s>>>(n, pv, static_cast<const float*>(out), st, first ? 1 : 0);
    check_launch("update_p");
}

void launch_sum_vec(int n, const double* a, pcg_state* st, double* part, hipStream_t s) {
    const int g = vec_blocks(n);
    sum_vec<<<g, kBlock, 0, s>>>(n, a, part);
    check_launch("sum_vec");
    finish(kFinOutSum, n, part, g, st, s);
}

void launch_gather_shift(int n, const unsigned* perm, const double* src, double* dst, const pcg_state* st,
                         bool use_mean, hipStream_t s) {
    gather_shift<<<vec_blocks(n), kBlock, 0, s>>>(n, perm, src, dst, st, use_mean ? 1 : 0);
    check_launch("gather_shift");
}

void launch_sum_shift(int n, const double* r, pcg_state* st, double* part, bool center, hipStream_t s) {
    // No centring: shift = 0, the finish of an empty reduction.
    const int g = center ? vec_blocks(n) : 0;
    if (center) {
        sum_shift<<<g, kBlock, 0, s>>>(n, r, part);
        check_launch("sum_shift");
    }
    finish(kFinShift, n, part, g, st, s);
}

void launch_apply_scatter(vec_prec p, int n, int m, const unsigned* perm, const double* r, double* rp, void* y,
                          void* z, hipStream_t s) {
    const int g = vec_blocks(n);
    if (p == vec_prec::fp64) apply_scatter<double><<<g, kBlock, 0, s>>>(n, m, perm, r, rp, static_cast<double*>(y), static_cast<double*>(z));
    else apply_scatter<float><<<g, kBlock, 0, s>>>(n, m, perm, r, rp, static_cast<float*>(y), static_cast<float*>(z));
    check_launch("apply_scatter");
}

void launch_sum_out(vec_prec p, int n, const void* a, pcg_state* st, double* part, hipStream_t s) {
    const int g = vec_blocks(n);
    if (p == vec_prec::fp64) sum_out<double><<<g, kBlock, 0, s>>>(n, static_cast<const double*>(a), part);
    else sum_out<float><<<g, kBlock, 0, s>>>(n, static_cast<const float*>(a), part);
    check_launch("sum_out");
    finish(kFinOutSum, n, part, g, st, s);
}

void launch_apply_gather(vec_prec p, int n, const unsigned* perm, const void* out, double* z, const pcg_state* st,
                         bool center, hipStream_t s) {
    const int g = vec_blocks(n);
    if (p == vec_prec::fp64) apply_gather<double><<<g, kBlock, 0, s>>>(n, perm, static_cast<const double*>(out), z, st, center ? 1 : 0);
    else apply_gather<float><<<g, kBlock, 0, s>>>(n, perm, static_cast<const float*>(out), z, st, center ? 1 : 0);
    check_launch("apply_gather");
}

void launch_fill_pending(vec_prec p, void* a, int n, hipStream_t s) {
    const int g = vec_blocks(n);
    if (p == vec_prec::fp64) fill_pending<double><<<g, kBlock, 0, s>>>(static_cast<double*>(a), n, nullptr);
    else fill_pending<float><<<g, kBlock, 0, s>>>(static_cast<float*>(a), n, nullptr);
    check_launch("fill_pending");
}

// ---- blocked solves ---------------------------------------------------------

namespace {

template <class TI, class TO, bool F16>
void sweep_k_t(const sweep_desc& d, const block_sched& b, const void* rhs, void* out, const pcg_state* st,
               const batch_state* bs, hipStream_t s) {
    auto level_launch = [&](int l) {
        const int w0 = b.level_first[l], w1 = b.level_first[l + 1];
        if (w1 <= w0) return;
        const int gmax = b.level_gmax ? b.level_gmax[l] : 0;
        if (gmax < 0) {
            // No entries in the level: one load per (row, column).
            const int p0 = b.level_pos[l], p1 = b.level_pos[l + 1];
            sweep_block_diag<TI, TO><<<vec_blocks(static_cast<long long>(p1 - p0) * b.kp), kBlock, 0, s>>>(
                d, p0, p1, b.kp, static_cast<const TI*>(rhs), static_cast<TO*>(out), st, bs);
            return;
        }
        if (gmax >= 1 && gmax <= b.rows_max_g) {
            // One thread per (row, column).
            const int p0 = b.level_pos[l], p1 = b.level_pos[l + 1];
            const long long threads = static_cast<long long>(p1 - p0) * b.kp;
            const dim3 g(static_cast<unsigned>((threads + kBlock - 1) / kBlock));
            const TI* in = static_cast<const TI*>(rhs);
            TO* o = static_cast<TO*>(out);
            if (gmax == 1) sweep_block_rows<TI, TO, F16, 1><<<g, kBlock, 0, s>>>(d, p0, p1, b.kp, in, o, st, bs);
            else if (gmax == 2) sweep_block_rows<TI, TO, F16, 2><<<g, kBlock, 0, s>>>(d, p0, p1, b.kp, in, o, st, bs);
            else if (gmax == 4) sweep_block_rows<TI, TO, F16, 4><<<g, kBlock, 0, s>>>(d, p0, p1, b.kp, in, o, st, bs);
            else if (gmax == 8) sweep_block_rows<TI, TO, F16, 8><<<g, kBlock, 0, s>>>(d, p0, p1, b.kp, in, o, st, bs);
            else sweep_block_rows<TI, TO, F16, 16><<<g, kBlock, 0, s>>>(d, p0, p1, b.kp, in, o, st, bs);
            return;
        }
        const int g = (w1 - w0 + kWavesPerBlock - 1) / kWavesPerBlock;
        sweep_block<TI, TO, F16><<<g, kBlock, 0, s>>>(d, b.pos, b.mask, b.col0, w0, w1, b.kp,
                                                      static_cast<const TI*>(rhs), static_cast<TO*>(out), st, bs);
    };
    if (b.cols) {
        // The single-vector schedule, column by column: its launched levels
        // as blocked level launches, each of its tail runs as one coupling
        // launch over (batch, column) and one tail workgroup per column.
        const sweep_sched& sc = *b.cols;
        static const bool attr = [] {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&sweep_tail<TO>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsBytes);
            return true;
        }();
        (void)attr;
        for (int k = 0; k < sc.nseg; ++k) {
            if (!sc.seg_tail[k]) {
                for (int l = sc.seg_l0[k]; l < sc.seg_l1[k]; ++l) level_launch(l);
                continue;
            }
            const int p0 = sc.seg_p0[k], p1 = sc.seg_p1[k];
            const int bf = sc.level_first[sc.seg_l0[k]], bl = sc.level_first[sc.seg_l1[k]];
            const size_t lds = tail_lds_bytes(p1 - p0, bl - bf, sizeof(TO));
            const int base = sc.seg_base[k];
            if (bl <= bf || base < 0 || base + (bl - bf) > sc.tail_batches || lds > static_cast<size_t>(kTailLdsBytes) ||
                static_cast<size_t>(bl - bf) * b.kp * kWave * sizeof(tail_val<TO>) > b.col_tail_val_bytes)
                throw std::runtime_error("apxchol_amd: column tail run outside its tables");
            tail_val<TO>* tv = static_cast<tail_val<TO>*>(b.col_tail_val);
            const int g = (bl - bf + kWavesPerBlock - 1) / kWavesPerBlock;
            sweep_couple<TI, TO, F16><<<dim3(g, b.kp), kBlock, 0, s>>>(d, bf, bl, 0, p0, 0, static_cast<const TI*>(rhs),
                                                                     static_cast<const TO*>(out), tv, st, b.kp, bs);
            sweep_tail<TO><<<b.kp, kTailBlock, lds, s>>>(sc.tail, base, bl - bf, sc.seg_phases[k], p1 - p0,
                                                         d.backward ? d.m - 1 - p0 : p0, static_cast<int>(d.backward),
                                                         static_cast<TO*>(out), static_cast<const tail_val<TO>*>(tv),
                                                         st, b.kp, bs);
        }
    } else if (b.nseg == 0) {
        for (int l = 0; l < b.nlevels; ++l) level_launch(l);
    } else {
        static const bool attr = [] {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&sweep_block_tail<TI, TO, F16>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsBytes);
            return true;
        }();
        (void)attr;
        for (int k = 0; k < b.nseg; ++k) {
            const int l0 = b.seg_l0[k], l1 = b.seg_l1[k];
            if (!b.seg_tail[k]) {
                for (int l = l0; l < l1; ++l) level_launch(l);
                continue;
            }
            const int p0 = b.level_pos[l0], p1 = b.level_pos[l1];
            const size_t lds = static_cast<size_t>(p1 - p0) * b.kp * sizeof(TO);
            if (lds > static_cast<size_t>(kTailLdsBytes)) throw std::runtime_error("apxchol_amd: blocked tail run exceeds LDS");
            for (int l = l0; l < l1; ++l)
                if (b.level_first[l + 1] - b.level_first[l] > kBlockTailMaxWaves)
                    throw std::runtime_error("apxchol_amd: blocked tail level wider than the workgroup");
            sweep_block_tail<TI, TO, F16><<<1, kTailBlock, lds, s>>>(
                d, b.pos, b.mask, b.col0, b.level_first_dev, l0, l1, p0, b.kp, static_cast<const TI*>(rhs),
                static_cast<TO*>(out), st, bs);
        }
    }
    check_launch("sweep_block");
}

void check_kp(int kp) {
    if (kp < 1 || kp > 64 || (kp & (kp - 1)) != 0) throw std::invalid_argument("apxchol_amd: block width must be a power of two <= 64");
}

// Stage 2 of a blocked reduction over the partials of an nb-workgroup producer.
void finish_k(int op, int n, int kp, const double* part, int nb, pcg_state* st, const batch_state* bs, hipStream_t s) {
    block_finish<<<1, kBlock, 0, s>>>(op, n, kp, part, nb, st, bs);
    check_launch("block_finish");
}

}  // namespace

int sweep_block_tail_max_rows(vec_prec p, int kp) {
    return kTailLdsBytes / (kp * static_cast<int>(p == vec_prec::fp64 ? sizeof(double) : sizeof(float)));
}

void launch_sweep_k(const sweep_desc& d, vec_prec p, bool fp16, const block_sched& b, const void* rhs, void* out,
                    const pcg_state* st, const batch_state* bs, hipStream_t s) {
    check_kp(b.kp);
    if (p == vec_prec::fp64) {
        if (fp16) sweep_k_t<double, double, true>(d, b, rhs, out, st, bs, s);
        else sweep_k_t<double, double, false>(d, b, rhs, out, st, bs, s);
    } else if (!d.backward) {
        if (fp16) sweep_k_t<double, float, true>(d, b, rhs, out, st, bs, s);
        else sweep_k_t<double, float, false>(d, b, rhs, out, st, bs, s);
    } else {
        if (fp16) sweep_k_t<float, float, true>(d, b, rhs, out, st, bs, s);
        else sweep_k_t<float, float, false>(d, b, rhs, out, st, bs, s);
    }
}

void launch_spmm_k(int n, int kp, int lanes, const unsigned* ptr, const int* idx, const void* val, bool val32,
                   const double* x, double* y, pcg_state* st, batch_state* bs, double* part, bool want_dot,
                   hipStream_t s) {
    check_kp(kp);
    if (lanes < 1 || lanes > 64 || (lanes & (lanes - 1))) throw std::invalid_argument("apxchol_amd: SpMM lanes");
    const int g = spmm_grid(n, kp, lanes);
    if (val32)
        spmm_pAp_k<float><<<g, kBlock, 0, s>>>(n, kp, lanes, ptr, idx, static_cast<const float*>(val), x, y, bs, part,
                                               want_dot ? 1 : 0);
    else
        spmm_pAp_k<double><<<g, kBlock, 0, s>>>(n, kp, lanes, ptr, idx, static_cast<const double*>(val), x, y, bs,
                                                part, want_dot ? 1 : 0);
    check_launch("spmm_pAp_k");
    if (want_dot) finish_k(kBFinPAp, n, kp, part, g, st, bs, s);
}

void launch_spmm_sell_k(const sell_desc& a, bool val32, int kp, const double* x, double* y, pcg_state* st,
                        batch_state* bs, double* part, bool want_dot, hipStream_t s) {
    check_kp(kp);
    const long long items = (static_cast<long long>(a.nslices) + a.nlong) * kp;
    if (items <= 0) return;
    long long g = (items + kWavesPerBlock - 1) / kWavesPerBlock;
    if (g > vec_block_cap()) g = vec_block_cap();
    g = std::max<long long>(kXcds, g / kXcds * kXcds);
    const int w = want_dot ? 1 : 0;
    const dim3 grid(static_cast<unsigned>(g));
    if (!a.rowmap) throw std::invalid_argument("apxchol_amd: sliced ELL without a row map");
    // From 4 columns on, a workgroup's four waves share each slice's entries
    // through LDS (APXCHOL_AMD_SPMM_STAGED=0: every wave loads them itself).
    static const bool staged = [] {
        const char* e = std::getenv("APXCHOL_AMD_SPMM_STAGED");
        return !(e && *e && std::atoi(e) == 0);
    }();
    if (staged && kp >= kWavesPerBlock) {
        if (val32) spmm_sell_staged_k<float><<<grid, kBlock, 0, s>>>(a, kp, x, y, bs, part, w);
        else spmm_sell_staged_k<double><<<grid, kBlock, 0, s>>>(a, kp, x, y, bs, part, w);
        check_launch("spmm_sell_staged_k");
    } else {
        if (val32) spmm_sell_k<float><<<grid, kBlock, 0, s>>>(a, kp, x, y, bs, part, w);
        else spmm_sell_k<double><<<grid, kBlock, 0, s>>>(a, kp, x, y, bs, part, w);
        check_launch("spmm_sell_k");
    }
    if (!want_dot) return;
    int nb = static_cast<int>(g);
    if (a.nlong > 0) {
        spmm_long_dot_k<<<1, kBlock, 0, s>>>(a, kp, x, y, bs, part, nb);
        check_launch("spmm_long_dot_k");
        ++nb;
    }
    finish_k(kBFinPAp, a.n, kp, part, nb, st, bs, s);
}

void launch_scatter_init_k(int n, int k, int kp, const unsigned* perm, const double* B, double* r, double* x,
                           hipStream_t s) {
    scatter_init_k<<<vec_blocks(static_cast<long long>(n) * kp), kBlock, 0, s>>>(n, k, kp, perm, B, r, x);
    check_launch("scatter_init_k");
}

void launch_init_residual_k(int n, int kp, const double* r, pcg_state* st, batch_state* bs, double* part,
                            hipStream_t s) {
    const int g = block_grid(n, kp);
    init_residual_k<<<g, kBlock, 0, s>>>(n, kp, r, part);
    check_launch("init_residual_k");
    finish_k(kBFinInit, n, kp, part, g, st, bs, s);
}

void launch_update_xr_k(int n, int kp, double* x, const double* p, double* r, const double* ap, pcg_state* st,
                        batch_state* bs, double* part, hipStream_t s) {
    const int g = block_grid(n, kp);
    update_xr_k<<<g, kBlock, 0, s>>>(n, kp, x, p, r, ap, st, bs, part);
    check_launch("update_xr_k");
    finish_k(kBFinXR, n, kp, part, g, st, bs, s);
}

void launch_refresh_active_k(int kp, const pcg_state* st, batch_state* bs, hipStream_t s) {
    refresh_active_k<<<1, 64, 0, s>>>(kp, st, bs);
    check_launch("refresh_active_k");
}

void launch_dot_center_k(vec_prec p, int n, int kp, const double* r, const void* out, pcg_state* st, batch_state* bs,
                         double* part, bool first, hipStream_t s) {
    const int g = block_grid(n, kp);
    if (p == vec_prec::fp64)
        dot_center_k<double><<<g, kBlock, 0, s>>>(n, kp, r, static_cast<const double*>(out), st, bs, part);
    else
        dot_center_k<float><<<g, kBlock, 0, s>>>(n, kp, r, static_cast<const float*>(out), st, bs, part);
    check_launch("dot_center_k");
    finish_k(first ? kBFinDotFirst : kBFinDot, n, kp, part, g, st, bs, s);
}

void launch_update_p_k(vec_prec p, int n, int kp, double* pv, const void* out, const pcg_state* st, batch_state* bs,
                       bool first, hipStream_t s) {
    const int g = block_grid(n, kp);
    if (p == vec_prec::fp64)
        update_p_k<double><<<g, kBlock, 0, s>>>(n, kp, pv, static_cast<const double*>(out), st, bs, first ? 1 : 0);
    else
        update_p_k<float><<<g, kBlock, 0, s>>>(n, kp, pv, static_cast<const float*>(out), st, bs, first ? 1 : 0);
    check_launch("update_p_k");
}

void launch_zero_tail_k(vec_prec p, int n, int m, int kp, void* a, hipStream_t s) {
    if (n <= m) return;
    const int g = vec_blocks(static_cast<long long>(n - m) * kp);
    if (p == vec_prec::fp64) zero_tail_k<double><<<g, kBlock, 0, s>>>(n, m, kp, static_cast<double*>(a));
    else zero_tail_k<float><<<g, kBlock, 0, s>>>(n, m, kp, static_cast<float*>(a));
    check_launch("zero_tail_k");
}

void launch_sum_cols_k(vec_prec p, int n, int kp, const void* x, pcg_state* st, double* part, hipStream_t s) {
    const int g = block_grid(n, kp);
    if (p == vec_prec::fp64) sum_cols_k<double><<<g, kBlock, 0, s>>>(n, kp, static_cast<const double*>(x), part);
    else sum_cols_k<float><<<g, kBlock, 0, s>>>(n, kp, static_cast<const float*>(x), part);
    check_launch("sum_cols_k");
    finish_k(kBFinOutSum, n, kp, part, g, st, nullptr, s);
}

void launch_permute(int m, const unsigned* Q, const double* src, double* dst, bool in, hipStream_t s) {
    permute_vec<<<vec_blocks(m), kBlock, 0, s>>>(m, Q, src, dst, in ? 1 : 0);
    check_launch("permute_vec");
}

void launch_permute_out(vec_prec p, int m, const unsigned* Q, const void* src, double* dst, hipStream_t s) {
    if (p == vec_prec::fp64) permute_out<double><<<vec_blocks(m), kBlock, 0, s>>>(m, Q, static_cast<const double*>(src), dst);
    else permute_out<float><<<vec_blocks(m), kBlock, 0, s>>>(m, Q, static_cast<const float*>(src), dst);
    check_launch("permute_out");
}

void launch_end_iteration_k(int kp, pcg_state* st, batch_state* bs, hipStream_t s) {
    end_iteration_k<<<1, 64, 0, s>>>(kp, st, bs);
    check_launch("end_iteration_k");
}

void launch_admit_scatter_k(int n, int kp, const unsigned* perm, int ncol, const int* cols,
                            const double* const* srcs, double* r, double* x, hipStream_t s) {
    if (ncol <= 0) return;
    admit_scatter_k<<<vec_blocks(static_cast<long long>(n) * ncol), kBlock, 0, s>>>(n, kp, perm, ncol, cols, srcs, r, x);
    check_launch("admit_scatter_k");
}

void launch_gather_cols_k(int n, int kp, const unsigned* perm, int ncol, const int* cols, double* const* dsts,
                          const double* x, const pcg_state* st, bool use_mean, hipStream_t s) {
    if (ncol <= 0) return;
    gather_cols_k<<<vec_blocks(static_cast<long long>(n) * ncol), kBlock, 0, s>>>(n, kp, perm, ncol, cols, dsts, x, st,
                                                                                  use_mean ? 1 : 0);
    check_launch("gather_cols_k");
}

void launch_shift_from_sum_k(int kp, pcg_state* st, bool center, hipStream_t s) {
    shift_from_sum_k<<<1, 64, 0, s>>>(kp, st, center ? 1 : 0);
    check_launch("shift_from_sum_k");
}

void launch_gather_k(vec_prec p, int n, int k, int kp, const unsigned* perm, const void* x, double* X,
                     const pcg_state* st, bool use_mean, hipStream_t s) {
    const int g = vec_blocks(static_cast<long long>(n) * k);
    if (p == vec_prec::fp64)
        gather_k<double><<<g, kBlock, 0, s>>>(n, k, kp, perm, static_cast<const double*>(x), X, st, use_mean ? 1 : 0);
    else
        gather_k<float><<<g, kBlock, 0, s>>>(n, k, kp, perm, static_cast<const float*>(x), X, st, use_mean ? 1 : 0);
    check_launch("gather_k");
}

}  // namespace apxchol_amd

// ---- stand-alone PCG pieces (caller-owned partials) -------------------------

namespace apxchol_amd {

int pcg_blocks(long long units) { return vec_blocks(units); }

namespace {
template <class VT>
void piece_spmv_t(int n, const unsigned* ptr, const int* idx, const VT* val, int lanes, const double* x, double* y,
                  double* part, hipStream_t s) {
    const int g = vec_blocks(static_cast<long long>(n) * lanes);
    const pcg_state* none = nullptr;
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, dim3(g), dim3(kBlock), 0, s, n, ptr, idx, val, x, y, none, part, 1); };
    switch (lanes) {
        case 1: go(spmv_pAp<VT, 1>); break;
        case 2: go(spmv_pAp<VT, 2>); break;
        case 4: go(spmv_pAp<VT, 4>); break;
        case 8: go(spmv_pAp<VT, 8>); break;
        case 16: go(spmv_pAp<VT, 16>); break;
        case 32: go(spmv_pAp<VT, 32>); break;
        case 64: go(spmv_pAp<VT, 64>); break;
        default: throw std::invalid_argument("apxchol_amd: SpMV lanes must be a power of two in [1, 64]");
    }
    check_launch("spmv_pAp");
}
}  // namespace

void launch_piece_spmv_pAp(int n, const unsigned* ptr, const int* idx, const void* val, bool val32, int lanes,
                           const double* x, double* y, double* part, hipStream_t s) {
    if (n <= 0) return;
    if (val32) piece_spmv_t<float>(n, ptr, idx, static_cast<const float*>(val), lanes, x, y, part, s);
    else piece_spmv_t<double>(n, ptr, idx, static_cast<const double*>(val), lanes, x, y, part, s);
}

void launch_piece_update_xr(int n, double* x, const double* p, double* r, const double* ap, double alpha, double* part,
                            hipStream_t s) {
    if (n <= 0) return;
    piece_update_xr<<<vec_blocks(n), kBlock, 0, s>>>(n, x, p, r, ap, alpha, part);
    check_launch("piece_update_xr");
}

void launch_piece_dot(int n, const double* a, const double* b, double* part, hipStream_t s) {
    if (n <= 0) return;
    piece_dot<<<vec_blocks(n), kBlock, 0, s>>>(n, a, b, part);
    check_launch("piece_dot");
}

void launch_piece_update_p(int n, double* p, const double* z, double beta, hipStream_t s) {
    if (n <= 0) return;
    piece_update_p<<<vec_blocks(n), kBlock, 0, s>>>(n, p, z, beta);
    check_launch("piece_update_p");
}

void launch_piece_reduce(const double* part, int nb, double* out, hipStream_t s) {
    piece_reduce<<<1, kBlock, 0, s>>>(part, nb, out);
    check_launch("piece_reduce");
}

}  // namespace apxchol_amd

// ---- allocation-free y = A x with x.y (library-owned partials) ---------------
//
// The partials of the SpMV's reduction live in `scratch`, a file-scope device
// array: every GPU the library runs on holds its own copy, from module load to
// the end of the process, so the entry point below allocates nothing. On the
// host the name `scratch` is only a stand-in for the symbol (passing it to a
// kernel would hand the kernel a host address): host code uses the address
// hipGetSymbolAddress returns for the current device, looked up once per
// device on the first call there. Every call on one device shares that
// device's array and leaves the last call's partials in it; calls that
// overlap on one GPU (unordered streams) race. Nothing in it needs to start
// at zero: each workgroup of the producer writes its own partial.

namespace apxchol_amd {

inline constexpr int kScratchDoubles = kMaxBlocks;   // 16 KiB: the largest producer grid's partials
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
inline constexpr int kScratchDevices = 64;

double* scratch_partials() {
    int dev = -1;
    const hipError_t e0 = hipGetDevice(&dev);
    if (e0 != hipSuccess) throw std::runtime_error(std::string("apxchol_amd: hipGetDevice failed: ") + hipGetErrorString(e0));
    if (dev < 0 || dev >= kScratchDevices)
        throw std::invalid_argument("apxchol_amd: the allocation-free entry point supports device ordinals below 64");
    static std::mutex mu;
    static double* addr[kScratchDevices] = {};
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("apxchol_amd: hipGetSymbolAddress(scratch) failed: ") +
                                     hipGetErrorString(e));
        addr[dev] = static_cast<double*>(p);
    }
    return addr[dev];
}
}  // namespace

void launch_piece_spmv_pAp_scratch(int n, const unsigned* ptr, const int* idx, const void* val, bool val32, int lanes,
                                   const double* x, double* y, double* pAp, hipStream_t s) {
    double* part = scratch_partials();
    if (n <= 0) {   // no rows: x.y = 0
        launch_piece_reduce(part, 0, pAp, s);
        return;
    }
    const int nb = pcg_blocks(static_cast<long long>(n) * lanes);
    if (nb > kScratchDoubles)
        throw std::invalid_argument("apxchol_amd: " + std::to_string(nb) + " partials exceed the scratch array's " +
                                    std::to_string(kScratchDoubles));
    launch_piece_spmv_pAp(n, ptr, idx, val, val32, lanes, x, y, part, s);
    launch_piece_reduce(part, nb, pAp, s);
}

}  // namespace apxchol_amd

// ---- grouped launches ----------------------------------------------------------

namespace apxchol_amd {

int group_vec_blocks(int n) { return vec_blocks(n); }

int group_spmv_blocks(const group_member& g) {
    if (!g.use_sell) return vec_blocks(static_cast<long long>(g.n) * g.lanes);
    // launch_spmv_sell's grid.
    const long long items = static_cast<long long>(g.sell.nslices) + g.sell.nlong;
    long long b = (items + kWavesPerBlock - 1) / kWavesPerBlock;
    if (b > g.sell.max_blocks) b = g.sell.max_blocks;
    return static_cast<int>(std::max<long long>(kXcds, b / kXcds * kXcds));
}

int group_fin_op(const char* name) {
    const std::string k(name);
    if (k == "pAp") return kFinPAp;
    if (k == "xr") return kFinXR;
    if (k == "dot") return kFinDot;
    if (k == "dot_first") return kFinDotFirst;
    throw std::invalid_argument("apxchol_amd: unknown reduction stage " + k);
}

void launch_group_spmv(const group_member* mem, const int* starts, int count, int blocks, hipStream_t s) {
    if (count <= 0 || blocks <= 0) return;
    group_spmv<<<blocks, kBlock, 0, s>>>(mem, starts, count);
    check_launch("group_spmv");
}

void launch_group_finish(int op, const group_member* mem, const int* starts, int count, hipStream_t s) {
    if (count <= 0) return;
    group_finish<<<count, kBlock, 0, s>>>(op, mem, starts);
    check_launch("group_finish");
}

void launch_group_update_xr(vec_prec p, const group_member* mem, const int* starts, int count, int blocks,
                            hipStream_t s) {
    if (count <= 0 || blocks <= 0) return;
    if (p == vec_prec::fp64) group_update_xr<double><<<blocks, kBlock, 0, s>>>(mem, starts, count);
    else group_update_xr<float><<<blocks, kBlock, 0, s>>>(mem, starts, count);
    check_launch("group_update_xr");
}

void launch_group_fill_outputs(vec_prec p, const group_member* mem, const int* starts, int count, int blocks,
                               hipStream_t s) {
    if (count <= 0 || blocks <= 0) return;
    if (p == vec_prec::fp64) group_fill_outputs<double><<<blocks, kBlock, 0, s>>>(mem, starts, count);
    else group_fill_outputs<float><<<blocks, kBlock, 0, s>>>(mem, starts, count);
    check_launch("group_fill_outputs");
}

void launch_group_dot_center(vec_prec p, const group_member* mem, const int* starts, int count, int blocks,
                             hipStream_t s) {
    if (count <= 0 || blocks <= 0) return;
    if (p == vec_prec::fp64) group_dot_center<double><<<blocks, kBlock, 0, s>>>(mem, starts, count);
    else group_dot_center<float><<<blocks, kBlock, 0, s>>>(mem, starts, count);
    check_launch("group_dot_center");
}

void launch_group_update_p(vec_prec p, const group_member* mem, const int* starts, int count, int blocks, bool first,
                           hipStream_t s) {
    if (count <= 0 || blocks <= 0) return;
    if (p == vec_prec::fp64) group_update_p<double><<<blocks, kBlock, 0, s>>>(mem, starts, count, first ? 1 : 0);
    else group_update_p<float><<<blocks, kBlock, 0, s>>>(mem, starts, count, first ? 1 : 0);
    check_launch("group_update_p");
}

void launch_group_sweep_level(vec_prec p, bool fp16, bool backward, const group_unit* units, const int* starts,
                              int count, int blocks, hipStream_t s) {
    if (count <= 0 || blocks <= 0) return;
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), 0, s, units, starts, count); };
    if (p == vec_prec::fp64) {
        if (fp16) go(group_sweep_level<double, double, true>);
        else go(group_sweep_level<double, double, false>);
    } else if (!backward) {
        if (fp16) go(group_sweep_level<double, float, true>);
        else go(group_sweep_level<double, float, false>);
    } else {
        if (fp16) go(group_sweep_level<float, float, true>);
        else go(group_sweep_level<float, float, false>);
    }
    check_launch("group_sweep_level");
}

void launch_group_sweep_couple(vec_prec p, bool fp16, bool backward, const group_unit* units, const int* starts,
                               int count, int blocks, hipStream_t s) {
    if (count <= 0 || blocks <= 0) return;
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kWave), 0, s, units, starts, count); };
    if (p == vec_prec::fp64) {
        if (fp16) go(group_sweep_couple<double, double, true>);
        else go(group_sweep_couple<double, double, false>);
    } else if (!backward) {
        if (fp16) go(group_sweep_couple<double, float, true>);
        else go(group_sweep_couple<double, float, false>);
    } else {
        if (fp16) go(group_sweep_couple<float, float, true>);
        else go(group_sweep_couple<float, float, false>);
    }
    check_launch("group_sweep_couple");
}

void launch_group_sweep_tail(vec_prec p, const group_unit* units, int count, std::size_t lds, hipStream_t s) {
    if (count <= 0) return;
    if (lds > static_cast<std::size_t>(kTailLdsBytes)) throw std::runtime_error("apxchol_amd: group tail run exceeds LDS");
    if (p == vec_prec::fp64) {
        static const bool attr = [] {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&group_sweep_tail<double>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsBytes);
            return true;
        }();
        (void)attr;
        group_sweep_tail<double><<<count, kTailBlock, static_cast<unsigned>(lds), s>>>(units);
    } else {
        static const bool attr = [] {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&group_sweep_tail<float>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsBytes);
            return true;
        }();
        (void)attr;
        group_sweep_tail<float><<<count, kTailBlock, static_cast<unsigned>(lds), s>>>(units);
    }
    check_launch("group_sweep_tail");
}

}  // namespace apxchol_amd
