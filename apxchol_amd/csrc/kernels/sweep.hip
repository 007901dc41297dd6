// Triangular sweeps of the preconditioner z = L^-T L^-1 r on gfx950.
//
// Compiled with -ffp-contract=off: every fused multiply-add is written out
// (fma), so the schedules below run literally the same arithmetic per row.
//
// One routine, sweep_batch, processes a BATCH of consecutive rows (in sweep
// order) with one wavefront: every row owns an aligned group of G lanes
// (G = pow2ceil(len / kEpl), 1..64), each lane keeps its kEpl (index, value)
// pairs and the gathered neighbour values in registers, and the row's sum is
// a fixed lane-strided order plus a fixed xor butterfly -- so a row's result
// depends on the row alone, never on which wave ran it or when.
//
// Readiness is carried by the output vector itself: before a sweep every
// output slot holds a NaN sentinel (kPending*), and a row is published by ONE
// agent-scope atomic store of its value. A lane polls the slots it gathers
// with agent-scope atomic loads until none is the sentinel. No flag array, no
// fence and no value atomics: the result is bit-deterministic.
//
// Two schedules share sweep_batch (and so produce identical bits):
//   * sweep_dataflow: one persistent launch; waves claim batches in sweep order
//     from an atomic ticket. A wave only ever waits on rows of smaller tickets,
//     which are held by running waves, so the sweep cannot deadlock; a bounded
//     spin (kSpinLimit polls per batch) turns a broken dependency into an
//     error flag instead of a hang.
//   * sweep_levels: the level-set fallback, one launch per level (a range of
//     batches; batches never straddle a level), static batch assignment.
//
// The reference's counterpart is the CUDA dataflow kernel
// (src/cuda_dataflow.cu:114-478) driven by cuda_sptrsv::solve_LLt_dev_impl
// (include/apxchol/solver/sptrsv/cuda.h:736-785); this design differs in the
// flag scheme (value sentinel instead of an epoch-tagged {fp32,epoch} word, so
// fp64 vectors work), in the 64-lane batches and in the fixed batch table.
#include "common.hpp"

namespace apxchol_amd {

// Polls of one batch before it gives up (error flag); every 1024 polls a
// waiting wave also checks whether another wave has already given up.
inline constexpr unsigned kSpinLimit = 1u << 20;

__device__ __forceinline__ bool sweep_failed(const int* err) {
    return __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

__device__ __forceinline__ int lanes_for(unsigned len) {
    unsigned g = (len + kEpl - 1) / kEpl;
    if (g <= 1) return 1;
    g = 1u << (32 - __builtin_clz(g - 1));
    return g > 64 ? 64 : static_cast<int>(g);
}

template <bool F16>
__device__ __forceinline__ void load_entry(const sweep_desc& d, unsigned p, int& j, float& v) {
    if constexpr (F16) {
        j = d.idx[p];
        v = static_cast<float>(__builtin_bit_cast(_Float16, d.h[p]));
    } else {
        const entry32 e = d.ent[p];
        j = e.idx;
        v = e.val;
    }
}

// One polling round over a lane's kEpl gathered slots: all loads are issued
// back to back and waited for once; a slot whose value has arrived leaves the
// pending mask. A lane's valid entries are a prefix of its slots, and the
// slots past it repeat the first entry's address, so no lane ever polls an
// address it does not need (no shared hot spot).
template <class TO>
__device__ __forceinline__ void poll(const TO* out, const int (&jj)[kEpl], TO (&xv)[kEpl], unsigned& pend) {
    typename flag_bits<TO>::type bits[kEpl];
#pragma unroll
    for (int k = 0; k < kEpl; ++k) bits[k] = load_flag(out + jj[k]);
#pragma unroll
    for (int k = 0; k < kEpl; ++k) {
        const bool arrived = ((pend >> k) & 1u) && bits[k] != flag_bits<TO>::pending;
        if (arrived) xv[k] = from_bits<TO>(bits[k]);
        pend &= arrived ? ~(1u << k) : ~0u;
    }
}

// Gather without waiting: plain (cacheable) loads of complete values; slots
// outside the lane's valid prefix stay 0 (their address may hold a value not
// computed yet, which must not reach the sum even times a zero weight).
template <class TO>
__device__ __forceinline__ void gather(const TO* out, const int (&jj)[kEpl], TO (&xv)[kEpl], unsigned valid) {
    TO g[kEpl];
#pragma unroll
    for (int k = 0; k < kEpl; ++k) g[k] = out[jj[k]];
#pragma unroll
    for (int k = 0; k < kEpl; ++k) xv[k] = ((valid >> k) & 1u) ? g[k] : static_cast<TO>(0);
}

// A long row's chain as sweep_batch walks it (lane l: entries l, l + 64,
// l + 128, ... in order, fma from 0), kLongUnroll chunks of 64 kEpl entries
// per round: a round's entry loads, then its gathers, are in flight together
// (the order of the fmas is the chunk-by-chunk one). value(j) reads column
// j; an entry with use(j) false adds nothing, and neither does a slot past
// the row's end (fma(0, 0, part) = part: a chain never holds -0).
// long_row_chain_from continues a chain: from `part`, starting at entry c_begin
// (a multiple of 64 kLongUnroll kEpl, so every entry keeps its lane and its
// place in the lane's order), and stops at the first round whose first entry
// satisfies past(j) (entries ascend in sweep position, so with past(j) = the
// entry lies at or after some position, nothing later is used either).
// Skipped entries are ones use(j) rejects, so the chain is the one a walk of
// the whole row gives.
inline constexpr int kLongUnroll = 4;
inline constexpr unsigned kLongRound = static_cast<unsigned>(kWave * kLongUnroll * kEpl);
template <bool F16, class TO, class V, class U, class P>
__device__ __forceinline__ TO long_row_chain_from(const sweep_desc& d, unsigned rb, unsigned len, V value, U use,
                                                  TO part, unsigned c_begin, P past) {
    const int lane = threadIdx.x & (kWave - 1);
    constexpr int R = kLongUnroll * kEpl;
    for (unsigned c0 = c_begin; c0 < len; c0 += kLongRound) {
        int jj[R];
        float vv[R];
        TO xv[R];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const unsigned q = c0 + static_cast<unsigned>(lane + i * kWave);
            load_entry<F16>(d, q < len ? rb + q : rb, jj[i], vv[i]);
        }
        // Lane 0's first slot holds the round's first entry (c0 < len).
        if (past(__builtin_amdgcn_readfirstlane(jj[0]))) break;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const bool ok = c0 + static_cast<unsigned>(lane + i * kWave) < len && use(jj[i]);
            xv[i] = ok ? value(jj[i]) : static_cast<TO>(0);
            if (!ok) vv[i] = 0.0f;
        }
#pragma unroll
        for (int i = 0; i < R; ++i) part = fma(static_cast<TO>(vv[i]), xv[i], part);
    }
    return part;
}
template <bool F16, class TO, class V, class U>
__device__ __forceinline__ TO long_row_chain(const sweep_desc& d, unsigned rb, unsigned len, V value, U use) {
    return long_row_chain_from<F16, TO>(d, rb, len, value, use, static_cast<TO>(0), 0u, [](int) { return false; });
}

// First entry of a row whose sweep position is at least plo (the row's
// entries ascend in sweep position), by a 64-ary search over the wave: each
// step probes 64 evenly spaced entries of the interval still open; len when
// there is none. Every lane of the wave takes part and gets the answer.
template <bool F16>
__device__ __forceinline__ unsigned first_entry_at(const sweep_desc& d, unsigned rb, unsigned len, int plo) {
    const unsigned lane = threadIdx.x & (kWave - 1);
    unsigned lo = 0, hi = len;   // entries before lo lie before plo; the answer is in [lo, hi]
    while (lo < hi) {
        const unsigned step = (hi - lo + kWave - 1) / kWave;
        const unsigned q = lo + lane * step;
        bool below = false;
        if (q < hi) {
            int j;
            float v;
            load_entry<F16>(d, rb + q, j, v);
            below = (d.backward ? d.m - 1 - j : j) < plo;
        }
        const unsigned k = static_cast<unsigned>(__popcll(__ballot(below)));
        if (k == 0) break;
        const unsigned last_below = lo + (k - 1) * step;
        const unsigned next = lo + k * step;
        lo = last_below + 1;
        hi = next < hi ? next : hi;
    }
    return lo;
}

// Publish a row: an agent-scope atomic store when other waves poll for it,
// a plain store when readers are ordered behind a launch or barrier.
template <class TO, bool WAIT>
__device__ __forceinline__ void publish(TO* p, TO v) {
    if constexpr (WAIT) store_flag(p, v);
    else *p = v;
}

// Right-hand side of `row`: forward reads r - shift (the Laplacian input
// centring), back reads y scaled by inv_scale^2 on the fp16 storage.
// `at` is the element of rhs that holds the row (row * k + column for a
// block of k vectors stored row by row).
template <class TI, class AT>
__device__ __forceinline__ AT row_rhs(const sweep_desc& d, const TI* rhs, int row, size_t at, double shift) {
    if (!d.backward) return static_cast<AT>(static_cast<double>(rhs[at]) - shift);
    if (d.inv_scale) {
        const double r = static_cast<double>(d.inv_scale[row]);
        return static_cast<AT>(static_cast<double>(rhs[at]) * (r * r));
    }
    return static_cast<AT>(rhs[at]);
}
template <class TI, class AT>
__device__ __forceinline__ AT row_rhs(const sweep_desc& d, const TI* rhs, int row, double shift) {
    return row_rhs<TI, AT>(d, rhs, row, static_cast<size_t>(row), shift);
}

// Process batch b. Returns false if a dependency never arrived within the
// spin limit (the row is then written from whatever was gathered and the
// caller raises the error flag).
// WAIT: poll the gathered slots until they leave the sentinel (dataflow
// schedule); otherwise every dependency is known to be complete (an earlier
// launch, or an earlier level behind a workgroup barrier) and one plain load
// per slot suffices. The arithmetic is the same either way.
template <class TI, class TO, bool F16, bool WAIT>
__device__ bool sweep_batch(const sweep_desc& d, int b, const TI* __restrict__ rhs, TO* out, double shift,
                            const int* err) {
    using AT = TO;
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned p0 = d.batch_pos[b], p1 = d.batch_pos[b + 1];
    const unsigned long long mask = d.batch_mask[b];
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long mu = mask & upto;
    const int rib = __popcll(mu) - 1;
    const int start = 63 - __clzll(mu);
    const int sub = lane - start;
    const unsigned pos = p0 + static_cast<unsigned>(rib);
    const bool has_row = pos < p1;
    const int row = has_row ? (d.backward ? d.m - 1 - static_cast<int>(pos) : static_cast<int>(pos)) : 0;
    const unsigned rb = has_row ? d.ptr[row] : 0u;
    const unsigned len = has_row ? d.ptr[row + 1] - rb : 0u;
    const int G = lanes_for(len);
    const bool active = has_row && sub < G;
    const unsigned long long gmask = (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << start;
    bool ok = true;

    if (__builtin_expect(len > static_cast<unsigned>(kWave * kEpl), 0)) {
        // A long row alone in its batch (G = 64): walk it in chunks of
        // 64 * kEpl entries; each lane accumulates its entries chunk by chunk.
        AT part = 0;
        if constexpr (!WAIT)
            part = long_row_chain<F16, AT>(d, rb, len, [&](int j) { return out[j]; }, [](int) { return true; });
        for (unsigned c0 = 0; WAIT && c0 < len; c0 += kWave * kEpl) {
            int jj[kEpl];
            float vv[kEpl];
            AT xv[kEpl];
            unsigned pend = 0;
#pragma unroll
            for (int k = 0; k < kEpl; ++k) {
                const unsigned q = c0 + static_cast<unsigned>(lane + k * kWave);
                const bool valid = q < len;
                load_entry<F16>(d, valid ? rb + q : rb + c0, jj[k], vv[k]);
                if (!valid) vv[k] = 0.0f;
                xv[k] = 0;
                pend |= valid ? (1u << k) : 0u;
            }
            if constexpr (WAIT) {
                unsigned spins = 0;
                for (;;) {
                    if (pend) poll<TO>(out, jj, xv, pend);
                    if (__ballot(pend != 0) == 0) break;
                    if (++spins > kSpinLimit || ((spins & 1023u) == 0 && sweep_failed(err))) { ok = false; break; }
                    __builtin_amdgcn_s_sleep(1);
                }
            } else {
                gather<TO>(out, jj, xv, pend);
            }
#pragma unroll
            for (int k = 0; k < kEpl; ++k) part = fma(static_cast<AT>(vv[k]), xv[k], part);
        }
        part = group_tree_sum(part, kWave);
        if (lane == 0) {
            const AT r = row_rhs<TI, AT>(d, rhs, row, shift);
            publish<TO, WAIT>(out + row, finish_row<AT>(r - part, diag_operand<AT>(d, row)));
        }
        return ok;
    }

    int jj[kEpl];
    float vv[kEpl];
    AT xv[kEpl];
    unsigned pend = 0;
#pragma unroll
    for (int k = 0; k < kEpl; ++k) {
        const unsigned q = static_cast<unsigned>(sub + k * G);
        const bool valid = active && q < len;
        load_entry<F16>(d, valid ? rb + q : (active && len > 0 ? rb : 0u), jj[k], vv[k]);
        if (!valid) vv[k] = 0.0f;
        xv[k] = 0;
        pend |= valid ? (1u << k) : 0u;
    }
    AT r = 0;
    float dg = 1.0f;
    if (active && sub == 0) {
        r = row_rhs<TI, AT>(d, rhs, row, shift);
        dg = diag_operand<AT>(d, row);
    }
    if constexpr (!WAIT) {
        // Dependencies complete: gather once, then the same per-row sum.
        if (active) gather<TO>(out, jj, xv, pend);
        AT part = 0;
#pragma unroll
        for (int k = 0; k < kEpl; ++k) part = fma(static_cast<AT>(vv[k]), xv[k], part);
        part = group_tree_sum(part, G);
        if (active && sub == 0) publish<TO, WAIT>(out + row, finish_row<AT>(r - part, dg));
        return true;
    }
    bool done = !active;
    unsigned spins = 0;
    for (;;) {
        if (pend) poll<TO>(out, jj, xv, pend);
        const unsigned long long ready = __ballot(pend == 0);
        const bool go = !done && (ready & gmask) == gmask;
        AT part = 0;
        if (go)
#pragma unroll
            for (int k = 0; k < kEpl; ++k) part = fma(static_cast<AT>(vv[k]), xv[k], part);
        part = group_tree_sum(part, G);
        if (go) {
            if (sub == 0) publish<TO, WAIT>(out + row, finish_row<AT>(r - part, dg));
            done = true;
        }
        if (__ballot(!done) == 0) break;
        if (++spins > kSpinLimit || ((spins & 1023u) == 0 && sweep_failed(err))) {
            ok = false;
            if (!done && sub == 0) publish<TO, WAIT>(out + row, static_cast<AT>(0));
            break;
        }
        if (!go) __builtin_amdgcn_s_sleep(1);
    }
    return ok;
}

template <class TI, class TO, bool F16>
__global__ __launch_bounds__(kBlock) void sweep_dataflow(sweep_desc d, const TI* __restrict__ rhs, TO* out,
                                                         const pcg_state* st, int total_waves, int* err) {
    if (st && st->done) return;
    const double shift = (st && !d.backward) ? st->shift : 0.0;
    const int lane = threadIdx.x & (kWave - 1);
    bool ok = true;
    // A wave claims d.chunk consecutive batches per ticket and runs them in
    // order, so it only ever waits on batches it already finished or on
    // smaller tickets held by running waves.
    for (;;) {
        int b0 = 0;
        if (lane == 0) b0 = atomicAdd(&d.ctrl[0], d.chunk);
        b0 = __shfl(b0, 0, kWave);
        if (b0 >= d.nb) break;
        const int b1 = min(b0 + d.chunk, d.nb);
        for (int b = b0; b < b1; ++b) {
            const bool bok = sweep_batch<TI, TO, F16, true>(d, b, rhs, out, shift, err);
            if (!bok && lane == 0) atomicOr(err, 1);
            ok = ok && bok;
        }
    }
    if (lane == 0) {
        const int prev = atomicAdd(&d.ctrl[1], 1);
        if (prev == total_waves - 1) {
            __hip_atomic_store(&d.ctrl[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&d.ctrl[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <class TI, class TO, bool F16>
__device__ __forceinline__ void sweep_levels_body(int vb, int vg, const sweep_desc& d, int b0, int b1,
                                                  const TI* __restrict__ rhs, TO* out, const pcg_state* st,
                                                  int* err) {
    if (st && st->done) return;
    const double shift = (st && !d.backward) ? st->shift : 0.0;
    const int waves = vg * kWavesPerBlock;
    const int w = vb * kWavesPerBlock + static_cast<int>(threadIdx.x) / kWave;
    for (int b = b0 + w; b < b1; b += waves) sweep_batch<TI, TO, F16, false>(d, b, rhs, out, shift, err);
}
template <class TI, class TO, bool F16>
__global__ __launch_bounds__(kBlock) void sweep_levels(sweep_desc d, int b0, int b1, const TI* __restrict__ rhs,
                                                       TO* out, const pcg_state* st, int* err) {
    sweep_levels_body<TI, TO, F16>(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), d, b0, b1, rhs, out,
                                   st, err);
}

// A launched level none of whose rows has an entry (the first forward level:
// the rows that depend on nothing): a row's sum is then the empty chain's +0,
// so its output is finish_row(rhs - 0) -- sweep_batch's bits for it, read
// with one load per row instead of the batch tables, row pointers, padding
// entries and gathers. Sweep positions [p0, p1), grid-strided.
template <class TI, class TO>
__global__ __launch_bounds__(kBlock) void sweep_diag_level(sweep_desc d, int p0, int p1, const TI* __restrict__ rhs,
                                                           TO* out, const pcg_state* st) {
    if (st && st->done) return;
    const double shift = (st && !d.backward) ? st->shift : 0.0;
    const int stride = static_cast<int>(gridDim.x) * kBlock;
    for (int t = p0 + static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x); t < p1; t += stride) {
        const int row = d.backward ? d.m - 1 - t : t;
        const TO r = row_rhs<TI, TO>(d, rhs, row, shift);
        out[row] = finish_row<TO>(r - static_cast<TO>(0), diag_operand<TO>(d, row));
    }
}

// One launched level from the static lane layout: wave w runs lane batches
// lb0 + w * BPW .. + BPW - 1 (BPW batches per wave: all their layout loads,
// then all their gathers, are in flight together, and a big level needs BPW
// times fewer waves). The layout loads are issued before the kernel reads the
// solve state, so the completion test does not add a dependent load; the
// per-row arithmetic is sweep_batch's (same slots, fma chain and butterfly),
// so the bits equal the other schedules'.
template <class TI, class TO, bool F16, int BS = kBlock, int BPW = 1>
__device__ __forceinline__ void sweep_lanes_body(int vb, const sweep_desc& d, int lb0, int nlb,
                                                 const lane_rec* __restrict__ rec, const entry32* __restrict__ ent,
                                                 const TI* __restrict__ rhs, TO* out, const pcg_state* st) {
    using AT = TO;
    const int w = vb * (BS / kWave) + static_cast<int>(threadIdx.x) / kWave;
    const int first = w * BPW;
    if (first >= nlb) return;
    const int lane = threadIdx.x & (kWave - 1);
    lane_rec rc[BPW];
    entry32 e[BPW][kEpl];
#pragma unroll
    for (int i = 0; i < BPW; ++i) {
        // A wave's batches past the level's end reload its first batch and
        // store nothing.
        const size_t b = static_cast<size_t>(lb0 + (first + i < nlb ? first + i : first));
        rc[i] = rec[b * kWave + lane];
#pragma unroll
        for (int k = 0; k < kEpl; ++k) e[i][k] = ent[(b * kEpl + k) * kWave + lane];
    }
    asm volatile("" ::: "memory");
    if (st && st->done) return;
    const double shift = (st && !d.backward) ? st->shift : 0.0;
    AT xv[BPW][kEpl];
#pragma unroll
    for (int i = 0; i < BPW; ++i) {
        const int nv = static_cast<int>((rc[i].info >> 9) & 7u);
#pragma unroll
        for (int k = 0; k < kEpl; ++k) xv[i][k] = k < nv ? out[e[i][k].idx] : static_cast<AT>(0);
    }
#pragma unroll
    for (int i = 0; i < BPW; ++i) {
        const int row = rc[i].row;
        const int G = 1 << (rc[i].info & 7u);
        const int sub = static_cast<int>((rc[i].info >> 3) & 63u);
        const bool head = row >= 0 && sub == 0 && first + i < nlb;
        AT r = 0;
        float dg = 1.0f;
        if (head) {
            r = row_rhs<TI, AT>(d, rhs, row, shift);
            dg = diag_operand<AT>(d, row);
        }
        AT part = 0;
#pragma unroll
        for (int k = 0; k < kEpl; ++k) part = fma(static_cast<AT>(e[i][k].val), xv[i][k], part);
        part = group_tree_sum(part, G);
        if (head) out[row] = finish_row<AT>(r - part, dg);
    }
}
template <class TI, class TO, bool F16, int BS = kBlock, int BPW = 1>
__global__ __launch_bounds__(BS) void sweep_lanes(sweep_desc d, int lb0, int nlb, const lane_rec* __restrict__ rec,
                                                  const entry32* __restrict__ ent, const TI* __restrict__ rhs,
                                                  TO* out, const pcg_state* st) {
    sweep_lanes_body<TI, TO, F16, BS, BPW>(static_cast<int>(blockIdx.x), d, lb0, nlb, rec, ent, rhs, out, st);
}

// A run of small levels [l0, l1) -- sweep positions [p0, p1) -- walked by ONE
// workgroup of kTailBlock threads. The run's outputs are kept in LDS as well as
// written out, so a dependency inside the run is an LDS read after the level
// barrier. Dependencies before the run are folded in beforehand by
// sweep_couple, which also stages the rows' right-hand sides; the rest of a
// batch comes from the static tail layout (tail_desc), whose addresses depend
// on (batch, lane) only. A wave therefore keeps its next two batches' loads in
// flight while it works, across the level barriers, which wait for LDS
// traffic only. The per-row arithmetic is sweep_batch's (same chain order, fma
// and butterfly), so the result is bit-identical to the other schedules.
inline constexpr int kTailLdsBytes = 160 * 1024;
// LDS of a tail run of `rows` rows and `batches` batches: the run's outputs
// and the zero slot (rounded up to 8 bytes), the run's walk (per batch: the
// same wave's next batch and that batch's level) and every wave's first
// batch and its level.
inline constexpr std::size_t tail_slot_bytes(int rows, std::size_t vsz) {
    return ((static_cast<std::size_t>(rows) + 1) * vsz + 7) / 8 * 8;
}
inline constexpr std::size_t tail_lds_bytes(int rows, int batches, std::size_t vsz) {
    return tail_slot_bytes(rows, vsz) + static_cast<std::size_t>(batches) * 2 * sizeof(int) +
           kTailWaves * 2 * sizeof(int);
}

// Workgroup barrier ordering LDS only: outstanding global loads stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Lane -> row mapping of batch b (as in sweep_batch).
struct lane_row {
    int row, pos, G, sub;
    unsigned len, rb;
    bool has_row, active;
};
__device__ __forceinline__ lane_row map_lane(const sweep_desc& d, int b) {
    const int lane = threadIdx.x & (kWave - 1);
    lane_row x;
    const unsigned q0 = d.batch_pos[b], q1 = d.batch_pos[b + 1];
    const unsigned long long mask = d.batch_mask[b];
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long mu = mask & upto;
    const int rib = __popcll(mu) - 1;
    x.sub = lane - (63 - __clzll(mu));
    const unsigned pos = q0 + static_cast<unsigned>(rib);
    x.has_row = pos < q1;
    x.pos = static_cast<int>(pos);
    x.row = x.has_row ? (d.backward ? d.m - 1 - static_cast<int>(pos) : static_cast<int>(pos)) : 0;
    x.rb = x.has_row ? d.ptr[x.row] : 0u;
    x.len = x.has_row ? d.ptr[x.row + 1] - x.rb : 0u;
    x.G = lanes_for(x.len);
    x.active = x.has_row && x.sub < x.G;
    return x;
}

// Coupling of a run (batches [bfirst, blast), positions from p0): for every
// lane of every batch, its fma chain over the entries that point before the
// run (complete when this kernel runs), in sweep_batch's slot order, and the
// row's right-hand side for the lane that writes it. Entries of a row are in
// ascending sweep position, so these are a prefix of every lane's chain.
// Column c = blockIdx.y of a block of kp vectors stored row by row (kp = 1:
// one vector; st then holds the one solve's state, else one per column, and
// bs, when given, stops the launch once no column is active). tv points at
// the run's first tail batch; column c's values start nb * 64 further on per
// column (nb = blast - bfirst).
// The chain may be computed in two launches: the first (cont = 0) over the
// entries before a position plo' < p0 with plo = 0, the second (cont = 1)
// continuing each lane's stored value over the entries in [plo, p0) with
// plo = plo'. A lane's entries before plo precede its entries in [plo, p0)
// (ascending sweep position), so the two give the one-launch chain.
template <class TI, class TO, bool F16, int BS = kBlock>
__device__ __forceinline__ void sweep_couple_body(int vb, int c, const sweep_desc& d, int bfirst, int blast, int plo,
                                                  int p0, int cont, const TI* __restrict__ rhs, const TO* out,
                                                  tail_val<TO>* tv, const pcg_state* st, int kp,
                                                  const batch_state* bs) {
    if (bs ? bs->active == 0 : (st && st->done)) return;
    const double shift = (st && !d.backward) ? st[c].shift : 0.0;
    const int lane = threadIdx.x & (kWave - 1);
    const int b = bfirst + vb * (BS / kWave) + static_cast<int>(threadIdx.x) / kWave;
    if (b >= blast) return;
    const lane_row x = map_lane(d, b);
    const auto value = [&](int j) { return out[static_cast<size_t>(j) * kp + c]; };
    const size_t at = (static_cast<size_t>(c) * (blast - bfirst) + (b - bfirst)) * kWave + lane;
    TO part = cont ? tv[at].part0 : static_cast<TO>(0);
    const auto pos_of = [&](int j) { return d.backward ? d.m - 1 - j : j; };
    const auto before = [&](int j) {
        const int q = pos_of(j);
        return q >= plo && q < p0;
    };
    if (p0 > plo && x.len > static_cast<unsigned>(kWave * kEpl)) {
        // Rounds wholly before plo are skipped, and the walk ends at the
        // first round that starts at or past p0.
        const unsigned e0 = plo > 0 ? first_entry_at<F16>(d, x.rb, x.len, plo) : 0u;
        part = long_row_chain_from<F16, TO>(d, x.rb, x.len, value, before, part, e0 / kLongRound * kLongRound,
                                            [&](int j) { return pos_of(j) >= p0; });
    } else if (p0 > plo) {
        int jj[kEpl];
        float vv[kEpl];
        TO g[kEpl];
        unsigned use = 0;
#pragma unroll
        for (int k = 0; k < kEpl; ++k) {
            const unsigned q = static_cast<unsigned>(x.sub + k * x.G);
            const bool v = x.active && q < x.len;
            load_entry<F16>(d, v ? x.rb + q : (x.len > 0 ? x.rb : 0u), jj[k], vv[k]);
            use |= (v && before(jj[k])) ? (1u << k) : 0u;
        }
#pragma unroll
        for (int k = 0; k < kEpl; ++k) g[k] = value(jj[k]);
#pragma unroll
        for (int k = 0; k < kEpl; ++k)
            if ((use >> k) & 1u) part = fma(static_cast<TO>(vv[k]), g[k], part);
    }
    TO r = 0;
    if (x.active && x.sub == 0)
        r = row_rhs<TI, TO>(d, rhs, x.row, static_cast<size_t>(x.row) * kp + c, shift);
    tail_val<TO> o;
    o.part0 = part;
    o.r = r;
    tv[at] = o;
}
template <class TI, class TO, bool F16, int BS = kBlock>
__global__ __launch_bounds__(BS) void sweep_couple(sweep_desc d, int bfirst, int blast, int plo, int p0, int cont,
                                                   const TI* __restrict__ rhs, const TO* out, tail_val<TO>* tv,
                                                   const pcg_state* st, int kp, const batch_state* bs) {
    sweep_couple_body<TI, TO, F16, BS>(static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), d, bfirst, blast,
                                       plo, p0, cont, rhs, out, tv, st, kp, bs);
}

typedef int i32x4 __attribute__((ext_vector_type(4)));
template <class T>
using vec2 = T __attribute__((ext_vector_type(2)));

// By value: a bit_cast of an ext-vector element lvalue is not reliable here
// (clang read element 0 for .y).
__device__ __forceinline__ float as_float(int v) { return __builtin_bit_cast(float, v); }

// One lane's share of one tail batch, as whole 16-byte vectors: each load
// lands in one register tuple that is consumed in place.
template <class TO>
struct tail_item {
    i32x4 ln;                // tail_lane
    i32x4 e0, e1, e2, e3;    // tail_pair x kTailPairs
    vec2<TO> v;              // tail_val
    int o;                   // the batch's first overflow row (tail_desc.ovf)
};

__device__ __forceinline__ size_t tail_at(int t, int k) {
    return (static_cast<size_t>(t) * kTailPairs + static_cast<size_t>(k)) * kWave + (threadIdx.x & (kWave - 1));
}

// t: the tail batch; tv: its run's coupling values, read at batch q of the run.
template <class TO>
__device__ __forceinline__ void tail_fetch(const tail_desc& td, const tail_val<TO>* tv, int t, int q,
                                           tail_item<TO>& it) {
    const size_t i = static_cast<size_t>(t) * kWave + (threadIdx.x & (kWave - 1));
    const i32x4* ent = reinterpret_cast<const i32x4*>(td.ent);
    // fp32 outputs take the lanes whose diagonal operand is 1 / diag (finish_row).
    it.ln = reinterpret_cast<const i32x4*>(sizeof(TO) == sizeof(float) ? td.lane_recip : td.lane)[i];
    it.e0 = ent[tail_at(t, 0)];
    it.e1 = ent[tail_at(t, 1)];
    it.e2 = ent[tail_at(t, 2)];
    it.e3 = ent[tail_at(t, 3)];
    it.v = reinterpret_cast<const vec2<TO>*>(tv)[static_cast<size_t>(q) * kWave + (threadIdx.x & (kWave - 1))];
    it.o = td.ovf[t];
}

template <class TO>
__device__ __forceinline__ TO tail_pair_fma(const i32x4 e, const TO* lds, TO part) {
    const TO x0 = lds[e.x], x1 = lds[e.z];
    part = fma(static_cast<TO>(as_float(e.y)), x0, part);
    return fma(static_cast<TO>(as_float(e.w)), x1, part);
}

// Row sums of tail batch t from the item and the run's LDS values. Padding
// entries are (zero slot, 0): fma(0, 0, part) is sweep_batch's empty slot, so
// every lane runs the same fmas.
template <class TO>
__device__ __forceinline__ void tail_step(const tail_desc& td, int t, const tail_item<TO>& it, TO* lds) {
    const int lane = threadIdx.x & (kWave - 1);
    const int c = it.ln.z >> 8;
    const int G = it.ln.z & 0xff;
    // A row longer than 64 * 2 * kTailPairs has further entries per lane,
    // read eight at a time from the overflow table: the first group's loads
    // are issued before the item's fmas, each next group's before the
    // current one is consumed. The fmas run in chain order (item, then
    // overflow).
    const int nov = c - 2 * kTailPairs;
    const size_t o = static_cast<size_t>(it.o);
    int s = 0;
    entry32 e[8];
    if (__builtin_expect(s < nov, 0))
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (s + u < nov) e[u] = td.ovf_ent[(o + static_cast<size_t>(s + u)) * kWave + lane];
    TO part = it.v.x;
    part = tail_pair_fma(it.e0, lds, part);
    part = tail_pair_fma(it.e1, lds, part);
    part = tail_pair_fma(it.e2, lds, part);
    part = tail_pair_fma(it.e3, lds, part);
    if (__builtin_expect(nov > 0, 0)) {
        while (s < nov) {
            const int sn = s + 8;
            entry32 nx[8];
            if (sn < nov)
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (sn + u < nov) nx[u] = td.ovf_ent[(o + static_cast<size_t>(sn + u)) * kWave + lane];
            TO g[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) g[u] = s + u < nov ? lds[e[u].idx] : static_cast<TO>(0);
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (s + u < nov) part = fma(static_cast<TO>(e[u].val), g[u], part);
#pragma unroll
            for (int u = 0; u < 8; ++u) e[u] = nx[u];
            s = sn;
        }
    }
    part = group_tree_sum(part, G);
    if (it.ln.x >= 0) lds[it.ln.y] = finish_row<TO>(it.v.y - part, as_float(it.ln.w));
}

// base: the run's first tail batch; nb: its batches; nlev: its barrier
// phases (a level may hold up to kTailMaxBatches batches; the run's walk
// table deals each wave its batches, in level order -- by default round
// robin over the whole run; a phase is one level, or a chain of one-batch
// levels that one wave walks with no barrier between them, its LDS accesses
// completing in order); zslot: the
// run's row count (its LDS holds zslot + 1 values); output row of slot s:
// row0 + s (forward) or row0 - s (back). Outputs reach HBM in one coalesced
// copy at the end.
// Column c = blockIdx.x of a block of kp vectors stored row by row (kp = 1:
// one vector); tv: the run's coupling values, nb * 64 per column; bs, when
// given, stops the launch once no column is active (else st->done does).
template <class TO>
__device__ __forceinline__ void sweep_tail_body(int col, unsigned char* tail_lds, const tail_desc& td, int base,
                                                int nb, int nlev, int zslot, int row0, int backward, TO* out,
                                                const tail_val<TO>* tv, const pcg_state* st, int kp,
                                                const batch_state* bs) {
    if (bs ? bs->active == 0 : (st && st->done)) return;
    tv += static_cast<size_t>(col) * nb * kWave;
    TO* lds = reinterpret_cast<TO*>(tail_lds);
    // Past the outputs and the zero slot: the run's walk, as (next batch of
    // the wave, its phase) pairs, and every wave's first batch and phase.
    // The walk is read from LDS, so its lookups never wait behind the item
    // loads in flight (global loads complete in order), and one read gives
    // both halves of a step.
    int2* walk = reinterpret_cast<int2*>(tail_lds + tail_slot_bytes(zslot, sizeof(TO)));
    int2* first = walk + nb;
    const int tid = static_cast<int>(threadIdx.x);
    if (tid < kTailWaves) first[tid] = make_int2(-1, nlev);
    // The zero slot past the run's outputs, read by padding entries; written
    // once, before the barrier that precedes the walk.
    if (tid == 0) lds[zslot] = TO(0);
    lds_barrier();
    for (int q = tid; q < nb; q += kTailBlock) {
        const tail_walk x = td.walk[base + q];
        const int nl = x.next < 0 ? nlev : (td.walk[base + x.next].info & (kTailWalkMaxLevels - 1));
        walk[q] = make_int2(x.next, nl);
        if (x.info & kTailWalkFirst)
            first[(x.info >> kTailWalkWaveShift) & (kTailWaves - 1)] = make_int2(q, x.info & (kTailWalkMaxLevels - 1));
    }
    lds_barrier();
    // Wave w walks its batches in order, passing the phase barriers between
    // them. Two items alternate: after the batch in one is computed, the one
    // two batches ahead is fetched into it, so each item's loads have the
    // next batch's step and every barrier in between to land (no register
    // copy of an item waits on its loads). Every fetch is issued (clamped to
    // the run's last batch), so the compiler's memory counters see the same
    // loads on every path. Cursors are batch indices inside the run (-1: the
    // walk is over) and wave-uniform (readfirstlane of LDS values read at one
    // address), so the walk stays scalar.
    const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
    // (next batch, its phase) after batch c (c = -1: the walk is over).
    const auto step_of = [&](int c) {
        if (c < 0) return make_int2(-1, nlev);
        const int2 e = walk[c];
        return make_int2(__builtin_amdgcn_readfirstlane(e.x), __builtin_amdgcn_readfirstlane(e.y));
    };
    const auto run_batch = [&](int c) { return c < 0 ? nb - 1 : c; };
    const auto batch_of = [&](int c) { return base + run_batch(c); };
    int2 f = first[w];
    int ca = __builtin_amdgcn_readfirstlane(f.x), la = __builtin_amdgcn_readfirstlane(f.y);
    int2 nb2 = step_of(ca);
    int cb = nb2.x, lb = nb2.y;
    tail_item<TO> ia, ib;
    tail_fetch<TO>(td, tv, batch_of(ca), run_batch(ca), ia);
    tail_fetch<TO>(td, tv, batch_of(cb), run_batch(cb), ib);
    int lvl = 0;   // the phase whose rows are being computed
    // Compute the batch in `it` (cursor a), shift the cursors and fetch the
    // batch two ahead into `it`. The walk entry is read after the step: read
    // before it, its result registers were shared with the overflow loads,
    // and the wait that sharing needs sat ahead of the barrier loop, so a
    // wave held every barrier up to its batch until its own loads landed.
    // One exit and every fetch on every pass (a finished walk stays at -1),
    // so the item loads in flight are the same wherever the passes meet.
    const auto turn = [&](tail_item<TO>& it) {
        if (la < nlev) {
            while (lvl < la) {
                lds_barrier();
                ++lvl;
            }
            tail_step<TO>(td, batch_of(ca), it, lds);
        }
        const int2 nx = step_of(cb);
        ca = cb;
        la = lb;
        cb = nx.x;
        lb = nx.y;
        tail_fetch<TO>(td, tv, batch_of(cb), run_batch(cb), it);
    };
    while (la < nlev) {
        turn(ia);
        turn(ib);
    }
    while (lvl < nlev) {
        lds_barrier();
        ++lvl;
    }
    for (int q = tid; q < zslot; q += kTailBlock)
        out[static_cast<size_t>(backward ? row0 - q : row0 + q) * kp + col] = lds[q];
}
template <class TO>
__global__ __launch_bounds__(kTailBlock) void sweep_tail(tail_desc td, int base, int nb, int nlev, int zslot,
                                                         int row0, int backward, TO* out, const tail_val<TO>* tv,
                                                         const pcg_state* st, int kp, const batch_state* bs) {
    extern __shared__ unsigned char tail_lds[];
    sweep_tail_body<TO>(static_cast<int>(blockIdx.x), tail_lds, td, base, nb, nlev, zslot, row0, backward, out, tv,
                        st, kp, bs);
}

// Fill a vector with the pending sentinel (the state every sweep output must
// start from).
template <class T>
__global__ __launch_bounds__(kBlock) void fill_pending(T* a, int n, const pcg_state* st) {
    if (st && st->done) return;
    using B = typename flag_bits<T>::type;
    B* ab = reinterpret_cast<B*>(a);
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) ab[i] = flag_bits<T>::pending;
}

}  // namespace apxchol_amd
