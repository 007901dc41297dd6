// Triangular sweeps of the preconditioner z = L^-T L^-1 r on gfx950.
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

// Right-hand side of `row`: forward reads r - shift (the Laplacian input
// centring), back reads y scaled by inv_scale^2 on the fp16 storage.
template <class TI, class AT>
__device__ __forceinline__ AT row_rhs(const sweep_desc& d, const TI* rhs, int row, double shift) {
    if (!d.backward) return static_cast<AT>(static_cast<double>(rhs[row]) - shift);
    if (d.inv_scale) {
        const double r = static_cast<double>(d.inv_scale[row]);
        return static_cast<AT>(static_cast<double>(rhs[row]) * (r * r));
    }
    return static_cast<AT>(rhs[row]);
}

// Process batch b. Returns false if a dependency never arrived within the
// spin limit (the row is then written from whatever was gathered and the
// caller raises the error flag).
template <class TI, class TO, bool F16>
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
        for (unsigned c0 = 0; c0 < len; c0 += kWave * kEpl) {
            int jj[kEpl];
            float vv[kEpl];
            AT xv[kEpl];
            unsigned pend = 0;
#pragma unroll
            for (int k = 0; k < kEpl; ++k) {
                const unsigned q = c0 + static_cast<unsigned>(lane + k * kWave);
                jj[k] = 0; vv[k] = 0.0f; xv[k] = 0;
                if (q < len) { load_entry<F16>(d, rb + q, jj[k], vv[k]); pend |= 1u << k; }
            }
            unsigned spins = 0;
            for (;;) {
#pragma unroll
                for (int k = 0; k < kEpl; ++k)
                    if (pend & (1u << k)) {
                        const auto bits = load_flag(out + jj[k]);
                        if (bits != flag_bits<TO>::pending) { xv[k] = from_bits<TO>(bits); pend &= ~(1u << k); }
                    }
                if (__ballot(pend != 0) == 0) break;
                if (++spins > kSpinLimit || ((spins & 1023u) == 0 && sweep_failed(err))) { ok = false; break; }
                __builtin_amdgcn_s_sleep(1);
            }
#pragma unroll
            for (int k = 0; k < kEpl; ++k) part += static_cast<AT>(vv[k]) * xv[k];
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) part += __shfl_xor(part, off, kWave);
        if (lane == 0) {
            const AT r = row_rhs<TI, AT>(d, rhs, row, shift);
            store_flag(out + row, (r - part) / static_cast<AT>(d.diag[row]));
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
        jj[k] = 0; vv[k] = 0.0f; xv[k] = 0;
        if (active && q < len) { load_entry<F16>(d, rb + q, jj[k], vv[k]); pend |= 1u << k; }
    }
    AT r = 0;
    AT dg = 1;
    if (active && sub == 0) {
        r = row_rhs<TI, AT>(d, rhs, row, shift);
        dg = static_cast<AT>(d.diag[row]);
    }
    bool done = !active;
    unsigned spins = 0;
    for (;;) {
        if (!done) {
#pragma unroll
            for (int k = 0; k < kEpl; ++k)
                if (pend & (1u << k)) {
                    const auto bits = load_flag(out + jj[k]);
                    if (bits != flag_bits<TO>::pending) { xv[k] = from_bits<TO>(bits); pend &= ~(1u << k); }
                }
        }
        const unsigned long long ready = __ballot(pend == 0);
        const bool go = !done && (ready & gmask) == gmask;
        AT part = 0;
        if (go)
#pragma unroll
            for (int k = 0; k < kEpl; ++k) part += static_cast<AT>(vv[k]) * xv[k];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const AT o = __shfl_xor(part, off, kWave);
            if (off < G) part += o;
        }
        if (go) {
            if (sub == 0) store_flag(out + row, (r - part) / dg);
            done = true;
        }
        if (__ballot(!done) == 0) break;
        if (++spins > kSpinLimit || ((spins & 1023u) == 0 && sweep_failed(err))) {
            ok = false;
            if (!done && sub == 0) store_flag(out + row, static_cast<AT>(0));
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
    for (;;) {
        int b = 0;
        if (lane == 0) b = atomicAdd(&d.ctrl[0], 1);
        b = __shfl(b, 0, kWave);
        if (b >= d.nb) break;
        const bool bok = sweep_batch<TI, TO, F16>(d, b, rhs, out, shift, err);
        if (!bok && lane == 0) atomicOr(err, 1);
        ok = ok && bok;
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
__global__ __launch_bounds__(kBlock) void sweep_levels(sweep_desc d, int b0, int b1, const TI* __restrict__ rhs,
                                                       TO* out, const pcg_state* st, int* err) {
    if (st && st->done) return;
    const double shift = (st && !d.backward) ? st->shift : 0.0;
    const int waves = static_cast<int>(gridDim.x) * kWavesPerBlock;
    const int w = static_cast<int>(blockIdx.x) * kWavesPerBlock + static_cast<int>(threadIdx.x) / kWave;
    bool ok = true;
    for (int b = b0 + w; b < b1; b += waves) ok = sweep_batch<TI, TO, F16>(d, b, rhs, out, shift, err) && ok;
    if ((threadIdx.x & (kWave - 1)) == 0 && !ok) atomicOr(err, 1);
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
