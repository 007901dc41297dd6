// Grouped launches: the PCG iteration of several factors ("members") in one
// launch per kernel kind. Every member keeps its own arrays (operator, factor
// layouts, vectors, partials) and its own solve state; a grouped launch is the
// concatenation of the members' own grids, and workgroup b runs workgroup
// b - starts[i] of member unit i (the unit whose range holds b) through the
// same kernel body the member's own launch runs, with the member's own grid
// size. A member's arithmetic -- every row, every partial, every fixed-order
// reduction -- is therefore the bits of its own solve, and members that have
// converged return at once from their units as they do from their own
// launches. Members run the hybrid schedule, so no sweep output is given the
// pending sentinel.
#include "common.hpp"

namespace apxchol_amd {

// The unit holding virtual workgroup b: starts[0..count] ascending,
// starts[0] = 0, starts[count] = the launch's grid.
__device__ __forceinline__ int group_unit_of(const int* __restrict__ starts, int count, int b) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class VT>
__device__ __forceinline__ void group_spmv_csr(int vb, int vg, const group_member& g) {
    const VT* val = static_cast<const VT*>(g.oval);
    switch (g.lanes) {
        case 1: spmv_pAp_body<VT, 1>(vb, vg, g.n, g.optr, g.oidx, val, g.p, g.ap, g.st, g.part, 1); break;
        case 2: spmv_pAp_body<VT, 2>(vb, vg, g.n, g.optr, g.oidx, val, g.p, g.ap, g.st, g.part, 1); break;
        case 4: spmv_pAp_body<VT, 4>(vb, vg, g.n, g.optr, g.oidx, val, g.p, g.ap, g.st, g.part, 1); break;
        case 8: spmv_pAp_body<VT, 8>(vb, vg, g.n, g.optr, g.oidx, val, g.p, g.ap, g.st, g.part, 1); break;
        case 16: spmv_pAp_body<VT, 16>(vb, vg, g.n, g.optr, g.oidx, val, g.p, g.ap, g.st, g.part, 1); break;
        case 32: spmv_pAp_body<VT, 32>(vb, vg, g.n, g.optr, g.oidx, val, g.p, g.ap, g.st, g.part, 1); break;
        default: spmv_pAp_body<VT, 64>(vb, vg, g.n, g.optr, g.oidx, val, g.p, g.ap, g.st, g.part, 1); break;
    }
}

// Ap = A p and the partials of p.Ap, every member with its own SpMV form.
__global__ __launch_bounds__(kBlock) void group_spmv(const group_member* __restrict__ mem,
                                                     const int* __restrict__ starts, int count) {
    const int b = static_cast<int>(blockIdx.x);
    const int i = group_unit_of(starts, count, b);
    const group_member& g = mem[i];
    const int vb = b - starts[i], vg = starts[i + 1] - starts[i];
    if (g.use_sell) {
        if (g.val32) spmv_sell_body<float>(vb, vg, g.sell, g.p, g.ap, g.st, g.part, 1);
        else spmv_sell_body<double>(vb, vg, g.sell, g.p, g.ap, g.st, g.part, 1);
    } else {
        if (g.val32) group_spmv_csr<float>(vb, vg, g);
        else group_spmv_csr<double>(vb, vg, g);
    }
}

// Stage 2 of a reduction for every member (workgroup i: member i; its
// producer's grid is starts[i + 1] - starts[i]).
__global__ __launch_bounds__(kBlock) void group_finish(int op, const group_member* __restrict__ mem,
                                                       const int* __restrict__ starts) {
    const int i = static_cast<int>(blockIdx.x);
    const group_member& g = mem[i];
    pcg_finish_body(op, g.n, g.part, starts[i + 1] - starts[i], g.st);
}

template <class TO>
__global__ __launch_bounds__(kBlock) void group_update_xr(const group_member* __restrict__ mem,
                                                          const int* __restrict__ starts, int count) {
    const int b = static_cast<int>(blockIdx.x);
    const int i = group_unit_of(starts, count, b);
    const group_member& g = mem[i];
    update_xr_body<TO>(b - starts[i], starts[i + 1] - starts[i], g.n, g.m, g.x, g.p, g.r, g.ap,
                       static_cast<TO*>(nullptr), static_cast<TO*>(g.z), g.st, g.part);
}

template <class TO>
__global__ __launch_bounds__(kBlock) void group_fill_outputs(const group_member* __restrict__ mem,
                                                             const int* __restrict__ starts, int count) {
    const int b = static_cast<int>(blockIdx.x);
    const int i = group_unit_of(starts, count, b);
    const group_member& g = mem[i];
    fill_outputs_body<TO>(b - starts[i], starts[i + 1] - starts[i], g.n, g.m, static_cast<TO*>(nullptr),
                          static_cast<TO*>(g.z), g.st);
}

template <class TO>
__global__ __launch_bounds__(kBlock) void group_dot_center(const group_member* __restrict__ mem,
                                                           const int* __restrict__ starts, int count) {
    const int b = static_cast<int>(blockIdx.x);
    const int i = group_unit_of(starts, count, b);
    const group_member& g = mem[i];
    dot_center_body<TO>(b - starts[i], starts[i + 1] - starts[i], g.n, g.r, static_cast<const TO*>(g.z), g.st,
                        g.part);
}

template <class TO>
__global__ __launch_bounds__(kBlock) void group_update_p(const group_member* __restrict__ mem,
                                                         const int* __restrict__ starts, int count, int first) {
    const int b = static_cast<int>(blockIdx.x);
    const int i = group_unit_of(starts, count, b);
    const group_member& g = mem[i];
    update_p_body<TO>(b - starts[i], starts[i + 1] - starts[i], g.n, g.p, static_cast<const TO*>(g.z), g.st, first);
}

// One step of the members' sweeps: launched levels from the lane layout or
// from the batch tables (kBlock threads), coupling launches (one wave per
// workgroup) and tail runs (one kTailBlock workgroup per unit).
template <class TI, class TO, bool F16>
__global__ __launch_bounds__(kBlock) void group_sweep_level(const group_unit* __restrict__ units,
                                                            const int* __restrict__ starts, int count) {
    const int b = static_cast<int>(blockIdx.x);
    const int i = group_unit_of(starts, count, b);
    const group_unit& u = units[i];
    const int vb = b - starts[i];
    if (u.kind == kGroupLane)
        sweep_lanes_body<TI, TO, F16>(vb, u.d, u.lb0, u.nlb, u.rec, u.ent, static_cast<const TI*>(u.rhs),
                                      static_cast<TO*>(u.out), u.st);
    else
        sweep_levels_body<TI, TO, F16>(vb, starts[i + 1] - starts[i], u.d, u.b0, u.b1, static_cast<const TI*>(u.rhs),
                                       static_cast<TO*>(u.out), u.st, u.err);
}

template <class TI, class TO, bool F16>
__global__ __launch_bounds__(kWave) void group_sweep_couple(const group_unit* __restrict__ units,
                                                            const int* __restrict__ starts, int count) {
    const int b = static_cast<int>(blockIdx.x);
    const int i = group_unit_of(starts, count, b);
    const group_unit& u = units[i];
    sweep_couple_body<TI, TO, F16, kWave>(b - starts[i], 0, u.d, u.b0, u.b1, 0, u.p0, 0, static_cast<const TI*>(u.rhs),
                                          static_cast<const TO*>(u.out), static_cast<tail_val<TO>*>(u.tv), u.st, 1,
                                          nullptr);
}

template <class TO>
__global__ __launch_bounds__(kTailBlock) void group_sweep_tail(const group_unit* __restrict__ units) {
    extern __shared__ unsigned char group_tail_lds[];
    const group_unit& u = units[blockIdx.x];
    sweep_tail_body<TO>(0, group_tail_lds, u.td, u.base, u.b1 - u.b0, u.nlev, u.zslot, u.row0, u.d.backward,
                        static_cast<TO*>(u.out), static_cast<const tail_val<TO>*>(u.tv), u.st, 1, nullptr);
}

}  // namespace apxchol_amd
