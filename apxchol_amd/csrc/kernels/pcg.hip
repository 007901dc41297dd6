// GPU-resident PCG kernels (gfx950). Vectors live in the factor's permuted
// space; every reduction is two-stage and fixed-order (block_partials in the
// producer, pcg_finish after it), and every scalar the recurrence needs stays
// on the device in pcg_state. Each kernel returns at once when the solve has
// finished (st->done), so the host can enqueue (or graph-replay) several
// iterations and test completion rarely.
//
// Reference counterparts: src/cuda_pcg_kernels.cu:18-152 (spmv_pAp,
// update_xr, dot, update_p, reduce_partials) driven by cuda_pcg::solve
// (include/apxchol/solver/pcg_cuda.h:157-261); the loop semantics -- centring
// schedule, stagnation stop, warm start -- are the CPU solver's
// (src/solve.cpp:478-654, include/apxchol/solver/preconditioner.h:344-529).
#include "common.hpp"

namespace apxchol_amd {

// y = A x over the permuted full-symmetric CSR, G lanes per row (lane-strided
// entries + xor butterfly). With want_dot, the partials of x.y go to part;
// pcg_finish (kFinPAp) sets st->pAp and ends the solve on a non-positive pAp
// (solve.cpp:596, breakdown).
// vb / vg: the workgroup's index and the grid (a grouped launch passes the
// workgroup's index inside its segment's own grid); the same in every kernel
// body below.
template <class VT, int G>
__device__ __forceinline__ void spmv_pAp_body(int vb, int vg, int n, const unsigned* __restrict__ ptr,
                                              const int* __restrict__ idx, const VT* __restrict__ val,
                                              const double* __restrict__ x, double* __restrict__ y,
                                              const pcg_state* st, double* part, int want_dot) {
    if (st && st->done) return;
    __shared__ double sh[kWavesPerBlock * 2];
    const int tid = vb * kBlock + threadIdx.x;
    const int sub = tid & (G - 1);
    const int stride = vg * kBlock / G;
    double acc = 0.0;
    for (int row = tid / G; row < n; row += stride) {
        const unsigned b = ptr[row], e = ptr[row + 1];
        double s = 0.0;
        for (unsigned p = b + sub; p < e; p += G) s = fma(static_cast<double>(val[p]), x[idx[p]], s);
#pragma unroll
        for (int off = G / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off, G);
        if (sub == 0) {
            y[row] = s;
            acc = fma(x[row], s, acc);
        }
    }
    if (!want_dot) return;
    double v[1] = {acc};
    block_partials<1>(v, part, sh, vb);
}
template <class VT, int G>
__global__ __launch_bounds__(kBlock) void spmv_pAp(int n, const unsigned* __restrict__ ptr,
                                                   const int* __restrict__ idx, const VT* __restrict__ val,
                                                   const double* __restrict__ x, double* __restrict__ y,
                                                   const pcg_state* st, double* part, int want_dot) {
    spmv_pAp_body<VT, G>(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), n, ptr, idx, val, x, y, st, part,
                         want_dot);
}

// y = A x over the sliced-ELL form (sell_desc): one lane per row, a slice's
// entries column-major so every entry column is one coalesced load, eight
// columns' loads and gathers in flight per lane; rows too long for a slice are
// walked by a whole wave from the CSR arrays (lane-strided + butterfly). A
// row's sum is its entries in CSR order (padding adds fma(0, x[row], s) = s).
// With want_dot, the partials of x.y go to part as in spmv_pAp.
template <class VT>
__device__ __forceinline__ void spmv_sell_body(int vb, int vg, const sell_desc& a, const double* __restrict__ x,
                                               double* __restrict__ y, const pcg_state* st, double* part,
                                               int want_dot) {
    if (st && st->done) return;
    __shared__ double sh[kWavesPerBlock * 2];
    const int lane = threadIdx.x & (kWave - 1);
    // XCD-aware split: the dispatcher places workgroup b on XCD b % kXcds, so
    // the items (slices, then long rows) are cut into kXcds contiguous bands
    // and band x is shared by the workgroups b = x (mod kXcds). With slices in
    // vertex order each XCD's L2 then serves one band of rows and the x
    // entries they gather. (The grid is a multiple of kXcds; the split only
    // places work, every item is done once whatever the placement.)
    const int band = vb % kXcds;
    const int per = vg / kXcds;
    const int items = a.nslices + a.nlong;
    const int i0 = static_cast<int>(static_cast<long long>(items) * band / kXcds);
    const int i1 = static_cast<int>(static_cast<long long>(items) * (band + 1) / kXcds);
    const int wave = __builtin_amdgcn_readfirstlane(vb / kXcds * kWavesPerBlock + static_cast<int>(threadIdx.x) / kWave);
    const int* __restrict__ sidx = a.idx;
    const VT* __restrict__ sval = static_cast<const VT*>(a.val);
    double acc = 0.0;
    for (int it = i0 + wave; it < i1; it += per * kWavesPerBlock) {
        if (it < a.nslices) {
            const unsigned c0 = a.soff[it], c1 = a.soff[it + 1];
            const int w = static_cast<int>(c1 - c0);
            const size_t base = static_cast<size_t>(c0) * kWave + lane;
            const int row = a.rowmap[static_cast<size_t>(it) * kWave + lane];
            // Columns in groups of eight (the last group predicated): all of a
            // group's loads, then its gathers, are in flight together.
            double s = 0.0;
            for (int k = 0; k < w; k += 8) {
                int j[8];
                double v[8], g[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const bool ok = k + u < w;
                    j[u] = ok ? sidx[base + static_cast<size_t>(k + u) * kWave] : 0;
                    v[u] = ok ? static_cast<double>(sval[base + static_cast<size_t>(k + u) * kWave]) : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) g[u] = k + u < w ? x[j[u]] : 0.0;
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (k + u < w) s = fma(v[u], g[u], s);
            }
            if (row >= 0 && row < a.n) {
                y[row] = s;
                acc = fma(x[row], s, acc);
            }
        } else {
            const int row = a.longrows[it - a.nslices];
            const unsigned b = a.ptr[row], e = a.ptr[row + 1];
            const VT* __restrict__ cval = static_cast<const VT*>(a.cval);
            double s = 0.0;
            for (unsigned p = b + lane; p < e; p += kWave) s = fma(static_cast<double>(cval[p]), x[a.cidx[p]], s);
            s = wave_sum(s);
            if (lane == 0) {
                y[row] = s;
                acc = fma(x[row], s, acc);
            }
        }
    }
    if (!want_dot) return;
    double v[1] = {acc};
    block_partials<1>(v, part, sh, vb);
}
template <class VT>
__global__ __launch_bounds__(kBlock) void spmv_sell(sell_desc a, const double* __restrict__ x, double* __restrict__ y,
                                                    const pcg_state* st, double* part, int want_dot) {
    spmv_sell_body<VT>(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), a, x, y, st, part, want_dot);
}

// r[perm[v]] = b[v]; x[perm[v]] = x0[v] (or 0).
__global__ __launch_bounds__(kBlock) void scatter_init(int n, const unsigned* __restrict__ perm,
                                                       const double* __restrict__ b, const double* __restrict__ x0,
                                                       double* __restrict__ r, double* __restrict__ x) {
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
        const unsigned q = perm[v];
        r[q] = b[v];
        x[q] = x0 ? x0[v] : 0.0;
    }
}

// Sum of squares of b (original order), for the warm-start ||b||.
__global__ __launch_bounds__(kBlock) void norm2_b(int n, const double* __restrict__ b, double* part) {
    __shared__ double sh[kWavesPerBlock * 2];
    double q = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) q = fma(b[i], b[i], q);
    double v[1] = {q};
    block_partials<1>(v, part, sh);
}

// Decide whether the NEXT preconditioner application centres (the center-k
// schedule, preconditioner.h:393-397) and set its input shift from sum(r).
__device__ __forceinline__ void next_application(pcg_state* st, double rs) {
    st->apply_count += 1;
    st->center_now = st->laplacian && (st->apply_count % st->center_k == 0);
    st->shift = st->center_now ? rs / static_cast<double>(st->n) : 0.0;
}

// r = r - Ax0 (Ax0 may be null); partials of rr = r.r, rs = sum r (the start
// of the solve, solve.cpp:529-566, is decided by kFinInit / kFinInitX0).
__global__ __launch_bounds__(kBlock) void init_residual(int n, double* __restrict__ r, const double* __restrict__ ax0,
                                                        double* part) {
    __shared__ double sh[kWavesPerBlock * 2];
    double q = 0.0, s = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double ri = r[i];
        if (ax0) { ri -= ax0[i]; r[i] = ri; }
        q = fma(ri, ri, q);
        s += ri;
    }
    double v[2] = {q, s};
    block_partials<2>(v, part, sh);
}

// x += alpha p; r -= alpha Ap; partials of rr, rs of the new r
// (solve.cpp:169-208), plus the sentinel fill of the next application's sweep
// outputs when y is given (the dataflow schedule polls them; the launched
// schedules never read an output before it is written) and the zero past the
// sweep rows of z. The loop's bookkeeping follows in pcg_finish (kFinXR).
template <class TO>
__device__ __forceinline__ void update_xr_body(int vb, int vg, int n, int m, double* __restrict__ x,
                                               const double* __restrict__ p, double* __restrict__ r,
                                               const double* __restrict__ ap, TO* __restrict__ y, TO* __restrict__ z,
                                               const pcg_state* st, double* part) {
    if (st->done) return;
    __shared__ double sh[kWavesPerBlock * 2];
    const double alpha = st->rz / st->pAp;
    using B = typename flag_bits<TO>::type;
    B* yb = reinterpret_cast<B*>(y);
    B* zb = reinterpret_cast<B*>(z);
    double q = 0.0, s = 0.0;
    for (int i = vb * kBlock + threadIdx.x; i < n; i += vg * kBlock) {
        x[i] = fma(alpha, p[i], x[i]);
        const double ri = fma(-alpha, ap[i], r[i]);
        r[i] = ri;
        q = fma(ri, ri, q);
        s += ri;
        if (i >= m) z[i] = static_cast<TO>(0);
        else if (yb) { yb[i] = flag_bits<TO>::pending; zb[i] = flag_bits<TO>::pending; }
    }
    double v[2] = {q, s};
    block_partials<2>(v, part, sh, vb);
}
template <class TO>
__global__ __launch_bounds__(kBlock) void update_xr(int n, int m, double* __restrict__ x, const double* __restrict__ p,
                                                    double* __restrict__ r, const double* __restrict__ ap,
                                                    TO* __restrict__ y, TO* __restrict__ z, const pcg_state* st,
                                                    double* part) {
    update_xr_body<TO>(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), n, m, x, p, r, ap, y, z, st, part);
}

// Sentinel fill for the first application of a solve, when y is given
// (update_xr does it for every later one), and the zero past the sweep rows
// of z.
template <class TO>
__device__ __forceinline__ void fill_outputs_body(int vb, int vg, int n, int m, TO* __restrict__ y,
                                                  TO* __restrict__ z, const pcg_state* st) {
    if (st && st->done) return;
    using B = typename flag_bits<TO>::type;
    B* yb = reinterpret_cast<B*>(y);
    B* zb = reinterpret_cast<B*>(z);
    for (int i = vb * kBlock + threadIdx.x; i < n; i += vg * kBlock) {
        if (i >= m) z[i] = static_cast<TO>(0);
        else if (yb) { yb[i] = flag_bits<TO>::pending; zb[i] = flag_bits<TO>::pending; }
    }
}
template <class TO>
__global__ __launch_bounds__(kBlock) void fill_outputs(int n, int m, TO* __restrict__ y, TO* __restrict__ z,
                                                       const pcg_state* st) {
    fill_outputs_body<TO>(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), n, m, y, z, st);
}

// Partials of r.out and sum(out) of the application's output; kFinDot then
// sets the output centring shift mu = mean(out) on centring applications
// (preconditioner.h:516-525), r.z = r.out - mu sum(r), beta = r.z / previous r.z.
template <class TO>
__device__ __forceinline__ void dot_center_body(int vb, int vg, int n, const double* __restrict__ r,
                                                const TO* __restrict__ out, const pcg_state* st, double* part) {
    if (st->done) return;
    __shared__ double sh[kWavesPerBlock * 2];
    double d = 0.0, s = 0.0;
    for (int i = vb * kBlock + threadIdx.x; i < n; i += vg * kBlock) {
        const double o = static_cast<double>(out[i]);
        d = fma(r[i], o, d);
        s += o;
    }
    double v[2] = {d, s};
    block_partials<2>(v, part, sh, vb);
}
template <class TO>
__global__ __launch_bounds__(kBlock) void dot_center(int n, const double* __restrict__ r, const TO* __restrict__ out,
                                                     const pcg_state* st, double* part) {
    dot_center_body<TO>(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), n, r, out, st, part);
}

// p = (out - mu) + beta p  (first application: p = out - mu).
template <class TO>
__device__ __forceinline__ void update_p_body(int vb, int vg, int n, double* __restrict__ p,
                                              const TO* __restrict__ out, const pcg_state* st, int first) {
    if (st->done) return;
    const double mu = st->mu, beta = st->beta;
    for (int i = vb * kBlock + threadIdx.x; i < n; i += vg * kBlock) {
        const double zi = static_cast<double>(out[i]) - mu;
        p[i] = first ? zi : fma(beta, p[i], zi);
    }
}
template <class TO>
__global__ __launch_bounds__(kBlock) void update_p(int n, double* __restrict__ p, const TO* __restrict__ out,
                                                   const pcg_state* st, int first) {
    update_p_body<TO>(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), n, p, out, st, first);
}

// Final min-norm centring (Laplacian, solve.cpp:262-269) folded into the
// gather back to the caller's order: x_out[v] = x[perm[v]] - mean(x).
__global__ __launch_bounds__(kBlock) void sum_vec(int n, const double* __restrict__ a, double* part) {
    __shared__ double sh[kWavesPerBlock * 2];
    double s = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += a[i];
    double v[1] = {s};
    block_partials<1>(v, part, sh);
}

__global__ __launch_bounds__(kBlock) void gather_shift(int n, const unsigned* __restrict__ perm,
                                                       const double* __restrict__ src, double* __restrict__ dst,
                                                       const pcg_state* st, int use_mean) {
    const double sh = use_mean ? st->out_sum / static_cast<double>(n) : 0.0;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) dst[v] = src[perm[v]] - sh;
}

// Stand-alone application z = M^-1 r in the caller's order
// (preconditioner.h:344-353, 403-529): scatter with the input shift, the two
// sweeps, then a gather with the output shift.
template <class TO>
__global__ __launch_bounds__(kBlock) void apply_scatter(int n, int m, const unsigned* __restrict__ perm,
                                                        const double* __restrict__ r, double* __restrict__ rp,
                                                        TO* __restrict__ y, TO* __restrict__ z) {
    using B = typename flag_bits<TO>::type;
    B* yb = reinterpret_cast<B*>(y);
    B* zb = reinterpret_cast<B*>(z);
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
        rp[perm[v]] = r[v];
        if (v < m) { yb[v] = flag_bits<TO>::pending; zb[v] = flag_bits<TO>::pending; }
        else z[v] = static_cast<TO>(0);
    }
}

// Partials of sum(r), for the input shift of a stand-alone centring
// application (kFinShift: shift = sum(r)/n).
__global__ __launch_bounds__(kBlock) void sum_shift(int n, const double* __restrict__ r, double* part) {
    __shared__ double sh[kWavesPerBlock * 2];
    double s = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += r[i];
    double v[1] = {s};
    block_partials<1>(v, part, sh);
}

template <class TO>
__global__ __launch_bounds__(kBlock) void sum_out(int n, const TO* __restrict__ a, double* part) {
    __shared__ double sh[kWavesPerBlock * 2];
    double s = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += static_cast<double>(a[i]);
    double v[1] = {s};
    block_partials<1>(v, part, sh);
}

template <class TO>
__global__ __launch_bounds__(kBlock) void apply_gather(int n, const unsigned* __restrict__ perm,
                                                       const TO* __restrict__ out, double* __restrict__ zdst,
                                                       const pcg_state* st, int center) {
    const double mu = center ? st->out_sum / static_cast<double>(n) : 0.0;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
        zdst[v] = static_cast<double>(out[perm[v]]) - mu;
}

// Stage 2 of the PCG reductions (one workgroup, launched right after the
// producer of `part`; nb = the producer's grid) and the bookkeeping that
// depends on the totals:
//   kFinPAp        st->pAp; a non-positive pAp ends the solve (breakdown)
//   kFinBnorm      st->bnorm = ||b|| (warm start)
//   kFinInit(X0)   rr, rs of the initial residual and the start decisions
//                  (solve.cpp:529-566): zero b, warm-start convergence, max_iter
//   kFinXR         rr, rs; iteration count, convergence, stagnation
//                  (solve.cpp:604-615), max_iter, the next application's centring
//   kFinDot(First) r.z, mu, beta of the application's output
//   kFinOutSum     st->out_sum (min-norm centring, application output shift)
//   kFinShift      st->shift = sum / n (input shift of a centring application)
// The in-loop ops (pAp, XR, Dot) do nothing once the solve has finished, as
// their producers do not run then.
enum : int { kFinPAp, kFinBnorm, kFinInit, kFinInitX0, kFinXR, kFinDot, kFinDotFirst, kFinOutSum, kFinShift };

__device__ __forceinline__ void pcg_finish_body(int op, int n, const double* __restrict__ part, int nb,
                                                pcg_state* st) {
    if ((op == kFinPAp || op == kFinXR || op == kFinDot || op == kFinDotFirst) && st->done) return;
    __shared__ double sh[kWavesPerBlock * 2];
    double v[2] = {0.0, 0.0};
    bool t0;
    if (op == kFinInit || op == kFinInitX0 || op == kFinXR || op == kFinDot || op == kFinDotFirst) {
        t0 = finish_sum<2>(v, part, nb, sh);
    } else {
        double w[1];
        t0 = finish_sum<1>(w, part, nb, sh);
        v[0] = w[0];
    }
    if (!t0) return;
    switch (op) {
        case kFinPAp:
            st->pAp = v[0];
            if (v[0] <= 0.0) { st->done = 1; st->status = 2; }
            break;
        case kFinBnorm:
            st->bnorm = sqrt(v[0]);
            break;
        case kFinInit:
        case kFinInitX0:
            st->rr = v[0];
            st->rs = v[1];
            st->iter = 0;
            st->prev_check = 1.0;
            st->apply_count = 0;
            if (op == kFinInit) {
                st->bnorm = sqrt(v[0]);
                st->residual = 1.0;
                if (st->bnorm == 0.0) { st->done = 1; st->status = 5; st->residual = 0.0; return; }
            } else {
                if (st->bnorm == 0.0) { st->done = 1; st->status = 5; st->residual = 0.0; return; }
                st->residual = sqrt(v[0]) / st->bnorm;
                if (st->residual < st->tol) { st->done = 1; st->status = 1; return; }
            }
            if (st->max_iter <= 0) { st->done = 1; st->status = 4; return; }
            next_application(st, v[1]);
            break;
        case kFinXR: {
            st->rr = v[0];
            st->rs = v[1];
            st->iter += 1;
            const double rn = sqrt(v[0]) / st->bnorm;
            st->residual = rn;
            if (rn < st->tol) { st->done = 1; st->status = 1; return; }
            if (st->stag_window > 0 && st->iter % st->stag_window == 0) {
                if (rn > st->prev_check * 0.5) { st->done = 1; st->status = 3; return; }
                st->prev_check = rn;
            }
            if (st->iter >= st->max_iter) { st->done = 1; st->status = 4; return; }
            next_application(st, v[1]);
            break;
        }
        case kFinDot:
        case kFinDotFirst: {
            const double mu = st->center_now ? v[1] / static_cast<double>(n) : 0.0;
            const double rz = v[0] - mu * st->rs;
            st->out_sum = v[1];
            st->mu = mu;
            if (op == kFinDotFirst) {
                st->rz = rz;
                st->beta = 0.0;
            } else {
                st->beta = rz / st->rz;
                st->rz = rz;
            }
            break;
        }
        case kFinOutSum:
            st->out_sum = v[0];
            break;
        case kFinShift:
            st->shift = v[0] / static_cast<double>(n);
            break;
        default:
            break;
    }
}
__global__ __launch_bounds__(kBlock) void pcg_finish(int op, int n, const double* __restrict__ part, int nb,
                                                     pcg_state* st) {
    pcg_finish_body(op, n, part, nb, st);
}

// Between factor order and storage order: dst[Q[q]] = src[q] (in) or
// dst[q] = src[Q[q]] (out), q < m; Q null: the identity.
__global__ __launch_bounds__(kBlock) void permute_vec(int m, const unsigned* __restrict__ Q,
                                                      const double* __restrict__ src, double* __restrict__ dst,
                                                      int in) {
    for (int q = blockIdx.x * kBlock + threadIdx.x; q < m; q += gridDim.x * kBlock) {
        const unsigned t = Q ? Q[q] : static_cast<unsigned>(q);
        if (in) dst[t] = src[q];
        else dst[q] = src[t];
    }
}

// dst[q] = double(src[Q[q]]): a sweep output (fp64 or fp32) back to factor
// order and fp64 (the boundary cast of cuda_cast.cu:24-34).
template <class TO>
__global__ __launch_bounds__(kBlock) void permute_out(int m, const unsigned* __restrict__ Q, const TO* __restrict__ src,
                                                      double* __restrict__ dst) {
    for (int q = blockIdx.x * kBlock + threadIdx.x; q < m; q += gridDim.x * kBlock)
        dst[q] = static_cast<double>(src[Q ? Q[q] : static_cast<unsigned>(q)]);
}

}  // namespace apxchol_amd
