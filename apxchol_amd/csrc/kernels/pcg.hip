// GPU-resident PCG kernels (gfx950). Vectors live in the factor's permuted
// space; every reduction is two-stage and fixed-order (grid_sum), and every
// scalar the recurrence needs stays on the device in pcg_state. Each kernel
// returns at once when the solve has finished (st->done), so the host can
// enqueue (or graph-replay) several iterations and test completion rarely.
//
// Reference counterparts: src/cuda_pcg_kernels.cu:18-152 (spmv_pAp,
// update_xr, dot, update_p, reduce_partials) driven by cuda_pcg::solve
// (include/apxchol/solver/pcg_cuda.h:157-261); the loop semantics -- centring
// schedule, stagnation stop, warm start -- are the CPU solver's
// (src/solve.cpp:478-654, include/apxchol/solver/preconditioner.h:344-529).
#include "common.hpp"

namespace apxchol_amd {

// y = A x over the permuted full-symmetric CSR, G lanes per row (lane-strided
// entries + xor butterfly). With want_dot, x.y is reduced into st->pAp and a
// non-positive pAp ends the solve (solve.cpp:596, breakdown).
template <class VT, int G>
__global__ __launch_bounds__(kBlock) void spmv_pAp(int n, const unsigned* __restrict__ ptr,
                                                   const int* __restrict__ idx, const VT* __restrict__ val,
                                                   const double* __restrict__ x, double* __restrict__ y,
                                                   pcg_state* st, double* part, unsigned* counter, int want_dot) {
    if (st && st->done) return;
    __shared__ double sh[kWavesPerBlock * 2];
    const int tid = blockIdx.x * kBlock + threadIdx.x;
    const int sub = tid & (G - 1);
    const int stride = gridDim.x * kBlock / G;
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
    if (grid_sum<1>(v, part, counter, sh)) {
        st->pAp = v[0];
        if (v[0] <= 0.0) { st->done = 1; st->status = 2; }
    }
}

// y = A x over the sliced-ELL form (sell_desc): one lane per row, a slice's
// entries column-major so every entry column is one coalesced load, eight
// columns' loads and gathers in flight per lane; rows too long for a slice are
// walked by a whole wave from the CSR arrays (lane-strided + butterfly). A
// row's sum is its entries in CSR order (padding adds fma(0, x[row], s) = s).
// With want_dot, x.y lands in st->pAp as in spmv_pAp.
template <class VT, bool MAP>
__global__ __launch_bounds__(kBlock) void spmv_sell(sell_desc a, const double* __restrict__ x, double* __restrict__ y,
                                                    pcg_state* st, double* part, unsigned* counter, int want_dot) {
    if (st && st->done) return;
    __shared__ double sh[kWavesPerBlock * 2];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.x) * kWavesPerBlock +
                                                    static_cast<int>(threadIdx.x) / kWave);
    const int waves = static_cast<int>(gridDim.x) * kWavesPerBlock;
    const int items = a.nslices + a.nlong;
    const int* __restrict__ sidx = a.idx;
    const VT* __restrict__ sval = static_cast<const VT*>(a.val);
    double acc = 0.0;
    for (int it = wave; it < items; it += waves) {
        if (it < a.nslices) {
            const unsigned c0 = a.soff[it], c1 = a.soff[it + 1];
            const int w = static_cast<int>(c1 - c0);
            const size_t base = static_cast<size_t>(c0) * kWave + lane;
            const int row = MAP ? a.rowmap[static_cast<size_t>(it) * kWave + lane] : it * kWave + lane;
            double s = 0.0;
            int k = 0;
            for (; k + 8 <= w; k += 8) {
                int j[8];
                double v[8], g[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    j[u] = sidx[base + static_cast<size_t>(k + u) * kWave];
                    v[u] = static_cast<double>(sval[base + static_cast<size_t>(k + u) * kWave]);
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) g[u] = x[j[u]];
#pragma unroll
                for (int u = 0; u < 8; ++u) s = fma(v[u], g[u], s);
            }
            for (; k < w; ++k) {
                const size_t q = base + static_cast<size_t>(k) * kWave;
                s = fma(static_cast<double>(sval[q]), x[sidx[q]], s);
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
    if (grid_sum<1>(v, part, counter, sh)) {
        st->pAp = v[0];
        if (v[0] <= 0.0) { st->done = 1; st->status = 2; }
    }
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
__global__ __launch_bounds__(kBlock) void norm2_b(int n, const double* __restrict__ b, pcg_state* st, double* part,
                                                  unsigned* counter) {
    __shared__ double sh[kWavesPerBlock * 2];
    double q = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) q = fma(b[i], b[i], q);
    double v[1] = {q};
    if (grid_sum<1>(v, part, counter, sh)) st->bnorm = sqrt(v[0]);
}

// Decide whether the NEXT preconditioner application centres (the center-k
// schedule, preconditioner.h:393-397) and set its input shift from sum(r).
__device__ __forceinline__ void next_application(pcg_state* st, double rs) {
    st->apply_count += 1;
    st->center_now = st->laplacian && (st->apply_count % st->center_k == 0);
    st->shift = st->center_now ? rs / static_cast<double>(st->n) : 0.0;
}

// r = r - Ax0 (Ax0 may be null); rr = r.r, rs = sum r; decides the start of
// the solve (solve.cpp:529-566).
__global__ __launch_bounds__(kBlock) void init_residual(int n, double* __restrict__ r, const double* __restrict__ ax0,
                                                        pcg_state* st, double* part, unsigned* counter) {
    __shared__ double sh[kWavesPerBlock * 2];
    double q = 0.0, s = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double ri = r[i];
        if (ax0) { ri -= ax0[i]; r[i] = ri; }
        q = fma(ri, ri, q);
        s += ri;
    }
    double v[2] = {q, s};
    if (grid_sum<2>(v, part, counter, sh)) {
        st->rr = v[0];
        st->rs = v[1];
        st->iter = 0;
        st->prev_check = 1.0;
        st->apply_count = 0;
        if (!ax0) {
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
    }
}

// x += alpha p; r -= alpha Ap; rr, rs of the new r (solve.cpp:169-208), plus
// the sentinel fill of the next application's sweep outputs. The last block
// runs the loop's bookkeeping: iteration count, convergence, stagnation
// (solve.cpp:604-615), max_iter, and the next application's centring.
template <class TO>
__global__ __launch_bounds__(kBlock) void update_xr(int n, int m, double* __restrict__ x, const double* __restrict__ p,
                                                    double* __restrict__ r, const double* __restrict__ ap,
                                                    TO* __restrict__ y, TO* __restrict__ z, pcg_state* st,
                                                    double* part, unsigned* counter) {
    if (st->done) return;
    __shared__ double sh[kWavesPerBlock * 2];
    const double alpha = st->rz / st->pAp;
    using B = typename flag_bits<TO>::type;
    B* yb = reinterpret_cast<B*>(y);
    B* zb = reinterpret_cast<B*>(z);
    double q = 0.0, s = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        x[i] = fma(alpha, p[i], x[i]);
        const double ri = fma(-alpha, ap[i], r[i]);
        r[i] = ri;
        q = fma(ri, ri, q);
        s += ri;
        if (i < m) { yb[i] = flag_bits<TO>::pending; zb[i] = flag_bits<TO>::pending; }
        else z[i] = static_cast<TO>(0);
    }
    double v[2] = {q, s};
    if (grid_sum<2>(v, part, counter, sh)) {
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
    }
}

// Sentinel fill for the first application of a solve (update_xr does it for
// every later one).
template <class TO>
__global__ __launch_bounds__(kBlock) void fill_outputs(int n, int m, TO* __restrict__ y, TO* __restrict__ z,
                                                       const pcg_state* st) {
    if (st && st->done) return;
    using B = typename flag_bits<TO>::type;
    B* yb = reinterpret_cast<B*>(y);
    B* zb = reinterpret_cast<B*>(z);
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        if (i < m) { yb[i] = flag_bits<TO>::pending; zb[i] = flag_bits<TO>::pending; }
        else z[i] = static_cast<TO>(0);
    }
}

// r.out and sum(out) of the application's output; the output centring
// shift mu = mean(out) on centring applications (preconditioner.h:516-525);
// r.z = r.out - mu sum(r); beta = r.z / previous r.z.
template <class TO>
__global__ __launch_bounds__(kBlock) void dot_center(int n, const double* __restrict__ r, const TO* __restrict__ out,
                                                     pcg_state* st, double* part, unsigned* counter, int first) {
    if (st->done) return;
    __shared__ double sh[kWavesPerBlock * 2];
    double d = 0.0, s = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double o = static_cast<double>(out[i]);
        d = fma(r[i], o, d);
        s += o;
    }
    double v[2] = {d, s};
    if (grid_sum<2>(v, part, counter, sh)) {
        const double mu = st->center_now ? v[1] / static_cast<double>(n) : 0.0;
        const double rz = v[0] - mu * st->rs;
        st->out_sum = v[1];
        st->mu = mu;
        if (first) {
            st->rz = rz;
            st->beta = 0.0;
        } else {
            st->beta = rz / st->rz;
            st->rz = rz;
        }
    }
}

// p = (out - mu) + beta p  (first application: p = out - mu).
template <class TO>
__global__ __launch_bounds__(kBlock) void update_p(int n, double* __restrict__ p, const TO* __restrict__ out,
                                                   const pcg_state* st, int first) {
    if (st->done) return;
    const double mu = st->mu, beta = st->beta;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double zi = static_cast<double>(out[i]) - mu;
        p[i] = first ? zi : fma(beta, p[i], zi);
    }
}

// Final min-norm centring (Laplacian, solve.cpp:262-269) folded into the
// gather back to the caller's order: x_out[v] = x[perm[v]] - mean(x).
__global__ __launch_bounds__(kBlock) void sum_vec(int n, const double* __restrict__ a, pcg_state* st, double* part,
                                                  unsigned* counter) {
    __shared__ double sh[kWavesPerBlock * 2];
    double s = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += a[i];
    double v[1] = {s};
    if (grid_sum<1>(v, part, counter, sh)) st->out_sum = v[0];
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

// Input shift of a stand-alone application: shift = center ? sum(r)/n : 0.
__global__ __launch_bounds__(kBlock) void sum_shift(int n, const double* __restrict__ r, pcg_state* st, double* part,
                                                    unsigned* counter, int center) {
    __shared__ double sh[kWavesPerBlock * 2];
    double s = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += r[i];
    double v[1] = {s};
    if (grid_sum<1>(v, part, counter, sh)) st->shift = center ? v[0] / static_cast<double>(n) : 0.0;
}

template <class TO>
__global__ __launch_bounds__(kBlock) void sum_out(int n, const TO* __restrict__ a, pcg_state* st, double* part,
                                                  unsigned* counter) {
    __shared__ double sh[kWavesPerBlock * 2];
    double s = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += static_cast<double>(a[i]);
    double v[1] = {s};
    if (grid_sum<1>(v, part, counter, sh)) st->out_sum = v[0];
}

template <class TO>
__global__ __launch_bounds__(kBlock) void apply_gather(int n, const unsigned* __restrict__ perm,
                                                       const TO* __restrict__ out, double* __restrict__ zdst,
                                                       const pcg_state* st, int center) {
    const double mu = center ? st->out_sum / static_cast<double>(n) : 0.0;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
        zdst[v] = static_cast<double>(out[perm[v]]) - mu;
}

// Between factor order and storage order: dst[Q[q]] = src[q] (in) or
// dst[q] = src[Q[q]] (out), q < m.
__global__ __launch_bounds__(kBlock) void permute_vec(int m, const unsigned* __restrict__ Q,
                                                      const double* __restrict__ src, double* __restrict__ dst,
                                                      int in) {
    for (int q = blockIdx.x * kBlock + threadIdx.x; q < m; q += gridDim.x * kBlock) {
        if (in) dst[Q[q]] = src[q];
        else dst[q] = src[Q[q]];
    }
}

}  // namespace apxchol_amd
