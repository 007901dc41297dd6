// Blocked k-right-hand-side PCG on gfx950: the SpMM, the two sweeps and the
// CG vector passes for k columns at once. Block vectors are row-major
// [rows][kp] (kp = the column count rounded up to a power of two, <= 64), so
// gathering one dependency brings all k values in one or two cache lines, and
// one thread works on one (row, column).
//
// Per-column arithmetic replays the single-right-hand-side kernels exactly:
// a row of `len` entries handled there by G lanes (lane s takes entries s,
// s + G, ... as an fma chain from 0, then an xor butterfly over the G lanes)
// is computed here by one thread as the same G chains combined in the same
// tree (bit-reversed pairwise order), so every column of a blocked sweep or
// SpMM equals the single-vector result bit for bit. The reductions are per
// column, two-stage and fixed-order (deterministic, not the single kernels'
// tree).
//
// Each column carries its own pcg_state (convergence, stagnation, centring
// schedule), so a column stops exactly as its own solve would; a batch_state
// counts the active columns so whole launches return once all are done.
#include "common.hpp"

namespace apxchol_amd {

__device__ __forceinline__ int bitrev(int v, int bits) {
    return bits == 0 ? 0 : static_cast<int>(__builtin_bitreverse32(static_cast<unsigned>(v)) >> (32 - bits));
}

// Pairwise combination stack of the butterfly replay: leaves pushed in
// bit-reversed lane order combine like the xor butterfly (off = G/2 .. 1).
template <class T>
struct pair_stack {
    T v[7];
    int top = 0;
    __device__ __forceinline__ void push(T x, int i) {
        // i: index of this leaf in visiting order; merge while the count of
        // leaves pushed so far (i + 1) has trailing zeros.
        unsigned cnt = static_cast<unsigned>(i) + 1u;
        while ((cnt & 1u) == 0u) {
            x = v[--top] + x;
            cnt >>= 1;
        }
        v[top++] = x;
    }
};

// The single-vector row sum of `len` entries at [rb, rb + len) over G lanes,
// for column c of a [*][kp] block `x`: lane s's fma chain over entries
// s, s + G, ... (ascending), then the butterfly tree.
template <class VT, class XT, class IT>
__device__ __forceinline__ XT replay_row(unsigned rb, unsigned len, int G, const IT* __restrict__ idx,
                                         const VT* __restrict__ val, const XT* x, int kp, int c) {
    int bits = 0;
    while ((1 << bits) < G) ++bits;
    pair_stack<XT> stk;
    for (int i = 0; i < G; ++i) {
        const int s = bitrev(i, bits);
        XT part = 0;
        for (unsigned q = static_cast<unsigned>(s); q < len; q += static_cast<unsigned>(G))
            part = fma(static_cast<XT>(val[rb + q]), x[static_cast<size_t>(idx[rb + q]) * kp + c], part);
        stk.push(part, i);
    }
    return stk.v[0];
}

// ---- sweeps ----------------------------------------------------------------

// One level of a blocked sweep: rows at sweep positions [p0, p1), kp threads
// per row. Forward input: fp64 r minus the column's shift; back input: y
// (times inv_scale^2 on the fp16 storage). Same lane grouping as sweep_batch:
// G = lanes_for(len), rows longer than 64 * kEpl use the whole wave (G = 64).
template <class TI, class TO, bool F16>
__global__ __launch_bounds__(kBlock) void sweep_levels_k(sweep_desc d, int p0, int p1, int kp,
                                                         const TI* __restrict__ rhs, TO* out,
                                                         const pcg_state* st, const batch_state* bs) {
    if (bs && bs->active == 0) return;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int c = t & (kp - 1);
    const int pos = p0 + t / kp;
    if (pos >= p1) return;
    if (st && st[c].done) return;
    const int row = d.backward ? d.m - 1 - pos : pos;
    const unsigned rb = d.ptr[row];
    const unsigned len = d.ptr[row + 1] - rb;
    int G = 1;
    {
        unsigned g = (len + kEpl - 1) / kEpl;
        if (g > 1) {
            g = 1u << (32 - __builtin_clz(g - 1));
            G = g > 64 ? 64 : static_cast<int>(g);
        }
    }
    TO part;
    if constexpr (F16) {
        int bits = 0;
        while ((1 << bits) < G) ++bits;
        pair_stack<TO> stk;
        for (int i = 0; i < G; ++i) {
            const int s = bitrev(i, bits);
            TO acc = 0;
            for (unsigned q = static_cast<unsigned>(s); q < len; q += static_cast<unsigned>(G)) {
                const float v = static_cast<float>(__builtin_bit_cast(_Float16, d.h[rb + q]));
                acc = fma(static_cast<TO>(v), out[static_cast<size_t>(d.idx[rb + q]) * kp + c], acc);
            }
            stk.push(acc, i);
        }
        part = stk.v[0];
    } else {
        int bits = 0;
        while ((1 << bits) < G) ++bits;
        pair_stack<TO> stk;
        for (int i = 0; i < G; ++i) {
            const int s = bitrev(i, bits);
            TO acc = 0;
            for (unsigned q = static_cast<unsigned>(s); q < len; q += static_cast<unsigned>(G)) {
                const entry32 e = d.ent[rb + q];
                acc = fma(static_cast<TO>(e.val), out[static_cast<size_t>(e.idx) * kp + c], acc);
            }
            stk.push(acc, i);
        }
        part = stk.v[0];
    }
    TO r;
    if (!d.backward) {
        const double shift = st ? st[c].shift : 0.0;
        r = static_cast<TO>(static_cast<double>(rhs[static_cast<size_t>(row) * kp + c]) - shift);
    } else if (d.inv_scale) {
        const double sc = static_cast<double>(d.inv_scale[row]);
        r = static_cast<TO>(static_cast<double>(rhs[static_cast<size_t>(row) * kp + c]) * (sc * sc));
    } else {
        r = static_cast<TO>(rhs[static_cast<size_t>(row) * kp + c]);
    }
    out[static_cast<size_t>(row) * kp + c] = (r - part) / static_cast<TO>(d.diag[row]);
}

// ---- per-column reductions -----------------------------------------------

// Sum N values per thread over the threads of one column in a block (kp
// columns, kBlock / kp rows per block), in row order, then over blocks in
// block order (last block). Thread c < kp of the last block returns true
// with the totals of column c in v.
template <int N>
__device__ __forceinline__ bool col_grid_sum(double (&v)[N], int kp, double* part, unsigned* counter, double* sh) {
    const int c = threadIdx.x & (kp - 1);
    const int rows = kBlock / kp;
    __shared__ double col[N * 64];
#pragma unroll
    for (int k = 0; k < N; ++k) sh[k * kBlock + threadIdx.x] = v[k];
    __syncthreads();
    __shared__ int is_last;
    if (threadIdx.x < kp) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double s = sh[k * kBlock + c];
            for (int j = 1; j < rows; ++j) s += sh[k * kBlock + j * kp + c];
            col[k * 64 + c] = s;
        }
    }
    __syncthreads();
    // One thread publishes the block's partials, then counts the block in
    // (the same store -> fence -> counter order as grid_sum).
    if (threadIdx.x == 0) {
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < kp; ++j)
                __hip_atomic_store(&part[(static_cast<size_t>(blockIdx.x) * N + k) * kp + j], col[k * 64 + j],
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const unsigned t = atomicAdd(counter, 1u);
        is_last = (t == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    if (threadIdx.x >= kp) return false;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double s = 0.0;
        for (unsigned b = 0; b < gridDim.x; ++b)
            s += __hip_atomic_load(&part[(static_cast<size_t>(b) * N + k) * kp + c], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        v[k] = s;
    }
    if (c == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// Workgroups of a blocked row-parallel kernel: one thread per (row, column)
// up to the vector kernels' cap (a pure function of rows and kp).
inline int block_grid(long long rows, int kp) { return vec_blocks(rows * kp); }

// ---- SpMM + p.Ap -----------------------------------------------------------

// Y = A P for kp columns, each row replaying spmv_pAp's G-lane order; with
// want_dot, p.Ap per column into st[c].pAp (breakdown ends that column).
template <class VT>
__global__ __launch_bounds__(kBlock) void spmm_pAp_k(int n, int kp, int G, const unsigned* __restrict__ ptr,
                                                     const int* __restrict__ idx, const VT* __restrict__ val,
                                                     const double* __restrict__ x, double* __restrict__ y,
                                                     pcg_state* st, batch_state* bs, double* part, unsigned* counter,
                                                     int want_dot) {
    if (bs->active == 0) return;
    __shared__ double sh[kBlock];
    const int c = threadIdx.x & (kp - 1);
    const bool live = !st[c].done;
    const int rows_per_grid = gridDim.x * (kBlock / kp);
    double acc = 0.0;
    if (live) {
        for (int row = (blockIdx.x * kBlock + threadIdx.x) / kp; row < n; row += rows_per_grid) {
            const unsigned b = ptr[row];
            const double s = replay_row<VT, double, int>(b, ptr[row + 1] - b, G, idx, val, x, kp, c);
            y[static_cast<size_t>(row) * kp + c] = s;
            acc = fma(x[static_cast<size_t>(row) * kp + c], s, acc);
        }
    }
    if (!want_dot) return;
    double v[1] = {acc};
    if (col_grid_sum<1>(v, kp, part, counter, sh) && live) {
        st[c].pAp = v[0];
        if (v[0] <= 0.0) { st[c].done = 1; st[c].status = 2; }
    }
}

// ---- vector passes ---------------------------------------------------------

// r[perm[v]][c] = B[c][v] (B is k columns of n, the caller's layout); x = 0;
// padded columns (c >= k) get r = 0.
__global__ __launch_bounds__(kBlock) void scatter_init_k(int n, int k, int kp, const unsigned* __restrict__ perm,
                                                         const double* __restrict__ B, double* __restrict__ r,
                                                         double* __restrict__ x) {
    const long long total = static_cast<long long>(n) * kp;
    for (long long t = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; t < total;
         t += static_cast<long long>(gridDim.x) * kBlock) {
        const int c = static_cast<int>(t & (kp - 1));
        const int v = static_cast<int>(t / kp);
        const size_t q = static_cast<size_t>(perm[v]) * kp + c;
        r[q] = c < k ? B[static_cast<size_t>(c) * n + v] : 0.0;
        x[q] = 0.0;
    }
}

__device__ __forceinline__ void next_application_k(pcg_state& s, double rs) {
    s.apply_count += 1;
    s.center_now = s.laplacian && (s.apply_count % s.center_k == 0);
    s.shift = s.center_now ? rs / static_cast<double>(s.n) : 0.0;
}

// Count of active columns (run by one block after the per-column updates).
__device__ __forceinline__ void count_active(int kp, const pcg_state* st, batch_state* bs) {
    if (threadIdx.x == 0) {
        int a = 0;
        for (int c = 0; c < kp; ++c) a += st[c].done ? 0 : 1;
        bs->active = a;
    }
}

// rr, rs per column of r; the start of each column's solve (zero right-hand
// side, max_iter, first centring decision).
__global__ __launch_bounds__(kBlock) void init_residual_k(int n, int kp, const double* __restrict__ r, pcg_state* st,
                                                          batch_state* bs, double* part, unsigned* counter) {
    __shared__ double sh[2 * kBlock];
    const int c = threadIdx.x & (kp - 1);
    const int rows_per_grid = gridDim.x * (kBlock / kp);
    double q = 0.0, s = 0.0;
    for (int row = (blockIdx.x * kBlock + threadIdx.x) / kp; row < n; row += rows_per_grid) {
        const double ri = r[static_cast<size_t>(row) * kp + c];
        q = fma(ri, ri, q);
        s += ri;
    }
    double v[2] = {q, s};
    const bool last = col_grid_sum<2>(v, kp, part, counter, sh);
    if (last) {
        pcg_state& t = st[c];
        t.rr = v[0];
        t.rs = v[1];
        t.iter = 0;
        t.prev_check = 1.0;
        t.apply_count = 0;
        t.bnorm = sqrt(v[0]);
        t.residual = 1.0;
        if (t.bnorm == 0.0) { t.done = 1; t.status = 5; t.residual = 0.0; }
        else if (t.max_iter <= 0) { t.done = 1; t.status = 4; }
        else next_application_k(t, v[1]);
    }
}

// x += alpha p; r -= alpha Ap per live column; rr, rs; each column's loop
// bookkeeping as update_xr's.
__global__ __launch_bounds__(kBlock) void update_xr_k(int n, int kp, double* __restrict__ x,
                                                      const double* __restrict__ p, double* __restrict__ r,
                                                      const double* __restrict__ ap, pcg_state* st, batch_state* bs,
                                                      double* part, unsigned* counter) {
    if (bs->active == 0) return;
    __shared__ double sh[2 * kBlock];
    const int c = threadIdx.x & (kp - 1);
    const bool live = !st[c].done;
    const double alpha = live ? st[c].rz / st[c].pAp : 0.0;
    const int rows_per_grid = gridDim.x * (kBlock / kp);
    double q = 0.0, s = 0.0;
    if (live) {
        for (int row = (blockIdx.x * kBlock + threadIdx.x) / kp; row < n; row += rows_per_grid) {
            const size_t i = static_cast<size_t>(row) * kp + c;
            x[i] = fma(alpha, p[i], x[i]);
            const double ri = fma(-alpha, ap[i], r[i]);
            r[i] = ri;
            q = fma(ri, ri, q);
            s += ri;
        }
    }
    double v[2] = {q, s};
    if (col_grid_sum<2>(v, kp, part, counter, sh) && live) {
        pcg_state& t = st[c];
        t.rr = v[0];
        t.rs = v[1];
        t.iter += 1;
        const double rn = sqrt(v[0]) / t.bnorm;
        t.residual = rn;
        bool stop = false;
        if (rn < t.tol) { t.done = 1; t.status = 1; stop = true; }
        if (!stop && t.stag_window > 0 && t.iter % t.stag_window == 0) {
            if (rn > t.prev_check * 0.5) { t.done = 1; t.status = 3; stop = true; }
            else t.prev_check = rn;
        }
        if (!stop && t.iter >= t.max_iter) { t.done = 1; t.status = 4; stop = true; }
        if (!stop) next_application_k(t, v[1]);
    }
}

// r.out, sum(out) per column; centring shift mu, rz, beta (dot_center's).
template <class TO>
__global__ __launch_bounds__(kBlock) void dot_center_k(int n, int kp, const double* __restrict__ r,
                                                       const TO* __restrict__ out, pcg_state* st, batch_state* bs,
                                                       double* part, unsigned* counter, int first) {
    if (bs->active == 0) return;
    __shared__ double sh[2 * kBlock];
    const int c = threadIdx.x & (kp - 1);
    const bool live = !st[c].done;
    const int rows_per_grid = gridDim.x * (kBlock / kp);
    double d = 0.0, s = 0.0;
    if (live) {
        for (int row = (blockIdx.x * kBlock + threadIdx.x) / kp; row < n; row += rows_per_grid) {
            const size_t i = static_cast<size_t>(row) * kp + c;
            const double o = static_cast<double>(out[i]);
            d = fma(r[i], o, d);
            s += o;
        }
    }
    double v[2] = {d, s};
    if (col_grid_sum<2>(v, kp, part, counter, sh) && live) {
        pcg_state& t = st[c];
        const double mu = t.center_now ? v[1] / static_cast<double>(n) : 0.0;
        const double rz = v[0] - mu * t.rs;
        t.out_sum = v[1];
        t.mu = mu;
        if (first) {
            t.rz = rz;
            t.beta = 0.0;
        } else {
            t.beta = rz / t.rz;
            t.rz = rz;
        }
    }
}

// p = (out - mu) + beta p per live column; the count of active columns is
// refreshed here (every column's state for this iteration is final).
template <class TO>
__global__ __launch_bounds__(kBlock) void update_p_k(int n, int kp, double* __restrict__ p, const TO* __restrict__ out,
                                                     const pcg_state* st, batch_state* bs, int first) {
    if (bs->active == 0) return;
    const int c = threadIdx.x & (kp - 1);
    if (st[c].done) return;
    const double mu = st[c].mu, beta = st[c].beta;
    const int rows_per_grid = gridDim.x * (kBlock / kp);
    for (int row = (blockIdx.x * kBlock + threadIdx.x) / kp; row < n; row += rows_per_grid) {
        const size_t i = static_cast<size_t>(row) * kp + c;
        const double zi = static_cast<double>(out[i]) - mu;
        p[i] = first ? zi : fma(beta, p[i], zi);
    }
}

// Active-column count after update_xr_k (one block, kp threads).
__global__ void refresh_active_k(int kp, const pcg_state* st, batch_state* bs) { count_active(kp, st, bs); }

// Zero the rows past m of a block vector (the Laplacian's dropped row) so the
// dot products see out[m] = 0.
template <class TO>
__global__ __launch_bounds__(kBlock) void zero_tail_k(int n, int m, int kp, TO* __restrict__ a) {
    const long long total = static_cast<long long>(n - m) * kp;
    for (long long t = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; t < total;
         t += static_cast<long long>(gridDim.x) * kBlock)
        a[static_cast<size_t>(m) * kp + t] = static_cast<TO>(0);
}

// Column sums into st[c].out_sum (the Laplacian min-norm centring of the
// result; the centring shifts of a blocked application).
template <class T>
__global__ __launch_bounds__(kBlock) void sum_cols_k(int n, int kp, const T* __restrict__ x, pcg_state* st,
                                                     double* part, unsigned* counter) {
    __shared__ double sh[kBlock];
    const int c = threadIdx.x & (kp - 1);
    const int rows_per_grid = gridDim.x * (kBlock / kp);
    double s = 0.0;
    for (int row = (blockIdx.x * kBlock + threadIdx.x) / kp; row < n; row += rows_per_grid)
        s += static_cast<double>(x[static_cast<size_t>(row) * kp + c]);
    double v[1] = {s};
    if (col_grid_sum<1>(v, kp, part, counter, sh)) st[c].out_sum = v[0];
}

// Input shift of a blocked application: shift = center ? out_sum / n : 0.
__global__ void shift_from_sum_k(int kp, pcg_state* st, int center) {
    const int c = threadIdx.x;
    if (c < kp) st[c].shift = center ? st[c].out_sum / static_cast<double>(st[c].n) : 0.0;
}

// X[c][v] = x[perm[v]][c] - (use_mean ? sum_c / n : 0) for c < k.
template <class T>
__global__ __launch_bounds__(kBlock) void gather_k(int n, int k, int kp, const unsigned* __restrict__ perm,
                                                   const T* __restrict__ x, double* __restrict__ X,
                                                   const pcg_state* st, int use_mean) {
    const long long total = static_cast<long long>(n) * k;
    for (long long t = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; t < total;
         t += static_cast<long long>(gridDim.x) * kBlock) {
        const int c = static_cast<int>(t / n);
        const int v = static_cast<int>(t - static_cast<long long>(c) * n);
        const double sh = use_mean ? st[c].out_sum / static_cast<double>(n) : 0.0;
        X[t] = static_cast<double>(x[static_cast<size_t>(perm[v]) * kp + c]) - sh;
    }
}

}  // namespace apxchol_amd
