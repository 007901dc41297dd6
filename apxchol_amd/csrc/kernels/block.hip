// Blocked k-right-hand-side PCG on gfx950: the SpMM, the two sweeps and the
// CG vector passes for k columns at once. Block vectors are row-major
// [rows][kp] (kp = the column count rounded up to a power of two, <= 64), so
// the values one gather needs for all columns are contiguous.
//
// The sweeps and the SpMM give every (row, column) the lanes the single-
// vector kernels give that row (lane = column * G + chain: the same
// lane-strided fma chains and the same xor butterfly), so every column of a
// blocked sweep or SpMM equals the single-vector result bit for bit. The
// per-column dot products are two-stage and fixed-order (deterministic, but
// not the single kernels' tree).
//
// Each column carries its own pcg_state (convergence, stagnation, centring
// schedule), so a column stops exactly as its own solve would; a batch_state
// counts the active columns so whole launches return once all are done.
#include "common.hpp"

namespace apxchol_amd {

// ---- sweeps ----------------------------------------------------------------

// One level of a blocked sweep: waves [w0, w1) of the blocked table (see
// block_batch_table). Lane = column * G + chain inside a row's group, so the
// arithmetic of (row, column) is sweep_batch's for that row, and the lanes
// of one chain index gather the row's dependency for consecutive columns
// from one contiguous span.
template <class TI, class TO, bool F16>
__global__ __launch_bounds__(kBlock) void sweep_block(sweep_desc d, const unsigned* __restrict__ wpos,
                                                      const unsigned long long* __restrict__ wmask,
                                                      const int* __restrict__ wcol0, int w0, int w1, int kp,
                                                      const TI* __restrict__ rhs, TO* out, const pcg_state* st,
                                                      const batch_state* bs) {
    if (bs && bs->active == 0) return;
    const int w = w0 + static_cast<int>(blockIdx.x) * kWavesPerBlock + static_cast<int>(threadIdx.x) / kWave;
    if (w >= w1) return;
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned p = wpos[w];
    const unsigned long long mask = wmask[w];
    int pos, sub0;
    if (mask) {
        const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        const unsigned long long mu = mask & upto;
        pos = static_cast<int>(p) + __popcll(mu) - 1;
        sub0 = lane - (63 - __clzll(mu));
    } else {
        pos = static_cast<int>(p);
        sub0 = lane;
    }
    const int row = d.backward ? d.m - 1 - pos : pos;
    const unsigned rb = d.ptr[row];
    const unsigned len = d.ptr[row + 1] - rb;
    const int G = lanes_for(len);
    const int c = mask ? sub0 / G : wcol0[w] + lane / G;
    const int s = sub0 & (G - 1);
    const bool active = mask ? sub0 < G * kp : true;
    // The writer lane's right-hand side, diagonal and shift, loaded with the
    // entries (independent of the gathers).
    const bool writer = active && s == 0;
    const size_t o = static_cast<size_t>(row) * kp + c;
    const double rin = writer ? static_cast<double>(rhs[o]) : 0.0;
    const float dg = writer ? diag_operand<TO>(d, row) : 1.0f;
    const double shift = (writer && !d.backward && st) ? st[c].shift : 0.0;
    const double sc = (writer && d.backward && d.inv_scale) ? static_cast<double>(d.inv_scale[row]) : 0.0;
    TO part = 0;
    if (__builtin_expect(len > static_cast<unsigned>(kWave * kEpl), 0)) {
        // Long row (G = 64, one column per wave): sweep_batch's chunked walk.
        part = long_row_chain<F16, TO>(d, rb, len, [&](int j) { return out[static_cast<size_t>(j) * kp + c]; },
                                       [](int) { return true; });
    } else {
        int jj[kEpl];
        float vv[kEpl];
        TO xv[kEpl];
#pragma unroll
        for (int k = 0; k < kEpl; ++k) {
            const unsigned q = static_cast<unsigned>(s + k * G);
            const bool valid = active && q < len;
            load_entry<F16>(d, valid ? rb + q : (len > 0 ? rb : 0u), jj[k], vv[k]);
            if (!valid) vv[k] = 0.0f;
            xv[k] = valid ? out[static_cast<size_t>(jj[k]) * kp + c] : static_cast<TO>(0);
        }
#pragma unroll
        for (int k = 0; k < kEpl; ++k) part = fma(static_cast<TO>(vv[k]), xv[k], part);
    }
    part = group_tree_sum(part, G);
    if (!writer) return;
    TO r;
    if (!d.backward) r = static_cast<TO>(rin - shift);
    else if (d.inv_scale) r = static_cast<TO>(rin * (sc * sc));
    else r = static_cast<TO>(rin);
    out[o] = finish_row<TO>(r - part, dg);
}

// One level of a blocked sweep with one thread per (row, column) (rows of at
// most GMAX lanes, i.e. <= 4 GMAX entries): thread t takes sweep position
// p0 + t / kp and column t % kp, runs the row's G lane chains itself (lane s:
// entries s + k G, fma from 0, padding slots fma(0, 0, .)) and then the xor
// butterfly's tree as lane 0 sees it (step off < G: v[s] += v[s + off] for
// s < off) -- the lane kernels' arithmetic, so the same bits, without the
// wave table, idle lanes or shuffles. The right-hand side, diagonal and shift
// loads are issued with the row pointers, so a level is three dependent
// loads (pointers, entries, gathered values).
template <class TI, class TO, bool F16, int GMAX>
__global__ __launch_bounds__(kBlock) void sweep_block_rows(sweep_desc d, int p0, int p1, int kp,
                                                           const TI* __restrict__ rhs, TO* out, const pcg_state* st,
                                                           const batch_state* bs) {
    if (bs && bs->active == 0) return;
    const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const long long q0 = t / kp;
    if (q0 >= p1 - p0) return;
    const int pos = p0 + static_cast<int>(q0);
    const int c = static_cast<int>(t & (kp - 1));
    const int row = d.backward ? d.m - 1 - pos : pos;
    const size_t o = static_cast<size_t>(row) * kp + c;
    const double shift = (!d.backward && st) ? st[c].shift : 0.0;
    const double rin = static_cast<double>(rhs[o]);
    const float dg = diag_operand<TO>(d, row);
    const double sc = (d.backward && d.inv_scale) ? static_cast<double>(d.inv_scale[row]) : 0.0;
    const unsigned rb = d.ptr[row];
    const unsigned len = d.ptr[row + 1] - rb;
    const int G = lanes_for(len);
    // Lane chains in chunks of CH (all of them up to 8 lanes): a chunk's
    // entry loads, then its gathers, are in flight together.
    constexpr int CH = GMAX < 8 ? GMAX : 8;
    TO v[GMAX];
#pragma unroll
    for (int s0 = 0; s0 < GMAX; s0 += CH) {
        int jj[CH][kEpl];
        float vv[CH][kEpl];
        TO xv[CH][kEpl];
#pragma unroll
        for (int u = 0; u < CH; ++u)
#pragma unroll
            for (int k = 0; k < kEpl; ++k) {
                const int s = s0 + u;
                const unsigned q = static_cast<unsigned>(s + k * G);
                const bool valid = s < G && q < len;
                load_entry<F16>(d, valid ? rb + q : (len > 0 ? rb : 0u), jj[u][k], vv[u][k]);
                if (!valid) vv[u][k] = 0.0f;
            }
#pragma unroll
        for (int u = 0; u < CH; ++u)
#pragma unroll
            for (int k = 0; k < kEpl; ++k) {
                const int s = s0 + u;
                const bool valid = s < G && static_cast<unsigned>(s + k * G) < len;
                xv[u][k] = valid ? out[static_cast<size_t>(jj[u][k]) * kp + c] : static_cast<TO>(0);
            }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            TO part = 0;
#pragma unroll
            for (int k = 0; k < kEpl; ++k) part = fma(static_cast<TO>(vv[u][k]), xv[u][k], part);
            v[s0 + u] = part;
        }
    }
#pragma unroll
    for (int off = GMAX / 2; off >= 1; off >>= 1)
        if (off < G)
#pragma unroll
            for (int s = 0; s < off; ++s) v[s] += v[s + off];
    TO r;
    if (!d.backward) r = static_cast<TO>(rin - shift);
    else if (d.inv_scale) r = static_cast<TO>(rin * (sc * sc));
    else r = static_cast<TO>(rin);
    out[o] = finish_row<TO>(r - v[0], dg);
}

// A blocked level none of whose rows has an entry: every (row, column) sum is
// the empty chain's +0, so the output is finish_row(rhs - 0) --
// sweep_block_rows' bits without its row pointers and padding entries.
// Thread t: sweep position p0 + t / kp, column t % kp, grid-strided.
template <class TI, class TO>
__global__ __launch_bounds__(kBlock) void sweep_block_diag(sweep_desc d, int p0, int p1, int kp,
                                                           const TI* __restrict__ rhs, TO* out, const pcg_state* st,
                                                           const batch_state* bs) {
    if (bs && bs->active == 0) return;
    const long long total = static_cast<long long>(p1 - p0) * kp;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; t < total; t += stride) {
        const int pos = p0 + static_cast<int>(t / kp);
        const int c = static_cast<int>(t & (kp - 1));
        const int row = d.backward ? d.m - 1 - pos : pos;
        const size_t o = static_cast<size_t>(row) * kp + c;
        const double rin = static_cast<double>(rhs[o]);
        TO r;
        if (!d.backward) r = static_cast<TO>(rin - ((st) ? st[c].shift : 0.0));
        else if (d.inv_scale) {
            const double sc = static_cast<double>(d.inv_scale[row]);
            r = static_cast<TO>(rin * (sc * sc));
        } else r = static_cast<TO>(rin);
        out[o] = finish_row<TO>(r - static_cast<TO>(0), diag_operand<TO>(d, row));
    }
}

// A run of small levels of a blocked sweep, walked by ONE workgroup of
// kTailBlock threads: levels [l0, l1) of the wave table (each at most
// kBlockTailMaxWaves table waves; wave w of the workgroup takes table waves
// w, w + kTailWaves, ...), sweep positions [p0, p1). The run's outputs live in LDS
// ([p1 - p0][kp]) as well as in `out`; a dependency inside the run is read
// from LDS after the level barrier, one before the run from `out` (final
// when the kernel starts). The per-(row, column) arithmetic is sweep_block's.
template <class TI, class TO, bool F16>
__global__ __launch_bounds__(kTailBlock) void sweep_block_tail(sweep_desc d, const unsigned* __restrict__ wpos,
                                                               const unsigned long long* __restrict__ wmask,
                                                               const int* __restrict__ wcol0,
                                                               const int* __restrict__ level_first, int l0, int l1,
                                                               int p0, int kp, const TI* __restrict__ rhs, TO* out,
                                                               const pcg_state* st, const batch_state* bs) {
    if (bs && bs->active == 0) return;
    extern __shared__ unsigned char btail_lds[];
    TO* lds = reinterpret_cast<TO*>(btail_lds);
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
    for (int l = l0; l < l1; ++l) {
        for (int w = level_first[l] + wv; w < level_first[l + 1]; w += kTailWaves) {
            const unsigned p = wpos[w];
            const unsigned long long mask = wmask[w];
            int pos, sub0;
            if (mask) {
                const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
                const unsigned long long mu = mask & upto;
                pos = static_cast<int>(p) + __popcll(mu) - 1;
                sub0 = lane - (63 - __clzll(mu));
            } else {
                pos = static_cast<int>(p);
                sub0 = lane;
            }
            const int row = d.backward ? d.m - 1 - pos : pos;
            const unsigned rb = d.ptr[row];
            const unsigned len = d.ptr[row + 1] - rb;
            const int G = lanes_for(len);
            const int c = mask ? sub0 / G : wcol0[w] + lane / G;
            const int s = sub0 & (G - 1);
            const bool active = mask ? sub0 < G * kp : true;
            auto value = [&](int j) -> TO {
                const int pj = d.backward ? d.m - 1 - j : j;
                return pj >= p0 ? lds[static_cast<size_t>(pj - p0) * kp + c] : out[static_cast<size_t>(j) * kp + c];
            };
            TO part = 0;
            if (__builtin_expect(len > static_cast<unsigned>(kWave * kEpl), 0)) {
                part = long_row_chain<F16, TO>(d, rb, len, value, [](int) { return true; });
            } else {
                int jj[kEpl];
                float vv[kEpl];
                TO xv[kEpl];
#pragma unroll
                for (int k = 0; k < kEpl; ++k) {
                    const unsigned q = static_cast<unsigned>(s + k * G);
                    const bool valid = active && q < len;
                    load_entry<F16>(d, valid ? rb + q : (len > 0 ? rb : 0u), jj[k], vv[k]);
                    if (!valid) vv[k] = 0.0f;
                    xv[k] = valid ? value(jj[k]) : static_cast<TO>(0);
                }
#pragma unroll
                for (int k = 0; k < kEpl; ++k) part = fma(static_cast<TO>(vv[k]), xv[k], part);
            }
            part = group_tree_sum(part, G);
            if (active && s == 0) {
                const size_t o = static_cast<size_t>(row) * kp + c;
                TO r;
                if (!d.backward) {
                    const double shift = st ? st[c].shift : 0.0;
                    r = static_cast<TO>(static_cast<double>(rhs[o]) - shift);
                } else if (d.inv_scale) {
                    const double sc = static_cast<double>(d.inv_scale[row]);
                    r = static_cast<TO>(static_cast<double>(rhs[o]) * (sc * sc));
                } else {
                    r = static_cast<TO>(rhs[o]);
                }
                const TO v = finish_row<TO>(r - part, diag_operand<TO>(d, row));
                out[o] = v;
                lds[static_cast<size_t>(pos - p0) * kp + c] = v;
            }
        }
        if (l + 1 < l1) lds_barrier();
    }
}

// ---- per-column reductions -----------------------------------------------

// Two-stage per-column reductions (kp columns, kBlock / kp rows per block),
// as block_partials / finish_sum for one vector. Stage 1: the block's column
// sums, in row order, to part[(block * N + k) * kp + j].
//
// Lane mapping: thread t of the grid serves column ((t / G) % kp) and only
// its chain-0 lanes (t % G == 0) carry values (the others hold exact zeros,
// which the block partial skips); G = 1 is one thread per (row, column).
template <int N>
__device__ __forceinline__ void col_block_partials(double (&v)[N], int kp, double* part, double* sh, int G = 1) {
    const int c = threadIdx.x & (kp - 1);
#pragma unroll
    for (int k = 0; k < N; ++k) sh[k * kBlock + threadIdx.x] = v[k];
    __syncthreads();
    if (threadIdx.x < kp) {
        const int U = G * kp;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double s = 0.0;
            if (U <= kBlock) {
                for (int r = 0; r < kBlock / U; ++r) s += sh[k * kBlock + r * U + c * G];
            } else {
                // The block spans kBlock / G consecutive columns of one row.
                const int first = static_cast<int>((static_cast<long long>(blockIdx.x) * kBlock / G) % kp);
                const int off = (c - first + kp) % kp;
                if (off < kBlock / G) s = sh[k * kBlock + off * G];
            }
            part[(static_cast<size_t>(blockIdx.x) * N + k) * kp + c] = s;
        }
    }
}

// Stage 2 (one workgroup) over the partials of nb blocks: thread (j, c) sums
// blocks j, j + rows, ... of column c (eight loads in flight at a time), then
// column c's `rows` sums are added in j order. Thread c < kp returns true
// with the totals of column c in v.
template <int N>
__device__ __forceinline__ bool col_finish_sum(double (&v)[N], int kp, const double* part, int nb, double* sh) {
    const int c = threadIdx.x & (kp - 1);
    const int rows = kBlock / kp;
    const int j = threadIdx.x / kp;
    double acc[N];
#pragma unroll
    for (int k = 0; k < N; ++k) acc[k] = 0.0;
    // Eight blocks' loads in flight at once, added in block order.
    int b = j;
    for (; b + 7 * rows < nb; b += 8 * rows) {
        double g[8][N];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int k = 0; k < N; ++k) g[u][k] = part[(static_cast<size_t>(b + u * rows) * N + k) * kp + c];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int k = 0; k < N; ++k) acc[k] += g[u][k];
    }
    for (; b < nb; b += rows)
#pragma unroll
        for (int k = 0; k < N; ++k) acc[k] += part[(static_cast<size_t>(b) * N + k) * kp + c];
#pragma unroll
    for (int k = 0; k < N; ++k) sh[k * kBlock + threadIdx.x] = acc[k];
    __syncthreads();
    if (threadIdx.x >= kp) return false;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double s = sh[k * kBlock + c];
        for (int i = 1; i < rows; ++i) s += sh[k * kBlock + i * kp + c];
        v[k] = s;
    }
    return true;
}

// Workgroups of a blocked row-parallel kernel: one thread per (row, column)
// up to the vector kernels' cap (a pure function of rows and kp).
inline int block_grid(long long rows, int kp) { return vec_blocks(rows * kp); }

// ---- SpMM + p.Ap -----------------------------------------------------------

// Y = A P for kp columns with spmv_pAp's lanes: thread t serves row
// t / (G kp), column (t / G) % kp, chain t % G (lane-strided entries, fma
// from 0, butterfly over the G lanes), so every column gets the single
// SpMV's bits. With want_dot, the per-column partials of p.Ap go to part
// (block_finish kBFinPAp: st[c].pAp; breakdown ends that column). The grid stride is a multiple of G kp (host), so a
// thread's column is fixed.
template <class VT>
__global__ __launch_bounds__(kBlock) void spmm_pAp_k(int n, int kp, int G, const unsigned* __restrict__ ptr,
                                                     const int* __restrict__ idx, const VT* __restrict__ val,
                                                     const double* __restrict__ x, double* __restrict__ y,
                                                     const batch_state* bs, double* part, int want_dot) {
    if (bs->active == 0) return;
    __shared__ double sh[kBlock];
    const long long U = static_cast<long long>(G) * kp;
    const long long t0 = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int c = static_cast<int>((t0 / G) % kp);
    const int s0 = static_cast<int>(t0 % G);
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    double acc = 0.0;
    for (long long t = t0; t / U < n; t += stride) {
        const int row = static_cast<int>(t / U);
        const unsigned b = ptr[row], e = ptr[row + 1];
        double sum = 0.0;
        for (unsigned q = b + static_cast<unsigned>(s0); q < e; q += static_cast<unsigned>(G))
            sum = fma(static_cast<double>(val[q]), x[static_cast<size_t>(idx[q]) * kp + c], sum);
        for (int off = G / 2; off >= 1; off >>= 1) sum += __shfl_xor(sum, off, G);
        if (s0 == 0) {
            const size_t o = static_cast<size_t>(row) * kp + c;
            y[o] = sum;
            acc = fma(x[o], sum, acc);
        }
    }
    if (!want_dot) return;
    double v[1] = {acc};
    col_block_partials<1>(v, kp, part, sh, G);
}

// Workgroups of the SpMM: a multiple of G kp / kBlock when one row's lanes
// span several workgroups (so the grid stride keeps every thread's column).
inline int spmm_grid(long long n, int kp, int G) {
    const long long U = static_cast<long long>(G) * kp;
    long long g = vec_blocks(n * U);
    const long long q = U > kBlock ? U / kBlock : 1;
    g = (g + q - 1) / q * q;
    return static_cast<int>(g);
}

// Y = A P for kp columns over the sliced-ELL form (sell_desc): wave q of a
// slice covers its rows [q 64/kp, (q+1) 64/kp) for all kp columns (thread:
// row lane t / kp, column t % kp), so a row's entry loads are shared by its kp
// threads and the gathers P[j][0..kp) are contiguous. A long row is walked by
// one wave per column (lane-strided + butterfly). Every column's sum is
// spmv_sell's for that row (same entries, padding and order), so each column
// gets the single sliced-ELL SpMV's bits. Items (slice waves, then long-row
// waves) are split into XCD bands as in spmv_sell. With want_dot, the
// per-column partials of p.Ap over the slices go to part.
template <class VT>
__global__ __launch_bounds__(kBlock) void spmm_sell_k(sell_desc a, int kp, const double* __restrict__ x,
                                                      double* __restrict__ y, const batch_state* bs, double* part,
                                                      int want_dot) {
    if (bs->active == 0) return;
    __shared__ double sh[kBlock];
    const int lane = threadIdx.x & (kWave - 1);
    const int c = lane & (kp - 1);
    const int band = static_cast<int>(blockIdx.x) % kXcds;
    const int per = static_cast<int>(gridDim.x) / kXcds;
    const long long items = (static_cast<long long>(a.nslices) + a.nlong) * kp;
    const long long i0 = items * band / kXcds, i1 = items * (band + 1) / kXcds;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.x) / kXcds * kWavesPerBlock +
                                                    static_cast<int>(threadIdx.x) / kWave);
    const int* __restrict__ sidx = a.idx;
    const VT* __restrict__ sval = static_cast<const VT*>(a.val);
    const long long slice_items = static_cast<long long>(a.nslices) * kp;
    double acc = 0.0;
    for (long long it = i0 + wave; it < i1; it += static_cast<long long>(per) * kWavesPerBlock) {
        if (it < slice_items) {
            const int sl = static_cast<int>(it / kp), q = static_cast<int>(it % kp);
            const int l = q * (kWave / kp) + lane / kp;
            const unsigned c0 = a.soff[sl], c1 = a.soff[sl + 1];
            const int w = static_cast<int>(c1 - c0);
            const size_t base = static_cast<size_t>(c0) * kWave + l;
            const int row = a.rowmap[static_cast<size_t>(sl) * kWave + l];
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
                for (int u = 0; u < 8; ++u) g[u] = k + u < w ? x[static_cast<size_t>(j[u]) * kp + c] : 0.0;
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (k + u < w) s = fma(v[u], g[u], s);
            }
            if (row >= 0 && row < a.n) {
                const size_t o = static_cast<size_t>(row) * kp + c;
                y[o] = s;
                acc = fma(x[o], s, acc);
            }
        } else {
            const long long li = it - slice_items;
            const int row = a.longrows[li / kp];
            const int col = static_cast<int>(li % kp);
            const unsigned b = a.ptr[row], e = a.ptr[row + 1];
            const VT* __restrict__ cval = static_cast<const VT*>(a.cval);
            double s = 0.0;
            for (unsigned p = b + lane; p < e; p += kWave)
                s = fma(static_cast<double>(cval[p]), x[static_cast<size_t>(a.cidx[p]) * kp + col], s);
            s = wave_sum(s);
            // The long rows' share of p.Ap comes from spmm_long_dot_k (its
            // waves sit in column-dependent workgroups).
            if (lane == 0) y[static_cast<size_t>(row) * kp + col] = s;
        }
    }
    if (!want_dot) return;
    double v[1] = {acc};
    col_block_partials<1>(v, kp, part, sh);
}

// Entries per row of one chunk of a slice staged in LDS by spmm_sell_staged_k.
inline constexpr int kSpmmChunk = 16;

// spmm_sell_k for kp >= 4 with each slice's entries read from HBM once per
// slice instead of once per item: a workgroup takes whole slices of its XCD
// band, copies kSpmmChunk entries per row (indices and values, coalesced) to
// LDS, and its four waves take the slice's kp items (wave w: items w, w + 4,
// ...; item q covers rows [q 64/kp, (q+1) 64/kp) for all kp columns, as in
// spmm_sell_k), reading the entries from LDS. An item's sums run over the
// same entries in the same order as spmm_sell_k's, so every column still gets
// spmv_sell's bits for its rows; the p.Ap partials are one per workgroup and
// grouped the same way for every column, so a column's total does not depend
// on its slot. The band's long rows follow, one wave per (row, column).
template <class VT>
__global__ __launch_bounds__(kBlock) void spmm_sell_staged_k(sell_desc a, int kp, const double* __restrict__ x,
                                                             double* __restrict__ y, const batch_state* bs,
                                                             double* part, int want_dot) {
    if (bs->active == 0) return;
    __shared__ double sh[kBlock];
    __shared__ int s_idx[kSpmmChunk * kWave];
    __shared__ VT s_val[kSpmmChunk * kWave];
    constexpr int kMaxItems = kWave / kWavesPerBlock;   // items per wave at kp = 64
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & (kWave - 1);
    const int wv = tid / kWave;
    const int c = lane & (kp - 1);
    const int nq = kp / kWavesPerBlock;
    const int rpi = kWave / kp;
    const int band = static_cast<int>(blockIdx.x) % kXcds;
    const int per = static_cast<int>(gridDim.x) / kXcds;
    const int slot = static_cast<int>(blockIdx.x) / kXcds;
    const int sl0 = static_cast<int>(static_cast<long long>(a.nslices) * band / kXcds);
    const int sl1 = static_cast<int>(static_cast<long long>(a.nslices) * (band + 1) / kXcds);
    const int* __restrict__ sidx = a.idx;
    const VT* __restrict__ sval = static_cast<const VT*>(a.val);
    double acc = 0.0;
    for (int sl = sl0 + slot; sl < sl1; sl += per) {
        const unsigned c0 = a.soff[sl], c1 = a.soff[sl + 1];
        const int w = static_cast<int>(c1 - c0);
        double s[kMaxItems];
#pragma unroll
        for (int i = 0; i < kMaxItems; ++i) s[i] = 0.0;
        for (int k0 = 0; k0 < w; k0 += kSpmmChunk) {
            const int cw = w - k0 < kSpmmChunk ? w - k0 : kSpmmChunk;
            const size_t g0 = static_cast<size_t>(c0 + static_cast<unsigned>(k0)) * kWave;
            for (int e = tid; e < cw * kWave; e += kBlock) {
                s_idx[e] = sidx[g0 + static_cast<size_t>(e)];
                s_val[e] = sval[g0 + static_cast<size_t>(e)];
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < kMaxItems; ++i) {
                if (i >= nq) break;
                const int l = (wv + i * kWavesPerBlock) * rpi + lane / kp;
                double t = s[i];
                for (int kk = 0; kk < cw; kk += 8) {
                    int j[8];
                    double v[8], g[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const bool ok = kk + u < cw;
                        j[u] = ok ? s_idx[(kk + u) * kWave + l] : 0;
                        v[u] = ok ? static_cast<double>(s_val[(kk + u) * kWave + l]) : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) g[u] = kk + u < cw ? x[static_cast<size_t>(j[u]) * kp + c] : 0.0;
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        if (kk + u < cw) t = fma(v[u], g[u], t);
                }
                s[i] = t;
            }
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < kMaxItems; ++i) {
            if (i >= nq) break;
            const int l = (wv + i * kWavesPerBlock) * rpi + lane / kp;
            const int row = a.rowmap[static_cast<size_t>(sl) * kWave + l];
            if (row >= 0 && row < a.n) {
                const size_t o = static_cast<size_t>(row) * kp + c;
                y[o] = s[i];
                acc = fma(x[o], s[i], acc);
            }
        }
    }
    const long long li0 = static_cast<long long>(a.nlong) * kp * band / kXcds;
    const long long li1 = static_cast<long long>(a.nlong) * kp * (band + 1) / kXcds;
    for (long long li = li0 + static_cast<long long>(slot) * kWavesPerBlock + wv; li < li1;
         li += static_cast<long long>(per) * kWavesPerBlock) {
        const int row = a.longrows[li / kp];
        const int col = static_cast<int>(li % kp);
        const unsigned b = a.ptr[row], e = a.ptr[row + 1];
        const VT* __restrict__ cval = static_cast<const VT*>(a.cval);
        double s = 0.0;
        for (unsigned p = b + lane; p < e; p += kWave)
            s = fma(static_cast<double>(cval[p]), x[static_cast<size_t>(a.cidx[p]) * kp + col], s);
        s = wave_sum(s);
        if (lane == 0) y[static_cast<size_t>(row) * kp + col] = s;
    }
    if (!want_dot) return;
    double v[1] = {acc};
    col_block_partials<1>(v, kp, part, sh);
}

// p.Ap over the long rows of the sliced-ELL form, per column in long-row
// order, stored as one more block of partials (block nb, after
// spmm_sell_k's): a column's total does not depend on its slot.
__global__ __launch_bounds__(kBlock) void spmm_long_dot_k(sell_desc a, int kp, const double* __restrict__ x,
                                                          const double* __restrict__ y, const batch_state* bs,
                                                          double* part, int nb) {
    if (bs->active == 0) return;
    __shared__ double sh[kBlock];
    const int c = threadIdx.x & (kp - 1);
    const int j = threadIdx.x / kp, rows = kBlock / kp;
    double s = 0.0;
    for (int i = j; i < a.nlong; i += rows) {
        const size_t o = static_cast<size_t>(a.longrows[i]) * kp + c;
        s = fma(x[o], y[o], s);
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < kp) {
        double t = sh[c];
        for (int r = 1; r < rows; ++r) t += sh[r * kp + c];
        part[static_cast<size_t>(nb) * kp + c] = t;
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

// Partials of rr, rs per column of r (block_finish kBFinInit starts each
// joining column's solve: zero right-hand side, max_iter, first centring
// decision; other columns are left as they are).
__global__ __launch_bounds__(kBlock) void init_residual_k(int n, int kp, const double* __restrict__ r, double* part) {
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
    col_block_partials<2>(v, kp, part, sh);
}

// x += alpha p; r -= alpha Ap per live column; partials of rr, rs (each
// column's loop bookkeeping, as update_xr's, in block_finish kBFinXR).
__global__ __launch_bounds__(kBlock) void update_xr_k(int n, int kp, double* __restrict__ x,
                                                      const double* __restrict__ p, double* __restrict__ r,
                                                      const double* __restrict__ ap, const pcg_state* st,
                                                      const batch_state* bs, double* part) {
    if (bs->active == 0) return;
    __shared__ double sh[2 * kBlock];
    const int c = threadIdx.x & (kp - 1);
    const bool live = !st[c].done && st[c].phase == 0;
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
    col_block_partials<2>(v, kp, part, sh);
}

// Partials of r.out, sum(out) per column (centring shift mu, rz, beta --
// dot_center's -- in block_finish kBFinDot).
template <class TO>
__global__ __launch_bounds__(kBlock) void dot_center_k(int n, int kp, const double* __restrict__ r,
                                                       const TO* __restrict__ out, const pcg_state* st,
                                                       const batch_state* bs, double* part) {
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
    col_block_partials<2>(v, kp, part, sh);
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
    const bool fresh = first || st[c].phase == 1;
    const int rows_per_grid = gridDim.x * (kBlock / kp);
    for (int row = (blockIdx.x * kBlock + threadIdx.x) / kp; row < n; row += rows_per_grid) {
        const size_t i = static_cast<size_t>(row) * kp + c;
        const double zi = static_cast<double>(out[i]) - mu;
        p[i] = fresh ? zi : fma(beta, p[i], zi);
    }
}

// End of an iteration: joining columns have had their first application.
__global__ void end_iteration_k(int kp, pcg_state* st, batch_state* bs) {
    if (bs->active == 0) return;
    const int c = threadIdx.x;
    if (c < kp && !st[c].done && st[c].phase == 1) st[c].phase = 0;
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

// Admission of columns into a running block: for j < ncol, column cols[j]
// gets r = the right-hand side srcs[j] (n doubles, caller order) permuted,
// and x = 0.
__global__ __launch_bounds__(kBlock) void admit_scatter_k(int n, int kp, const unsigned* __restrict__ perm, int ncol,
                                                          const int* __restrict__ cols,
                                                          const double* const* __restrict__ srcs,
                                                          double* __restrict__ r, double* __restrict__ x) {
    const long long total = static_cast<long long>(n) * ncol;
    for (long long t = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; t < total;
         t += static_cast<long long>(gridDim.x) * kBlock) {
        const int j = static_cast<int>(t / n);
        const int v = static_cast<int>(t - static_cast<long long>(j) * n);
        const size_t q = static_cast<size_t>(perm[v]) * kp + cols[j];
        r[q] = srcs[j][v];
        x[q] = 0.0;
    }
}

// Collection of finished columns: dsts[j][v] = x[perm[v]][cols[j]] - mean
// (the Laplacian's min-norm centring, from st[c].out_sum, when use_mean).
__global__ __launch_bounds__(kBlock) void gather_cols_k(int n, int kp, const unsigned* __restrict__ perm, int ncol,
                                                        const int* __restrict__ cols, double* const* __restrict__ dsts,
                                                        const double* __restrict__ x, const pcg_state* st,
                                                        int use_mean) {
    const long long total = static_cast<long long>(n) * ncol;
    for (long long t = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; t < total;
         t += static_cast<long long>(gridDim.x) * kBlock) {
        const int j = static_cast<int>(t / n);
        const int v = static_cast<int>(t - static_cast<long long>(j) * n);
        const int c = cols[j];
        const double sh = use_mean ? st[c].out_sum / static_cast<double>(n) : 0.0;
        dsts[j][v] = x[static_cast<size_t>(perm[v]) * kp + c] - sh;
    }
}

// Partials of the column sums (block_finish kBFinOutSum: st[c].out_sum, for
// the Laplacian min-norm centring of the result and the centring shifts of a
// blocked application).
template <class T>
__global__ __launch_bounds__(kBlock) void sum_cols_k(int n, int kp, const T* __restrict__ x, double* part) {
    __shared__ double sh[kBlock];
    const int c = threadIdx.x & (kp - 1);
    const int rows_per_grid = gridDim.x * (kBlock / kp);
    double s = 0.0;
    for (int row = (blockIdx.x * kBlock + threadIdx.x) / kp; row < n; row += rows_per_grid)
        s += static_cast<double>(x[static_cast<size_t>(row) * kp + c]);
    double v[1] = {s};
    col_block_partials<1>(v, kp, part, sh);
}

// Stage 2 of the blocked reductions (one workgroup, launched right after the
// producer of `part`; nb = the producer's grid), per column c < kp:
//   kBFinPAp        st[c].pAp of an iterating column; breakdown ends it
//   kBFinInit       start of each joining column (init_residual's decisions)
//   kBFinXR         each live column's loop bookkeeping (update_xr's)
//   kBFinDot(First) r.z, mu, beta (dot_center's; joining columns start fresh)
//   kBFinOutSum     st[c].out_sum
// The in-loop ops do nothing when no column is active, as their producers.
enum : int { kBFinPAp, kBFinInit, kBFinXR, kBFinDot, kBFinDotFirst, kBFinOutSum };

__global__ __launch_bounds__(kBlock) void block_finish(int op, int n, int kp, const double* __restrict__ part, int nb,
                                                       pcg_state* st, const batch_state* bs) {
    if ((op == kBFinPAp || op == kBFinXR || op == kBFinDot || op == kBFinDotFirst) && bs->active == 0) return;
    __shared__ double sh[2 * kBlock];
    double v[2] = {0.0, 0.0};
    bool mine;
    if (op == kBFinPAp || op == kBFinOutSum) {
        double w[1];
        mine = col_finish_sum<1>(w, kp, part, nb, sh);
        v[0] = w[0];
    } else {
        mine = col_finish_sum<2>(v, kp, part, nb, sh);
    }
    if (!mine) return;
    pcg_state& t = st[threadIdx.x];
    switch (op) {
        case kBFinPAp:
            if (!t.done && t.phase == 0) {
                t.pAp = v[0];
                if (v[0] <= 0.0) { t.done = 1; t.status = 2; }
            }
            break;
        case kBFinInit:
            if (!t.done && t.phase == 1) {
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
            break;
        case kBFinXR:
            if (!t.done && t.phase == 0) {
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
            break;
        case kBFinDot:
        case kBFinDotFirst:
            if (!t.done) {
                const double mu = t.center_now ? v[1] / static_cast<double>(n) : 0.0;
                const double rz = v[0] - mu * t.rs;
                t.out_sum = v[1];
                t.mu = mu;
                if (op == kBFinDotFirst || t.phase == 1) {
                    t.rz = rz;
                    t.beta = 0.0;
                } else {
                    t.beta = rz / t.rz;
                    t.rz = rz;
                }
            }
            break;
        case kBFinOutSum:
            t.out_sum = v[0];
            break;
        default:
            break;
    }
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
