// Wave64 / workgroup reductions with a FIXED combination order (the result is a
// function of the inputs and the launch shape only, never of scheduling), and
// the flagged loads/stores the sync-free sweeps use.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "../device.hpp"

namespace apxchol_amd {

inline constexpr int kBlock = 256;          // threads per workgroup of the vector kernels
inline constexpr int kXcds = 8;             // MI355X: accelerator dies, each with its own L2
inline constexpr int kWavesPerBlock = kBlock / kWave;

// Butterfly over the whole wave: every lane ends with the same total.
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Lane i's copy of lane i + OFF of its row of 16 lanes (OFF < 16), by a DPP
// row shift; lanes whose source lies past the row read 0.
template <int OFF>
__device__ __forceinline__ int dpp_from_right(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x100 + OFF, 0xF, 0xF, false);
}
template <int OFF>
__device__ __forceinline__ float dpp_from_right(float v) {
    return __builtin_bit_cast(float, dpp_from_right<OFF>(__builtin_bit_cast(int, v)));
}
template <int OFF>
__device__ __forceinline__ double dpp_from_right(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = static_cast<unsigned>(dpp_from_right<OFF>(static_cast<int>(b & 0xFFFFFFFFull)));
    const unsigned hi = static_cast<unsigned>(dpp_from_right<OFF>(static_cast<int>(b >> 32)));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// Lane i's copy of lane i ^ OFF (OFF = 16 or 32) by gfx950's VALU row swaps:
// permlane16_swap(x, x) leaves in its first result the even row of each row
// pair in both rows, in its second the odd row; permlane32_swap does the same
// for the two halves. The partner of lane i is the other row (half). Lane i
// is threadIdx.x mod 64: every kernel here runs one-dimensional workgroups.
template <int OFF>
__device__ __forceinline__ unsigned swap_partner(unsigned v) {
    const auto r = OFF == 16 ? __builtin_amdgcn_permlane16_swap(v, v, false, false)
                             : __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (threadIdx.x & OFF) ? r[0] : r[1];
}
template <int OFF>
__device__ __forceinline__ float swap_partner(float v) {
    return __builtin_bit_cast(float, swap_partner<OFF>(__builtin_bit_cast(unsigned, v)));
}
template <int OFF>
__device__ __forceinline__ double swap_partner(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = swap_partner<OFF>(static_cast<unsigned>(b & 0xFFFFFFFFull));
    const unsigned hi = swap_partner<OFF>(static_cast<unsigned>(b >> 32));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// The butterfly of a lane group's partial sums as the group's first lane
// sees it: for off = 32, 16, ..., 1 with off < G, lane s adds lane s + off
// (groups are aligned to their power-of-two size G). Only the first lane of
// each group holds the group's sum afterwards; its value is the xor
// butterfly's (v[s] += v[s ^ off] gives lane s < off the same operand).
// Steps below 16 read the partner with a DPP row shift, the two wider ones
// with a row swap (both stay in the VALU, no LDS round trip); a step no lane
// needs is skipped as a whole.
template <int OFF, class T>
__device__ __forceinline__ T tree_step(T part, int G) {
    if (__ballot(OFF < G) == 0) return part;
    T o;
    if constexpr (OFF >= 16) o = swap_partner<OFF>(part);
    else o = dpp_from_right<OFF>(part);
    return OFF < G ? part + o : part;
}
template <class T>
__device__ __forceinline__ T group_tree_sum(T part, int G) {
    part = tree_step<32>(part, G);
    part = tree_step<16>(part, G);
    part = tree_step<8>(part, G);
    part = tree_step<4>(part, G);
    part = tree_step<2>(part, G);
    return tree_step<1>(part, G);
}

// A sweep row's output from its numerator (right-hand side minus the row's
// sum) and the row's diagonal operand: fp32 vectors multiply by the fp32
// reciprocal of the diagonal (sweep_desc::inv_diag; one dependent multiply
// instead of the dozen-instruction IEEE division, at most one more rounding),
// fp64 vectors divide by the diagonal. Every sweep kernel forms its outputs
// through these two, so all schedules and layouts keep the same bits.
template <class AT>
__device__ __forceinline__ float diag_operand(const sweep_desc& d, int row) {
    if constexpr (sizeof(AT) == sizeof(float)) return d.inv_diag[row];
    else return d.diag[row];
}
template <class AT>
__device__ __forceinline__ AT finish_row(AT num, float operand) {
    if constexpr (sizeof(AT) == sizeof(float)) return num * operand;
    else return num / static_cast<AT>(operand);
}

// Sum over a workgroup of kBlock threads: per-wave butterfly, then wave 0..3
// in order. Valid in thread 0 only. `sh` holds kWavesPerBlock * N values.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double* sh) {
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) sh[w * N + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double s = sh[k];
            for (int j = 1; j < kWavesPerBlock; ++j) s += sh[j * N + k];
            v[k] = s;
        }
}

// Deterministic grid reductions in two launches. Stage 1 runs in every
// workgroup of the producing kernel: the workgroup's totals go to
// part[block * N + k] with plain stores. Stage 2 (finish_sum) runs in a
// separate one-workgroup launch that the stream orders after the producer;
// the kernel boundary makes the partials visible. The producer therefore
// needs no fence: an agent-scope fence in every workgroup writes back and
// invalidates its XCD's L2 each time, which halved the bandwidth of the
// streaming kernels that reduce.
// blk: the producer's workgroup index (its slot in part); a grouped launch
// passes the workgroup's index inside its segment's own grid.
template <int N>
__device__ __forceinline__ void block_partials(double (&v)[N], double* part, double* sh, int blk) {
    block_sum<N>(v, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) part[blk * N + k] = v[k];
}
template <int N>
__device__ __forceinline__ void block_partials(double (&v)[N], double* part, double* sh) {
    block_partials<N>(v, part, sh, static_cast<int>(blockIdx.x));
}

// Stage 2 over the partials of nb workgroups, in block order: strided
// per-thread sums, then block_sum. True in thread 0, with the totals in v.
template <int N>
__device__ __forceinline__ bool finish_sum(double (&v)[N], const double* part, int nb, double* sh) {
    double acc[N];
#pragma unroll
    for (int k = 0; k < N; ++k) acc[k] = 0.0;
    // Eight blocks' loads in flight at once, added in block order.
    int b = static_cast<int>(threadIdx.x);
    for (; b + 7 * kBlock < nb; b += 8 * kBlock) {
        double g[8][N];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int k = 0; k < N; ++k) g[u][k] = part[(b + u * kBlock) * N + k];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int k = 0; k < N; ++k) acc[k] += g[u][k];
    }
    for (; b < nb; b += kBlock)
#pragma unroll
        for (int k = 0; k < N; ++k) acc[k] += part[b * N + k];
    block_sum<N>(acc, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = acc[k];
    return threadIdx.x == 0;
}

// Grid shape of the vector kernels: a pure function of n and of the
// process's cap (so reductions are reproducible), at most kMaxBlocks
// workgroups (APXCHOL_AMD_VEC_BLOCKS lowers the cap).
inline constexpr int kMaxBlocks = 2048;
int vec_block_cap();
inline int vec_blocks(long long n) {
    long long b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    const int cap = vec_block_cap();
    return static_cast<int>(b > cap ? cap : b);
}

template <class T> struct flag_bits;
template <> struct flag_bits<double> { using type = unsigned long long; static constexpr type pending = kPending64; };
template <> struct flag_bits<float> { using type = unsigned int; static constexpr type pending = kPending32; };

template <class T>
__device__ __forceinline__ typename flag_bits<T>::type load_flag(const T* p) {
    using B = typename flag_bits<T>::type;
    return __hip_atomic_load(reinterpret_cast<const B*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__device__ __forceinline__ void store_flag(T* p, T v) {
    using B = typename flag_bits<T>::type;
    __hip_atomic_store(reinterpret_cast<B*>(p), __builtin_bit_cast(B, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__device__ __forceinline__ T from_bits(typename flag_bits<T>::type b) { return __builtin_bit_cast(T, b); }

template <class T>
__device__ __forceinline__ T pending_value() { return __builtin_bit_cast(T, flag_bits<T>::pending); }

}  // namespace apxchol_amd
