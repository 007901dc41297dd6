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
template <int N>
__device__ __forceinline__ void block_partials(double (&v)[N], double* part, double* sh) {
    block_sum<N>(v, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) part[blockIdx.x * N + k] = v[k];
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
