// Stand-alone PCG pieces on caller-owned device arrays: the device boundary of
// the reference's GPU PCG (include/apxchol/solver/pcg_cuda_kernels.h:43-93,
// src/cuda_pcg_kernels.cu:18-152) -- y = A x with the partials of p.Ap, the
// fused x / r update with the partials of r.r, a dot product, the p update
// and the fixed-order reduction of the partials. Scalars come from the host
// (alpha, beta), the partials buffer from the caller: one double per
// workgroup of the producer, whose grid is a pure function of the size
// (vec_blocks), so every reduction is bit-identical run to run.
//
// The SpMV piece is spmv_pAp (pcg.hip) without a solve state.
#include "common.hpp"

namespace apxchol_amd {

// x += alpha p; r -= alpha Ap; partials of r.r on the updated r.
__global__ __launch_bounds__(kBlock) void piece_update_xr(int n, double* __restrict__ x, const double* __restrict__ p,
                                                          double* __restrict__ r, const double* __restrict__ ap,
                                                          double alpha, double* part) {
    __shared__ double sh[kWavesPerBlock * 2];
    double q = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        x[i] = fma(alpha, p[i], x[i]);
        const double ri = fma(-alpha, ap[i], r[i]);
        r[i] = ri;
        q = fma(ri, ri, q);
    }
    double v[1] = {q};
    block_partials<1>(v, part, sh);
}

// Partials of a.b.
__global__ __launch_bounds__(kBlock) void piece_dot(int n, const double* __restrict__ a, const double* __restrict__ b,
                                                    double* part) {
    __shared__ double sh[kWavesPerBlock * 2];
    double q = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) q = fma(a[i], b[i], q);
    double v[1] = {q};
    block_partials<1>(v, part, sh);
}

// p = z + beta p.
__global__ __launch_bounds__(kBlock) void piece_update_p(int n, double* __restrict__ p, const double* __restrict__ z,
                                                         double beta) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) p[i] = fma(beta, p[i], z[i]);
}

// out[0] = the partials part[0..nb) summed in block order (finish_sum's fixed
// order), one workgroup.
__global__ __launch_bounds__(kBlock) void piece_reduce(const double* __restrict__ part, int nb, double* out) {
    __shared__ double sh[kWavesPerBlock * 2];
    double v[1];
    if (finish_sum<1>(v, part, nb, sh)) out[0] = v[0];
}

}  // namespace apxchol_amd
