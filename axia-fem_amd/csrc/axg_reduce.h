// Deterministic device reductions shared by the PCG kernels.
//
// A kernel reduces its values to one partial per workgroup (fixed tree) and
// stores it; a one-workgroup finalize kernel launched right after sums the
// partials with a fixed tree and updates the solver scalars. The kernel
// boundary orders the partial stores before the finalize reads (no
// per-workgroup device-scope fences, which on gfx950 write back the L2 of the
// XCD). Results do not depend on scheduling: same inputs and launch shape,
// same bits.
#pragma once
#include <hip/hip_runtime.h>

namespace axg {

__device__ inline double wave_sum64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Workgroup sum of NC values per thread (BSZ threads); result in thread 0.
template <int NC, int BSZ>
__device__ inline void group_sums(double (&v)[NC]) {
    __shared__ double sh[NC][BSZ / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NC; j++) v[j] = wave_sum64(v[j]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < NC; j++) sh[j][wid] = v[j];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < NC; j++) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < BSZ / 64; w++) t += sh[j][w];
            v[j] = t;
        }
}

// Sum over b < nblk of part[b * NC + j] for every j, by the whole workgroup
// (fixed strided order, then a fixed tree); result valid in every thread.
template <int NC, int BSZ>
__device__ inline void sum_partials(const double *part, int nblk, double (&out)[NC]) {
    __shared__ double t[NC][BSZ];
    double s[NC];
#pragma unroll
    for (int j = 0; j < NC; j++) s[j] = 0.0;
    for (int i = threadIdx.x; i < nblk; i += BSZ)
#pragma unroll
        for (int j = 0; j < NC; j++) s[j] += part[(size_t)i * NC + j];
#pragma unroll
    for (int j = 0; j < NC; j++) t[j][threadIdx.x] = s[j];
    __syncthreads();
    for (int off = BSZ / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
#pragma unroll
            for (int j = 0; j < NC; j++) t[j][threadIdx.x] += t[j][threadIdx.x + off];
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < NC; j++) out[j] = t[j][0];
    __syncthreads();
}

}  // namespace axg
