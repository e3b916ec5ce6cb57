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

// Workgroups of one kernel resident per CU (its occupancy), at least 1.
inline int resident_per_cu(const void *kernel, int bsz) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, bsz, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return per_cu;
}

// Rows of a streaming operator kernel, by workgroup. Workgroups are dealt to
// the 8 XCDs round-robin (workgroup b starts on XCD b % 8), each XCD with its
// own L2. Renumbering them so that one XCD's workgroups hold consecutive
// logical ids, and giving each logical id one contiguous run of row blocks,
// hands every XCD one band of rows: the x gathers of a banded matrix then hit
// that XCD's L2 instead of being fetched once per XCD. The logical id also
// indexes the workgroup's reduction partial, so sums keep a fixed order.
struct RowBand {
    uint32_t id;          // logical workgroup id
    int64_t first, end;   // rows [first, end)
    int64_t step;
};

__device__ inline RowBand row_band(int64_t nrows, int bsz) {
    const uint32_t g = gridDim.x, b = blockIdx.x;
    RowBand r;
    r.id = (g % 8 == 0) ? (b % 8) * (g / 8) + b / 8 : b;
    r.first = (int64_t)r.id * bsz + threadIdx.x;
    r.end = nrows;
    r.step = (int64_t)g * bsz;
    return r;
}

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
