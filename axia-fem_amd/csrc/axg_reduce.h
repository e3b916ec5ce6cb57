// Deterministic device reductions shared by the PCG kernels.
//
// A kernel reduces its values to one partial per workgroup (fixed tree) and
// stores it; a one-workgroup finalize kernel launched right after sums the
// partials in a fixed order and updates the solver scalars. The kernel
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

// Threads of a finalize kernel (one or a few workgroups that sum partials).
constexpr int FIN_TB = 1024;

// Sum over b < nblk of part[b * NC + j] for every j, by a workgroup of FIN_TB
// threads; result valid in every thread. The records are read as one flat
// array: thread t takes elements t, t + FIN_TB, ... (consecutive lanes on
// consecutive doubles, eight loads in flight per thread), all of which belong
// to column t % NC since FIN_TB is a multiple of NC; lanes of equal column are
// then folded by shuffles and the waves through LDS. The order is fixed for a
// given nblk. The partials were written by other XCDs, so the reads miss this
// XCD's L2: the cost is a few round trips, not bandwidth.
// TB: threads of the workgroup that sums (FIN_TB for the finalize kernels;
// a vector kernel's own workgroup size when every workgroup sums the partials
// of the previous kernel itself, all of them in the same order).
template <int NC, int TB = FIN_TB>
__device__ inline void fin_sums(const double *__restrict__ part, int nblk, double (&out)[NC]) {
    static_assert(NC >= 1 && NC <= 64 && (NC & (NC - 1)) == 0, "NC must be a power of two");
    static_assert(TB % 64 == 0 && TB >= 64, "whole wavefronts");
    const int64_t nflat = (int64_t)nblk * NC;
    double a[8];
#pragma unroll
    for (int u = 0; u < 8; u++) a[u] = 0.0;
    int64_t f = threadIdx.x;
    for (; f + 7 * TB < nflat; f += 8 * TB)
#pragma unroll
        for (int u = 0; u < 8; u++) a[u] += part[f + u * TB];
#pragma unroll
    for (int u = 0; u < 7; u++)
        if (f + u * TB < nflat) a[u] += part[f + u * TB];
    double s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
#pragma unroll
    for (int off = 32; off >= NC; off >>= 1) s += __shfl_down(s, off, 64);
    __shared__ double sh[TB / 64][NC];
    __shared__ double tot[NC];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane < NC) sh[w][lane] = s;
    __syncthreads();
    if (threadIdx.x < NC) {
        double t = 0.0;
#pragma unroll
        for (int v = 0; v < TB / 64; v++) t += sh[v][threadIdx.x];
        tot[threadIdx.x] = t;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NC; j++) out[j] = tot[j];
}

}  // namespace axg
