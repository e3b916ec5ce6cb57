// Batched solves: many meshes, several right-hand sides each, advanced
// together by one sequence of launches per CG iteration (captured once into a
// hipGraph and replayed).
//
// The meshes' FP64 node-block operators (3x3 blocks, block-sliced ELL) are
// concatenated into one block-diagonal operator. Each mesh is a segment whose
// block rows start on a 768-row boundary, so every SpMM slice and every vector
// workgroup works on exactly one segment and its partial sums belong to one
// mesh; the finalize kernels run one workgroup per segment. Scalars (alpha, beta, r.z, ||r||, done) are kept per (segment,
// right-hand side): every solve follows the reference's Jacobi PCG
// (linalg.rs:194-259) on its own, converging at its own iteration, and its
// convergence time is stamped from the device's constant-rate clock.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include <hip/hip_ext.h>

#include "axg_internal.h"
#include "axg_mixed.h"
#include "axg_pcg.h"
#include "axg_reduce.h"

namespace axg {

constexpr int TB = 256;

struct SegScalars {
    double rz, alpha, beta, res, tol_abs, bnorm;
    int64_t iters;
    int64_t t_done;   // s_memrealtime stamp at convergence
    int32_t done, status;
};

// Row structure of a system of segments, for the vector and finalize kernels.
struct BatchDev {
    const double *binv;     // 3x3 inverse per block row (block-Jacobi), or null (Jacobi)
    int32_t rows_wg;        // rows per vector workgroup: TB, or 255 (85 whole nodes) with binv
    const double *dsel;     // Jacobi divisor per row (1 on padding rows)
    const int32_t *wg_seg;  // segment of each workgroup
    const int64_t *wg_row;  // first row of each workgroup
    const int64_t *seg_row; // seg_row[m + 1]: end of the real rows of segment m
    const int32_t *seg_wg;  // first workgroup per segment (nseg + 1)
    int32_t *seg_live;      // right-hand sides still iterating per segment
};

__device__ inline int64_t now_ticks() { return (int64_t)__builtin_amdgcn_s_memrealtime(); }

__global__ void stamp_kernel(int64_t *t) { *t = now_ticks(); }

// One workgroup per segment: sums of its workgroups' partials (fixed order).
template <int NC>
__device__ inline void seg_sums(const BatchDev &B, const double *part, int seg, double (&out)[NC]) {
    const int w0 = B.seg_wg[seg], w1 = B.seg_wg[seg + 1];
    fin_sums<NC>(part + (size_t)w0 * NC, w1 - w0, out);
}

// The finalize kernels load the scalars they need before the partials and
// test the stop flags after the sum, so the reads overlap instead of costing
// one memory round trip each.
template <int K>
__global__ void __launch_bounds__(FIN_TB) batch_fin_pap_kernel(BatchDev B, const double *part,
                                                          SegScalars *sc, int32_t *live_total) {
    const int seg = blockIdx.x;
    const int live = B.seg_live[seg];
    int done[K];
    double rz[K];
#pragma unroll
    for (int j = 0; j < K; j++) done[j] = sc[seg * K + j].done, rz[j] = sc[seg * K + j].rz;
    double pap[K];
    seg_sums<K>(B, part, seg, pap);
    if (live == 0 || threadIdx.x != 0) return;
    for (int j = 0; j < K; j++) {
        SegScalars &s = sc[seg * K + j];
        if (done[j]) continue;
        if (fabs(pap[j]) < 1e-30) {  // linalg.rs:231-233
            s.status = AXG_E_NOT_SPD;
            s.done = 1;
            s.t_done = now_ticks();
            B.seg_live[seg]--;
            atomicSub(live_total, 1);
        } else {
            s.alpha = rz[j] / pap[j];
        }
    }
}

// The vector kernels work on a workgroup's TB rows x K right-hand sides as one
// flat range: element t = (row - row0) K + j, thread tid takes t = tid + q TB
// (q < K). Consecutive lanes touch consecutive doubles, and since TB is a
// multiple of K every element of a thread belongs to right-hand side
// j = tid % K, so a thread accumulates two sums and the per-right-hand-side
// totals are shuffles over the lanes of equal j.
template <int K>
__device__ inline void flat_sums_out(double rr, double rz, double *__restrict__ part) {
    for (int off = 32; off >= K; off >>= 1) {
        rr += __shfl_down(rr, off, 64);
        rz += __shfl_down(rz, off, 64);
    }
    __shared__ double sh[TB / 64][2 * K];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane < K) {
        sh[w][lane] = rr;
        sh[w][K + lane] = rz;
    }
    __syncthreads();
    if (threadIdx.x < 2 * K) {
        double t = 0.0;
#pragma unroll
        for (int v = 0; v < TB / 64; v++) t += sh[v][threadIdx.x];
        part[(size_t)blockIdx.x * 2 * K + threadIdx.x] = t;
    }
}

// z for the elements of one workgroup from its residual slab rl (rows row0 ..,
// [row][K] in LDS): Jacobi z = r / d, or block-Jacobi, row d of the node's 3x3
// inverse times the node's three residuals (the slab holds whole nodes).
template <int K, bool BLK>
__device__ inline double precond_z(const BatchDev &B, const double *rl, int t, int64_t row,
                                   int64_t row0, double r) {
    if (!BLK) return r / B.dsel[row];
    const int lr = t / K, j = t % K, d = (int)(row % 3);
    const double *Mi = B.binv + 9 * (row / 3) + 3 * d;
    const double *rn = rl + (lr - d) * K + j;  // the node's first row
    return Mi[0] * rn[0] + Mi[1] * rn[K] + Mi[2] * rn[2 * K];
}

// x += alpha p, r -= alpha Ap, z = M^-1 r, ||r||^2, r.z (per right-hand side).
// With BLK the updated residuals go through LDS first: z of a row needs the
// other rows of its node, which other threads hold.
template <int K, bool BLK>
__global__ void __launch_bounds__(TB) batch_update_kernel(BatchDev B, double *__restrict__ X,
                                                         double *__restrict__ R,
                                                         const double *__restrict__ P,
                                                         const double *__restrict__ AP,
                                                         double *__restrict__ Z,
                                                         double *__restrict__ part,
                                                         const SegScalars *sc) {
    static_assert(TB % K == 0, "flat layout needs TB % K == 0");
    __shared__ double rl[BLK ? TB * K : 1];
    const int seg = B.wg_seg[blockIdx.x];
    if (B.seg_live[seg] == 0) return;
    const int64_t row0 = B.wg_row[blockIdx.x], rend = B.seg_row[seg + 1];
    const int tend = B.rows_wg * K;
    const int j = threadIdx.x % K;
    const SegScalars &s = sc[seg * K + j];
    const bool live = !s.done;
    const double alpha = s.alpha;
    double rr = 0.0, rz = 0.0;
    double rk[K];
#pragma unroll
    for (int q = 0; q < K; q++) {
        const int t = threadIdx.x + q * TB;
        const int64_t row = row0 + t / K;
        rk[q] = 0.0;
        if (t >= tend || row >= rend) break;  // t / K grows with q
        const int64_t i = row0 * K + t;
        double r = R[i];
        if (live) {
            X[i] += alpha * P[i];
            r -= alpha * AP[i];
            R[i] = r;
        }
        rk[q] = r;
        rr += r * r;
        if (BLK) {
            rl[t] = r;
        } else {
            const double z = precond_z<K, false>(B, nullptr, t, row, row0, r);
            Z[i] = z;
            rz += r * z;
        }
    }
    if (BLK) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < K; q++) {
            const int t = threadIdx.x + q * TB;
            const int64_t row = row0 + t / K;
            if (t >= tend || row >= rend) break;
            const double z = precond_z<K, true>(B, rl, t, row, row0, rk[q]);
            Z[row0 * K + t] = z;
            rz += rk[q] * z;
        }
    }
    flat_sums_out<K>(rr, rz, part);
}

template <int K>
__global__ void __launch_bounds__(FIN_TB) batch_fin_update_kernel(BatchDev B, const double *part,
                                                             SegScalars *sc, int init,
                                                             double tol, const double *tol_rel,
                                                             int64_t max_iter,
                                                             int32_t *live_total) {
    const int seg = blockIdx.x;
    const int live = B.seg_live[seg];
    int done[K];
    int64_t iters[K];
    double rz[K], tol_abs[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        const SegScalars &s = sc[seg * K + j];
        done[j] = s.done, iters[j] = s.iters, rz[j] = s.rz, tol_abs[j] = s.tol_abs;
    }
    double tot[2 * K];
    seg_sums<2 * K>(B, part, seg, tot);
    if ((!init && live == 0) || threadIdx.x != 0) return;
    int finished = 0;
    for (int j = 0; j < K; j++) {
        SegScalars &s = sc[seg * K + j];
        if (init) {
            const double bn = sqrt(tot[j]);
            s.bnorm = bn;
            s.res = bn;
            s.tol_abs = (tol_rel ? tol_rel[seg * K + j] : tol) * bn;
            s.rz = tot[K + j];
            if (bn < 1e-30) {  // zero right-hand side: x = 0 (linalg.rs:206-214)
                s.done = 1;
                s.t_done = now_ticks();
                finished++;
            }
            continue;
        }
        if (done[j]) continue;
        const double res = sqrt(tot[j]);
        s.res = res;
        s.iters = iters[j] + 1;
        if (res <= tol_abs[j] || iters[j] + 1 >= max_iter) {
            s.done = 1;
            if (res > tol_abs[j] && res > 10.0 * tol_abs[j]) s.status = AXG_E_NOT_CONVERGED;
            s.t_done = now_ticks();
            finished++;
        } else {
            s.beta = tot[K + j] / rz[j];
            s.rz = tot[K + j];
        }
    }
    if (finished) {
        B.seg_live[seg] -= finished;
        atomicSub(live_total, finished);
    }
}

template <int K>
__global__ void __launch_bounds__(TB) batch_update_p_kernel(BatchDev B, double *__restrict__ P,
                                                           const double *__restrict__ Z,
                                                           const SegScalars *sc, int init) {
    const int seg = B.wg_seg[blockIdx.x];
    if (!init && B.seg_live[seg] == 0) return;
    const int64_t row0 = B.wg_row[blockIdx.x], rend = B.seg_row[seg + 1];
    const int j = threadIdx.x % K;
    const SegScalars &s = sc[seg * K + j];
    if (!init && s.done) return;
    const double beta = s.beta;
#pragma unroll
    for (int q = 0; q < K; q++) {
        const int t = threadIdx.x + q * TB;
        if (t >= B.rows_wg * K || row0 + t / K >= rend) break;
        const int64_t i = row0 * K + t;
        P[i] = init ? Z[i] : Z[i] + beta * P[i];
    }
}

// x = 0, r = b, z = M^-1 r and the sums for ||b|| and r.z.
template <int K, bool BLK>
__global__ void __launch_bounds__(TB) batch_init_kernel(BatchDev B, const double *__restrict__ Bv,
                                                       double *__restrict__ X,
                                                       double *__restrict__ R,
                                                       double *__restrict__ Z,
                                                       double *__restrict__ part) {
    __shared__ double rl[BLK ? TB * K : 1];
    const int seg = B.wg_seg[blockIdx.x];
    const int64_t row0 = B.wg_row[blockIdx.x], rend = B.seg_row[seg + 1];
    const int tend = B.rows_wg * K;
    double rr = 0.0, rz = 0.0;
    double bk[K];
#pragma unroll
    for (int q = 0; q < K; q++) {
        const int t = threadIdx.x + q * TB;
        const int64_t row = row0 + t / K;
        bk[q] = 0.0;
        if (t >= tend || row >= rend) break;
        const int64_t i = row0 * K + t;
        const double b = Bv[i];
        X[i] = 0.0;
        R[i] = b;
        bk[q] = b;
        rr += b * b;
        if (BLK) {
            rl[t] = b;
        } else {
            const double z = precond_z<K, false>(B, nullptr, t, row, row0, b);
            Z[i] = z;
            rz += b * z;
        }
    }
    if (BLK) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < K; q++) {
            const int t = threadIdx.x + q * TB;
            const int64_t row = row0 + t / K;
            if (t >= tend || row >= rend) break;
            const double z = precond_z<K, true>(B, rl, t, row, row0, bk[q]);
            Z[row0 * K + t] = z;
            rz += bk[q] * z;
        }
    }
    flat_sums_out<K>(rr, rz, part);
}

// a[i] = v
__global__ void fill_kernel(int64_t n, double *__restrict__ a, double v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// ------------------------------------------------- one mesh, device vectors
// The block PCG of one mesh runs in node-block space: row 3 I + d is DOF d of
// block row I (a node with a free DOF); rows of prescribed DOFs carry b = 0,
// divisor 1 and zero matrix entries, so they stay exactly zero. Vectors are
// row-major [row][K]. The operator is the FP64 block-sliced-ELL copy of K_ff.

// right-hand side columns (vector-major, n each) -> block rows, zero on
// prescribed and padding rows and on columns beyond kk
template <int K>
__global__ void pack_blocks_kernel(int64_t npad, int64_t rows, int kk,
                                   const int32_t *__restrict__ fob, int64_t n,
                                   const double *__restrict__ b, double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npad * K) return;
    const int64_t r = t / K;
    const int j = (int)(t % K);
    const int32_t f = r < rows ? fob[r] : -1;
    out[t] = (f >= 0 && j < kk) ? b[(int64_t)j * n + f] : 0.0;
}

template <int K>
__global__ void unpack_blocks_kernel(int64_t n, int kk, const int32_t *__restrict__ bof,
                                     const double *__restrict__ in, double *__restrict__ x) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * kk) return;
    const int j = (int)(t / n);
    const int64_t i = t % n;
    x[t] = in[(int64_t)bof[i] * K + j];
}

// Jacobi divisor per block row: the free DOF's, 1 on prescribed and padding rows
__global__ void block_dsel_kernel(int64_t npad, int64_t rows, const int32_t *__restrict__ fob,
                                  const double *__restrict__ dsel, double *__restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= npad) return;
    const int32_t f = r < rows ? fob[r] : -1;
    out[r] = f >= 0 ? dsel[f] : 1.0;
}

typedef double d2v __attribute__((ext_vector_type(2)));

// FP64 block-sliced-ELL SpMM: LPR lanes per block row (each takes KL = K / LPR
// right-hand sides), the 64 block rows of a slice are one workgroup. Block k of
// a row is five coalesced 16-byte loads (entries 0-7 and entry 8 with the
// column's bits; the LPR lanes of a row load the same addresses); the 3 KL
// values of the column block's x that a lane needs are gathered once and
// shared by the block's three rows. U blocks are loaded before their gathers;
// with GF a batch's x gathers are issued before the next batch's values (loads
// return in issue order, so the multiply waits for the gathers only, and the
// next values arrive meanwhile). With DOT the p.Ap partial sums of the
// workgroup go to part.
template <int K, int LPR, int U, bool DOT, bool GF = false>
__global__ void __launch_bounds__(64 * LPR) bsell64_spmm_kernel(
    int64_t nb, const int64_t *__restrict__ sptr, const int32_t *__restrict__ width,
    const double *__restrict__ bv, int64_t nsl, const double *__restrict__ P,
    double *__restrict__ AP, double *__restrict__ part, const int32_t *__restrict__ live,
    const int32_t *__restrict__ slice_seg) {
    constexpr int KL = K / LPR;
    // XCD-aware order: the grid is a multiple of 8 and workgroup b runs on XCD
    // b % 8, so XCD x takes the x-th contiguous eighth of the slices and the
    // neighbouring slices that share x gathers share its L2
    const int64_t g = gridDim.x, b = blockIdx.x;
    const int64_t sl = (b % 8) * (g / 8) + b / 8;
    // a slice of a finished system does nothing (its partials are not read)
    if (DOT && live[slice_seg ? slice_seg[sl] : 0] == 0) return;
    const int lane = threadIdx.x / LPR, h = threadIdx.x % LPR;
    const int64_t I = sl * 64 + lane;
    double dot[KL];
#pragma unroll
    for (int k = 0; k < KL; k++) dot[k] = 0.0;
    if (I < nb) {
        const d2v *__restrict__ q0 = reinterpret_cast<const d2v *>(bv);
        const d2v *__restrict__ q1 = reinterpret_cast<const d2v *>(bv + 2 * nsl);
        const d2v *__restrict__ q2 = reinterpret_cast<const d2v *>(bv + 4 * nsl);
        const d2v *__restrict__ q3 = reinterpret_cast<const d2v *>(bv + 6 * nsl);
        const d2v *__restrict__ q4 = reinterpret_cast<const d2v *>(bv + 8 * nsl);
        const int64_t base = sptr[sl] + lane;
        const int w = width[sl];
        double acc[3][KL];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int k = 0; k < KL; k++) acc[d][k] = 0.0;
        // the U blocks from block j0 (clamped to the row's last block), their
        // column blocks' x, and their products
        auto load_vals = [&](int j0, d2v (&a)[U][5]) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int64_t pos = base + 64 * (int64_t)(j0 + u < w ? j0 + u : j0);
                a[u][0] = __builtin_nontemporal_load(&q0[pos]);
                a[u][1] = __builtin_nontemporal_load(&q1[pos]);
                a[u][2] = __builtin_nontemporal_load(&q2[pos]);
                a[u][3] = __builtin_nontemporal_load(&q3[pos]);
                a[u][4] = __builtin_nontemporal_load(&q4[pos]);
            }
        };
        auto gather = [&](const d2v (&a)[U][5], double (&x)[U][3][KL]) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int64_t J = (int64_t)__double_as_longlong(a[u][4].y) & 0xffffffffll;
                const double *xb = P + J * 3 * K + h * KL;
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    if constexpr (KL % 2 == 0) {  // KL doubles at a 16-byte aligned offset
                        const d2v *xp = reinterpret_cast<const d2v *>(xb + e * K);
#pragma unroll
                        for (int t = 0; t < KL / 2; t++) {
                            const d2v v2 = xp[t];
                            x[u][e][2 * t] = v2.x;
                            x[u][e][2 * t + 1] = v2.y;
                        }
                    } else {
#pragma unroll
                        for (int t = 0; t < KL; t++) x[u][e][t] = xb[e * K + t];
                    }
                }
            }
        };
        auto fma = [&](int j0, const d2v (&a)[U][5], const double (&x)[U][3][KL]) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (j0 + u >= w) break;
                const double v[9] = {a[u][0].x, a[u][0].y, a[u][1].x, a[u][1].y, a[u][2].x,
                                     a[u][2].y, a[u][3].x, a[u][3].y, a[u][4].x};
#pragma unroll
                for (int d = 0; d < 3; d++)
#pragma unroll
                    for (int k = 0; k < KL; k++)
                        acc[d][k] += v[3 * d] * x[u][0][k] + v[3 * d + 1] * x[u][1][k] +
                                     v[3 * d + 2] * x[u][2][k];
            }
        };
        if constexpr (GF) {
            d2v a[U][5];
            if (w > 0) load_vals(0, a);
            for (int j0 = 0; j0 < w; j0 += U) {
                double x[U][3][KL];
                gather(a, x);
                __builtin_amdgcn_sched_barrier(0);
                d2v an[U][5];
                const bool more = j0 + U < w;
                if (more) load_vals(j0 + U, an);
                __builtin_amdgcn_sched_barrier(0);
                fma(j0, a, x);
                if (more) {
#pragma unroll
                    for (int u = 0; u < U; u++)
#pragma unroll
                        for (int q = 0; q < 5; q++) a[u][q] = an[u][q];
                }
            }
        } else {
            for (int j0 = 0; j0 < w; j0 += U) {
                d2v a[U][5];
                load_vals(j0, a);
                double x[U][3][KL];
                gather(a, x);
                fma(j0, a, x);
            }
        }
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int k = 0; k < KL; k++) {
                const int64_t i = (3 * I + d) * K + h * KL + k;
                AP[i] = acc[d][k];
                if (DOT) dot[k] += P[i] * acc[d][k];
            }
    }
    if (!DOT) return;
    // per right-hand side: lanes of the same h summed within the wavefront,
    // then the LPR wavefronts of the workgroup through LDS
#pragma unroll
    for (int k = 0; k < KL; k++)
#pragma unroll
        for (int off = 32; off >= LPR; off >>= 1) dot[k] += __shfl_down(dot[k], off, 64);
    __shared__ double sh[LPR][K];
    const int wl = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (wl < LPR)
#pragma unroll
        for (int k = 0; k < KL; k++) sh[wv][wl * KL + k] = dot[k];
    __syncthreads();
    if (threadIdx.x < K) {
        double t = 0.0;
#pragma unroll
        for (int w2 = 0; w2 < LPR; w2++) t += sh[w2][threadIdx.x];
        part[(size_t)sl * K + threadIdx.x] = t;  // by slice: a system's partials are contiguous
    }
}

struct BlockPcg {
    axg_mesh *m = nullptr;
    int64_t n = 0, nb = 0, rows = 0, npad = 0, max_iter = 0;
    int nwg = 0, nsl_wg = 0;
    int U[4] = {2, 2, 2, 2};  // blocks in flight per lane, tuned per K = 1, 2, 4, 8
    int L[4] = {1, 1, 1, 2};  // lanes per block row, tuned per K = 2, 4, 8
    DevBuf<int64_t> wg_row, seg_end;
    DevBuf<int32_t> wg_seg, segwg, segwg_mv, seg_live, live_total;
    DevBuf<double> dsel, Bv, X, R, P, AP, Z, part, tolr;
    DevBuf<SegScalars> sc;
    BatchDev B, Bmv;  // Bmv: the SpMM's partial sums (one per slice)
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};  // K = 1, 2, 4, 8
    bool blk3 = false;                 // block-Jacobi (3x3 node inverses) instead of Jacobi
    const double *graph_bv = nullptr;  // operator values the graphs were captured with
    const double *graph_binv = nullptr;
    void drop_graphs() {
        for (auto &g : graph)
            if (g) (void)hipGraphExecDestroy(g), g = nullptr;
    }
    ~BlockPcg() { drop_graphs(); }
};

static constexpr int kGraphIters = 32;

// The form code of the FP64 block SpMM: blocks per batch (1, 2, 4), plus
// kBsellGF for the gathers-first form; the tuners try every one.
constexpr int kBsellGF = 10;
static const int kBsellForms[] = {1, 2, 4, 1 + kBsellGF, 2 + kBsellGF, 4 + kBsellGF};

// One launch of the FP64 block SpMM over nsl_wg slices (a multiple of 8).
// With events, the launch itself stamps e0 / e1 (the SpMM's device time).
struct Bsell64Args {
    int64_t nb;
    const int64_t *sptr;
    const int32_t *width;
    const double *bv;
    int64_t nsl;
    int nsl_wg;
    double *ap, *part;
    const int32_t *live, *slice_seg;
};

template <int K, int LPR>
static void launch_bsell64(const Bsell64Args &a, hipStream_t s, bool dot, int u, const double *x,
                           hipEvent_t e0, hipEvent_t e1) {
    const dim3 g(a.nsl_wg), b(64 * LPR);
#define AXG_BSELL64(UU, DD, GG)                                                                     \
    if (e0)                                                                                             \
        hipExtLaunchKernelGGL((bsell64_spmm_kernel<K, LPR, UU, DD, GG>), g, b, 0, s, e0, e1, 0, a.nb,  \
                              a.sptr, a.width, a.bv, a.nsl, x, a.ap, a.part, a.live, a.slice_seg);     \
    else                                                                                                \
        hipLaunchKernelGGL((bsell64_spmm_kernel<K, LPR, UU, DD, GG>), g, b, 0, s, a.nb, a.sptr,        \
                           a.width, a.bv, a.nsl, x, a.ap, a.part, a.live, a.slice_seg)
    // u: blocks per batch (1, 2, 4), plus kBsellGF for the gathers-first form
    const bool gf = u >= kBsellGF;
    const int uu = gf ? u - kBsellGF : u;
    if (dot) {
        if (uu == 1) {
            if (gf) AXG_BSELL64(1, true, true); else AXG_BSELL64(1, true, false);
        } else if (uu == 2) {
            if (gf) AXG_BSELL64(2, true, true); else AXG_BSELL64(2, true, false);
        } else {
            if (gf) AXG_BSELL64(4, true, true); else AXG_BSELL64(4, true, false);
        }
    } else {
        if (uu == 1) {
            if (gf) AXG_BSELL64(1, false, true); else AXG_BSELL64(1, false, false);
        } else if (uu == 2) {
            if (gf) AXG_BSELL64(2, false, true); else AXG_BSELL64(2, false, false);
        } else {
            if (gf) AXG_BSELL64(4, false, true); else AXG_BSELL64(4, false, false);
        }
    }
#undef AXG_BSELL64
}

// Lanes per block row: one for a single right-hand side; 1 or 2 for 2 and 4
// (two lanes halve each lane's x gathers and registers, the block's values are
// loaded by both from the same lines); 2 or 4 for 8.
template <int K>
static void bsell64_spmm(const Bsell64Args &a, hipStream_t s, bool dot, int u, int l,
                         const double *x, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    if constexpr (K == 1) {
        launch_bsell64<1, 1>(a, s, dot, u, x, e0, e1);
    } else if constexpr (K == 8) {
        if (l == 4) launch_bsell64<8, 4>(a, s, dot, u, x, e0, e1);
        else launch_bsell64<8, 2>(a, s, dot, u, x, e0, e1);
    } else {
        if (l == 2) launch_bsell64<K, 2>(a, s, dot, u, x, e0, e1);
        else launch_bsell64<K, 1>(a, s, dot, u, x, e0, e1);
    }
}

// Candidate lanes per block row for K right-hand sides.
static std::vector<int> bsell64_lanes(int K) {
    if (K == 1) return {1};
    if (K == 8) return {2, 4};
    return {1, 2};
}

template <int K>
static void block_spmv(BlockPcg &bp, hipStream_t s, bool dot, const double *x,
                       hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const int slot = K == 1 ? 0 : K == 2 ? 1 : K == 4 ? 2 : 3;
    axg_mesh *m = bp.m;
    const Bsell64Args a{bp.nb,     m->bs_sptr.p, m->bs_width.p, m->bval64.p,   m->bs_nslots,
                        bp.nsl_wg, bp.AP.p,      bp.part.p,     bp.seg_live.p, nullptr};
    bsell64_spmm<K>(a, s, dot, bp.U[slot], bp.L[slot], x, e0, e1);
}

template <int K>
static void block_iter(BlockPcg &bp, hipStream_t s, hipEvent_t e0 = nullptr,
                       hipEvent_t e1 = nullptr) {
    block_spmv<K>(bp, s, true, bp.P.p, e0, e1);
    hipLaunchKernelGGL((batch_fin_pap_kernel<K>), dim3(1), dim3(FIN_TB), 0, s, bp.Bmv, bp.part.p,
                       bp.sc.p, bp.live_total.p);
    if (bp.blk3)
        hipLaunchKernelGGL((batch_update_kernel<K, true>), dim3(bp.nwg), dim3(TB), 0, s, bp.B, bp.X.p,
                           bp.R.p, bp.P.p, bp.AP.p, bp.Z.p, bp.part.p, bp.sc.p);
    else
        hipLaunchKernelGGL((batch_update_kernel<K, false>), dim3(bp.nwg), dim3(TB), 0, s, bp.B, bp.X.p,
                           bp.R.p, bp.P.p, bp.AP.p, bp.Z.p, bp.part.p, bp.sc.p);
    hipLaunchKernelGGL((batch_fin_update_kernel<K>), dim3(1), dim3(FIN_TB), 0, s, bp.B, bp.part.p,
                       bp.sc.p, 0, 0.0, bp.tolr.p, bp.max_iter, bp.live_total.p);
    hipLaunchKernelGGL((batch_update_p_kernel<K>), dim3(bp.nwg), dim3(TB), 0, s, bp.B, bp.P.p, bp.Z.p,
                       bp.sc.p, 0);
}

// Lanes per block row and blocks in flight per lane, timed on this operator.
template <int K>
static void tune_u(BlockPcg &bp, int slot) {
    axg_ctx *ctx = bp.m->ctx;
    hipStream_t s = ctx->stream;
    const std::vector<int> lanes = bsell64_lanes(K);
    float best = 1e30f;
    int best_u = 2, best_l = lanes[0];
    for (int l : lanes)
        for (int cand : kBsellForms) {
            bp.U[slot] = cand;
            bp.L[slot] = l;
            block_spmv<K>(bp, s, false, bp.Bv.p);
            AXG_HIP(hipEventRecord(ctx->ev0, s));
            for (int r = 0; r < 3; r++) block_spmv<K>(bp, s, false, bp.Bv.p);
            AXG_HIP(hipEventRecord(ctx->ev1, s));
            AXG_HIP(hipEventSynchronize(ctx->ev1));
            float ms = 0.f;
            AXG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            if (ms < best) best = ms, best_u = cand, best_l = l;
        }
    bp.U[slot] = best_u;
    bp.L[slot] = best_l;
}

BlockPcg *block_pcg_create(axg_mesh *m, int64_t max_iter, bool blk3) {
    auto bp = std::make_unique<BlockPcg>();
    axg_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    ensure_bsell64(m);
    if (blk3) ensure_block_inverse64(m);
    bp->m = m;
    bp->blk3 = blk3;
    const int64_t n = m->n_free, nb = m->nb, rows = 3 * nb;
    const int64_t npad = std::max<int64_t>(((rows + TB - 1) / TB) * TB, TB);
    AXG_REQUIRE(npad * 8 < INT32_MAX, AXG_E_INVALID, "block PCG exceeds 2^31 vector entries");
    // block-Jacobi needs whole nodes in a vector workgroup: 255 rows
    const int rows_wg = blk3 ? 255 : TB;
    bp->n = n;
    bp->nb = nb;
    bp->rows = rows;
    bp->npad = npad;
    bp->max_iter = max_iter > 0 ? max_iter : std::max<int64_t>(4 * n, 200);
    bp->nwg = (int)std::max<int64_t>((rows + rows_wg - 1) / rows_wg, 1);
    bp->nsl_wg = (int)(((std::max<int64_t>((nb + 63) / 64, 1) + 7) / 8) * 8);  // multiple of 8
    std::vector<int64_t> h_wg_row(bp->nwg);
    std::vector<int32_t> h_wg_seg(bp->nwg, 0);
    for (int w = 0; w < bp->nwg; w++) h_wg_row[w] = (int64_t)w * rows_wg;
    bp->wg_row.alloc(bp->nwg);
    bp->wg_seg.alloc(bp->nwg);
    bp->seg_end.alloc(2);
    bp->segwg.alloc(2);
    bp->segwg_mv.alloc(2);
    bp->seg_live.alloc(1);
    bp->live_total.alloc(1);
    bp->dsel.alloc(npad);
    const int64_t h_end[2] = {0, rows};
    const int32_t h_segwg[2] = {0, bp->nwg}, h_segwg_mv[2] = {0, bp->nsl_wg};
    AXG_HIP(hipMemcpyAsync(bp->wg_row.p, h_wg_row.data(), 8 * bp->nwg, hipMemcpyHostToDevice, s));
    AXG_HIP(hipMemcpyAsync(bp->wg_seg.p, h_wg_seg.data(), 4 * bp->nwg, hipMemcpyHostToDevice, s));
    AXG_HIP(hipMemcpyAsync(bp->seg_end.p, h_end, 16, hipMemcpyHostToDevice, s));
    AXG_HIP(hipMemcpyAsync(bp->segwg.p, h_segwg, 8, hipMemcpyHostToDevice, s));
    AXG_HIP(hipMemcpyAsync(bp->segwg_mv.p, h_segwg_mv, 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(block_dsel_kernel, dim3((unsigned)((npad + 255) / 256)), dim3(256), 0, s,
                       npad, rows, m->free_of_blk.p, m->dsel.p, bp->dsel.p);
    const size_t nv = (size_t)npad * 8;
    bp->Bv.alloc(nv);
    bp->X.alloc(nv);
    bp->R.alloc(nv);
    bp->P.alloc(nv);
    bp->AP.alloc(nv);
    bp->Z.alloc(nv);
    bp->part.alloc((size_t)std::max(bp->nwg * 16, bp->nsl_wg * 8) + 16);
    bp->tolr.alloc(8);
    bp->sc.alloc(8);
    BatchDev &B = bp->B;
    B.binv = blk3 ? m->binv64.p : nullptr;
    B.rows_wg = rows_wg;
    B.dsel = bp->dsel.p;
    B.wg_seg = bp->wg_seg.p;
    B.wg_row = bp->wg_row.p;
    B.seg_row = bp->seg_end.p;
    B.seg_wg = bp->segwg.p;
    B.seg_live = bp->seg_live.p;
    bp->Bmv = B;
    bp->Bmv.seg_wg = bp->segwg_mv.p;
    hipLaunchKernelGGL(fill_kernel, dim3(256), dim3(256), 0, s, (int64_t)nv, bp->Bv.p, 0.0);
    // blocks in flight (U) tuned on this operator for each width
    tune_u<1>(*bp, 0);
    tune_u<2>(*bp, 1);
    tune_u<4>(*bp, 2);
    tune_u<8>(*bp, 3);
    AXG_HIP(hipGetLastError());
    return bp.release();
}

void block_pcg_destroy(BlockPcg *bp) { delete bp; }

// With timing, the iterations are launched one by one (no graph) and every
// SpMM stamps a pair of events; the launches that did work are counted from
// the iteration counters (the SpMM returns at once when no column is live).
template <int K>
static void block_group(BlockPcg &bp, const double *d_b, int kk, const double *tol_rel,
                        double *d_x, PcgResult &res, bool timing) {
    axg_ctx *ctx = bp.m->ctx;
    hipStream_t s = ctx->stream;
    const int slot = K == 1 ? 0 : K == 2 ? 1 : K == 4 ? 2 : 3;
    if (!timing && !bp.graph[slot]) {
        hipGraph_t g;
        AXG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < kGraphIters; i++) block_iter<K>(bp, s);
        AXG_HIP(hipStreamEndCapture(s, &g));
        AXG_HIP(hipGraphInstantiate(&bp.graph[slot], g, nullptr, nullptr, 0));
        AXG_HIP(hipGraphDestroy(g));
        bp.graph_bv = bp.m->bval64.p;
        bp.graph_binv = bp.B.binv;
    }
    if (timing)
        while (ctx->timing_events.size() < 2 * (size_t)kGraphIters) {
            hipEvent_t e;
            AXG_HIP(hipEventCreate(&e));
            ctx->timing_events.push_back(e);
        }
    double h_tol[8];
    for (int j = 0; j < K; j++) h_tol[j] = j < kk ? tol_rel[j] : 1.0;
    std::vector<SegScalars> hs(K);
    std::memset(hs.data(), 0, sizeof(SegScalars) * K);
    const int32_t live0 = K;
    AXG_HIP(hipMemcpyAsync(bp.tolr.p, h_tol, 8 * K, hipMemcpyHostToDevice, s));
    AXG_HIP(hipMemcpyAsync(bp.sc.p, hs.data(), sizeof(SegScalars) * K, hipMemcpyHostToDevice, s));
    AXG_HIP(hipMemcpyAsync(bp.seg_live.p, &live0, 4, hipMemcpyHostToDevice, s));
    AXG_HIP(hipMemcpyAsync(bp.live_total.p, &live0, 4, hipMemcpyHostToDevice, s));
    const int64_t tot = bp.npad * K;
    hipLaunchKernelGGL((pack_blocks_kernel<K>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       bp.npad, bp.rows, kk, bp.m->free_of_blk.p, bp.n, d_b, bp.Bv.p);
    AXG_HIP(hipEventRecord(ctx->ev0, s));
    if (bp.blk3)
        hipLaunchKernelGGL((batch_init_kernel<K, true>), dim3(bp.nwg), dim3(TB), 0, s, bp.B, bp.Bv.p,
                           bp.X.p, bp.R.p, bp.Z.p, bp.part.p);
    else
        hipLaunchKernelGGL((batch_init_kernel<K, false>), dim3(bp.nwg), dim3(TB), 0, s, bp.B, bp.Bv.p,
                           bp.X.p, bp.R.p, bp.Z.p, bp.part.p);
    hipLaunchKernelGGL((batch_fin_update_kernel<K>), dim3(1), dim3(FIN_TB), 0, s, bp.B, bp.part.p,
                       bp.sc.p, 1, 0.0, bp.tolr.p, bp.max_iter, bp.live_total.p);
    hipLaunchKernelGGL((batch_update_p_kernel<K>), dim3(bp.nwg), dim3(TB), 0, s, bp.B, bp.P.p,
                       bp.Z.p, bp.sc.p, 1);
    AXG_HIP(hipGetLastError());
    int32_t live = live0;
    int64_t launched = 0, done_iters = 0;
    for (;;) {
        AXG_HIP(hipMemcpyAsync(&live, bp.live_total.p, 4, hipMemcpyDeviceToHost, s));
        AXG_HIP(hipStreamSynchronize(s));
        if (live <= 0 || launched >= bp.max_iter) break;
        if (timing) {
            for (int i = 0; i < kGraphIters; i++)
                block_iter<K>(bp, s, ctx->timing_events[2 * i], ctx->timing_events[2 * i + 1]);
            AXG_HIP(hipMemcpyAsync(hs.data(), bp.sc.p, sizeof(SegScalars) * K,
                                   hipMemcpyDeviceToHost, s));
            AXG_HIP(hipStreamSynchronize(s));
            int64_t most = 0;
            for (int j = 0; j < kk; j++) most = std::max<int64_t>(most, hs[j].iters);
            for (int64_t i = 0; i < most - done_iters && i < kGraphIters; i++) {
                float ms = 0.f;
                AXG_HIP(hipEventElapsedTime(&ms, ctx->timing_events[2 * i],
                                            ctx->timing_events[2 * i + 1]));
                res.spmv_ms += ms;
                res.spmv_launches++;
            }
            done_iters = most;
        } else {
            AXG_HIP(hipGraphLaunch(bp.graph[slot], s));
        }
        launched += kGraphIters;
    }
    AXG_HIP(hipEventRecord(ctx->ev1, s));
    if (bp.n)
        hipLaunchKernelGGL((unpack_blocks_kernel<K>), dim3((unsigned)((bp.n * kk + 255) / 256)),
                           dim3(256), 0, s, bp.n, kk, bp.m->blk_of_free.p, bp.X.p, d_x);
    AXG_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    AXG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    AXG_HIP(hipMemcpy(hs.data(), bp.sc.p, sizeof(SegScalars) * K, hipMemcpyDeviceToHost));
    res.solve_ms += ms;
    for (int j = 0; j < kk; j++) {
        const SegScalars &r = hs[j];
        res.iters += r.iters;
        const int st = r.status != AXG_OK ? r.status : r.done ? AXG_OK : AXG_E_NOT_CONVERGED;
        if (st != AXG_OK && res.status == AXG_OK) {
            res.status = st;
            res.fail_iters = r.iters;
            res.fail_residual = r.res;
        }
        if (r.bnorm > 0 && r.res / r.bnorm >= res.residual / std::max(res.bnorm, 1e-300))
            res.residual = r.res, res.bnorm = r.bnorm;
    }
}

// ------------------------------------------------------- many meshes
// Every mesh is a segment of one node-block system: its block rows (3 rows per
// node with a free DOF) padded to a multiple of 768 rows, so that segments
// start on whole 64-block-row slices of the SpMM and on whole 256-row
// workgroups of the vector kernels. The meshes' FP64 block planes are copied
// side by side with their columns shifted to the segment's first block row;
// a slice of a mesh whose right-hand sides have all stopped does nothing.
constexpr int64_t kSegRows = 768;
constexpr int kMaxGroups = 16;      // size groups of one batched solve
constexpr int kMeshesPerGroup = 6;  // group size when the caller leaves the choice

// column bits of plane 4 (the low 32 bits of entry 8's neighbour) + off
__global__ void rebase_block_cols_kernel(int64_t nslots, long long *__restrict__ plane4,
                                         int64_t off) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nslots;
         p += (int64_t)gridDim.x * blockDim.x) {
        const long long bits = plane4[2 * p + 1];
        plane4[2 * p + 1] = (long long)((bits & 0xffffffffll) + off);
    }
}

// One group of meshes: the segments of its meshes, its vectors and scalars,
// and the captured iteration sequence.
template <int K>
struct BatchGroup {
    hipStream_t s = nullptr;
    std::vector<int> mesh;  // indices into the caller's mesh list
    std::vector<int64_t> row0;
    std::vector<std::vector<int32_t>> bof;
    int64_t nrow = 0, max_iter = 0, launched = 0;
    int nseg = 0, nwg = 0;
    DevBuf<int64_t> sptr, wg_row, seg_end;
    DevBuf<int32_t> width, slice_seg, wg_seg, segwg, segsl, seg_live, live_total;
    DevBuf<double> bv, dsel, Bv, X, R, P, AP, Z, part;
    DevBuf<SegScalars> sc;
    BatchDev B, Bmv;
    Bsell64Args op{};
    int U = 2, L = 1;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t ev = nullptr;  // the live count of the last replay has arrived
    bool running = false;

    ~BatchGroup() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (ev) (void)hipEventDestroy(ev);
    }

    void one_iter() {
        bsell64_spmm<K>(op, s, true, U, L, P.p);
        hipLaunchKernelGGL((batch_fin_pap_kernel<K>), dim3(nseg), dim3(FIN_TB), 0, s, Bmv, part.p,
                           sc.p, live_total.p);
        hipLaunchKernelGGL((batch_update_kernel<K, false>), dim3(nwg), dim3(TB), 0, s, B, X.p,
                           R.p, P.p, AP.p, Z.p, part.p, sc.p);
        hipLaunchKernelGGL((batch_fin_update_kernel<K>), dim3(nseg), dim3(FIN_TB), 0, s, B, part.p,
                           sc.p, 0, tol, nullptr, max_iter, live_total.p);
        hipLaunchKernelGGL((batch_update_p_kernel<K>), dim3(nwg), dim3(TB), 0, s, B, P.p, Z.p,
                           sc.p, 0);
    }
    double tol = 1e-8;

    // segments, operator planes, right-hand sides, tuning and the captured
    // iteration sequence (nothing iterates yet)
    void setup(axg_ctx *ctx, axg_mesh *const *meshes, const double *const *rhs,
               const int32_t *nrhs, const axg_solve_opts &o) {
        nseg = (int)mesh.size();
        tol = o.tol;
        row0.assign(nseg + 1, 0);
        std::vector<int64_t> slot0(nseg + 1, 0), slice0(nseg + 1, 0);
        std::vector<int32_t> seg_wg(nseg + 1, 0);
        int64_t max_n = 0;
        for (int g = 0; g < nseg; g++) {
            axg_mesh *me = meshes[mesh[g]];
            if (me->n_free > 0) {  // a mesh without free DOFs is an empty segment
                ensure_block_maps(me);
                ensure_bsell64(me);
            }
            const int64_t rows = 3 * me->nb;
            const int64_t rpad =
                std::max<int64_t>(kSegRows, (rows + kSegRows - 1) / kSegRows * kSegRows);
            max_n = std::max(max_n, me->n_free);
            row0[g + 1] = row0[g] + rpad;
            slice0[g + 1] = slice0[g] + rpad / 192;  // 64 block rows per slice
            slot0[g + 1] = slot0[g] + me->bs_nslots;
            seg_wg[g + 1] = seg_wg[g] + (int32_t)(rpad / TB);
        }
        nrow = row0[nseg];
        const int64_t nslot = slot0[nseg];
        const int64_t nslice = slice0[nseg], nsl_wg = (nslice + 7) / 8 * 8;
        nwg = seg_wg[nseg];
        AXG_REQUIRE(nrow * K < INT32_MAX && nslot < INT32_MAX, AXG_E_INVALID,
                    "batch exceeds 2^31 vector entries or slots");
        sptr.alloc(nsl_wg + 1), wg_row.alloc(nwg), width.alloc(nsl_wg), slice_seg.alloc(nsl_wg);
        wg_seg.alloc(nwg), segwg.alloc(nseg + 1), segsl.alloc(nseg + 1), seg_live.alloc(nseg);
        bv.alloc(std::max<int64_t>(10 * nslot, 2)), dsel.alloc(nrow);
        std::vector<int64_t> h_sptr(nsl_wg + 1, nslot), h_wg_row(nwg);
        std::vector<int32_t> h_width(nsl_wg, 0), h_slice_seg(nsl_wg, nseg - 1), h_wg_seg(nwg),
            h_live(nseg, K), h_segsl(nseg + 1);
        bof.assign(nseg, {});
        for (int g = 0; g < nseg; g++) {
            axg_mesh *me = meshes[mesh[g]];
            const int64_t ns = (me->nb + 63) / 64, nsl = me->bs_nslots;
            std::vector<int64_t> sp(ns + 1, 0);
            std::vector<int32_t> wd(ns);
            if (ns) {
                AXG_HIP(hipMemcpy(sp.data(), me->bs_sptr.p, 8 * (ns + 1), hipMemcpyDeviceToHost));
                AXG_HIP(hipMemcpy(wd.data(), me->bs_width.p, 4 * ns, hipMemcpyDeviceToHost));
            }
            bof[g].resize(me->n_free);
            if (me->n_free)
                AXG_HIP(hipMemcpy(bof[g].data(), me->blk_of_free.p, 4 * me->n_free,
                                  hipMemcpyDeviceToHost));
            for (int64_t k = 0; k < slice0[g + 1] - slice0[g]; k++) {
                const int64_t q = slice0[g] + k;
                h_width[q] = k < ns ? wd[k] : 0;
                h_sptr[q] = slot0[g] + (k < ns ? sp[k] : sp[ns]);
                h_slice_seg[q] = g;
            }
            h_segsl[g] = (int32_t)slice0[g];
            for (int32_t wgi = seg_wg[g]; wgi < seg_wg[g + 1]; wgi++) {
                h_wg_seg[wgi] = g;
                h_wg_row[wgi] = row0[g] + (int64_t)(wgi - seg_wg[g]) * TB;
            }
            // the five value planes side by side, columns shifted to this segment
            if (nsl) {
                for (int q = 0; q < 5; q++)
                    AXG_HIP(hipMemcpyAsync(bv.p + 2 * q * nslot + 2 * slot0[g],
                                           me->bval64.p + 2 * q * nsl, 16 * nsl,
                                           hipMemcpyDeviceToDevice, s));
                hipLaunchKernelGGL(rebase_block_cols_kernel, dim3(1024), dim3(256), 0, s, nsl,
                                   reinterpret_cast<long long *>(bv.p + 8 * nslot + 2 * slot0[g]),
                                   row0[g] / 3);
            }
            const int64_t rpad = row0[g + 1] - row0[g];
            hipLaunchKernelGGL(block_dsel_kernel, dim3((unsigned)((rpad + 255) / 256)), dim3(256),
                               0, s, rpad, 3 * me->nb, me->free_of_blk.p, me->dsel.p,
                               dsel.p + row0[g]);
        }
        h_segsl[nseg] = (int32_t)nslice;
        seg_end.alloc(nseg + 1);
        std::vector<int64_t> h_end(nseg + 1, 0);
        for (int g = 0; g < nseg; g++) h_end[g + 1] = row0[g] + 3 * meshes[mesh[g]]->nb;
        AXG_HIP(hipMemcpyAsync(sptr.p, h_sptr.data(), 8 * (nsl_wg + 1), hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(width.p, h_width.data(), 4 * nsl_wg, hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(slice_seg.p, h_slice_seg.data(), 4 * nsl_wg, hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(wg_seg.p, h_wg_seg.data(), 4 * nwg, hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(wg_row.p, h_wg_row.data(), 8 * nwg, hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(seg_end.p, h_end.data(), 8 * (nseg + 1), hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(segwg.p, seg_wg.data(), 4 * (nseg + 1), hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(segsl.p, h_segsl.data(), 4 * (nseg + 1), hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(seg_live.p, h_live.data(), 4 * nseg, hipMemcpyHostToDevice, s));
        B.binv = nullptr;
        B.rows_wg = TB;
        B.dsel = dsel.p;
        B.wg_seg = wg_seg.p;
        B.wg_row = wg_row.p;
        B.seg_row = seg_end.p;
        B.seg_wg = segwg.p;
        B.seg_live = seg_live.p;
        Bmv = B;  // the SpMM's partials: one per slice
        Bmv.seg_wg = segsl.p;
        // right-hand sides in block rows: row-major [row][K], zero on prescribed,
        // padding rows and missing columns
        std::vector<double> hb((size_t)nrow * K, 0.0);
        for (int g = 0; g < nseg; g++) {
            const int m = mesh[g];
            const int64_t n = meshes[m]->n_free;
            for (int j = 0; j < std::min(K, (int)nrhs[m]); j++)
                for (int64_t i = 0; i < n; i++) hb[(row0[g] + bof[g][i]) * K + j] = rhs[m][j * n + i];
        }
        Bv.alloc(nrow * K), X.alloc(nrow * K), R.alloc(nrow * K), P.alloc(nrow * K);
        AP.alloc(nrow * K), Z.alloc(nrow * K);
        part.alloc((size_t)std::max<int64_t>(nwg * 2 * K, nsl_wg * K) + 8);
        sc.alloc(nseg * K);
        live_total.alloc(1);
        // padding rows are read by the SpMM's p.Ap (times zero rows) and never
        // written by the vector kernels: they must hold zeros, not stale bits
        for (double *v : {X.p, R.p, P.p, AP.p, Z.p})
            hipLaunchKernelGGL(fill_kernel, dim3(1024), dim3(256), 0, s, nrow * K, v, 0.0);
        AXG_HIP(hipMemcpyAsync(Bv.p, hb.data(), 8 * hb.size(), hipMemcpyHostToDevice, s));
        std::vector<SegScalars> hs(nseg * K);
        std::memset(hs.data(), 0, sizeof(SegScalars) * hs.size());
        AXG_HIP(hipMemcpyAsync(sc.p, hs.data(), sizeof(SegScalars) * hs.size(),
                               hipMemcpyHostToDevice, s));
        const int32_t live0 = nseg * K;
        AXG_HIP(hipMemcpyAsync(live_total.p, &live0, 4, hipMemcpyHostToDevice, s));
        max_iter = o.max_iter > 0 ? o.max_iter : std::max<int64_t>(4 * max_n, 200);
        op = Bsell64Args{nrow / 3, sptr.p, width.p,     bv.p,          nslot,
                         (int)nsl_wg, AP.p, part.p, seg_live.p, slice_seg.p};
        // lanes per block row and blocks in flight, timed on this operator
        L = bsell64_lanes(K)[0];
        {
            float best = 1e30f;
            for (int l : bsell64_lanes(K))
                for (int cand : kBsellForms) {
                    bsell64_spmm<K>(op, s, false, cand, l, Bv.p);
                    AXG_HIP(hipEventRecord(ctx->ev0, s));
                    for (int r = 0; r < 3; r++) bsell64_spmm<K>(op, s, false, cand, l, Bv.p);
                    AXG_HIP(hipEventRecord(ctx->ev1, s));
                    AXG_HIP(hipEventSynchronize(ctx->ev1));
                    float ms = 0.f;
                    AXG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
                    if (ms < best) best = ms, U = cand, L = l;
                }
        }
        hipGraph_t graph;
        AXG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < kReplay; i++) one_iter();
        AXG_HIP(hipStreamEndCapture(s, &graph));
        AXG_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        AXG_HIP(hipGraphDestroy(graph));
        AXG_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        AXG_HIP(hipStreamSynchronize(s));
    }

    // x = 0, r = b, z = M^-1 r, first direction; then the live count
    void start(int32_t *h_live) {
        hipLaunchKernelGGL((batch_init_kernel<K, false>), dim3(nwg), dim3(TB), 0, s, B, Bv.p, X.p,
                           R.p, Z.p, part.p);
        hipLaunchKernelGGL((batch_fin_update_kernel<K>), dim3(nseg), dim3(FIN_TB), 0, s, B, part.p,
                           sc.p, 1, tol, nullptr, max_iter, live_total.p);
        hipLaunchKernelGGL((batch_update_p_kernel<K>), dim3(nwg), dim3(TB), 0, s, B, P.p, Z.p,
                           sc.p, 1);
        AXG_HIP(hipGetLastError());
        post(h_live);
        running = true;
    }
    void post(int32_t *h_live) {
        AXG_HIP(hipMemcpyAsync(h_live, live_total.p, 4, hipMemcpyDeviceToHost, s));
        AXG_HIP(hipEventRecord(ev, s));
    }
    // after the last live count has arrived: another replay while any of the
    // group's right-hand sides iterates; false when the group is finished
    bool step(int32_t *h_live) {
        if (!running) return false;
        AXG_HIP(hipEventSynchronize(ev));
        if (*h_live <= 0 || launched >= max_iter) {
            running = false;
            return false;
        }
        AXG_HIP(hipGraphLaunch(gexec, s));
        launched += kReplay;
        post(h_live);
        return true;
    }
    static constexpr int kReplay = 32;  // iterations per graph replay
};

// Many meshes in one call, in G groups of meshes by size, each a
// block-diagonal system of segments with its own captured iteration sequence.
// The groups run one after another, smallest meshes first: a small mesh
// iterates at the pace of its own group's launches instead of the largest
// meshes' (launches of concurrent streams wait behind the large SpMMs), and
// the largest group, which sets the total time, starts after the small groups'
// few milliseconds. G = 1 advances every mesh in one sequence.
template <int K>
static void run_batch(axg_ctx *ctx, int nmesh, axg_mesh *const *meshes, const double *const *rhs,
                      const int32_t *nrhs, const axg_solve_opts &o, double *const *x,
                      int64_t *iters_out, double *ms_out, axg_solve_info *info) {
    int G = o.batch_groups > 0 ? o.batch_groups : std::max(1, nmesh / kMeshesPerGroup);
    G = std::max(1, std::min({G, kMaxGroups, nmesh}));
    std::vector<int> order(nmesh);
    for (int m = 0; m < nmesh; m++) order[m] = m;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return meshes[a]->n_free < meshes[b]->n_free; });
    std::vector<std::unique_ptr<BatchGroup<K>>> grp;
    for (int g = 0; g < G; g++) {
        auto bg = std::make_unique<BatchGroup<K>>();
        bg->s = ctx->stream;
        // meshes in ascending size, split into G runs of (nearly) equal count
        for (int r = (int)((int64_t)nmesh * g / G); r < (int)((int64_t)nmesh * (g + 1) / G); r++)
            bg->mesh.push_back(order[r]);
        bg->setup(ctx, meshes, rhs, nrhs, o);
        grp.push_back(std::move(bg));
    }
    hipStream_t s = ctx->stream;
    ctx->pinned.ensure(sizeof(int32_t));
    int32_t *h_live = (int32_t *)ctx->pinned.p;
    *h_live = 0;
    DevBuf<int64_t> t0(1);
    AXG_HIP(hipEventRecord(ctx->ev0, s));
    hipLaunchKernelGGL(stamp_kernel, dim3(1), dim3(1), 0, s, t0.p);  // start of the solves
    for (int g = 0; g < G; g++) {
        grp[g]->start(h_live);
        while (grp[g]->step(h_live)) {
        }
    }
    AXG_HIP(hipEventRecord(ctx->ev1, s));
    AXG_HIP(hipEventSynchronize(ctx->ev1));
    float total_ms = 0.f;
    AXG_HIP(hipEventElapsedTime(&total_ms, ctx->ev0, ctx->ev1));
    int64_t h_t0 = 0;
    AXG_HIP(hipMemcpy(&h_t0, t0.p, 8, hipMemcpyDeviceToHost));
    // per-solve results in the caller's order
    std::vector<int64_t> qbase(nmesh + 1, 0);
    for (int m = 0; m < nmesh; m++) qbase[m + 1] = qbase[m] + nrhs[m];
    int status = AXG_OK;
    int64_t total_iters = 0, launches = 0;
    double worst = 0.0, worst_b = 1.0, spmv_ms = 0.0;
    for (int g = 0; g < G; g++) {
        BatchGroup<K> &bg = *grp[g];
        std::vector<SegScalars> hs(bg.nseg * K);
        AXG_HIP(hipMemcpy(hs.data(), bg.sc.p, sizeof(SegScalars) * hs.size(), hipMemcpyDeviceToHost));
        std::vector<double> hx((size_t)bg.nrow * K);
        AXG_HIP(hipMemcpy(hx.data(), bg.X.p, 8 * hx.size(), hipMemcpyDeviceToHost));
        int64_t most_iters = 0;
        for (int sg = 0; sg < bg.nseg; sg++) {
            const int m = bg.mesh[sg];
            const int64_t n = meshes[m]->n_free;
            for (int j = 0; j < nrhs[m]; j++) {
                const int64_t q = qbase[m] + j;
                if (j >= K) {  // more right-hand sides than this call handles
                    status = AXG_E_INVALID;
                    continue;
                }
                const SegScalars &r = hs[sg * K + j];
                for (int64_t i = 0; i < n; i++) x[m][j * n + i] = hx[(bg.row0[sg] + bg.bof[sg][i]) * K + j];
                if (iters_out) iters_out[q] = r.iters;
                // s_memrealtime runs at 100 MHz
                if (ms_out) ms_out[q] = r.done ? (double)(r.t_done - h_t0) / 1e5 : (double)total_ms;
                total_iters += r.iters;
                most_iters = std::max<int64_t>(most_iters, r.iters);
                if (!r.done) status = AXG_E_NOT_CONVERGED;
                else if (r.status != AXG_OK && status == AXG_OK) status = r.status;
                if (r.bnorm > 0 && r.res / r.bnorm > worst / worst_b) worst = r.res, worst_b = r.bnorm;
            }
        }
        // one launch of the group's operator, timed on the solve's buffers (every
        // segment active) right after the solve
        const int reps = 10;
        AXG_HIP(hipEventRecord(ctx->ev0, bg.s));
        for (int r = 0; r < reps; r++) bsell64_spmm<K>(bg.op, bg.s, false, bg.U, bg.L, bg.P.p);
        AXG_HIP(hipEventRecord(ctx->ev1, bg.s));
        AXG_HIP(hipEventSynchronize(ctx->ev1));
        float op_ms = 0.f;
        AXG_HIP(hipEventElapsedTime(&op_ms, ctx->ev0, ctx->ev1));
        spmv_ms += (double)op_ms / reps * (double)most_iters;
        launches += most_iters;
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->iters = total_iters;
        info->outer_iters = 1;
        info->nrhs = K;
        info->residual = worst;
        info->rel_residual = worst / worst_b;
        info->solve_ms = total_ms;
        // SpMM launches that did work (each group's longest solve) at the timed launch cost
        info->spmv_launches = launches;
        info->spmv_ms = spmv_ms;
        std::snprintf(info->name, sizeof(info->name),
                      "GPU batched PCG (FP64 node blocks, Jacobi, %d meshes, %d groups)", nmesh, G);
    }
    if (status != AXG_OK)
        throw Error{status, status == AXG_E_NOT_SPD
                                ? "CG: unerwartetes Null-Pivot — Matrix nicht positiv definit."
                                : "CG konvergierte nicht (batched solve)."};
}

template <int K>
static double time_block_spmm(BlockPcg &bp, int reps) {
    axg_ctx *ctx = bp.m->ctx;
    hipStream_t s = ctx->stream;
    block_spmv<K>(bp, s, false, bp.Bv.p);
    AXG_HIP(hipEventRecord(ctx->ev0, s));
    for (int r = 0; r < reps; r++) block_spmv<K>(bp, s, false, bp.Bv.p);
    AXG_HIP(hipEventRecord(ctx->ev1, s));
    AXG_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    AXG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    return (double)ms / reps;
}

void bench_block_spmm(axg_mesh *m, int nrhs, int reps, double *avg_ms) {
    std::unique_ptr<BlockPcg> bp(block_pcg_create(m, 0));
    *avg_ms = nrhs == 8   ? time_block_spmm<8>(*bp, reps)
              : nrhs == 4 ? time_block_spmm<4>(*bp, reps)
              : nrhs == 2 ? time_block_spmm<2>(*bp, reps)
                          : time_block_spmm<1>(*bp, reps);
}

void block_pcg_refresh(BlockPcg *bp, int64_t max_iter) {
    axg_mesh *m = bp->m;
    hipStream_t s = m->ctx->stream;
    ensure_bsell64(m);
    if (bp->blk3) {
        ensure_block_inverse64(m);
        bp->B.binv = bp->Bmv.binv = m->binv64.p;
    }
    const int64_t mi = max_iter > 0 ? max_iter : std::max<int64_t>(4 * bp->n, 200);
    // the graphs hold max_iter and the operator's and inverses' addresses as
    // kernel arguments
    if (mi != bp->max_iter || m->bval64.p != bp->graph_bv || bp->B.binv != bp->graph_binv)
        bp->drop_graphs();
    bp->max_iter = mi;
    hipLaunchKernelGGL(block_dsel_kernel, dim3((unsigned)((bp->npad + 255) / 256)), dim3(256), 0, s,
                       bp->npad, bp->rows, m->free_of_blk.p, m->dsel.p, bp->dsel.p);
    AXG_HIP(hipGetLastError());
}

void block_pcg_solve(BlockPcg *bp, const double *d_b, int nrhs, const double *tol_rel,
                     double *d_x, PcgResult &res, bool timing) {
    int j0 = 0;
    while (j0 < nrhs) {
        const int left = nrhs - j0;
        const int kk = left >= 8 ? 8 : left;
        const double *b = d_b + (int64_t)j0 * bp->n;
        double *x = d_x + (int64_t)j0 * bp->n;
        if (kk > 4) block_group<8>(*bp, b, kk, tol_rel + j0, x, res, timing);
        else if (kk > 2) block_group<4>(*bp, b, kk, tol_rel + j0, x, res, timing);
        else if (kk == 2) block_group<2>(*bp, b, kk, tol_rel + j0, x, res, timing);
        else block_group<1>(*bp, b, kk, tol_rel + j0, x, res, timing);
        j0 += kk;
    }
}

}  // namespace axg

using namespace axg;

extern "C" int axg_solve_batch(axg_ctx *ctx, int32_t nmesh, axg_mesh *const *meshes,
                               const double *const *rhs, const int32_t *nrhs,
                               const axg_solve_opts *opts, double *const *x, int64_t *iters,
                               double *solve_ms, axg_solve_info *info) {
    try {
        clear_error();
        AXG_REQUIRE(ctx && nmesh > 0 && meshes && rhs && nrhs && x, AXG_E_INVALID,
                    "axg_solve_batch: invalid argument");
        axg_solve_opts o;
        axg_default_opts(&o);
        if (opts) o = *opts;
        // the batched solver is the FP64 Jacobi PCG: other forms are refused, not
        // replaced (checked before the meshes or the device are touched)
        AXG_REQUIRE(o.precision == AXG_PREC_FP64, AXG_E_INVALID,
                    "axg_solve_batch: precision must be AXG_PREC_FP64 (the batched solver is FP64)");
        AXG_REQUIRE(o.precond == AXG_PRECOND_JACOBI, AXG_E_INVALID,
                    "axg_solve_batch: precond must be AXG_PRECOND_JACOBI (the batched solver is "
                    "Jacobi PCG)");
        AXG_HIP(hipSetDevice(ctx->device));
        int kmax = 0;
        for (int m = 0; m < nmesh; m++) {
            AXG_REQUIRE(meshes[m] && meshes[m]->ctx == ctx && meshes[m]->assembled64,
                        AXG_E_INVALID, "axg_solve_batch: every mesh must be assembled on ctx");
            AXG_REQUIRE(nrhs[m] >= 1 && nrhs[m] <= 8 && rhs[m] && x[m], AXG_E_INVALID,
                        "axg_solve_batch: 1 to 8 right-hand sides per mesh");
            kmax = std::max(kmax, (int)nrhs[m]);
        }
        if (kmax == 1) run_batch<1>(ctx, nmesh, meshes, rhs, nrhs, o, x, iters, solve_ms, info);
        else if (kmax == 2) run_batch<2>(ctx, nmesh, meshes, rhs, nrhs, o, x, iters, solve_ms, info);
        else if (kmax <= 4) run_batch<4>(ctx, nmesh, meshes, rhs, nrhs, o, x, iters, solve_ms, info);
        else run_batch<8>(ctx, nmesh, meshes, rhs, nrhs, o, x, iters, solve_ms, info);
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    }
    return AXG_OK;
}
