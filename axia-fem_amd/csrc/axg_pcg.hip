// Preconditioned conjugate gradients on gfx950.
//
// FP64 path: the reference's Jacobi PCG (linalg.rs:194-259) — x0 = 0,
// recursive residual, stop when ||r|| <= tol ||b||, "not SPD" when
// |p.Ap| < 1e-30, "not converged" when the final ||r|| > 10 tol ||b|| — on a
// CSR matrix. Each iteration is three kernels:
//   spmv_dot   Ap = A p and p.Ap (last workgroup forms alpha)
//   update_xr  x += alpha p, r -= alpha Ap, ||r||^2, r.z (last workgroup forms
//              beta and the stop decision)
//   update_p   p = z + beta p
// Scalars live in device memory (PcgScalars), so a batch of iterations runs
// without host round trips; it is captured once into a hipGraph and replayed
// until the device-side "done" flag is set. Reductions are deterministic:
// fixed grid, per-workgroup partials, summed in index order by the last
// workgroup to finish.
//
// Mixed path: the same iteration on an FP32 BSR-3 operator with k
// right-hand sides at once (SpMM reuses each block for all k), 3x3
// block-Jacobi or Jacobi preconditioner, inside an FP64 iterative refinement
// whose true residual f - K u uses the FP64 CSR.
#include <cmath>

#include "axg_internal.h"
#include "axg_pcg.h"

namespace axg {

constexpr int BS = 256;  // threads per workgroup for vector kernels

// ------------------------------------------------------------ reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum over the workgroup; result valid in thread 0. sh: BS/64 doubles.
__device__ inline double block_sum(double v, double *sh) {
    v = wave_sum(v);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += sh[w];
    return t;
}

// Called by every thread after thread 0 stored this workgroup's partial(s).
// Returns true in the workgroup that finished last; it may then read all
// partials. The ticket is reset for the next launch.
__device__ inline bool last_block(uint32_t *ticket) {
    __shared__ bool is_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = atomicAdd(ticket, 1u);
        is_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (is_last) __threadfence();
    return is_last;
}

// Thread 0 of the last workgroup sums nblk partials (stride `stride`,
// component `comp`) in index order, in chunks handled by the whole group.
__device__ inline double reduce_partials(const double *part, int nblk, int stride, int comp,
                                         double *sh) {
    // each thread sums a contiguous range; then thread 0 sums thread sums in order
    int per = (nblk + blockDim.x - 1) / blockDim.x;
    int a = threadIdx.x * per, b = min(nblk, a + per);
    double s = 0.0;
    for (int i = a; i < b; i++) s += part[(size_t)i * stride + comp];
    __shared__ double tsum[BS];
    tsum[threadIdx.x] = s;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)blockDim.x; i++) t += tsum[i];
    __syncthreads();
    (void)sh;
    return t;
}

// ------------------------------------------------------------- FP64 CSR
template <int L>
__global__ void __launch_bounds__(BS) spmv_dot_kernel(int64_t n, const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ colind,
                                                     const double *__restrict__ val,
                                                     const double *__restrict__ p,
                                                     double *__restrict__ ap,
                                                     double *__restrict__ part, PcgScalars *sc) {
    if (sc->done) return;
    __shared__ double sh[BS / 64];
    const int lane = threadIdx.x % L;
    const int64_t group = ((int64_t)blockIdx.x * BS + threadIdx.x) / L;
    const int64_t ngroups = (int64_t)gridDim.x * BS / L;
    double dot = 0.0;
    for (int64_t row = group; row < n; row += ngroups) {
        int32_t k0 = rowptr[row], k1 = rowptr[row + 1];
        double s = 0.0;
        for (int32_t k = k0 + lane; k < k1; k += L) s += val[k] * p[colind[k]];
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_down(s, off, L);
        if (lane == 0) {
            ap[row] = s;
            dot += p[row] * s;
        }
    }
    double t = block_sum(dot, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
    if (last_block(&sc->ticket[0])) {
        double pap = reduce_partials(part, gridDim.x, 1, 0, sh);
        if (threadIdx.x == 0) {
            sc->pap = pap;
            if (fabs(pap) < 1e-30) {
                sc->status = AXG_E_NOT_SPD;
                sc->done = 1;
            } else {
                sc->alpha = sc->rz / pap;
            }
            sc->ticket[0] = 0;
        }
    }
}

// Plain y = A x (no reduction), used for true residuals and reactions.
template <int L>
__global__ void __launch_bounds__(BS) spmv_kernel(int64_t n, const int32_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ colind,
                                                 const double *__restrict__ val,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ y) {
    const int lane = threadIdx.x % L;
    const int64_t group = ((int64_t)blockIdx.x * BS + threadIdx.x) / L;
    const int64_t ngroups = (int64_t)gridDim.x * BS / L;
    for (int64_t row = group; row < n; row += ngroups) {
        int32_t k0 = rowptr[row], k1 = rowptr[row + 1];
        double s = 0.0;
        for (int32_t k = k0 + lane; k < k1; k += L) s += val[k] * x[colind[k]];
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_down(s, off, L);
        if (lane == 0) y[row] = s;
    }
}

// z = M^-1 r for one DOF: Jacobi (r / d, linalg.rs:213-216) or the 3x3
// node-block inverse (block of DOF i, row d).
struct Precond64 {
    const double *dsel;       // Jacobi divisor
    const double *binv;       // 9 per block, or null
    const int32_t *blk_of_free;  // 3*blockrow + d
    const int32_t *free_of_blk;  // free index or -1
    __device__ inline double apply(const double *r, int64_t i) const {
        if (!binv) return r[i] / dsel[i];
        int32_t bd = blk_of_free[i];
        int32_t b = bd / 3, d = bd % 3;
        const double *B = binv + 9 * (int64_t)b + 3 * d;
        double z = 0.0;
#pragma unroll
        for (int e = 0; e < 3; e++) {
            int32_t f = free_of_blk[3 * b + e];
            if (f >= 0) z += B[e] * r[f];
        }
        return z;
    }
};

// x += alpha p, r -= alpha Ap, ||r||^2; with Jacobi also z = r / d and r.z
// (FUSED). The last workgroup forms ||r||, counts the iteration, decides
// convergence (linalg.rs:241-245) and, with FUSED, beta.
template <bool FUSED>
__global__ void __launch_bounds__(BS) update_xr_kernel(int64_t n, double *__restrict__ x,
                                                      double *__restrict__ r,
                                                      const double *__restrict__ p,
                                                      const double *__restrict__ ap,
                                                      double *__restrict__ z,
                                                      const double *__restrict__ dsel,
                                                      double *__restrict__ part, PcgScalars *sc) {
    if (sc->done) return;
    __shared__ double sh[BS / 64];
    const double alpha = sc->alpha;
    double rr = 0.0, rz = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        x[i] += alpha * p[i];
        double ri = r[i] - alpha * ap[i];
        r[i] = ri;
        rr += ri * ri;
        if (FUSED) {
            double zi = ri / dsel[i];
            z[i] = zi;
            rz += ri * zi;
        }
    }
    double t = block_sum(rr, sh);
    if (threadIdx.x == 0) part[2 * blockIdx.x] = t;
    if (FUSED) {
        t = block_sum(rz, sh);
        if (threadIdx.x == 0) part[2 * blockIdx.x + 1] = t;
    }
    if (last_block(&sc->ticket[1])) {
        double rrs = reduce_partials(part, gridDim.x, 2, 0, sh);
        double rzs = FUSED ? reduce_partials(part, gridDim.x, 2, 1, sh) : 0.0;
        if (threadIdx.x == 0) {
            double res = sqrt(rrs);
            sc->residual = res;
            sc->iters += 1;
            if (res <= sc->tol_abs) {
                sc->done = 1;
            } else if (sc->iters >= sc->max_iter) {
                sc->done = 1;
                sc->status = AXG_E_NOT_CONVERGED;
            } else if (FUSED) {
                sc->beta = rzs / sc->rz;
                sc->rz = rzs;
            }
            sc->ticket[1] = 0;
        }
    }
}

// z = M^-1 r, r.z, then (last workgroup) beta.
__global__ void __launch_bounds__(BS) precond_dot_kernel(int64_t n, const double *__restrict__ r,
                                                        double *__restrict__ z, Precond64 M,
                                                        double *__restrict__ part,
                                                        PcgScalars *sc, int init) {
    if (sc->done) return;
    __shared__ double sh[BS / 64];
    double rz = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        double zi = M.apply(r, i);
        z[i] = zi;
        rz += r[i] * zi;
    }
    double t = block_sum(rz, sh);
    if (threadIdx.x == 0) part[2 * blockIdx.x + 1] = t;
    if (last_block(&sc->ticket[2])) {
        double rzs = reduce_partials(part, gridDim.x, 2, 1, sh);
        if (threadIdx.x == 0) {
            if (!init) sc->beta = rzs / sc->rz;
            sc->rz = rzs;
            sc->ticket[2] = 0;
        }
    }
}

__global__ void __launch_bounds__(BS) update_p_kernel(int64_t n, double *__restrict__ p,
                                                     const double *__restrict__ z, PcgScalars *sc,
                                                     int init) {
    if (sc->done) return;
    const double beta = init ? 0.0 : sc->beta;
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
        p[i] = init ? z[i] : z[i] + beta * p[i];
}

// x = 0, r = b, ||b|| (last workgroup sets tol_abs and the zero-rhs exit).
__global__ void __launch_bounds__(BS) init_kernel(int64_t n, const double *__restrict__ b,
                                                 double *__restrict__ x, double *__restrict__ r,
                                                 double *__restrict__ part, PcgScalars *sc,
                                                 double tol) {
    __shared__ double sh[BS / 64];
    double bb = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        double bi = b[i];
        x[i] = 0.0;
        r[i] = bi;
        bb += bi * bi;
    }
    double t = block_sum(bb, sh);
    if (threadIdx.x == 0) part[2 * blockIdx.x] = t;
    if (last_block(&sc->ticket[1])) {
        double s = reduce_partials(part, gridDim.x, 2, 0, sh);
        if (threadIdx.x == 0) {
            double bn = sqrt(s);
            sc->bnorm = bn;
            sc->tol_abs = tol * bn;
            sc->residual = bn;
            if (bn < 1e-30) sc->done = 1;
            sc->ticket[1] = 0;
        }
    }
}

// dsel[i] = first (i, i) entry of row i if |d| > 1e-30, else 1.0 (Csr::diag,
// linalg.rs:52-63, and the divisor rule of linalg.rs:213-216).
__global__ void diag_kernel(int64_t n, const int32_t *__restrict__ rowptr,
                            const int32_t *__restrict__ colind, const double *__restrict__ val,
                            double *__restrict__ dsel, double *__restrict__ diag_raw) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; k++)
        if (colind[k] == i) {
            d = val[k];
            break;
        }
    if (diag_raw) diag_raw[i] = d;
    dsel[i] = fabs(d) > 1e-30 ? d : 1.0;
}

// ------------------------------------------------------------ host driver
static int pick_lanes(int64_t n, int64_t nnz) {
    double avg = n > 0 ? (double)nnz / (double)n : 1.0;
    if (avg <= 12) return 4;
    if (avg <= 24) return 8;
    if (avg <= 48) return 16;
    if (avg <= 160) return 32;
    return 64;
}

static int grid_for(axg_ctx *ctx, int64_t work_threads, int per_cu) {
    int64_t g = (work_threads + BS - 1) / BS;
    int64_t cap = (int64_t)ctx->cu_count * per_cu;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

template <int L>
static void launch_spmv_dot_l(hipStream_t s, int grid, const CsrView &A, const double *p,
                              double *ap, double *part, PcgScalars *sc) {
    hipLaunchKernelGGL((spmv_dot_kernel<L>), dim3(grid), dim3(BS), 0, s, A.n, A.rowptr, A.colind,
                       A.val, p, ap, part, sc);
}

static void launch_spmv_dot(int L, hipStream_t s, int grid, const CsrView &A, const double *p,
                            double *ap, double *part, PcgScalars *sc) {
    switch (L) {
        case 4: launch_spmv_dot_l<4>(s, grid, A, p, ap, part, sc); break;
        case 8: launch_spmv_dot_l<8>(s, grid, A, p, ap, part, sc); break;
        case 16: launch_spmv_dot_l<16>(s, grid, A, p, ap, part, sc); break;
        case 32: launch_spmv_dot_l<32>(s, grid, A, p, ap, part, sc); break;
        default: launch_spmv_dot_l<64>(s, grid, A, p, ap, part, sc); break;
    }
}

void spmv_fp64(axg_ctx *ctx, const CsrView &A, const double *x, double *y) {
    int L = pick_lanes(A.n, A.nnz);
    int grid = grid_for(ctx, A.n * L, 16);
    hipStream_t s = ctx->stream;
    switch (L) {
        case 4: hipLaunchKernelGGL((spmv_kernel<4>), dim3(grid), dim3(BS), 0, s, A.n, A.rowptr, A.colind, A.val, x, y); break;
        case 8: hipLaunchKernelGGL((spmv_kernel<8>), dim3(grid), dim3(BS), 0, s, A.n, A.rowptr, A.colind, A.val, x, y); break;
        case 16: hipLaunchKernelGGL((spmv_kernel<16>), dim3(grid), dim3(BS), 0, s, A.n, A.rowptr, A.colind, A.val, x, y); break;
        case 32: hipLaunchKernelGGL((spmv_kernel<32>), dim3(grid), dim3(BS), 0, s, A.n, A.rowptr, A.colind, A.val, x, y); break;
        default: hipLaunchKernelGGL((spmv_kernel<64>), dim3(grid), dim3(BS), 0, s, A.n, A.rowptr, A.colind, A.val, x, y); break;
    }
    AXG_HIP(hipGetLastError());
}

void diag_fp64(axg_ctx *ctx, const CsrView &A, double *dsel, double *diag_raw) {
    if (A.n == 0) return;
    hipLaunchKernelGGL(diag_kernel, dim3((unsigned)((A.n + 255) / 256)), dim3(256), 0, ctx->stream,
                       A.n, A.rowptr, A.colind, A.val, dsel, diag_raw);
    AXG_HIP(hipGetLastError());
}

// One FP64 PCG solve. b, x: device vectors of length A.n.
void pcg_fp64(axg_ctx *ctx, const CsrView &A, const Precond64Args &pa, const double *b,
              double *x, double tol, int64_t max_iter, bool timing, PcgResult &res) {
    hipStream_t s = ctx->stream;
    const int64_t n = A.n;
    Workspace &w = ctx->ws;
    w.r.ensure(n);
    w.p.ensure(n);
    w.ap.ensure(n);
    w.z.ensure(n);
    const int L = pick_lanes(n, A.nnz);
    const int grid_mv = grid_for(ctx, n * L, 8);
    const int grid_v = grid_for(ctx, n, 4);
    w.part.ensure(2 * (size_t)std::max(grid_mv, grid_v) + 2);
    w.sc.ensure(1);
    Precond64 M{pa.dsel, pa.binv, pa.blk_of_free, pa.free_of_blk};

    PcgScalars init{};
    init.max_iter = max_iter;
    ctx->pinned.ensure(sizeof(PcgScalars) * 2);
    PcgScalars *h = (PcgScalars *)ctx->pinned.p;
    *h = init;
    AXG_HIP(hipMemcpyAsync(w.sc.p, h, sizeof(PcgScalars), hipMemcpyHostToDevice, s));
    AXG_HIP(hipEventRecord(ctx->ev0, s));
    hipLaunchKernelGGL(init_kernel, dim3(grid_v), dim3(BS), 0, s, n, b, x, w.r.p, w.part.p, w.sc.p,
                       tol);
    hipLaunchKernelGGL(precond_dot_kernel, dim3(grid_v), dim3(BS), 0, s, n, w.r.p, w.z.p, M,
                       w.part.p, w.sc.p, 1);
    hipLaunchKernelGGL(update_p_kernel, dim3(grid_v), dim3(BS), 0, s, n, w.p.p, w.z.p, w.sc.p, 1);
    AXG_HIP(hipGetLastError());

    auto one_iter = [&](hipEvent_t e0, hipEvent_t e1) {
        if (e0) AXG_HIP(hipEventRecord(e0, s));
        launch_spmv_dot(L, s, grid_mv, A, w.p.p, w.ap.p, w.part.p, w.sc.p);
        if (e1) AXG_HIP(hipEventRecord(e1, s));
        if (M.binv == nullptr) {
            hipLaunchKernelGGL((update_xr_kernel<true>), dim3(grid_v), dim3(BS), 0, s, n, x, w.r.p,
                               w.p.p, w.ap.p, w.z.p, M.dsel, w.part.p, w.sc.p);
        } else {
            hipLaunchKernelGGL((update_xr_kernel<false>), dim3(grid_v), dim3(BS), 0, s, n, x,
                               w.r.p, w.p.p, w.ap.p, w.z.p, M.dsel, w.part.p, w.sc.p);
            hipLaunchKernelGGL(precond_dot_kernel, dim3(grid_v), dim3(BS), 0, s, n, w.r.p, w.z.p,
                               M, w.part.p, w.sc.p, 0);
        }
        hipLaunchKernelGGL(update_p_kernel, dim3(grid_v), dim3(BS), 0, s, n, w.p.p, w.z.p, w.sc.p,
                           0);
    };

    res.spmv_ms = 0.0;
    res.spmv_launches = 0;
    const int batch = n > 200000 ? 16 : 64;
    hipGraphExec_t gexec = nullptr;
    if (!timing) {
        hipGraph_t graph;
        AXG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < batch; i++) one_iter(nullptr, nullptr);
        AXG_HIP(hipStreamEndCapture(s, &graph));
        AXG_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        AXG_HIP(hipGraphDestroy(graph));
    } else {
        size_t need = 2 * (size_t)batch;
        while (ctx->timing_events.size() < need) {
            hipEvent_t e;
            AXG_HIP(hipEventCreate(&e));
            ctx->timing_events.push_back(e);
        }
    }
    int64_t launched = 0;
    for (;;) {
        AXG_HIP(hipMemcpyAsync(h, w.sc.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
        AXG_HIP(hipStreamSynchronize(s));
        if (h->done || launched >= max_iter) break;
        if (timing) {
            for (int i = 0; i < batch; i++)
                one_iter(ctx->timing_events[2 * i], ctx->timing_events[2 * i + 1]);
            AXG_HIP(hipMemcpyAsync(h, w.sc.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
            AXG_HIP(hipStreamSynchronize(s));
            // SpMV launches that did work: the iterations counted in this batch
            int64_t did = h->iters - res.spmv_launches;
            for (int64_t i = 0; i < did && i < batch; i++) {
                float ms = 0.f;
                AXG_HIP(hipEventElapsedTime(&ms, ctx->timing_events[2 * i],
                                            ctx->timing_events[2 * i + 1]));
                res.spmv_ms += ms;
            }
            res.spmv_launches = h->iters;
        } else {
            AXG_HIP(hipGraphLaunch(gexec, s));
        }
        launched += batch;
    }
    AXG_HIP(hipEventRecord(ctx->ev1, s));
    AXG_HIP(hipEventSynchronize(ctx->ev1));
    if (gexec) AXG_HIP(hipGraphExecDestroy(gexec));
    float ms = 0.f;
    AXG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    res.solve_ms = ms;
    res.iters = h->iters;
    res.residual = h->residual;
    res.bnorm = h->bnorm;
    res.status = h->status;
    if (res.status == AXG_E_NOT_CONVERGED && h->residual <= 10.0 * tol * h->bnorm)
        res.status = AXG_OK;  // linalg.rs:253: only > 10 tol ||b|| is an error
    if (res.status == AXG_OK && !h->done && h->bnorm >= 1e-30) res.status = AXG_E_NOT_CONVERGED;
}

}  // namespace axg
