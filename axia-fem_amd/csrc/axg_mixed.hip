// Node-block (3x3) structures, the FP32 BSR-3 operator and the
// mixed-precision solve: an FP32 multi-right-hand-side PCG (BSR-3 SpMM,
// 3x3 block-Jacobi or Jacobi) inside an FP64 iterative refinement
//   r = f - K u (FP64 CSR), solve K32 d = r / |r| (FP32), u += |r| d (FP64)
// until |f - K u| <= tol |f| for every right-hand side.
//
// Vectors of the inner solve live in block space: 3 entries per node that
// keeps a free DOF (constrained entries stay zero), k right-hand sides
// interleaved (entry (3 b + d) * K + j). BSR values are stored as nine planes
// (plane 3 d + e holds entry (d, e) of every block) so that consecutive lanes
// read consecutive floats.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "axg_internal.h"
#include "axg_mixed.h"

namespace axg {


constexpr int MBS = 256;

// ---------------------------------------------------------- block maps
void ensure_block_maps(axg_mesh *m) {
    if (m->nb > 0 || m->n_free == 0) return;
    hipStream_t s = m->ctx->stream;
    std::vector<int32_t> ind(m->ndof);
    AXG_HIP(hipMemcpyAsync(ind.data(), m->ind_of.p, 4 * m->ndof, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> fob, bof(m->n_free), node;
    for (int64_t a = 0; a < m->nnode; a++) {
        bool any = ind[3 * a] >= 0 || ind[3 * a + 1] >= 0 || ind[3 * a + 2] >= 0;
        if (!any) continue;
        int32_t b = (int32_t)node.size();
        node.push_back((int32_t)a);
        for (int d = 0; d < 3; d++) {
            fob.push_back(ind[3 * a + d]);
            if (ind[3 * a + d] >= 0) bof[ind[3 * a + d]] = 3 * b + d;
        }
    }
    m->nb = (int64_t)node.size();
    m->free_of_blk.alloc(fob.size());
    m->blk_of_free.alloc(bof.size());
    m->b_node.alloc(node.size());
    AXG_HIP(hipMemcpyAsync(m->free_of_blk.p, fob.data(), 4 * fob.size(), hipMemcpyHostToDevice, s));
    AXG_HIP(hipMemcpyAsync(m->blk_of_free.p, bof.data(), 4 * bof.size(), hipMemcpyHostToDevice, s));
    AXG_HIP(hipMemcpyAsync(m->b_node.p, node.data(), 4 * node.size(), hipMemcpyHostToDevice, s));
    AXG_HIP(hipStreamSynchronize(s));
}

// Inverse of the free-DOF part of each node's 3x3 diagonal block of K_ff,
// formed in FP64 (cofactors). Constrained positions are identity in the
// block and zero in the inverse. A block with |det| <= 1e-300 falls back to
// Jacobi on its diagonal.
__global__ void block_inverse_kernel(int64_t nb, const int32_t *__restrict__ rowptr,
                                     const int32_t *__restrict__ colind,
                                     const double *__restrict__ val,
                                     const int32_t *__restrict__ free_of_blk,
                                     double *__restrict__ binv, float *__restrict__ binv32,
                                     float *__restrict__ dinv32) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int32_t f[3] = {free_of_blk[3 * b], free_of_blk[3 * b + 1], free_of_blk[3 * b + 2]};
    double a[3][3];
    for (int d = 0; d < 3; d++)
        for (int e = 0; e < 3; e++) a[d][e] = (d == e) ? 1.0 : 0.0;
    for (int d = 0; d < 3; d++) {
        if (f[d] < 0) continue;
        a[d][d] = 0.0;
        for (int32_t k = rowptr[f[d]]; k < rowptr[f[d] + 1]; k++) {
            int32_t c = colind[k];
            for (int e = 0; e < 3; e++)
                if (f[e] >= 0 && c == f[e]) a[d][e] = val[k];
        }
    }
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                 a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    double inv[3][3];
    if (fabs(det) > 1e-300) {
        double id = 1.0 / det;
        inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) * id;
        inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
        inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
        inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * id;
        inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
        inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
        inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) * id;
        inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
        inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
    } else {
        for (int d = 0; d < 3; d++)
            for (int e = 0; e < 3; e++)
                inv[d][e] = (d == e && fabs(a[d][d]) > 1e-30) ? 1.0 / a[d][d] : 0.0;
    }
    for (int d = 0; d < 3; d++) {
        for (int e = 0; e < 3; e++) {
            double v = (f[d] >= 0 && f[e] >= 0) ? inv[d][e] : 0.0;
            binv[9 * b + 3 * d + e] = v;
            if (binv32) binv32[9 * b + 3 * d + e] = (float)v;
        }
        if (dinv32)
            dinv32[3 * b + d] = (f[d] >= 0 && fabs(a[d][d]) > 1e-30) ? (float)(1.0 / a[d][d]) : 0.f;
    }
}

void ensure_block_inverse64(axg_mesh *m) {
    ensure_block_maps(m);
    if (m->binv_valid) return;
    m->binv64.ensure(9 * std::max<int64_t>(m->nb, 1));
    m->binv32.ensure(9 * std::max<int64_t>(m->nb, 1));
    m->dinv32.ensure(3 * std::max<int64_t>(m->nb, 1));
    CsrView A = active_csr(m);
    if (m->nb > 0)
        hipLaunchKernelGGL(block_inverse_kernel, dim3((unsigned)((m->nb + 127) / 128)), dim3(128),
                           0, m->ctx->stream, m->nb, A.rowptr, A.colind, A.val, m->free_of_blk.p,
                           m->binv64.p, m->binv32.p, m->dinv32.p);
    AXG_HIP(hipGetLastError());
    AXG_HIP(hipStreamSynchronize(m->ctx->stream));
    m->binv_valid = true;
}

// ------------------------------------------------------- BSR-3 pattern
// Block (I, J) exists when any free DOF of node I couples with any free DOF
// of node J in the topological K_ff pattern. entry_of_slot maps each scalar
// slot to its plane position (3 d + e) * nnzb + block.
static void build_bsr(axg_mesh *m) {
    ensure_block_maps(m);
    hipStream_t s = m->ctx->stream;
    const int64_t n = m->n_free, nnz = m->nnz_topo;
    std::vector<int32_t> rp(n + 1), ci(nnz), bof(n), fob(3 * m->nb);
    AXG_HIP(hipMemcpyAsync(rp.data(), m->rowptr_t.p, 4 * (n + 1), hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(ci.data(), m->colind_t.p, 4 * nnz, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(bof.data(), m->blk_of_free.p, 4 * n, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(fob.data(), m->free_of_blk.p, 4 * 3 * m->nb, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    const int64_t nb = m->nb;
    std::vector<int32_t> brp(nb + 1, 0), bci;
    bci.reserve(nnz / 3 + 16);
    // block columns of block row I = block rows of the columns of I's first free row
    // (all free rows of a node share the same node neighbourhood)
    for (int64_t I = 0; I < nb; I++) {
        int32_t r = -1;
        for (int d = 0; d < 3 && r < 0; d++) r = fob[3 * I + d];
        int32_t last = -1;
        for (int32_t k = rp[r]; k < rp[r + 1]; k++) {
            int32_t J = bof[ci[k]] / 3;
            if (J != last) {
                bci.push_back(J);
                last = J;
            }
        }
        brp[I + 1] = (int32_t)bci.size();
    }
    const int64_t nnzb = (int64_t)bci.size();
    std::vector<int32_t> eos(nnz);
    for (int64_t i = 0; i < n; i++) {
        int32_t I = bof[i] / 3, d = bof[i] % 3;
        int32_t k0 = brp[I], k1 = brp[I + 1];
        for (int32_t k = rp[i]; k < rp[i + 1]; k++) {
            int32_t J = bof[ci[k]] / 3, e = bof[ci[k]] % 3;
            int32_t pos = (int32_t)(std::lower_bound(bci.begin() + k0, bci.begin() + k1, J) -
                                    bci.begin());
            eos[k] = (int32_t)((3 * d + e) * nnzb + pos);
        }
    }
    AXG_REQUIRE(9 * nnzb < INT32_MAX, AXG_E_INVALID, "BSR-3 operator exceeds 2^31 entries");
    m->nnzb = nnzb;
    m->b_rowptr.alloc(nb + 1);
    m->b_colind.alloc(std::max<int64_t>(nnzb, 1));
    m->entry_of_slot.alloc(std::max<int64_t>(nnz, 1));
    AXG_HIP(hipMemcpyAsync(m->b_rowptr.p, brp.data(), 4 * (nb + 1), hipMemcpyHostToDevice, s));
    if (nnzb)
        AXG_HIP(hipMemcpyAsync(m->b_colind.p, bci.data(), 4 * nnzb, hipMemcpyHostToDevice, s));
    if (nnz)
        AXG_HIP(hipMemcpyAsync(m->entry_of_slot.p, eos.data(), 4 * nnz, hipMemcpyHostToDevice, s));
    AXG_HIP(hipStreamSynchronize(s));
}

void assemble_mixed(axg_mesh *m, double *elapsed_ms) {
    axg_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    if (m->nnzb == 0) build_bsr(m);
    ensure_block_inverse64(m);
    m->ke32.ensure((size_t)m->nelem * m->m * m->m);
    m->bval.ensure(std::max<int64_t>(9 * m->nnzb, 1));
    Workspace &w = ctx->ws;
    AXG_HIP(hipMemsetAsync(w.err.p, 0, 8, s));
    AXG_HIP(hipMemsetAsync(m->bval.p, 0, sizeof(float) * 9 * m->nnzb, s));
    AXG_HIP(hipEventRecord(ctx->ev0, s));
    launch_ke_fp32(m, s, w.err.p);
    AXG_HIP(hipEventRecord(ctx->evA, s));
    launch_gather_bsr32(m, s, m->entry_of_slot.p);
    AXG_HIP(hipEventRecord(ctx->ev1, s));
    int32_t err[2];
    AXG_HIP(hipMemcpyAsync(err, w.err.p, 8, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    if (err[0]) throw Error{err[0], "FP32 integration: negative Jakobideterminante (element index " +
                                        std::to_string(err[1]) + ")"};
    float ms = 0.f, ms_a = 0.f;
    AXG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    AXG_HIP(hipEventElapsedTime(&ms_a, ctx->ev0, ctx->evA));
    m->t_ke32 = ms_a;
    m->t_gather32 = ms - ms_a;
    if (elapsed_ms) *elapsed_ms = ms;
    m->assembled32 = true;
}

// ---------------------------------------------------------- reductions
__device__ inline double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// K workgroup sums of v[j]; results valid in thread 0.
template <int K>
__device__ inline void block_sums(double (&v)[K], double *sh /* K * MBS/64 */) {
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < K; j++) v[j] = wsum(v[j]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < K; j++) sh[j * (MBS / 64) + wid] = v[j];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < K; j++) {
            double t = 0.0;
            for (int w = 0; w < MBS / 64; w++) t += sh[j * (MBS / 64) + w];
            v[j] = t;
        }
}

__device__ inline bool last_block_m(uint32_t *ticket) {
    __shared__ bool is_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) is_last = (atomicAdd(ticket, 1u) == gridDim.x - 1);
    __syncthreads();
    if (is_last) __threadfence();
    return is_last;
}

// Sum of part[b * stride + comp] over b < nblk in index order (whole group).
__device__ inline double ordered_sum(const double *part, int nblk, int stride, int comp) {
    __shared__ double t[MBS];
    int per = (nblk + MBS - 1) / MBS;
    int a = threadIdx.x * per, b = min(nblk, a + per);
    double s = 0.0;
    for (int i = a; i < b; i++) s += part[(size_t)i * stride + comp];
    t[threadIdx.x] = s;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < MBS; i++) r += t[i];
    __syncthreads();
    return r;
}

// Per right-hand-side scalars of the inner solve (device).
struct InnerScalars {
    double rz[8], alpha[8], beta[8], res[8], tol_abs[8], bnorm[8];
    int32_t done[8];
    int32_t iters[8];
    int32_t all_done;
    int32_t max_iter;
    uint32_t ticket[4];
};

// ------------------------------------------------------ BSR-3 SpMM (FP32)
// G lanes per block row; lane l takes blocks k0 + l, k0 + l + G, ... and
// accumulates 3 x K partial outputs, reduced over the G lanes.
template <int K, int G>
__global__ void __launch_bounds__(MBS) bsr_spmm_dot_kernel(
    int64_t nb, const int32_t *__restrict__ brp, const int32_t *__restrict__ bci,
    const float *__restrict__ bval, int64_t nnzb, const float *__restrict__ P,
    float *__restrict__ AP, double *__restrict__ part, InnerScalars *sc) {
    if (sc->all_done) return;
    __shared__ double sh[K * MBS / 64];
    const int lane = threadIdx.x % G;
    const int64_t group = ((int64_t)blockIdx.x * MBS + threadIdx.x) / G;
    const int64_t ngroups = (int64_t)gridDim.x * MBS / G;
    double dot[K];
#pragma unroll
    for (int j = 0; j < K; j++) dot[j] = 0.0;
    for (int64_t I = group; I < nb; I += ngroups) {
        float acc[3][K];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int j = 0; j < K; j++) acc[d][j] = 0.f;
        int32_t k1 = brp[I + 1];
        for (int32_t k = brp[I] + lane; k < k1; k += G) {
            int32_t J = bci[k];
            float v[9];
#pragma unroll
            for (int q = 0; q < 9; q++) v[q] = bval[(int64_t)q * nnzb + k];
            const float *x = P + (int64_t)3 * J * K;
#pragma unroll
            for (int e = 0; e < 3; e++)
#pragma unroll
                for (int j = 0; j < K; j++) {
                    float xe = x[e * K + j];
                    acc[0][j] += v[e] * xe;
                    acc[1][j] += v[3 + e] * xe;
                    acc[2][j] += v[6 + e] * xe;
                }
        }
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int j = 0; j < K; j++)
#pragma unroll
                for (int off = G / 2; off > 0; off >>= 1)
                    acc[d][j] += __shfl_down(acc[d][j], off, G);
        if (lane == 0) {
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
                for (int j = 0; j < K; j++) {
                    int64_t idx = (3 * I + d) * K + j;
                    AP[idx] = acc[d][j];
                    dot[j] += (double)P[idx] * (double)acc[d][j];
                }
        }
    }
    block_sums<K>(dot, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < K; j++) part[(size_t)blockIdx.x * K + j] = dot[j];
    if (last_block_m(&sc->ticket[0])) {
        for (int j = 0; j < K; j++) {
            double pap = ordered_sum(part, gridDim.x, K, j);
            if (threadIdx.x == 0 && !sc->done[j]) {
                if (fabs(pap) < 1e-300 || !(pap > 0.0)) {
                    sc->done[j] = 1;  // breakdown: the sweep ends here, refinement continues
                    sc->alpha[j] = 0.0;
                } else {
                    sc->alpha[j] = sc->rz[j] / pap;
                }
            }
        }
        if (threadIdx.x == 0) sc->ticket[0] = 0;
    }
}

// x += alpha p, r -= alpha Ap, z = M r (3x3 block or diagonal), ||r||^2, r.z;
// the last workgroup decides per right-hand side. One thread per block row.
template <int K, bool BLOCK>
__global__ void __launch_bounds__(MBS) inner_update_kernel(
    int64_t nb, float *__restrict__ X, float *__restrict__ R, const float *__restrict__ P,
    const float *__restrict__ AP, float *__restrict__ Z, const float *__restrict__ binv,
    const float *__restrict__ dinv, double *__restrict__ part, InnerScalars *sc) {
    if (sc->all_done) return;
    __shared__ double sh[K * MBS / 64];
    float alpha[K];
    bool live[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        live[j] = !sc->done[j];
        alpha[j] = live[j] ? (float)sc->alpha[j] : 0.f;
    }
    double rr[K], rz[K];
#pragma unroll
    for (int j = 0; j < K; j++) rr[j] = rz[j] = 0.0;
    for (int64_t b = (int64_t)blockIdx.x * MBS + threadIdx.x; b < nb;
         b += (int64_t)gridDim.x * MBS) {
        float r[3][K];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int j = 0; j < K; j++) {
                int64_t idx = (3 * b + d) * K + j;
                float rv = R[idx];
                if (live[j]) {
                    X[idx] += alpha[j] * P[idx];
                    rv -= alpha[j] * AP[idx];
                    R[idx] = rv;
                }
                r[d][j] = rv;
            }
#pragma unroll
        for (int d = 0; d < 3; d++) {
            float m0 = 0.f, m1 = 0.f, m2 = 0.f, dd = 0.f;
            if (BLOCK) {
                m0 = binv[9 * b + 3 * d], m1 = binv[9 * b + 3 * d + 1], m2 = binv[9 * b + 3 * d + 2];
            } else {
                dd = dinv[3 * b + d];
            }
#pragma unroll
            for (int j = 0; j < K; j++) {
                float z = BLOCK ? (m0 * r[0][j] + m1 * r[1][j] + m2 * r[2][j]) : dd * r[d][j];
                Z[(3 * b + d) * K + j] = z;
                rr[j] += (double)r[d][j] * r[d][j];
                rz[j] += (double)r[d][j] * z;
            }
        }
    }
    block_sums<K>(rr, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < K; j++) part[(size_t)blockIdx.x * 2 * K + j] = rr[j];
    block_sums<K>(rz, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < K; j++) part[(size_t)blockIdx.x * 2 * K + K + j] = rz[j];
    if (last_block_m(&sc->ticket[1])) {
        for (int j = 0; j < K; j++) {
            double rrs = ordered_sum(part, gridDim.x, 2 * K, j);
            double rzs = ordered_sum(part, gridDim.x, 2 * K, K + j);
            if (threadIdx.x == 0 && !sc->done[j]) {
                double res = sqrt(rrs);
                sc->res[j] = res;
                sc->iters[j] += 1;
                if (res <= sc->tol_abs[j] || sc->iters[j] >= sc->max_iter) {
                    sc->done[j] = 1;
                } else {
                    sc->beta[j] = rzs / sc->rz[j];
                    sc->rz[j] = rzs;
                }
            }
        }
        if (threadIdx.x == 0) {
            int all = 1;
            for (int j = 0; j < K; j++) all &= sc->done[j];
            sc->all_done = all;
            sc->ticket[1] = 0;
        }
    }
}

template <int K>
__global__ void __launch_bounds__(MBS) inner_update_p_kernel(int64_t nv, float *__restrict__ P,
                                                            const float *__restrict__ Z,
                                                            InnerScalars *sc, int init) {
    if (sc->all_done) return;
    float beta[K];
    bool live[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        live[j] = !sc->done[j];
        beta[j] = init ? 0.f : (float)sc->beta[j];
    }
    for (int64_t i = (int64_t)blockIdx.x * MBS + threadIdx.x; i < nv;
         i += (int64_t)gridDim.x * MBS) {
#pragma unroll
        for (int j = 0; j < K; j++)
            if (live[j]) P[i * K + j] = init ? Z[i * K + j] : Z[i * K + j] + beta[j] * P[i * K + j];
    }
}

// Inner start: X = 0, R = B, Z = M R, norms; the last workgroup sets
// tol_abs = inner_tol |B_j|, rz and the zero-rhs exits.
template <int K, bool BLOCK>
__global__ void __launch_bounds__(MBS) inner_init_kernel(
    int64_t nb, const float *__restrict__ B, float *__restrict__ X, float *__restrict__ R,
    float *__restrict__ Z, const float *__restrict__ binv, const float *__restrict__ dinv,
    double *__restrict__ part, InnerScalars *sc, double inner_tol) {
    __shared__ double sh[K * MBS / 64];
    double bb[K], rz[K];
#pragma unroll
    for (int j = 0; j < K; j++) bb[j] = rz[j] = 0.0;
    for (int64_t b = (int64_t)blockIdx.x * MBS + threadIdx.x; b < nb;
         b += (int64_t)gridDim.x * MBS) {
        float r[3][K];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int j = 0; j < K; j++) {
                int64_t idx = (3 * b + d) * K + j;
                r[d][j] = B[idx];
                R[idx] = r[d][j];
                X[idx] = 0.f;
            }
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int j = 0; j < K; j++) {
                float z = BLOCK ? (binv[9 * b + 3 * d] * r[0][j] + binv[9 * b + 3 * d + 1] * r[1][j] +
                                   binv[9 * b + 3 * d + 2] * r[2][j])
                                : dinv[3 * b + d] * r[d][j];
                Z[(3 * b + d) * K + j] = z;
                bb[j] += (double)r[d][j] * r[d][j];
                rz[j] += (double)r[d][j] * z;
            }
    }
    block_sums<K>(bb, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < K; j++) part[(size_t)blockIdx.x * 2 * K + j] = bb[j];
    block_sums<K>(rz, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < K; j++) part[(size_t)blockIdx.x * 2 * K + K + j] = rz[j];
    if (last_block_m(&sc->ticket[1])) {
        for (int j = 0; j < K; j++) {
            double bbs = ordered_sum(part, gridDim.x, 2 * K, j);
            double rzs = ordered_sum(part, gridDim.x, 2 * K, K + j);
            if (threadIdx.x == 0) {
                double bn = sqrt(bbs);
                sc->bnorm[j] = bn;
                sc->res[j] = bn;
                sc->tol_abs[j] = inner_tol * bn;
                sc->rz[j] = rzs;
                sc->iters[j] = 0;
                sc->done[j] = (bn < 1e-30) ? 1 : 0;
            }
        }
        if (threadIdx.x == 0) {
            int all = 1;
            for (int j = 0; j < K; j++) all &= sc->done[j];
            sc->all_done = all;
            sc->ticket[1] = 0;
        }
    }
}

// ------------------------------------------------- FP64 residual pieces
// Y[:, j] = F[:, j] - A X[:, j] for K right-hand sides (column-major n x K),
// matrix streamed once. G lanes per row.
template <int K, int G>
__global__ void __launch_bounds__(MBS) csr_residual_kernel(
    int64_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colind,
    const double *__restrict__ val, const double *__restrict__ X, const double *__restrict__ F,
    double *__restrict__ Y) {
    const int lane = threadIdx.x % G;
    const int64_t group = ((int64_t)blockIdx.x * MBS + threadIdx.x) / G;
    const int64_t ngroups = (int64_t)gridDim.x * MBS / G;
    for (int64_t row = group; row < n; row += ngroups) {
        double acc[K];
#pragma unroll
        for (int j = 0; j < K; j++) acc[j] = 0.0;
        for (int32_t k = rowptr[row] + lane; k < rowptr[row + 1]; k += G) {
            double v = val[k];
            int64_t c = colind[k];
#pragma unroll
            for (int j = 0; j < K; j++) acc[j] += v * X[c + j * n];
        }
#pragma unroll
        for (int j = 0; j < K; j++)
#pragma unroll
            for (int off = G / 2; off > 0; off >>= 1) acc[j] += __shfl_down(acc[j], off, G);
        if (lane == 0)
#pragma unroll
            for (int j = 0; j < K; j++) Y[row + j * n] = F[row + j * n] - acc[j];
    }
}

// Column norms of an n x K column-major FP64 array (deterministic).
template <int K>
__global__ void __launch_bounds__(MBS) col_norms_kernel(int64_t n, const double *__restrict__ Y,
                                                       double *__restrict__ part,
                                                       double *__restrict__ out,
                                                       uint32_t *ticket) {
    __shared__ double sh[K * MBS / 64];
    double s[K];
#pragma unroll
    for (int j = 0; j < K; j++) s[j] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * MBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * MBS)
#pragma unroll
        for (int j = 0; j < K; j++) s[j] += Y[i + j * n] * Y[i + j * n];
    block_sums<K>(s, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < K; j++) part[(size_t)blockIdx.x * K + j] = s[j];
    if (last_block_m(ticket)) {
        for (int j = 0; j < K; j++) {
            double t = ordered_sum(part, gridDim.x, K, j);
            if (threadIdx.x == 0) out[j] = sqrt(t);
        }
        if (threadIdx.x == 0) *ticket = 0;
    }
}

// FP64 free-space residual -> FP32 block space, scaled by 1/|r_j|.
template <int K>
__global__ void pack_kernel(int64_t n, const double *__restrict__ Y,
                            const int32_t *__restrict__ blk_of_free,
                            const double *__restrict__ scale, float *__restrict__ B) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t bi = blk_of_free[i];
#pragma unroll
    for (int j = 0; j < K; j++) {
        double s = scale[j];
        B[bi * K + j] = s > 0.0 ? (float)(Y[i + j * n] / s) : 0.f;
    }
}

// u += |r_j| d (FP64 accumulation of the FP32 correction).
template <int K>
__global__ void unpack_add_kernel(int64_t n, const float *__restrict__ D,
                                  const int32_t *__restrict__ blk_of_free,
                                  const double *__restrict__ scale, double *__restrict__ U) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t bi = blk_of_free[i];
#pragma unroll
    for (int j = 0; j < K; j++) U[i + j * n] += scale[j] * (double)D[bi * K + j];
}

// ------------------------------------------------------------ driver
static int lanes_for(double avg) {
    if (avg <= 6) return 4;
    if (avg <= 12) return 8;
    if (avg <= 24) return 16;
    if (avg <= 48) return 32;
    return 64;
}

template <int K>
struct MixedSolver {
    static void run(axg_mesh *m, const double *d_f, const axg_solve_opts &o, double *d_u,
                    MixedResult &res) {
        axg_ctx *ctx = m->ctx;
        hipStream_t s = ctx->stream;
        const int64_t n = m->n_free, nb = m->nb, nv = 3 * nb;
        const bool block = o.precond == AXG_PRECOND_BLOCK3;
        CsrView A = active_csr(m);
        // workspace
        Workspace &w = ctx->ws;
        w.f0.ensure(nv * K);  // B (inner rhs)
        w.f1.ensure(nv * K);  // X
        w.f2.ensure(nv * K);  // R
        w.f3.ensure(nv * K);  // P
        w.f4.ensure(nv * K);  // AP
        DevBuf<float> Z(nv * K);
        w.t0.ensure(n * K);   // r64
        w.t1.ensure(16);      // norms / scales
        const int gb = lanes_for(m->nb ? (double)m->nnzb / (double)m->nb : 1.0);
        const int grid_mv = std::max(1, (int)std::min<int64_t>((nb * gb + MBS - 1) / MBS,
                                                               (int64_t)ctx->cu_count * 8));
        const int grid_v = std::max(1, (int)std::min<int64_t>((nb + MBS - 1) / MBS,
                                                              (int64_t)ctx->cu_count * 4));
        const int gr = lanes_for(n ? (double)A.nnz / (double)n : 1.0);
        const int grid_r = std::max(1, (int)std::min<int64_t>((n * gr + MBS - 1) / MBS,
                                                              (int64_t)ctx->cu_count * 8));
        w.part.ensure((size_t)std::max({grid_mv, grid_v, grid_r}) * 2 * K + 8);
        DevBuf<InnerScalars> dsc(1);
        DevBuf<uint32_t> ticket(1);
        ctx->pinned.ensure(sizeof(InnerScalars) + 256);
        InnerScalars *h = (InnerScalars *)ctx->pinned.p;
        double *hn = (double *)((char *)ctx->pinned.p + sizeof(InnerScalars));
        std::memset(h, 0, sizeof(InnerScalars));
        AXG_HIP(hipMemcpyAsync(ticket.p, h, 4, hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemsetAsync(w.f0.p, 0, sizeof(float) * nv * K, s));
        AXG_HIP(hipMemsetAsync(d_u, 0, sizeof(double) * n * K, s));
        AXG_HIP(hipEventRecord(ctx->ev0, s));

        auto launch_spmm = [&](hipEvent_t e0, hipEvent_t e1) {
            if (e0) AXG_HIP(hipEventRecord(e0, s));
#define AXG_SPMM(GG)                                                                          \
    hipLaunchKernelGGL((bsr_spmm_dot_kernel<K, GG>), dim3(grid_mv), dim3(MBS), 0, s, nb,        \
                       m->b_rowptr.p, m->b_colind.p, m->bval.p, m->nnzb, w.f3.p, w.f4.p, w.part.p, \
                       dsc.p)
            switch (gb) {
                case 4: AXG_SPMM(4); break;
                case 8: AXG_SPMM(8); break;
                case 16: AXG_SPMM(16); break;
                case 32: AXG_SPMM(32); break;
                default: AXG_SPMM(64); break;
            }
#undef AXG_SPMM
            if (e1) AXG_HIP(hipEventRecord(e1, s));
        };
        auto launch_update = [&]() {
            if (block)
                hipLaunchKernelGGL((inner_update_kernel<K, true>), dim3(grid_v), dim3(MBS), 0, s,
                                   nb, w.f1.p, w.f2.p, w.f3.p, w.f4.p, Z.p, m->binv32.p,
                                   m->dinv32.p, w.part.p, dsc.p);
            else
                hipLaunchKernelGGL((inner_update_kernel<K, false>), dim3(grid_v), dim3(MBS), 0, s,
                                   nb, w.f1.p, w.f2.p, w.f3.p, w.f4.p, Z.p, m->binv32.p,
                                   m->dinv32.p, w.part.p, dsc.p);
            hipLaunchKernelGGL((inner_update_p_kernel<K>), dim3(grid_v), dim3(MBS), 0, s, nv,
                               w.f3.p, Z.p, dsc.p, 0);
        };
        auto residual = [&]() {
#define AXG_RES(GG)                                                                              \
    hipLaunchKernelGGL((csr_residual_kernel<K, GG>), dim3(grid_r), dim3(MBS), 0, s, n, A.rowptr,   \
                       A.colind, A.val, d_u, d_f, w.t0.p)
            switch (gr) {
                case 4: AXG_RES(4); break;
                case 8: AXG_RES(8); break;
                case 16: AXG_RES(16); break;
                case 32: AXG_RES(32); break;
                default: AXG_RES(64); break;
            }
#undef AXG_RES
            hipLaunchKernelGGL((col_norms_kernel<K>), dim3(grid_v), dim3(MBS), 0, s, n, w.t0.p,
                               w.part.p, w.t1.p, ticket.p);
            AXG_HIP(hipMemcpyAsync(hn, w.t1.p, sizeof(double) * K, hipMemcpyDeviceToHost, s));
            AXG_HIP(hipStreamSynchronize(s));
        };

        // |f_j| from the first residual (u = 0)
        residual();
        double fnorm[K], rnorm[K];
        for (int j = 0; j < K; j++) fnorm[j] = rnorm[j] = hn[j];
        const int inner_batch = nb > 100000 ? 16 : 64;
        const int64_t inner_cap = o.max_iter > 0 ? o.max_iter : std::max<int64_t>(4 * n, 200);
        int64_t total_inner = 0;
        int sweeps = 0;
        res.spmv_ms = 0.0;
        res.spmv_launches = 0;
        hipGraphExec_t gexec = nullptr;
        if (!o.timing) {
            hipGraph_t graph;
            AXG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < inner_batch; i++) {
                launch_spmm(nullptr, nullptr);
                launch_update();
            }
            AXG_HIP(hipStreamEndCapture(s, &graph));
            AXG_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
            AXG_HIP(hipGraphDestroy(graph));
        } else {
            while (ctx->timing_events.size() < (size_t)(2 * inner_batch)) {
                hipEvent_t e;
                AXG_HIP(hipEventCreate(&e));
                ctx->timing_events.push_back(e);
            }
        }
        int status = AXG_OK;
        for (;;) {
            bool conv = true;
            double scale[K];
            for (int j = 0; j < K; j++) {
                bool cj = rnorm[j] <= o.tol * fnorm[j] || fnorm[j] < 1e-30;
                conv &= cj;
                scale[j] = cj ? 0.0 : rnorm[j];
            }
            if (conv) break;
            if (sweeps >= o.max_outer || total_inner >= inner_cap) {
                status = AXG_E_NOT_CONVERGED;
                break;
            }
            sweeps++;
            std::memcpy(hn, scale, sizeof(double) * K);
            AXG_HIP(hipMemcpyAsync(w.t1.p + 8, hn, sizeof(double) * K, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL((pack_kernel<K>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                               n, w.t0.p, m->blk_of_free.p, w.t1.p + 8, w.f0.p);
            // inner scalars
            std::memset(h, 0, sizeof(InnerScalars));
            h->max_iter = (int32_t)std::min<int64_t>(inner_cap - total_inner, INT32_MAX);
            AXG_HIP(hipMemcpyAsync(dsc.p, h, sizeof(InnerScalars), hipMemcpyHostToDevice, s));
            if (block)
                hipLaunchKernelGGL((inner_init_kernel<K, true>), dim3(grid_v), dim3(MBS), 0, s, nb,
                                   w.f0.p, w.f1.p, w.f2.p, Z.p, m->binv32.p, m->dinv32.p, w.part.p,
                                   dsc.p, o.inner_tol);
            else
                hipLaunchKernelGGL((inner_init_kernel<K, false>), dim3(grid_v), dim3(MBS), 0, s,
                                   nb, w.f0.p, w.f1.p, w.f2.p, Z.p, m->binv32.p, m->dinv32.p,
                                   w.part.p, dsc.p, o.inner_tol);
            hipLaunchKernelGGL((inner_update_p_kernel<K>), dim3(grid_v), dim3(MBS), 0, s, nv,
                               w.f3.p, Z.p, dsc.p, 1);
            AXG_HIP(hipGetLastError());
            int32_t prev_it[K];
            for (int j = 0; j < K; j++) prev_it[j] = 0;
            for (;;) {
                AXG_HIP(hipMemcpyAsync(h, dsc.p, sizeof(InnerScalars), hipMemcpyDeviceToHost, s));
                AXG_HIP(hipStreamSynchronize(s));
                if (h->all_done) break;
                if (o.timing) {
                    for (int i = 0; i < inner_batch; i++) {
                        launch_spmm(ctx->timing_events[2 * i], ctx->timing_events[2 * i + 1]);
                        launch_update();
                    }
                    AXG_HIP(hipMemcpyAsync(h, dsc.p, sizeof(InnerScalars), hipMemcpyDeviceToHost,
                                           s));
                    AXG_HIP(hipStreamSynchronize(s));
                    int32_t mx = 0;
                    for (int j = 0; j < K; j++) mx = std::max(mx, h->iters[j] - prev_it[j]);
                    for (int j = 0; j < K; j++) prev_it[j] = h->iters[j];
                    for (int i = 0; i < mx && i < inner_batch; i++) {
                        float ms = 0.f;
                        AXG_HIP(hipEventElapsedTime(&ms, ctx->timing_events[2 * i],
                                                    ctx->timing_events[2 * i + 1]));
                        res.spmv_ms += ms;
                        res.spmv_launches++;
                    }
                } else {
                    AXG_HIP(hipGraphLaunch(gexec, s));
                }
            }
            int32_t mx = 0;
            for (int j = 0; j < K; j++) mx = std::max(mx, h->iters[j]);
            total_inner += mx;
            hipLaunchKernelGGL((unpack_add_kernel<K>), dim3((unsigned)((n + 255) / 256)), dim3(256),
                               0, s, n, w.f1.p, m->blk_of_free.p, w.t1.p + 8, d_u);
            residual();
            for (int j = 0; j < K; j++) rnorm[j] = hn[j];
        }
        AXG_HIP(hipEventRecord(ctx->ev1, s));
        AXG_HIP(hipEventSynchronize(ctx->ev1));
        if (gexec) AXG_HIP(hipGraphExecDestroy(gexec));
        float ms = 0.f;
        AXG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        res.solve_ms = ms;
        res.status = status;
        res.inner_iters = total_inner;
        res.outer_iters = sweeps;
        int worst = 0;
        for (int j = 1; j < K; j++)
            if (rnorm[j] / std::max(fnorm[j], 1e-300) > rnorm[worst] / std::max(fnorm[worst], 1e-300))
                worst = j;
        res.residual = rnorm[worst];
        res.bnorm = fnorm[worst];
    }
};

void solve_mixed(axg_mesh *m, const double *d_rhs, int32_t nrhs, const axg_solve_opts &o,
                 double *d_x, MixedResult &res) {
    AXG_REQUIRE(m->assembled32, AXG_E_INVALID, "mixed operator not assembled");
    const int64_t n = m->n_free;
    // right-hand sides go through in groups of 8, 4, 2, 1
    int32_t k0 = 0;
    MixedResult acc;
    while (k0 < nrhs) {
        int32_t left = nrhs - k0;
        MixedResult r;
        int32_t kk = left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
        const double *f = d_rhs + (int64_t)k0 * n;
        double *u = d_x + (int64_t)k0 * n;
        switch (kk) {
            case 8: MixedSolver<8>::run(m, f, o, u, r); break;
            case 4: MixedSolver<4>::run(m, f, o, u, r); break;
            case 2: MixedSolver<2>::run(m, f, o, u, r); break;
            default: MixedSolver<1>::run(m, f, o, u, r); break;
        }
        if (r.status != AXG_OK && acc.status == AXG_OK) acc.status = r.status;
        acc.inner_iters += r.inner_iters;
        acc.outer_iters = std::max(acc.outer_iters, r.outer_iters);
        if (k0 == 0 || r.residual / std::max(r.bnorm, 1e-300) > acc.residual / std::max(acc.bnorm, 1e-300)) {
            acc.residual = r.residual;
            acc.bnorm = r.bnorm;
        }
        acc.solve_ms += r.solve_ms;
        acc.spmv_ms += r.spmv_ms;
        acc.spmv_launches += r.spmv_launches;
        k0 += kk;
    }
    res = acc;
}

}  // namespace axg

extern "C" int axg_solve_batch(axg_ctx *ctx, int32_t nmesh, axg_mesh *const *meshes,
                               const double *const *rhs, const int32_t *nrhs,
                               const axg_solve_opts *opts, double *const *x, int64_t *iters,
                               double *solve_ms, axg_solve_info *info) {
    (void)ctx, (void)nmesh, (void)meshes, (void)rhs, (void)nrhs, (void)opts, (void)x, (void)iters;
    (void)solve_ms, (void)info;
    axg::set_error("batched solve not built yet");
    return AXG_E_INVALID;
}
