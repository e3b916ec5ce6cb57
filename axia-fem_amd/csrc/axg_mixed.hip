// Node-block (3x3) structures: the block-Jacobi inverse in FP64 and, for the
// mixed-precision path, the FP32 BSR-3 operator.
#include <cmath>

#include "axg_internal.h"
#include "axg_mixed.h"

namespace axg {

// Block rows are the nodes that keep at least one free DOF, in node order;
// DOF d of block row b has free index free_of_blk[3b + d] (or -1).
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
// block and zero in the inverse. A block with |det| < 1e-300 falls back to
// Jacobi on its diagonal.
__global__ void block_inverse_kernel(int64_t nb, const int32_t *__restrict__ rowptr,
                                     const int32_t *__restrict__ colind,
                                     const double *__restrict__ val,
                                     const int32_t *__restrict__ free_of_blk,
                                     double *__restrict__ binv) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int32_t f[3] = {free_of_blk[3 * b], free_of_blk[3 * b + 1], free_of_blk[3 * b + 2]};
    double a[3][3];
    for (int d = 0; d < 3; d++)
        for (int e = 0; e < 3; e++) a[d][e] = (d == e) ? 1.0 : 0.0;
    for (int d = 0; d < 3; d++) {
        if (f[d] < 0) continue;
        for (int e = 0; e < 3; e++) a[d][e] = (d == e) ? 0.0 : a[d][e];
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
    for (int d = 0; d < 3; d++)
        for (int e = 0; e < 3; e++)
            binv[9 * b + 3 * d + e] = (f[d] >= 0 && f[e] >= 0) ? inv[d][e] : 0.0;
}

void ensure_block_inverse64(axg_mesh *m) {
    ensure_block_maps(m);
    if (m->binv64.p && m->binv64.n >= (size_t)(9 * m->nb)) return;
    m->binv64.alloc(9 * std::max<int64_t>(m->nb, 1));
    CsrView A = active_csr(m);
    if (m->nb > 0)
        hipLaunchKernelGGL(block_inverse_kernel, dim3((unsigned)((m->nb + 127) / 128)), dim3(128),
                           0, m->ctx->stream, m->nb, A.rowptr, A.colind, A.val, m->free_of_blk.p,
                           m->binv64.p);
    AXG_HIP(hipGetLastError());
    AXG_HIP(hipStreamSynchronize(m->ctx->stream));
}

void assemble_mixed(axg_mesh *m, double *elapsed_ms) {
    (void)m;
    (void)elapsed_ms;
    throw Error{AXG_E_INVALID, "mixed precision path not built yet"};
}

void solve_mixed(axg_mesh *m, const double *d_rhs, int32_t nrhs, const axg_solve_opts &o,
                 double *d_x, MixedResult &res) {
    (void)m, (void)d_rhs, (void)nrhs, (void)o, (void)d_x, (void)res;
    throw Error{AXG_E_INVALID, "mixed precision path not built yet"};
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
