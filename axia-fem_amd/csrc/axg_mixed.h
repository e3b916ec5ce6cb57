// Mixed precision (FP32 BSR-3 inner PCG inside FP64 iterative refinement),
// node-block structures and batched solves.
#pragma once
#include "axg_internal.h"
#include "axg_pcg.h"

namespace axg {

struct MixedResult {
    int status = AXG_OK;
    int64_t inner_iters = 0;   // summed over right-hand sides and sweeps
    int32_t outer_iters = 0;   // refinement sweeps (max over right-hand sides)
    double residual = 0.0;     // worst final FP64 ||f - K u||
    double bnorm = 0.0;        // ||f|| of that right-hand side
    double solve_ms = 0.0, spmv_ms = 0.0;
    int64_t spmv_launches = 0;
};

CsrView active_csr(axg_mesh *m);
double true_residual(axg_ctx *ctx, const CsrView &A, const double *d_x, const double *d_b);
void ensure_block_maps(axg_mesh *m);
void ensure_block_inverse64(axg_mesh *m);
void assemble_mixed(axg_mesh *m, double *elapsed_ms);
// FP64 block-sliced-ELL values of K_ff (bval64) in the FP32 operator's block
// layout; built on first use after each FP64 assembly.
void ensure_bsell64(axg_mesh *m);
// Number of FP32 SpMM shapes the autotuner chooses from (mesh->spmm_shape).
int spmm_shape_count();
// Whether SpMM shape `shape` exists for K right-hand sides per pass.
bool spmm_shape_ok(int shape, int K);
void solve_mixed(axg_mesh *m, const double *d_rhs, int32_t nrhs, const axg_solve_opts &o,
                 double *d_x, MixedResult &res);

}  // namespace axg
