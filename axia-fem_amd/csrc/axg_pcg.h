// Internal PCG interfaces (device-resident vectors).
#pragma once
#include "axg_internal.h"

namespace axg {

struct CsrView {
    int64_t n = 0, nnz = 0;
    const int32_t *rowptr = nullptr;
    const int32_t *colind = nullptr;
    const double *val = nullptr;
};

struct Precond64Args {
    const double *dsel = nullptr;         // Jacobi divisor
    const double *binv = nullptr;         // 3x3 block inverses (null: Jacobi)
    const int32_t *blk_of_free = nullptr;
    const int32_t *free_of_blk = nullptr;
};

struct PcgResult {
    int status = AXG_OK;
    int64_t iters = 0;
    double residual = 0.0, bnorm = 0.0;
    double solve_ms = 0.0, spmv_ms = 0.0;
    int64_t spmv_launches = 0;
};

void spmv_fp64(axg_ctx *ctx, const CsrView &A, const double *x, double *y);
void diag_fp64(axg_ctx *ctx, const CsrView &A, double *dsel, double *diag_raw);
void pcg_fp64(axg_ctx *ctx, const CsrView &A, const Precond64Args &pa, const double *b, double *x,
              double tol, int64_t max_iter, bool timing, PcgResult &res);

}  // namespace axg
