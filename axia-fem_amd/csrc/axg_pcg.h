// Internal PCG interfaces (device-resident vectors).
#pragma once
#include "axg_internal.h"

namespace axg {

struct CsrView {
    int64_t n = 0, nnz = 0;
    const int32_t *rowptr = nullptr;
    const int32_t *colind = nullptr;
    const double *val = nullptr;
    int *shape = nullptr;  // tuned SpMV shape cache (-1: not tuned yet), owned by the caller
    // optional sliced-ELL copy (64 rows per slice); used by the solver when set
    const int64_t *sell_sptr = nullptr;
    const int32_t *sell_width = nullptr;
    const int32_t *sell_col = nullptr;
    const double *sell_val = nullptr;
};

// Sliced ELL: slice s holds rows 64 s .. 64 s + 63; entry j of row r sits at
// sptr[s] + 64 j + (r % 64); rows are padded to the slice's longest row with
// column = the row itself and value 0.
struct SellHost {
    int64_t nslots = 0;
    std::vector<int64_t> sptr;
    std::vector<int32_t> width;
};
void sell_layout(const std::vector<int64_t> &row_len, SellHost &out);
void ensure_sell(axg_mesh *m);
CsrView solver_csr(axg_mesh *m);
CsrView active_csr(axg_mesh *m);

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
    int spmv_shape_L = 0, spmv_shape_U = 0;
    // multi-right-hand-side solves: the first failing column's count and residual
    int64_t fail_iters = -1;
    double fail_residual = 0.0;
};

void spmv_fp64(axg_ctx *ctx, const CsrView &A, const double *x, double *y);
void diag_fp64(axg_ctx *ctx, const CsrView &A, double *dsel, double *diag_raw);
void pcg_fp64(axg_ctx *ctx, const CsrView &A, const Precond64Args &pa, const double *b, double *x,
              double tol, int64_t max_iter, bool timing, PcgResult &res);

// Partials a solve's reductions need (2 per workgroup of the largest grid, + 2).
size_t pcg_part_count(axg_ctx *ctx);
size_t pcg_noalloc_doubles(axg_ctx *ctx, int64_t n);
// linalg::pcg (Jacobi) on a device CSR without allocating: every intermediate in
// scratch (pcg_noalloc_doubles(ctx, n) doubles), plain launches (no graph),
// the CSR SpMV shape from the mean row length. Returns ||b - A x||_2 (FP64).
double pcg_fp64_noalloc(axg_ctx *ctx, const CsrView &A, const double *b, double *x, double tol,
                        int64_t max_iter, double *scratch, PcgResult &res);

// Looks up this device's copy of the no-allocation solve's fixed scratch array
// (axg_noalloc.hip); called once per device by axg_create, device current.
void noalloc_setup(int device);

// The reference's CG error text for a failed solve (linalg.rs:231-257).
std::string pcg_error(const PcgResult &r);

// Multi-right-hand-side FP64 PCG of one assembled mesh on device vectors, in
// node-block space on the FP64 block-sliced-ELL operator, preconditioned by
// Jacobi or (blk3) the 3x3 node-block inverses (the batched solver's vector
// kernels, 8, 4, 2 or 1 right-hand sides per pass, one graph per width,
// reused across calls). Column j of d_b (n_free each) is solved to
// ||r|| <= tol_rel[j] ||b_j||. block_pcg_refresh re-reads the mesh's operator
// and preconditioner after an assembly and sets max_iter (0: max(4 n, 200));
// with timing every SpMM is timed by events.
struct BlockPcg;
BlockPcg *block_pcg_create(axg_mesh *m, int64_t max_iter, bool blk3 = false);
void block_pcg_destroy(BlockPcg *bp);
void block_pcg_refresh(BlockPcg *bp, int64_t max_iter);
void block_pcg_solve(BlockPcg *bp, const double *d_b, int nrhs, const double *tol_rel,
                     double *d_x, PcgResult &res, bool timing = false);
// Average device time of one block-PCG SpMM launch (nrhs = 1, 2, 4 or 8).
void bench_block_spmm(axg_mesh *m, int nrhs, int reps, double *avg_ms);

}  // namespace axg
