/*
 * axia_gpu — C ABI of the MI355X (gfx950) linear-static hot path of axia-fem.
 *
 * Plain C types only: pointers, sizes, integer status codes. Every call is
 * synchronous with respect to the host buffers it receives. A context owns one
 * HIP stream and the device memory of the meshes created on it; a context is
 * not shared across threads without external locking, separate contexts may
 * run concurrently.
 *
 * The reference boundary these entry points replace (solver/src/...):
 *   linalg::solve_kff            linalg.rs:305-404   -> axg_solve_kff
 *   linalg::csr_from_triplets    linalg.rs:10-37     -> axg_csr_from_triplets
 *   the per-element element_ke + triplet loop + DofMap::reduce +
 *   csr_from_triplets of solve_linear (analysis.rs:574-668, 1018-1020;
 *   elem.rs:199-219; quadratic.rs:122-146; constraint.rs:138-173)
 *                                                    -> axg_mesh_create + axg_assemble
 *   element_nodal_stress + nodal averaging (analysis.rs:1061-1162,
 *   elem.rs:743-779, quadratic.rs:148-244)           -> axg_nodal_stress
 * Error strings mirror the reference's FemError texts (error.rs:3-18).
 */
#ifndef AXIA_GPU_H
#define AXIA_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum axg_status {
    AXG_OK = 0,
    AXG_E_INVALID = 1,        /* bad argument, unsupported mesh, size limit     */
    AXG_E_JAC_SINGULAR = 2,   /* "Jakobimatrix ist singulär (entartetes Element)." elem.rs:14 */
    AXG_E_NEG_JACOBIAN = 3,   /* "C3D8: negative Jakobideterminante." elem.rs:210 */
    AXG_E_MATERIAL = 4,       /* "Ungültige Querkontraktion" / "Ungültiger E-Modul" elem.rs:43-48 */
    AXG_E_NOT_SPD = 5,        /* "CG: unerwartetes Null-Pivot ..." linalg.rs:231-233 */
    AXG_E_NOT_CONVERGED = 6,  /* "CG konvergierte nicht (...)" linalg.rs:253-257 */
    AXG_E_SINGULAR = 7,       /* "Steifigkeitsmatrix ist singulär ..." linalg.rs:89-92 */
    AXG_E_DEVICE = 8,         /* HIP runtime error                              */
    AXG_E_OOM = 9,            /* device allocation failed                       */
    AXG_E_INVERSION = 10      /* "NLGEOM: det(F) ≤ 0 (Element inversion)." nlgeom.rs:93 */
};

enum axg_precision {
    AXG_PREC_FP64 = 0,   /* FP64 integration, FP64 CSR, FP64 PCG (linalg::pcg)      */
    AXG_PREC_MIXED = 1   /* FP32 integration + FP32 BSR-3 PCG inside FP64 refinement */
};

enum axg_precond {
    AXG_PRECOND_JACOBI = 0,  /* z = r / diag (linalg.rs:213-216)                   */
    AXG_PRECOND_BLOCK3 = 1   /* 3x3 node-block Jacobi (inverse formed in FP64)      */
};

/* Mixed precision: how the FP64 outer loop uses the FP32 inner solves. */
enum axg_refinement {
    AXG_REFINE_FCG = 0,      /* flexible CG in FP64, the inner solve as preconditioner */
    AXG_REFINE_PLAIN = 1,    /* classical refinement u += M32(f - K u)                 */
    AXG_REFINE_RELIABLE = 2  /* one FP32 CG whose search direction runs on across the
                                FP64 updates: u += correction, the FP32 residual
                                replaced by f - K u (FP64) each time it has dropped by
                                inner_tol ("reliable updates")                        */
};

typedef struct axg_solve_opts {
    int32_t precision;   /* enum axg_precision                                       */
    int32_t precond;     /* enum axg_precond                                         */
    double tol;          /* stop when ||f - K u||_2 <= tol * ||f||_2 (default 1e-8)  */
    int64_t max_iter;    /* total PCG iterations; <= 0 means max(4n, 200)            */
    double inner_tol;    /* mixed: relative residual drop per inner solve (1e-3)    */
    int32_t max_outer;   /* mixed: refinement sweeps / inner solves (default 30)     */
    int32_t timing;      /* 1: record per-kernel HIP events (SpMV durations in info) */
    int32_t refinement;  /* mixed: enum axg_refinement (default FCG)                 */
    int32_t batch_groups;/* axg_solve_batch: G > 1 groups of meshes by size, solved
                            one after another, smallest first (<= 16); 1 every mesh
                            in one launch sequence; 0 one group per 6 meshes      */
} axg_solve_opts;

typedef struct axg_solve_info {
    int64_t iters;        /* PCG iterations (FP64) or total inner iterations (mixed) */
    int32_t outer_iters;  /* refinement sweeps (mixed), 1 for FP64                  */
    int32_t nrhs;
    double residual;      /* final FP64 ||f - K u||_2 (max over right-hand sides)   */
    double rel_residual;  /* residual / ||f||_2                                     */
    double solve_ms;      /* device time of the solve (HIP events)                  */
    double spmv_ms;       /* summed SpMV kernel time when opts.timing = 1; batched
                             solves: launches that did work x one launch timed on
                             the same operator after the solve                     */
    int64_t spmv_launches;
    char name[64];        /* e.g. "GPU PCG (FP64, Jacobi)" — SparseResult.name      */
} axg_solve_info;

typedef struct axg_ctx axg_ctx;
typedef struct axg_mesh axg_mesh;
typedef struct axg_deck axg_deck;

/* ------------------------------------------------------------- context */
/* Thread-local text of the last error (FemError message), "" if none. */
const char *axg_last_error(void);
const char *axg_version(void);
void axg_default_opts(axg_solve_opts *opts);
int axg_create(int device, axg_ctx **out);
/* Drops the caller's reference; the context (stream, workspace) is freed once
 * every mesh created on it has been destroyed as well, in either order. */
int axg_destroy(axg_ctx *ctx);
/* The context's HIP stream (hipStream_t), for callers that enqueue around it. */
void *axg_stream(axg_ctx *ctx);

/* --------------------------------------------------- solve_kff drop-in */
/* linalg::solve_kff (linalg.rs:311): triplets (i, j, v) in free-DOF space, any
 * order, duplicates summed in the given order, exact zeros dropped
 * (csr_from_triplets rules), then GPU PCG. x_out has n entries. */
int axg_solve_kff(axg_ctx *ctx, int64_t n, int64_t ntrip, const int64_t *ti, const int64_t *tj,
                  const double *tv, const double *rhs, const axg_solve_opts *opts, double *x_out,
                  axg_solve_info *info);

/* linalg::csr_from_triplets (linalg.rs:10-37) on the device. Two-call
 * protocol: with indices == NULL only indptr (n+1) is written and
 * indptr[n] = nnz; then call again with indices/data of nnz entries. */
int axg_csr_from_triplets(axg_ctx *ctx, int64_t n, int64_t ntrip, const int64_t *ti,
                          const int64_t *tj, const double *tv, int64_t *indptr, int64_t *indices,
                          double *data);

/* Solve K x = rhs for a host CSR matrix (int64 indptr, int64 indices). */
int axg_solve_csr(axg_ctx *ctx, int64_t n, const int64_t *indptr, const int64_t *indices,
                  const double *data, const double *rhs, const axg_solve_opts *opts,
                  double *x_out, axg_solve_info *info);

/* No counterpart in the reference. The solve of axg_solve_csr with FP64 +
 * Jacobi (linalg::pcg) for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc):
 * the CSR (int32 row pointer of n+1 and columns, FP64 values), rhs and x are
 * device pointers owned by the caller; every intermediate lives in a fixed
 * library array of 40 MiB per device, set up by axg_create, so n is at most
 * 1048576 (larger n: AXG_E_INVALID). Other precisions or preconditioners:
 * AXG_E_INVALID. Calls that overlap on one device race on that array; the
 * array keeps the last call's contents. Synchronous on the context's stream. */
int axg_solve_csr_device_noalloc(axg_ctx *ctx, int64_t n, const int32_t *d_rowptr,
                                 const int32_t *d_colind, const double *d_val, const double *d_rhs,
                                 const axg_solve_opts *opts, double *d_x, axg_solve_info *info);

/* ------------------------------------------------------------ assembly */
/* An all-C3D8 (nn = 8) or all-C3D20 (nn = 20) isotropic mesh.
 * xyz: nnode x 3; conn: nelem x nn, 0-based node indices in the reference's
 * node order (model.rs:1234-1280 sorts by id); e, nu: per element;
 * ind_of: 3*nnode entries, free index of DOF 3*node+d or -1 if prescribed
 * (DofMap::build, constraint.rs:99-108). Builds the K_ff CSR pattern and the
 * deterministic scatter map on the host (one-time, symbolic). */
int axg_mesh_create(axg_ctx *ctx, int32_t nn, int64_t nnode, const double *xyz, int64_t nelem,
                    const int32_t *conn, const double *e, const double *nu, const int64_t *ind_of,
                    axg_mesh **out);
int axg_mesh_destroy(axg_mesh *mesh);
/* n_free, nnz of K_ff, number of scatter contributions, seconds of the symbolic phase. */
int axg_mesh_info(axg_mesh *mesh, int64_t *n_free, int64_t *nnz, int64_t *ncontrib,
                  double *symbolic_s);
/* Replace per-element moduli (e.g. SIMP E = Emin + rho^3 E0); call axg_assemble again. */
int axg_mesh_set_modulus(axg_mesh *mesh, const double *e);
/* Numeric assembly on the device: Ke of every element and the gather into the
 * CSR slots, summed in ascending element order (csr_from_triplets order).
 * precision AXG_PREC_FP64 builds the FP64 CSR; AXG_PREC_MIXED also builds the
 * FP32 BSR-3 inner operator, integrated in FP32 and assembled block row by
 * block row in one fused kernel. elapsed_ms (may be
 * NULL) receives the device time of the numeric assembly. */
int axg_assemble(axg_mesh *mesh, int32_t precision, double *elapsed_ms);
/* Device times (ms) of the last assembly: FP64 integration, FP64 gather,
 * FP32 fused integration + BSR-3 assembly, and 0 (no separate FP32 gather). */
int axg_assembly_times(axg_mesh *mesh, double *ke64_ms, double *gather64_ms, double *ke32_ms,
                       double *gather32_ms);
/* Copy the FP64 K_ff CSR to the host (indptr n+1, indices nnz, data nnz). */
int axg_csr_export(axg_mesh *mesh, int64_t *indptr, int64_t *indices, double *data);
/* Copy the scatter map: slot_ptr (nnz+1), contrib (ncontrib) where each
 * contribution is el * (3nn)^2 + i * 3nn + j (element, local row, local col). */
int axg_scatter_map_export(axg_mesh *mesh, int64_t *slot_ptr, int64_t *contrib);
/* Copy the FP64 element matrices of elements [first, first+count) (row-major). */
int axg_ke_export(axg_mesh *mesh, int64_t first, int64_t count, double *ke);

/* The FP32 operator of the mixed solve (after axg_assemble(..., AXG_PREC_MIXED))
 * as triplets over free DOFs: every entry of every stored 3x3 block whose row
 * and column DOF are free. With rows, cols and vals NULL, *count receives the
 * number of triplets; otherwise *count gives their capacity on entry and the
 * number written on return. No reference counterpart (a test and inspection
 * export of the inner operator). */
int axg_op32_export(axg_mesh *mesh, int64_t *count, int32_t *rows, int32_t *cols, float *vals);

/* Solve K_ff x = rhs for nrhs right-hand sides (column k at rhs + k*n_free). */
int axg_mesh_solve(axg_mesh *mesh, const double *rhs, int32_t nrhs, const axg_solve_opts *opts,
                   double *x_out, axg_solve_info *info);
/* Same with device pointers (HBM-resident inputs and outputs). */
int axg_mesh_solve_device(axg_mesh *mesh, const double *d_rhs, int32_t nrhs,
                          const axg_solve_opts *opts, double *d_x, axg_solve_info *info);

/* Stress recovery (analysis.rs:1061-1162): per-element nodal stresses
 * (elem_nodal: nelem x nn x 6, may be NULL), their plain mean per node
 * (stress: nnode x 6) and von Mises per node (vm: nnode, may be NULL).
 * u_full: 3*nnode displacements in the full DOF space. */
int axg_nodal_stress(axg_mesh *mesh, const double *u_full, double *stress, double *vm,
                     double *elem_nodal);

/* Reactions rf = K_full u_full - f_full over all 3*nnode DOFs, constrained
 * ones included (analysis.rs:1027-1035), from the FP64 element matrices of
 * the last assembly, summed in the reference's triplet order. */
int axg_reactions(axg_mesh *mesh, const double *u_full, const double *f_full, double *rf);

/* Reduced right-hand side (DofMap::reduce, constraint.rs:138-173):
 * rhs[ind_of[i]] = f_full[i] - sum_j K_ij u0_j over prescribed j with
 * u0_j != 0, for every free DOF i (n_free entries). */
int axg_reduce_rhs(axg_mesh *mesh, const double *f_full, const double *u0, double *rhs);

/* Diagonal of the full stiffness, all 3*nnode DOFs: the sum of the diagonal
 * triplets per DOF in element order (autospc_null_diag, analysis.rs:2164-2228). */
int axg_full_diagonal(axg_mesh *mesh, double *diag);

/* Compliance adjoint (no counterpart in the reference; new path). With
 * K = sum_e E_e K1_e and c = f^T u, the system is self-adjoint, so
 * dc/dE_e = -u_e^T K1_e u_e. Computes c from f_full and u_full and writes
 * dc_de (nelem). u_full, f_full: 3*nnode entries. */
int axg_compliance_adjoint(axg_mesh *mesh, const double *u_full, const double *f_full, double *c,
                           double *dc_de);

/* ------------------------------------------------- eigen (*FREQUENCY) */
/* Element volumes (nelem) of the last FP64 assembly: the sum over the Gauss
 * points of weight x det J, in point order, as the stiffness loop of
 * element_ke accumulates it (elem.rs:199-219, quadratic.rs:122-146). */
int axg_element_volumes(axg_mesh *mesh, double *vol);
/* Lumped mass (analysis.rs:669-717) over the full DOF space (3*nnode):
 * density[e] x |vol_e| / nn on each translational DOF of each node of e,
 * summed per node in element order. density: nelem entries. */
int axg_lumped_mass(axg_mesh *mesh, const double *density, double *m_full);
/* eigen::subspace_gen (eigen.rs:12-111): the lowest nmodes of
 * K_ff phi = lambda M phi with M = diag(m_diag) (n_free entries), by subspace
 * iteration over nmodes + 4 vectors with the solve of opts (GPU PCG) in place
 * of the direct factorisation. values: ascending (nmodes); vectors: nmodes x
 * n_free, M-orthonormal (may be NULL); nfound: number written; sweeps: the
 * subspace sweeps taken (may be NULL); info: summed solver work (may be NULL). */
int axg_eigen(axg_mesh *mesh, const double *m_diag, int32_t nmodes, const axg_solve_opts *opts,
              double *values, double *vectors, int32_t *nfound, int32_t *sweeps,
              axg_solve_info *info);

/* --------------------------------------------------- buckling (*BUCKLE) */
/* Geometric stiffness for the displacements u_full (3*nnode): element mean
 * stresses stress_gp (analysis.rs:1061-1151; mean: nelem x 6, may be NULL)
 * and Kg_ff assembled from hex8_kg / hex20_kg (eigen.rs:299-360) at those
 * stresses, in element order (analysis.rs:1176-1247). kg_val (may be NULL)
 * holds one value per slot of the topological K_ff pattern (the pattern of
 * axg_pattern_host), zero where Kg has no term. */
int axg_geometric_stiffness(axg_mesh *mesh, const double *u_full, double *mean, double *kg_val);
/* eigen::subspace_ab (eigen.rs:113-179) on K_ff phi = lambda (-Kg_ff) phi
 * with Kg_ff from u_full: nmodes + 4 vectors, 30 sweeps of multi-right-hand-
 * side solves (opts) with Euclidean Gram-Schmidt, Rayleigh quotients sorted
 * by |lambda|. values: nmodes; vectors (may be NULL): nmodes x n_free of unit
 * length; nfound: number written. */
int axg_buckle(axg_mesh *mesh, const double *u_full, int32_t nmodes, const axg_solve_opts *opts,
               double *values, double *vectors, int32_t *nfound, axg_solve_info *info);

/* ------------------------------------------------------ NLGEOM (Newton) */
/* The element loop of solve_continuum_newton (analysis.rs:4474-4545) on the
 * device: per element the total-Lagrange St. Venant-Kirchhoff tangent and
 * internal force at u_full (hex8_nl / hex20_nl, nlgeom.rs:127-273), the
 * tangent gathered into the mesh's K_ff like K (exportable with
 * axg_csr_export / axg_ke_export, solvable with axg_mesh_solve), f_int
 * (3*nnode) summed in element order, element means of Cauchy stress and
 * Green-Lagrange strain (nelem x 6). Any output pointer may be NULL. */
int axg_nl_assemble(axg_mesh *mesh, const double *u_full, double *f_int, double *cauchy,
                    double *gl);

/* ------------------------------------------------------- deck ingestion */
/* Parses an INP deck (the keyword subset of inp.rs:383-1061: *NODE, *ELEMENT
 * TYPE=C3D8|C3D20, *NSET and *ELSET (GENERATE, nested names), *MATERIAL, *ELASTIC,
 * *DENSITY, *SOLID SECTION, *BOUNDARY, *CLOAD, *DLOAD Pn, *STEP (NLGEOM, INC=),
 * *STATIC, *FREQUENCY, *BUCKLE) into the compacted model
 * (model.rs:1234-1309) and solve_linear's setup on the host
 * (analysis.rs:548-946): arrays, prescribed values (last BC wins), isolated
 * nodes pinned, load vector, free-DOF numbering (DofMap::build). text need not
 * be NUL-terminated. */
int axg_deck_parse(const char *text, int64_t len, axg_deck **out);
int axg_deck_destroy(axg_deck *deck);
int axg_deck_info(axg_deck *deck, int32_t *nn, int64_t *nnode, int64_t *nelem, int64_t *n_free);
/* Copies of the setup (any pointer may be NULL): node ids (nnode, ascending),
 * element ids (nelem, ascending), xyz (nnode x 3), conn (nelem x nn, 0-based),
 * e, nu (nelem), ind_of, u0, f_full (3*nnode). */
int axg_deck_arrays(axg_deck *deck, int32_t *node_ids, int32_t *elem_ids, double *xyz,
                    int32_t *conn, double *e, double *nu, int64_t *ind_of, double *u0,
                    double *f_full);
/* solve_linear (analysis.rs:548-1162) on the device: assembly, autospc of
 * null diagonals (analysis.rs:2164-2228), reduced rhs, PCG (opts), u_full
 * (3*nnode), reactions rf (3*nnode, may be NULL), nodal stress (nnode x 6)
 * and von Mises (nnode) (both may be NULL), element mean stresses stress_gp
 * (nelem x 6, analysis.rs:1145-1151, may be NULL). */
int axg_deck_solve(axg_ctx *ctx, axg_deck *deck, const axg_solve_opts *opts, double *u_full,
                   double *rf, double *stress, double *vm, double *stress_gp,
                   axg_solve_info *info);

/* The deck's procedure: modes of its *FREQUENCY step (inp.rs:1062-1071),
 * buckling factors of its *BUCKLE step (inp.rs:1072-1081), 0 for a step that
 * is absent; nlgeom = 1 for *STEP, NLGEOM (inp.rs:953-982) and the static
 * step's increments (*STATIC round(period / initial increment), capped by
 * INC=; inp.rs:983-1061). Any pointer may be NULL. */
int axg_deck_procedure(axg_deck *deck, int32_t *nmodes, int32_t *nbuckle, int32_t *nlgeom,
                       int32_t *ninc);
/* The *FREQUENCY branch of solve_linear (analysis.rs:955-1005) without a
 * preload: assembly and autospc as in axg_deck_solve, lumped mass from the
 * materials' *DENSITY, axg_eigen with opts, freqs = sqrt(lambda) / (2 pi)
 * (0 for lambda <= 0). freqs, values (may be NULL): nmodes entries;
 * modes_full (may be NULL): nmodes x 3*nnode mass-normalised mode shapes;
 * nfound: modes written; sweeps, info as for axg_eigen. */
int axg_deck_frequency(axg_ctx *ctx, axg_deck *deck, const axg_solve_opts *opts, double *freqs,
                       double *values, double *modes_full, int32_t *nfound, int32_t *sweeps,
                       axg_solve_info *info);

/* The *BUCKLE branch of solve_linear (analysis.rs:1176-1260): assembly and
 * autospc as in axg_deck_solve, the static solution of the step's loads
 * (u_static, 3*nnode, may be NULL), then axg_buckle. factors: nbuckle
 * entries; modes_full (may be NULL): nbuckle x 3*nnode. */
int axg_deck_buckle(axg_ctx *ctx, axg_deck *deck, const axg_solve_opts *opts, double *factors,
                    double *modes_full, double *u_static, int32_t *nfound, axg_solve_info *info);

/* solve_continuum_newton (analysis.rs:4313-4860) for an all-hex St. Venant-
 * Kirchhoff deck under *STEP, NLGEOM (no contact, pretension, rigid bodies,
 * plasticity or amplitudes): load steps of 1 / ninc with the reference's
 * growth and cutback, full Newton with the tangent and internal force
 * assembled on the device (axg_nl_assemble), follower pressure, the GPU PCG
 * of opts for each correction (FP64: opts->precision other than
 * AXG_PREC_FP64 is AXG_E_INVALID). u_full (3*nnode); rf = f_int - f_ext, nodal
 * Cauchy stress (nnode x 6), von Mises (nnode), nodal Green-Lagrange strain
 * (nnode x 6), element mean Cauchy stress_gp (nelem x 6) may be NULL;
 * ninc_done, newton_iters (the last increment's iterations) may be NULL. */
int axg_deck_nlgeom(axg_ctx *ctx, axg_deck *deck, const axg_solve_opts *opts, double *u_full,
                    double *rf, double *stress, double *vm, double *strain, double *stress_gp,
                    int32_t *ninc_done, int32_t *newton_iters, axg_solve_info *info);

/* ------------------------------------------------------------- writers */
/* The reference's DAT text (dat.rs:3-94) for a deck's results, byte for byte:
 * displacements, forces (|fx|+|fy|+|fz| > 1e-14), element stresses stress_gp
 * (nelem x 6, may be NULL), nodal stresses (nnode x 6, may be NULL) and
 * frequencies (nfreq, may be 0). Two-call protocol: with buf == NULL only
 * *len (bytes, no terminating NUL) is set; then pass a buffer of cap >= *len. */
int axg_deck_write_dat(axg_deck *deck, const double *u_full, const double *rf,
                       const double *stress_gp, const double *nodal_stress, const double *freqs,
                       int32_t nfreq, char *buf, int64_t cap, int64_t *len);
/* The reference's FRD text (frd.rs:173-452, one static frame), byte for byte:
 * heading and header records, nodes used by elements, elements (C3D20 in FRD
 * node order), then DISP / STRESS / FORC / TOSTRAIN as the deck's *NODE FILE
 * and *EL FILE select them. stress, rf, strain may be NULL (zeros; without a
 * strain, TOSTRAIN follows the linear path's rule from the nodal stress).
 * Same two-call protocol. */
int axg_deck_write_frd(axg_deck *deck, const double *u_full, const double *stress,
                       const double *rf, const double *strain, char *buf, int64_t cap,
                       int64_t *len);

/* Batched solves: nmesh meshes (assembled), each with nrhs[m] right-hand
 * sides rhs[m] (n_free[m] * nrhs[m], column-major by right-hand side),
 * solved in one graph-captured launch sequence with per-(mesh, rhs)
 * convergence. x[m] receives the solutions, iters (sum nrhs entries) the
 * iteration count of each solve, solve_ms[] (sum nrhs entries, may be NULL)
 * the device time at which each solve converged. The batched solver is the
 * FP64 Jacobi PCG: opts->precision other than AXG_PREC_FP64 or opts->precond
 * other than AXG_PRECOND_JACOBI is refused with AXG_E_INVALID. */
int axg_solve_batch(axg_ctx *ctx, int32_t nmesh, axg_mesh *const *meshes,
                    const double *const *rhs, const int32_t *nrhs, const axg_solve_opts *opts,
                    double *const *x, int64_t *iters, double *solve_ms, axg_solve_info *info);

/* ------------------------------------------------------ measurement */
/* STREAM-style copy of n doubles, reps launches: achieved GB/s (read+write). */
int axg_bench_stream(axg_ctx *ctx, int64_t n, int32_t reps, double *copy_gbs);
/* Average device time of one launch of a solver operator kernel on an
 * assembled mesh: which = 0 FP64 SpMV (solver layout), 1 FP64 CSR SpMV,
 * 2 FP32 block SpMM with nrhs right-hand sides (1, 2, 4 or 8), 3 FP64
 * sliced-ELL SpMM of the multi-right-hand-side block PCG (the eigen and
 * buckling solves) with nrhs right-hand sides (1, 2, 4 or 8). */
int axg_bench_operator(axg_mesh *mesh, int32_t which, int32_t nrhs, int32_t reps, double *avg_ms);
/* The FP32 block SpMM variant of a mesh's mixed solves with nrhs (1..8)
 * right-hand sides: *shape >= 0 pins it (instead of the autotuner's pick at
 * the first solve), then *shape receives the active one (-1: not tuned yet). */
int axg_mesh_spmm_shape(axg_mesh *mesh, int32_t nrhs, int32_t *shape);
/* The FP32 operator's assembly form of a mesh's mixed solves (the element
 * integration of elem.rs:199-219 / quadratic.rs:122-146 and the scatter of
 * analysis.rs:651-668, in FP32): 0 = integration fused with the block-row
 * assembly (no element matrix stored), 1 = element blocks on the matrix cores
 * (P = W W^T by MFMA, the upper 3x3 blocks of each element stored), then the
 * block rows gathered from them. *mode >= 0 pins it; *mode then receives the
 * active one (the default: 0 for C3D8, 1 for C3D20). Both give the same
 * operator up to FP32 rounding; the next mixed assembly uses it. */
int axg_mesh_asm32_mode(axg_mesh *mesh, int32_t *mode);

/* ------------------------------------------------- host-only helpers */
/* The symbolic phase alone, on the host (no device): K_ff pattern and scatter
 * map as built by axg_mesh_create, before any exact-zero slot is dropped.
 * Two-call protocol: pass NULL arrays to get n_free, nnz, ncontrib; then
 * rowptr (n_free+1), colind (nnz), slot_ptr (nnz+1), contrib (ncontrib). */
int axg_pattern_host(int32_t nn, int64_t nnode, int64_t nelem, const int32_t *conn,
                     const int64_t *ind_of, int64_t *n_free, int64_t *nnz, int64_t *ncontrib,
                     int64_t *rowptr, int64_t *colind, int64_t *slot_ptr, int64_t *contrib);

/* Face-pressure loads (elem.rs:892-950, quadratic.rs:266-315; applied by
 * apply_pressure/scatter_fe, analysis.rs:723-760, 1365-1450, 1546-1553):
 * for each loaded face k, elem_conn[k*nn..] are the element's node indices,
 * face[k] in 1..6 (CalculiX Pn), mag[k] the pressure (positive into the
 * element). Accumulated into f_full (3*nnode) in the given order. */
int axg_pressure_loads(int32_t nn, int64_t nnode, const double *xyz, int64_t nface,
                       const int32_t *elem_conn, const int32_t *face, const double *mag,
                       double *f_full);

#ifdef __cplusplus
}
#endif
#endif /* AXIA_GPU_H */
