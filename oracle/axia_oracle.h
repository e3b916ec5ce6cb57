/*
 * CPU FP64 oracle for the axia linear-static hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * Rust routines (solver/src/elem.rs, quadratic.rs, linalg.rs, constraint.rs,
 * analysis.rs), written in the reference's operation order and compiled with
 * -ffp-contract=off so that FP64 results follow the Rust bits. Only tests/,
 * the smoke check and the CPU-baseline leg of the benchmark may load it; the
 * GPU product never links or calls it.
 *
 * Pinning (what each group of functions is checked against) is listed in the
 * header comment of axia_oracle.c.
 */
#ifndef AXIA_ORACLE_H
#define AXIA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OR_OK = 0,
    OR_E_INVALID = 1,        /* bad argument / unsupported input            */
    OR_E_JAC_SINGULAR = 2,   /* "Jakobimatrix ist singulär (entartetes ...)"  */
    OR_E_NEG_JACOBIAN = 3,   /* "C3D8: negative Jakobideterminante."          */
    OR_E_MATERIAL = 4,       /* "Ungültige Querkontraktion" / "Ungültiger E"  */
    OR_E_NOT_SPD = 5,        /* "CG: unerwartetes Null-Pivot ..."             */
    OR_E_NOT_CONVERGED = 6,  /* "CG konvergierte nicht ..."                   */
    OR_E_INDEFINITE = 7,     /* "Steifigkeitsmatrix ist indefinit."           */
    OR_E_SINGULAR = 8,       /* "Steifigkeitsmatrix ist singulär ..."         */
    OR_E_NOMEM = 9,
    OR_E_INVERSION = 10      /* "NLGEOM: det(F) <= 0 (Element inversion)."     */
};

/* elem.rs:41-59 */
int or_d_iso_3d(double e, double nu, double d[36]);
/* elem.rs:10-29 */
int or_invert3(const double a[9], double inv[9], double *det);
/* elem.rs:166-178 (N and dN/dxi at one point) */
void or_hex8_shape(double xi, double eta, double zeta, double n[8], double dn[24]);
/* quadratic.rs:36-73 */
void or_hex20_shape(double xi, double eta, double zeta, double n[20], double dn[60]);
/* elem.rs:180-192, quadratic.rs:75-98 (nn = 8 or 20) */
int or_hex_dndx(int nn, const double *xyz, double xi, double eta, double zeta,
                double *dndx, double *det, double *n);
/* elem.rs:199-219 (nn=8) and quadratic.rs:122-146 (nn=20, full 27-point rule) */
int or_hex_stiffness(int nn, const double *xyz, const double d[36], double *ke, double *vol);
/* elem.rs:892-950, quadratic.rs:266-315 ; face 1..6, fe has 3*nn entries */
int or_hex_face_pressure(int nn, const double *xyz, int face, double p, double *fe);
/* elem.rs:743-779, quadratic.rs:148-172,208-244 ; out nn x 6 */
int or_hex_kg(int nn, const double *xyz, const double stress[6], double *kg);
/* nlgeom.rs:127-273 (hex8_nl / hex20_nl, St. Venant-Kirchhoff): tangent ke
 * ((3nn)^2), internal force fe (3nn), volume-averaged Cauchy and
 * Green-Lagrange (6 each), reference volume. */
int or_hex_nl(int nn, const double *xyz0, const double *ue, double e, double nu, double *ke,
              double *fe, double cauchy[6], double gl[6], double *vol);
/* nlgeom.rs:274-308: hex20 nodal Cauchy and Green-Lagrange (20 x 6 each) */
int or_hex20_nodal_nl(const double *xyz0, const double *ue, double e, double nu, double *cs,
                      double *gs);
int or_hex_nodal_stress(int nn, const double *xyz, const double *ue, double e, double nu,
                        double *out);
/* elem.rs:879-889 */
double or_von_mises(const double s[6]);

/* analysis.rs:574-668 restricted to all-Hex8 / all-Hex20 isotropic meshes:
 * per element (in order) Ke, then push (gdof_i, gdof_j, v) for every entry
 * with |v| > 0 in row-major order. ntrip_max must be nelem*(3nn)^2.
 * Returns the number of triplets in *ntrip. Full (unreduced) DOF space. */
int or_assemble_triplets(int nn, int64_t nnode, const double *xyz, int64_t nelem,
                         const int32_t *conn, const double *e, const double *nu,
                         int64_t *ti, int64_t *tj, double *tv, int64_t *ntrip,
                         int64_t *bad_elem);

/* constraint.rs:138-173 for a map with prescribed DOFs only (t = identity on
 * free DOFs). ind_of[d] = free index or -1; u0 = prescribed values (ndof).
 * Output triplets in free index space (capacity ntrip), rhs (n_free). */
int or_reduce(int64_t ndof, const int64_t *ind_of, const double *u0, int64_t n_free,
              int64_t ntrip, const int64_t *ti, const int64_t *tj, const double *tv,
              const double *f_full, int64_t *oi, int64_t *oj, double *ov,
              int64_t *nout, double *rhs);

/* linalg.rs:10-37. Two-call protocol: call with indices==NULL to get nnz in
 * indptr[n]; then call again with buffers of size nnz. */
int or_csr_from_triplets(int64_t n, int64_t ntrip, const int64_t *ti, const int64_t *tj,
                         const double *tv, int64_t *indptr, int64_t *indices, double *data);

/* linalg.rs:41-50 */
void or_csr_matvec(int64_t n, const int64_t *indptr, const int64_t *indices,
                   const double *data, const double *x, double *y);

/* linalg.rs:194-259 (Jacobi PCG). */
int or_pcg(int64_t n, const int64_t *indptr, const int64_t *indices, const double *data,
           const double *b, double tol, int64_t max_iter, double *x, int64_t *iters,
           double *residual);

/* linalg.rs:78-119 (dense lower Cholesky, in place on a, n x n row-major) */
int or_chol_solve(int64_t n, double *a, const double *b, double *x);

#ifdef __cplusplus
}
#endif
#endif
