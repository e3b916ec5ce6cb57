/*
 * CPU FP64 oracle for the axia linear-static hot path — TEST INFRASTRUCTURE.
 *
 * Plain-C restatement of the reference's Rust code, one function per
 * reference routine, in the reference's operation order. Build with
 * -ffp-contract=off (oracle/Makefile): Rust never fuses a*b+c unless
 * mul_add is written, and none of the cited routines writes it, so the FP64
 * bits of Ke, loads, CSR values and stresses follow the reference.
 *
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load this library. The GPU product never calls it.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   or_d_iso_3d            closed-form Lame constants; material errors.
 *   shape functions        partition of unity (reference test
 *                          quadratic.rs:890-894), Kronecker delta at nodes,
 *                          sum of dN/dxi = 0, finite differences of N.
 *   or_hex_stiffness       symmetry, 6 rigid-body null modes, rank 18 / 54,
 *                          unit-cube hex8 Ke vs the closed-form integral
 *                          (independent sympy-free polynomial integration in
 *                          the test), volume = sum of weights*det, affine
 *                          invariance, one-element patch tests.
 *   or_hex_face_pressure   total force = p * face area * inward normal.
 *   or_hex_nodal_stress    constant-strain exactness; patch tests
 *                          (lib.rs:429-462, 700-774); NAFEMS LE10 sigma_yy(D)
 *                          = -5.51 / -5.50 MPa (examples/nafems/README.md).
 *   or_csr_from_triplets   brute-force dict summation on random triplets
 *                          (same stable order), closed-form structured nnz.
 *   or_pcg / or_chol_solve numpy.linalg.solve on random SPD systems; error
 *                          paths (indefinite, singular, not converged).
 *   or_hex_kg              symmetry, rigid translations in the null space,
 *                          closed form for a uniaxial stress on a box
 *                          (sigma A / L string stiffness), Euler buckling.
 */
#include "axia_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double G2 = 0.5773502691896257; /* elem.rs:8 */
static const double G3[3] = {-0.7745966692414834, 0.0, 0.7745966692414834}; /* quadratic.rs:10 */
static const double W3[3] = {0.5555555555555556, 0.8888888888888888, 0.5555555555555556};

/* elem.rs:147-156 */
static const double HEX_XI[8][3] = {
    {-1.0, -1.0, -1.0}, {1.0, -1.0, -1.0}, {1.0, 1.0, -1.0}, {-1.0, 1.0, -1.0},
    {-1.0, -1.0, 1.0},  {1.0, -1.0, 1.0},  {1.0, 1.0, 1.0},  {-1.0, 1.0, 1.0}};

/* quadratic.rs:13-34 */
static const double HEX20_XI[20][3] = {
    {-1.0, -1.0, -1.0}, {1.0, -1.0, -1.0}, {1.0, 1.0, -1.0}, {-1.0, 1.0, -1.0},
    {-1.0, -1.0, 1.0},  {1.0, -1.0, 1.0},  {1.0, 1.0, 1.0},  {-1.0, 1.0, 1.0},
    {0.0, -1.0, -1.0},  {1.0, 0.0, -1.0},  {0.0, 1.0, -1.0}, {-1.0, 0.0, -1.0},
    {0.0, -1.0, 1.0},   {1.0, 0.0, 1.0},   {0.0, 1.0, 1.0},  {-1.0, 0.0, 1.0},
    {-1.0, -1.0, 0.0},  {1.0, -1.0, 0.0},  {1.0, 1.0, 0.0},  {-1.0, 1.0, 0.0}};

/* elem.rs:250 */
static const double QUAD_XI[4][2] = {{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0}};
/* quadratic.rs:317-326 */
static const double QUAD8_XI[8][2] = {{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0},
                                      {0.0, -1.0},  {1.0, 0.0},  {0.0, 1.0}, {-1.0, 0.0}};
/* elem.rs:894-901 */
static const int HEX8_FACE[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 4, 5, 1},
                                    {1, 5, 6, 2}, {2, 6, 7, 3}, {3, 7, 4, 0}};
/* quadratic.rs:257-264 */
static const int HEX20_FACE[6][8] = {{0, 1, 2, 3, 8, 9, 10, 11},   {4, 7, 6, 5, 15, 14, 13, 12},
                                     {0, 4, 5, 1, 16, 12, 17, 8},  {1, 5, 6, 2, 17, 13, 18, 9},
                                     {2, 6, 7, 3, 18, 14, 19, 10}, {3, 7, 4, 0, 19, 15, 16, 11}};

/* ------------------------------------------------------------------ */
/* elem.rs:10-29 */
int or_invert3(const double a[9], double b[9], double *det_out) {
#define A(i, j) a[(i)*3 + (j)]
    double det = A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) -
                 A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) +
                 A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
    if (fabs(det) < 1e-18) return OR_E_JAC_SINGULAR;
    double inv_det = 1.0 / det;
    b[0] = (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) * inv_det;
    b[1] = (A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2)) * inv_det;
    b[2] = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) * inv_det;
    b[3] = (A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2)) * inv_det;
    b[4] = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) * inv_det;
    b[5] = (A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2)) * inv_det;
    b[6] = (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0)) * inv_det;
    b[7] = (A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1)) * inv_det;
    b[8] = (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) * inv_det;
#undef A
    *det_out = det;
    return OR_OK;
}

/* elem.rs:41-59 */
int or_d_iso_3d(double e, double nu, double d[36]) {
    if (nu >= 0.5 || nu <= -1.0) return OR_E_MATERIAL;
    if (e <= 0.0) return OR_E_MATERIAL;
    double lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = e / (2.0 * (1.0 + nu));
    for (int i = 0; i < 36; i++) d[i] = 0.0;
    double diag = lam + 2.0 * mu;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) d[i * 6 + j] = (i == j) ? diag : lam;
        d[(i + 3) * 6 + (i + 3)] = mu;
    }
    return OR_OK;
}

/* elem.rs:90-112 */
static void gemm_bt_d_b(double *ke, int n, const double *b, int nrow, const double *d, double w,
                        double *tmp) {
    for (int i = 0; i < nrow; i++)
        for (int j = 0; j < n; j++) {
            double s = 0.0;
            for (int k = 0; k < nrow; k++) s += d[i * nrow + k] * b[k * n + j];
            tmp[i * n + j] = s;
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0.0;
            for (int k = 0; k < nrow; k++) s += b[k * n + i] * tmp[k * n + j];
            ke[i * n + j] += w * s;
        }
}

/* elem.rs:114-132 */
static void fill_b3(double *b, int nnode, const double *dndx) {
    int n = 3 * nnode;
    for (int i = 0; i < 6 * n; i++) b[i] = 0.0;
    for (int i = 0; i < nnode; i++) {
        int c = 3 * i;
        double dx = dndx[3 * i], dy = dndx[3 * i + 1], dz = dndx[3 * i + 2];
        b[0 * n + c] = dx;
        b[1 * n + c + 1] = dy;
        b[2 * n + c + 2] = dz;
        b[3 * n + c] = dy;
        b[3 * n + c + 1] = dx;
        b[4 * n + c + 1] = dz;
        b[4 * n + c + 2] = dy;
        b[5 * n + c] = dz;
        b[5 * n + c + 2] = dx;
    }
}

/* elem.rs:158-171 */
void or_hex8_shape(double xi, double eta, double zeta, double n[8], double dn[24]) {
    for (int i = 0; i < 8; i++) {
        double x = HEX_XI[i][0], e = HEX_XI[i][1], z = HEX_XI[i][2];
        n[i] = 0.125 * (1.0 + x * xi) * (1.0 + e * eta) * (1.0 + z * zeta);
        dn[3 * i + 0] = 0.125 * x * (1.0 + e * eta) * (1.0 + z * zeta);
        dn[3 * i + 1] = 0.125 * e * (1.0 + x * xi) * (1.0 + z * zeta);
        dn[3 * i + 2] = 0.125 * z * (1.0 + x * xi) * (1.0 + e * eta);
    }
}

/* quadratic.rs:36-73 */
void or_hex20_shape(double xi, double eta, double zeta, double n[20], double dn[60]) {
    for (int i = 0; i < 8; i++) {
        double x = HEX20_XI[i][0], e = HEX20_XI[i][1], z = HEX20_XI[i][2];
        double a = 1.0 + x * xi, b = 1.0 + e * eta, c = 1.0 + z * zeta;
        n[i] = 0.125 * a * b * c * (x * xi + e * eta + z * zeta - 2.0);
        dn[3 * i + 0] = 0.125 * x * b * c * (2.0 * x * xi + e * eta + z * zeta - 1.0);
        dn[3 * i + 1] = 0.125 * e * a * c * (x * xi + 2.0 * e * eta + z * zeta - 1.0);
        dn[3 * i + 2] = 0.125 * z * a * b * (x * xi + e * eta + 2.0 * z * zeta - 1.0);
    }
    for (int i = 8; i < 20; i++) {
        double x = HEX20_XI[i][0], e = HEX20_XI[i][1], z = HEX20_XI[i][2];
        if (fabs(x) < 0.5) {
            n[i] = 0.25 * (1.0 - xi * xi) * (1.0 + e * eta) * (1.0 + z * zeta);
            dn[3 * i + 0] = 0.25 * (-2.0 * xi) * (1.0 + e * eta) * (1.0 + z * zeta);
            dn[3 * i + 1] = 0.25 * (1.0 - xi * xi) * e * (1.0 + z * zeta);
            dn[3 * i + 2] = 0.25 * (1.0 - xi * xi) * (1.0 + e * eta) * z;
        } else if (fabs(e) < 0.5) {
            n[i] = 0.25 * (1.0 - eta * eta) * (1.0 + x * xi) * (1.0 + z * zeta);
            dn[3 * i + 0] = 0.25 * (1.0 - eta * eta) * x * (1.0 + z * zeta);
            dn[3 * i + 1] = 0.25 * (-2.0 * eta) * (1.0 + x * xi) * (1.0 + z * zeta);
            dn[3 * i + 2] = 0.25 * (1.0 - eta * eta) * (1.0 + x * xi) * z;
        } else {
            n[i] = 0.25 * (1.0 - zeta * zeta) * (1.0 + x * xi) * (1.0 + e * eta);
            dn[3 * i + 0] = 0.25 * (1.0 - zeta * zeta) * x * (1.0 + e * eta);
            dn[3 * i + 1] = 0.25 * (1.0 - zeta * zeta) * (1.0 + x * xi) * e;
            dn[3 * i + 2] = 0.25 * (-2.0 * zeta) * (1.0 + x * xi) * (1.0 + e * eta);
        }
    }
}

/* elem.rs:173-192 and quadratic.rs:75-98 */
int or_hex_dndx(int nn, const double *xyz, double xi, double eta, double zeta, double *dndx,
                double *det, double *n) {
    double dn[60], nloc[20];
    if (nn == 8)
        or_hex8_shape(xi, eta, zeta, nloc, dn);
    else if (nn == 20)
        or_hex20_shape(xi, eta, zeta, nloc, dn);
    else
        return OR_E_INVALID;
    double j[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < nn; a++)
        for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++) j[q * 3 + p] += dn[3 * a + p] * xyz[3 * a + q];
    double inv[9];
    int st = or_invert3(j, inv, det);
    if (st) return st;
    for (int a = 0; a < nn; a++)
        for (int i = 0; i < 3; i++)
            dndx[3 * a + i] =
                inv[0 * 3 + i] * dn[3 * a + 0] + inv[1 * 3 + i] * dn[3 * a + 1] + inv[2 * 3 + i] * dn[3 * a + 2];
    if (n)
        for (int a = 0; a < nn; a++) n[a] = nloc[a];
    return OR_OK;
}

/* elem.rs:199-219 (nn = 8, 2x2x2 Gauss, weight det) and
 * quadratic.rs:100-146 (nn = 20, 3x3x3 Gauss, weight W_i W_j W_k det). */
int or_hex_stiffness(int nn, const double *xyz, const double d[36], double *ke, double *vol_out) {
    int n = 3 * nn;
    double b[6 * 60], tmp[6 * 60], dndx[60], det;
    for (int i = 0; i < n * n; i++) ke[i] = 0.0;
    double vol = 0.0;
    if (nn == 8) {
        const double pts[2] = {-G2, G2};
        for (int a = 0; a < 2; a++)
            for (int c = 0; c < 2; c++)
                for (int e = 0; e < 2; e++) {
                    int st = or_hex_dndx(8, xyz, pts[a], pts[c], pts[e], dndx, &det, NULL);
                    if (st) return st;
                    if (det <= 0.0) return OR_E_NEG_JACOBIAN;
                    fill_b3(b, 8, dndx);
                    gemm_bt_d_b(ke, n, b, 6, d, det, tmp);
                    vol += det;
                }
    } else if (nn == 20) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) {
                    double w = W3[i] * W3[j] * W3[k];
                    int st = or_hex_dndx(20, xyz, G3[i], G3[j], G3[k], dndx, &det, NULL);
                    if (st) return st;
                    if (det <= 0.0) return OR_E_NEG_JACOBIAN;
                    fill_b3(b, 20, dndx);
                    gemm_bt_d_b(ke, n, b, 6, d, w * det, tmp);
                    vol += w * det;
                }
    } else {
        return OR_E_INVALID;
    }
    if (vol_out) *vol_out = vol;
    return OR_OK;
}

/* eigen.rs:299-360: continuum geometric stiffness from one constant stress
 * (Voigt xx yy zz xy yz zx). At each Gauss point with det > 0,
 * Kg[3a+d][3b+d] += (grad N_a . sigma . grad N_b) w for d = 0, 1, 2
 * (add_continuum_kg); w = det (hex8, 2x2x2) or W_i W_j W_k det (hex20, 3x3x3).
 * Points with det <= 0 are skipped. */
static void add_continuum_kg(double *kg, int nnode, const double *dndx, const double s[6],
                             double w) {
    int n = 3 * nnode;
    double sxx = s[0], syy = s[1], szz = s[2], sxy = s[3], syz = s[4], szx = s[5];
    for (int a = 0; a < nnode; a++)
        for (int b = 0; b < nnode; b++) {
            const double *ga = dndx + 3 * a, *gb = dndx + 3 * b;
            double gtg = ga[0] * (sxx * gb[0] + sxy * gb[1] + szx * gb[2]) +
                         ga[1] * (sxy * gb[0] + syy * gb[1] + syz * gb[2]) +
                         ga[2] * (szx * gb[0] + syz * gb[1] + szz * gb[2]);
            double v = gtg * w;
            for (int d = 0; d < 3; d++) kg[(3 * a + d) * n + (3 * b + d)] += v;
        }
}

int or_hex_kg(int nn, const double *xyz, const double stress[6], double *kg) {
    int n = 3 * nn;
    double dndx[60], det;
    for (int i = 0; i < n * n; i++) kg[i] = 0.0;
    if (nn == 8) {
        const double pts[2] = {-G2, G2};
        for (int a = 0; a < 2; a++)
            for (int c = 0; c < 2; c++)
                for (int e = 0; e < 2; e++) {
                    int st = or_hex_dndx(8, xyz, pts[a], pts[c], pts[e], dndx, &det, NULL);
                    if (st) return st;
                    if (det <= 0.0) continue;
                    add_continuum_kg(kg, 8, dndx, stress, det);
                }
    } else if (nn == 20) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) {
                    double w = W3[i] * W3[j] * W3[k];
                    int st = or_hex_dndx(20, xyz, G3[i], G3[j], G3[k], dndx, &det, NULL);
                    if (st) return st;
                    if (det <= 0.0) continue;
                    add_continuum_kg(kg, 20, dndx, stress, w * det);
                }
    } else {
        return OR_E_INVALID;
    }
    return OR_OK;
}

/* quadratic.rs:328-354 */
static void quad8_shape(double xi, double eta, double n[8], double dn[16]) {
    for (int i = 0; i < 4; i++) {
        double x = QUAD8_XI[i][0], e = QUAD8_XI[i][1];
        double a = 1.0 + x * xi, b = 1.0 + e * eta;
        n[i] = 0.25 * a * b * (x * xi + e * eta - 1.0);
        dn[2 * i + 0] = 0.25 * x * b * (2.0 * x * xi + e * eta);
        dn[2 * i + 1] = 0.25 * e * a * (x * xi + 2.0 * e * eta);
    }
    for (int i = 4; i < 8; i++) {
        double x = QUAD8_XI[i][0], e = QUAD8_XI[i][1];
        if (fabs(x) < 0.5) {
            n[i] = 0.5 * (1.0 - xi * xi) * (1.0 + e * eta);
            dn[2 * i + 0] = -xi * (1.0 + e * eta);
            dn[2 * i + 1] = 0.5 * (1.0 - xi * xi) * e;
        } else {
            n[i] = 0.5 * (1.0 - eta * eta) * (1.0 + x * xi);
            dn[2 * i + 0] = 0.5 * (1.0 - eta * eta) * x;
            dn[2 * i + 1] = -eta * (1.0 + x * xi);
        }
    }
}

/* elem.rs:892-950 (nn = 8) and quadratic.rs:266-315 (nn = 20). */
int or_hex_face_pressure(int nn, const double *xyz, int face, double p, double *fe) {
    if (face < 1 || face > 6) return OR_E_INVALID;
    int fi = face - 1;
    for (int i = 0; i < 3 * nn; i++) fe[i] = 0.0;
    if (nn == 8) {
        const double pts[2] = {-G2, G2};
        for (int s = 0; s < 2; s++)
            for (int t = 0; t < 2; t++) {
                double xi = pts[s], eta = pts[t];
                double n[4], dnxi[4], dneta[4];
                for (int i = 0; i < 4; i++) {
                    double x = QUAD_XI[i][0], e = QUAD_XI[i][1];
                    n[i] = 0.25 * (1.0 + x * xi) * (1.0 + e * eta);
                    dnxi[i] = 0.25 * x * (1.0 + e * eta);
                    dneta[i] = 0.25 * e * (1.0 + x * xi);
                }
                double rxi[3] = {0, 0, 0}, reta[3] = {0, 0, 0};
                for (int a = 0; a < 4; a++) {
                    const double *q = xyz + 3 * HEX8_FACE[fi][a];
                    for (int k = 0; k < 3; k++) {
                        rxi[k] += dnxi[a] * q[k];
                        reta[k] += dneta[a] * q[k];
                    }
                }
                double nx = rxi[1] * reta[2] - rxi[2] * reta[1];
                double ny = rxi[2] * reta[0] - rxi[0] * reta[2];
                double nz = rxi[0] * reta[1] - rxi[1] * reta[0];
                double tx = p * nx, ty = p * ny, tz = p * nz;
                for (int a = 0; a < 4; a++) {
                    int gd = 3 * HEX8_FACE[fi][a];
                    fe[gd] += n[a] * tx;
                    fe[gd + 1] += n[a] * ty;
                    fe[gd + 2] += n[a] * tz;
                }
            }
    } else if (nn == 20) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double xi = G3[i], eta = G3[j], w = W3[i] * W3[j];
                double ns[8], dn[16];
                quad8_shape(xi, eta, ns, dn);
                double rxi[3] = {0, 0, 0}, reta[3] = {0, 0, 0};
                for (int a = 0; a < 8; a++) {
                    const double *q = xyz + 3 * HEX20_FACE[fi][a];
                    for (int k = 0; k < 3; k++) {
                        rxi[k] += dn[2 * a] * q[k];
                        reta[k] += dn[2 * a + 1] * q[k];
                    }
                }
                double nx = rxi[1] * reta[2] - rxi[2] * reta[1];
                double ny = rxi[2] * reta[0] - rxi[0] * reta[2];
                double nz = rxi[0] * reta[1] - rxi[1] * reta[0];
                double tx = p * nx * w, ty = p * ny * w, tz = p * nz * w;
                for (int a = 0; a < 8; a++) {
                    int gd = 3 * HEX20_FACE[fi][a];
                    fe[gd] += ns[a] * tx;
                    fe[gd + 1] += ns[a] * ty;
                    fe[gd + 2] += ns[a] * tz;
                }
            }
    } else {
        return OR_E_INVALID;
    }
    return OR_OK;
}

/* elem.rs:723-741 */
static void sigma_from_b(const double *b, int n, const double *d, const double *ue, double sig[6]) {
    double eps[6];
    for (int r = 0; r < 6; r++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += b[r * n + j] * ue[j];
        eps[r] = s;
    }
    for (int r = 0; r < 6; r++) {
        double s = 0.0;
        for (int k = 0; k < 6; k++) s += d[r * 6 + k] * eps[k];
        sig[r] = s;
    }
}

/* quadratic.rs:232-244 */
static void lagrange1d(const double *pts, int n, double x, double *w) {
    for (int i = 0; i < n; i++) {
        w[i] = 1.0;
        for (int j = 0; j < n; j++) {
            if (i == j) continue;
            w[i] *= (x - pts[j]) / (pts[i] - pts[j]);
        }
    }
}

/* quadratic.rs:208-230: 27 Gauss-point values to the 20 nodes by tensor
 * Lagrange interpolation through the Gauss abscissae. */
static void hex20_extrapolate(const double g[27][6], double *out) {
    for (int a = 0; a < 20; a++) {
        double lx[3], ly[3], lz[3];
        lagrange1d(G3, 3, HEX20_XI[a][0], lx);
        lagrange1d(G3, 3, HEX20_XI[a][1], ly);
        lagrange1d(G3, 3, HEX20_XI[a][2], lz);
        for (int c = 0; c < 6; c++) out[a * 6 + c] = 0.0;
        int idx = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) {
                    double w = lx[i] * ly[j] * lz[k];
                    for (int c = 0; c < 6; c++) out[a * 6 + c] += w * g[idx][c];
                    idx++;
                }
    }
}

/* elem.rs:743-779 (nn = 8) and quadratic.rs:148-172 + 208-230 (nn = 20,
 * 27-point rule, tensor Lagrange extrapolation). */
int or_hex_nodal_stress(int nn, const double *xyz, const double *ue, double e, double nu,
                        double *out) {
    double d[36];
    int st = or_d_iso_3d(e, nu, d);
    if (st) return st;
    int n = 3 * nn;
    double b[6 * 60], dndx[60], det;
    if (nn == 8) {
        double gsig[8][6];
        for (int i = 0; i < 8; i++) {
            st = or_hex_dndx(8, xyz, HEX_XI[i][0] * G2, HEX_XI[i][1] * G2, HEX_XI[i][2] * G2, dndx,
                             &det, NULL);
            if (st) return st;
            fill_b3(b, 8, dndx);
            sigma_from_b(b, n, d, ue, gsig[i]);
        }
        double s3 = sqrt(3.0);
        for (int a = 0; a < 8; a++) {
            double nshp[8], dn[24];
            or_hex8_shape(HEX_XI[a][0] * s3, HEX_XI[a][1] * s3, HEX_XI[a][2] * s3, nshp, dn);
            for (int c = 0; c < 6; c++) {
                double v = 0.0;
                for (int g = 0; g < 8; g++) v += nshp[g] * gsig[g][c];
                out[a * 6 + c] = v;
            }
        }
    } else if (nn == 20) {
        double gsig[27][6];
        int idx = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) {
                    st = or_hex_dndx(20, xyz, G3[i], G3[j], G3[k], dndx, &det, NULL);
                    if (st) return st;
                    fill_b3(b, 20, dndx);
                    sigma_from_b(b, n, d, ue, gsig[idx]);
                    idx++;
                }
        hex20_extrapolate(gsig, out);
    } else {
        return OR_E_INVALID;
    }
    return OR_OK;
}

/* elem.rs:879-889 */
double or_von_mises(const double s[6]) {
    double sx = s[0], sy = s[1], sz = s[2], txy = s[3], tyz = s[4], tzx = s[5];
    double a = sx - sy, b = sy - sz, c = sz - sx;
    return sqrt(0.5 * (a * a + b * b + c * c) + 3.0 * (txy * txy + tyz * tyz + tzx * tzx));
}

/* analysis.rs:574-668 for pure Hex8 / Hex20 meshes. */
int or_assemble_triplets(int nn, int64_t nnode, const double *xyz, int64_t nelem,
                         const int32_t *conn, const double *e, const double *nu, int64_t *ti,
                         int64_t *tj, double *tv, int64_t *ntrip, int64_t *bad_elem) {
    if (nn != 8 && nn != 20) return OR_E_INVALID;
    int m = 3 * nn;
    double exyz[60], ke[3600], d[36], vol;
    int64_t gd[60];
    int64_t k = 0;
    for (int64_t el = 0; el < nelem; el++) {
        for (int a = 0; a < nn; a++) {
            int32_t ni = conn[el * nn + a];
            if (ni < 0 || ni >= nnode) {
                if (bad_elem) *bad_elem = el;
                return OR_E_INVALID;
            }
            for (int q = 0; q < 3; q++) exyz[3 * a + q] = xyz[3 * (int64_t)ni + q];
            for (int q = 0; q < 3; q++) gd[3 * a + q] = 3 * (int64_t)ni + q;
        }
        int st = or_d_iso_3d(e[el], nu[el], d);
        if (!st) st = or_hex_stiffness(nn, exyz, d, ke, &vol);
        if (st) {
            if (bad_elem) *bad_elem = el;
            return st;
        }
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
                double v = ke[i * m + j];
                if (fabs(v) > 0.0) {
                    ti[k] = gd[i];
                    tj[k] = gd[j];
                    tv[k] = v;
                    k++;
                }
            }
    }
    *ntrip = k;
    return OR_OK;
}

/* constraint.rs:138-173 with t_row = {(ind_of[d], 1.0)} for free d. */
int or_reduce(int64_t ndof, const int64_t *ind_of, const double *u0, int64_t n_free,
              int64_t ntrip, const int64_t *ti, const int64_t *tj, const double *tv,
              const double *f_full, int64_t *oi, int64_t *oj, double *ov, int64_t *nout,
              double *rhs) {
    for (int64_t a = 0; a < n_free; a++) rhs[a] = 0.0;
    for (int64_t i = 0; i < ndof; i++) {
        double fi = f_full[i];
        if (fabs(fi) == 0.0) continue;
        if (ind_of[i] >= 0) rhs[ind_of[i]] += 1.0 * fi;
    }
    int64_t k = 0;
    for (int64_t t = 0; t < ntrip; t++) {
        int64_t i = ti[t], j = tj[t];
        double v = tv[t];
        if (i >= ndof || j >= ndof || fabs(v) == 0.0) continue;
        double u0j = u0[j];
        if (ind_of[i] >= 0) {
            int64_t a = ind_of[i];
            if (fabs(u0j) > 0.0) rhs[a] -= v * 1.0 * u0j;
            if (ind_of[j] >= 0) {
                oi[k] = a;
                oj[k] = ind_of[j];
                ov[k] = v * 1.0 * 1.0;
                k++;
            }
        }
    }
    *nout = k;
    return OR_OK;
}

/* linalg.rs:10-37. The stable sort by (i, j) is done as two stable counting
 * passes (by j, then by i), which yields exactly the stable-sorted order. */
int or_csr_from_triplets(int64_t n, int64_t ntrip, const int64_t *ti, const int64_t *tj,
                         const double *tv, int64_t *indptr, int64_t *indices, double *data) {
    int64_t *keep = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ntrip > 0 ? ntrip : 1));
    int64_t *tmp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ntrip > 0 ? ntrip : 1));
    int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    if (!keep || !tmp || !cnt) {
        free(keep), free(tmp), free(cnt);
        return OR_E_NOMEM;
    }
    int64_t m = 0;
    for (int64_t t = 0; t < ntrip; t++)
        if (ti[t] >= 0 && ti[t] < n && tj[t] >= 0 && tj[t] < n && fabs(tv[t]) > 0.0) keep[m++] = t;
    /* pass 1: stable by j */
    for (int64_t t = 0; t < m; t++) cnt[tj[keep[t]] + 1]++;
    for (int64_t c = 0; c < n; c++) cnt[c + 1] += cnt[c];
    for (int64_t t = 0; t < m; t++) tmp[cnt[tj[keep[t]]]++] = keep[t];
    /* pass 2: stable by i */
    memset(cnt, 0, sizeof(int64_t) * ((size_t)n + 1));
    for (int64_t t = 0; t < m; t++) cnt[ti[tmp[t]] + 1]++;
    for (int64_t c = 0; c < n; c++) cnt[c + 1] += cnt[c];
    for (int64_t t = 0; t < m; t++) keep[cnt[ti[tmp[t]]]++] = tmp[t];
    /* left-to-right duplicate sums */
    for (int64_t i = 0; i <= n; i++) indptr[i] = 0;
    int64_t k = 0, nnz = 0;
    while (k < m) {
        int64_t i = ti[keep[k]], j = tj[keep[k]];
        double v = 0.0;
        while (k < m && ti[keep[k]] == i && tj[keep[k]] == j) {
            v += tv[keep[k]];
            k++;
        }
        if (indices) {
            indices[nnz] = j;
            data[nnz] = v;
        }
        nnz++;
        indptr[i + 1] += 1;
    }
    for (int64_t i = 0; i < n; i++) indptr[i + 1] += indptr[i];
    free(keep), free(tmp), free(cnt);
    return OR_OK;
}

/* linalg.rs:41-50 */
void or_csr_matvec(int64_t n, const int64_t *indptr, const int64_t *indices, const double *data,
                   const double *x, double *y) {
    for (int64_t i = 0; i < n; i++) {
        double s = 0.0;
        for (int64_t k = indptr[i]; k < indptr[i + 1]; k++) s += data[k] * x[indices[k]];
        y[i] = s;
    }
}

/* linalg.rs:52-63: first (i,i) entry found, else 0. */
static void csr_diag(int64_t n, const int64_t *indptr, const int64_t *indices, const double *data,
                     double *d) {
    for (int64_t i = 0; i < n; i++) {
        d[i] = 0.0;
        for (int64_t k = indptr[i]; k < indptr[i + 1]; k++)
            if (indices[k] == i) {
                d[i] = data[k];
                break;
            }
    }
}

/* linalg.rs:194-259 */
int or_pcg(int64_t n, const int64_t *indptr, const int64_t *indices, const double *data,
           const double *b, double tol, int64_t max_iter, double *x, int64_t *iters_out,
           double *residual_out) {
    double *r = (double *)malloc(sizeof(double) * (size_t)n * 5 + 8);
    if (!r) return OR_E_NOMEM;
    double *z = r + n, *p = r + 2 * n, *ap = r + 3 * n, *diag = r + 4 * n;
    for (int64_t i = 0; i < n; i++) {
        x[i] = 0.0;
        r[i] = b[i];
    }
    double bnorm = 0.0;
    for (int64_t i = 0; i < n; i++) bnorm += b[i] * b[i];
    bnorm = sqrt(bnorm);
    if (bnorm < 1e-30) {
        *iters_out = 0;
        *residual_out = 0.0;
        free(r);
        return OR_OK;
    }
    csr_diag(n, indptr, indices, data, diag);
    for (int64_t i = 0; i < n; i++) z[i] = r[i] / (fabs(diag[i]) > 1e-30 ? diag[i] : 1.0);
    for (int64_t i = 0; i < n; i++) p[i] = z[i];
    double rz = -0.0;
    for (int64_t i = 0; i < n; i++) rz += r[i] * z[i];
    int64_t iters = 0;
    double residual = bnorm;
    for (int64_t it = 0; it < max_iter; it++) {
        iters = it + 1;
        or_csr_matvec(n, indptr, indices, data, p, ap);
        double pap = -0.0;
        for (int64_t i = 0; i < n; i++) pap += p[i] * ap[i];
        if (fabs(pap) < 1e-30) {
            free(r);
            *iters_out = iters;
            *residual_out = residual;
            return OR_E_NOT_SPD;
        }
        double alpha = rz / pap;
        for (int64_t i = 0; i < n; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * ap[i];
        }
        double rr = -0.0;
        for (int64_t i = 0; i < n; i++) rr += r[i] * r[i];
        residual = sqrt(rr);
        if (residual <= tol * bnorm) break;
        for (int64_t i = 0; i < n; i++) z[i] = r[i] / (fabs(diag[i]) > 1e-30 ? diag[i] : 1.0);
        double rz_new = -0.0;
        for (int64_t i = 0; i < n; i++) rz_new += r[i] * z[i];
        double beta = rz_new / rz;
        for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        rz = rz_new;
    }
    free(r);
    *iters_out = iters;
    *residual_out = residual;
    if (residual > tol * bnorm * 10.0) return OR_E_NOT_CONVERGED;
    return OR_OK;
}

/* linalg.rs:78-119 */
int or_chol_solve(int64_t n, double *a, const double *b, double *x) {
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j <= i; j++) {
            double s = a[i * n + j];
            for (int64_t k = 0; k < j; k++) s -= a[i * n + k] * a[j * n + k];
            if (i == j) {
                if (s < 0.0) return OR_E_INDEFINITE;
                if (s <= 1e-30) return OR_E_SINGULAR;
                a[i * n + i] = sqrt(s);
            } else {
                a[i * n + j] = s / a[j * n + j];
            }
        }
    double *y = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!y) return OR_E_NOMEM;
    for (int64_t i = 0; i < n; i++) {
        double s = b[i];
        for (int64_t k = 0; k < i; k++) s -= a[i * n + k] * y[k];
        y[i] = s / a[i * n + i];
    }
    for (int64_t i = n - 1; i >= 0; i--) {
        double s = y[i];
        for (int64_t k = i + 1; k < n; k++) s -= a[k * n + i] * x[k];
        x[i] = s / a[i * n + i];
    }
    free(y);
    return OR_OK;
}


/* ------------------------------------------------------------- NLGEOM
 * nlgeom.rs:22-127, 197-273: total-Lagrange St. Venant-Kirchhoff for hex8
 * (2x2x2, weight det) and hex20 (3x3x3, weight W_i W_j W_k det). */
static int lam_mu(double e, double nu, double *lam, double *mu) {
    if (e <= 0.0 || nu >= 0.5 || nu <= -1.0) return OR_E_MATERIAL;
    *lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    *mu = e / (2.0 * (1.0 + nu));
    return OR_OK;
}

/* nlgeom.rs:50-64 */
static void deformation_gradient(const double *dndx, const double *ue, int nn, double f[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) f[i][j] = (i == j) ? 1.0 : 0.0;
    for (int a = 0; a < nn; a++)
        for (int i = 0; i < 3; i++) {
            double u = ue[3 * a + i];
            for (int j = 0; j < 3; j++) f[i][j] += u * dndx[3 * a + j];
        }
}

/* nlgeom.rs:66-70 */
static double det3(const double a[3][3]) {
    return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
           a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}

/* nlgeom.rs:34-48 */
static void green_lagrange(const double f[3][3], double e[3][3]) {
    double c[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c[i][j] = f[0][i] * f[0][j] + f[1][i] * f[1][j] + f[2][i] * f[2][j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) e[i][j] = 0.5 * (c[i][j] - (i == j ? 1.0 : 0.0));
}

/* nlgeom.rs:22-32 */
static void voigt_s(const double e[3][3], double lam, double mu, double s[6]) {
    double tr = e[0][0] + e[1][1] + e[2][2];
    s[0] = lam * tr + 2.0 * mu * e[0][0];
    s[1] = lam * tr + 2.0 * mu * e[1][1];
    s[2] = lam * tr + 2.0 * mu * e[2][2];
    s[3] = 2.0 * mu * e[0][1];
    s[4] = 2.0 * mu * e[1][2];
    s[5] = 2.0 * mu * e[2][0];
}

/* nlgeom.rs:72-99 */
static int pk2_to_cauchy(const double f[3][3], const double s[6], double out[6]) {
    double sm[3][3] = {{s[0], s[3], s[5]}, {s[3], s[1], s[4]}, {s[5], s[4], s[2]}};
    double fs[3][3], c[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) fs[i][j] = f[i][0] * sm[0][j] + f[i][1] * sm[1][j] + f[i][2] * sm[2][j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c[i][j] = fs[i][0] * f[j][0] + fs[i][1] * f[j][1] + fs[i][2] * f[j][2];
    double j = det3(f);
    if (j <= 1e-18) return OR_E_INVERSION;
    double invj = 1.0 / j;
    out[0] = c[0][0] * invj;
    out[1] = c[1][1] * invj;
    out[2] = c[2][2] * invj;
    out[3] = c[0][1] * invj;
    out[4] = c[1][2] * invj;
    out[5] = c[2][0] * invj;
    return OR_OK;
}

/* nlgeom.rs:101-115 */
static void fill_b_nl(double *b, int nn, const double f[3][3], const double *dndx) {
    int n = 3 * nn;
    for (int a = 0; a < nn; a++) {
        const double *g = dndx + 3 * a;
        for (int i = 0; i < 3; i++) {
            int c = 3 * a + i;
            b[0 * n + c] = f[i][0] * g[0];
            b[1 * n + c] = f[i][1] * g[1];
            b[2 * n + c] = f[i][2] * g[2];
            b[3 * n + c] = f[i][0] * g[1] + f[i][1] * g[0];
            b[4 * n + c] = f[i][1] * g[2] + f[i][2] * g[1];
            b[5 * n + c] = f[i][2] * g[0] + f[i][0] * g[2];
        }
    }
}

/* nlgeom.rs:197-238 (assemble_gp) for an isotropic elastic material */
static int assemble_gp_nl(double *ke, double *fe, double *vol, double cau[6], double gl[6],
                          int nn, const double *dndx, const double *ue, double lam, double mu,
                          const double *cmat, double w) {
    int nd = 3 * nn;
    double f[3][3], egl[3][3], s[6], cs[6];
    deformation_gradient(dndx, ue, nn, f);
    if (det3(f) <= 1e-18) return OR_E_INVERSION;
    green_lagrange(f, egl);
    voigt_s(egl, lam, mu, s);
    double b[6 * 60], tmp[6 * 60];
    fill_b_nl(b, nn, f, dndx);
    gemm_bt_d_b(ke, nd, b, 6, cmat, w, tmp);
    add_continuum_kg(ke, nn, dndx, s, w);
    for (int j = 0; j < nd; j++) {
        double q = 0.0;
        for (int i = 0; i < 6; i++) q += b[i * nd + j] * s[i];
        fe[j] += q * w;
    }
    int st = pk2_to_cauchy(f, s, cs);
    if (st) return st;
    double glv[6] = {egl[0][0], egl[1][1], egl[2][2], 2.0 * egl[0][1], 2.0 * egl[1][2],
                     2.0 * egl[2][0]};
    for (int i = 0; i < 6; i++) {
        cau[i] += cs[i] * w;
        gl[i] += glv[i] * w;
    }
    *vol += w;
    return OR_OK;
}

/* nlgeom.rs:127-273: tangent ke ((3nn)^2), internal force fe (3nn), volume-
 * averaged Cauchy and Green-Lagrange (6 each) of one hex at displacements ue. */
int or_hex_nl(int nn, const double *xyz0, const double *ue, double e, double nu, double *ke,
              double *fe, double cauchy[6], double gl[6], double *vol_out) {
    double lam, mu, cmat[36];
    int st = lam_mu(e, nu, &lam, &mu);
    if (st) return st;
    st = or_d_iso_3d(e, nu, cmat);
    if (st) return st;
    int nd = 3 * nn;
    for (int i = 0; i < nd * nd; i++) ke[i] = 0.0;
    for (int i = 0; i < nd; i++) fe[i] = 0.0;
    double cau[6] = {0, 0, 0, 0, 0, 0}, g[6] = {0, 0, 0, 0, 0, 0}, vol = 0.0, dndx[60], det;
    if (nn == 8) {
        const double pts[2] = {-G2, G2};
        for (int a = 0; a < 2; a++)
            for (int c = 0; c < 2; c++)
                for (int k = 0; k < 2; k++) {
                    st = or_hex_dndx(8, xyz0, pts[a], pts[c], pts[k], dndx, &det, NULL);
                    if (st) return st;
                    if (det <= 0.0) return OR_E_NEG_JACOBIAN;
                    st = assemble_gp_nl(ke, fe, &vol, cau, g, 8, dndx, ue, lam, mu, cmat, det);
                    if (st) return st;
                }
    } else if (nn == 20) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) {
                    double w = W3[i] * W3[j] * W3[k];
                    st = or_hex_dndx(20, xyz0, G3[i], G3[j], G3[k], dndx, &det, NULL);
                    if (st) return st;
                    if (det <= 0.0) return OR_E_NEG_JACOBIAN;
                    st = assemble_gp_nl(ke, fe, &vol, cau, g, 20, dndx, ue, lam, mu, cmat, w * det);
                    if (st) return st;
                }
    } else {
        return OR_E_INVALID;
    }
    /* GpAcc::finish (nlgeom.rs:168-188) */
    double inv = fabs(vol) > 0.0 ? 1.0 / vol : 0.0;
    for (int i = 0; i < 6; i++) {
        cauchy[i] = cau[i] * inv;
        gl[i] = g[i] * inv;
    }
    if (vol_out) *vol_out = vol;
    return OR_OK;
}

/* nlgeom.rs:274-308: Cauchy and Green-Lagrange at the 27 points of a hex20,
 * extrapolated to its 20 nodes (20 x 6 each). */
int or_hex20_nodal_nl(const double *xyz0, const double *ue, double e, double nu, double *cs,
                      double *gs) {
    double lam, mu, dndx[60], det, f[3][3], egl[3][3], s[6];
    int st = lam_mu(e, nu, &lam, &mu);
    if (st) return st;
    double gc[27][6], ge[27][6];
    int idx = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                st = or_hex_dndx(20, xyz0, G3[i], G3[j], G3[k], dndx, &det, NULL);
                if (st) return st;
                if (det <= 0.0) return OR_E_NEG_JACOBIAN;
                deformation_gradient(dndx, ue, 20, f);
                if (det3(f) <= 1e-18) return OR_E_INVERSION;
                green_lagrange(f, egl);
                voigt_s(egl, lam, mu, s);
                st = pk2_to_cauchy(f, s, gc[idx]);
                if (st) return st;
                ge[idx][0] = egl[0][0];
                ge[idx][1] = egl[1][1];
                ge[idx][2] = egl[2][2];
                ge[idx][3] = 2.0 * egl[0][1];
                ge[idx][4] = 2.0 * egl[1][2];
                ge[idx][5] = 2.0 * egl[2][0];
                idx++;
            }
    hex20_extrapolate(gc, cs);
    hex20_extrapolate(ge, gs);
    return OR_OK;
}
