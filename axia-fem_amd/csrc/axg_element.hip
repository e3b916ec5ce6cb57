// Element kernels for gfx950: C3D8 / C3D20 stiffness integration, the
// deterministic gather into CSR slots, stress recovery and the element
// strain energies used by the compliance adjoint.
//
// FP64 kernels reproduce the reference's operation order (elem.rs:10-219,
// 723-779; quadratic.rs:36-244) and this file is compiled with
// -ffp-contract=off, so Ke entries, K_ff values and nodal stresses carry the
// same FP64 bits as the CPU restatement. The dense B^T D B of gemm_bt_d_b
// (elem.rs:90-112) multiplies many exact zeros; a product with an exact zero
// added to a running sum that started at +0.0 never changes the sum, so the
// kernels evaluate only the nonzero products, in the reference's k order.
#include <cmath>

#include "axg_internal.h"

namespace axg {

// ------------------------------------------------------------ shape tables
// dN/dxi at the integration points, built on the host with the reference
// formulas (the values depend only on the rule, not on the element).
struct ShapeTables {
    double h8_dn[8][8][3];     // stiffness rule, elem.rs:204-207 loop order
    double h8s_dn[8][8][3];    // stress points HEX_XI[i] * G2, elem.rs:751-756
    double h8_ext[8][8];       // N_g(HEX_XI[a] * sqrt 3), elem.rs:766-777
    double h20_dn[27][20][3];  // 27-point rule, quadratic.rs:100-120
    double h20_w[27];          // W_i W_j W_k
    double h20_ext[20][27];    // lx[i] ly[j] lz[k], quadratic.rs:208-230
};
__constant__ ShapeTables c_tab;

static const double G2 = 0.5773502691896257;
static const double G3[3] = {-0.7745966692414834, 0.0, 0.7745966692414834};
static const double W3[3] = {0.5555555555555556, 0.8888888888888888, 0.5555555555555556};
static const double HEX_XI[8][3] = {{-1.0, -1.0, -1.0}, {1.0, -1.0, -1.0}, {1.0, 1.0, -1.0},
                                    {-1.0, 1.0, -1.0},  {-1.0, -1.0, 1.0}, {1.0, -1.0, 1.0},
                                    {1.0, 1.0, 1.0},    {-1.0, 1.0, 1.0}};
static const double HEX20_XI[20][3] = {
    {-1.0, -1.0, -1.0}, {1.0, -1.0, -1.0}, {1.0, 1.0, -1.0}, {-1.0, 1.0, -1.0},
    {-1.0, -1.0, 1.0},  {1.0, -1.0, 1.0},  {1.0, 1.0, 1.0},  {-1.0, 1.0, 1.0},
    {0.0, -1.0, -1.0},  {1.0, 0.0, -1.0},  {0.0, 1.0, -1.0}, {-1.0, 0.0, -1.0},
    {0.0, -1.0, 1.0},   {1.0, 0.0, 1.0},   {0.0, 1.0, 1.0},  {-1.0, 0.0, 1.0},
    {-1.0, -1.0, 0.0},  {1.0, -1.0, 0.0},  {1.0, 1.0, 0.0},  {-1.0, 1.0, 0.0}};

static void hex8_shape_h(double xi, double eta, double zeta, double n[8], double dn[8][3]) {
    for (int i = 0; i < 8; i++) {
        double x = HEX_XI[i][0], e = HEX_XI[i][1], z = HEX_XI[i][2];
        n[i] = 0.125 * (1.0 + x * xi) * (1.0 + e * eta) * (1.0 + z * zeta);
        dn[i][0] = 0.125 * x * (1.0 + e * eta) * (1.0 + z * zeta);
        dn[i][1] = 0.125 * e * (1.0 + x * xi) * (1.0 + z * zeta);
        dn[i][2] = 0.125 * z * (1.0 + x * xi) * (1.0 + e * eta);
    }
}

static void hex20_shape_h(double xi, double eta, double zeta, double dn[20][3]) {
    for (int i = 0; i < 8; i++) {
        double x = HEX20_XI[i][0], e = HEX20_XI[i][1], z = HEX20_XI[i][2];
        double a = 1.0 + x * xi, b = 1.0 + e * eta, c = 1.0 + z * zeta;
        dn[i][0] = 0.125 * x * b * c * (2.0 * x * xi + e * eta + z * zeta - 1.0);
        dn[i][1] = 0.125 * e * a * c * (x * xi + 2.0 * e * eta + z * zeta - 1.0);
        dn[i][2] = 0.125 * z * a * b * (x * xi + e * eta + 2.0 * z * zeta - 1.0);
    }
    for (int i = 8; i < 20; i++) {
        double x = HEX20_XI[i][0], e = HEX20_XI[i][1], z = HEX20_XI[i][2];
        if (std::fabs(x) < 0.5) {
            dn[i][0] = 0.25 * (-2.0 * xi) * (1.0 + e * eta) * (1.0 + z * zeta);
            dn[i][1] = 0.25 * (1.0 - xi * xi) * e * (1.0 + z * zeta);
            dn[i][2] = 0.25 * (1.0 - xi * xi) * (1.0 + e * eta) * z;
        } else if (std::fabs(e) < 0.5) {
            dn[i][0] = 0.25 * (1.0 - eta * eta) * x * (1.0 + z * zeta);
            dn[i][1] = 0.25 * (-2.0 * eta) * (1.0 + x * xi) * (1.0 + z * zeta);
            dn[i][2] = 0.25 * (1.0 - eta * eta) * (1.0 + x * xi) * z;
        } else {
            dn[i][0] = 0.25 * (1.0 - zeta * zeta) * x * (1.0 + e * eta);
            dn[i][1] = 0.25 * (1.0 - zeta * zeta) * (1.0 + x * xi) * e;
            dn[i][2] = 0.25 * (-2.0 * zeta) * (1.0 + x * xi) * (1.0 + e * eta);
        }
    }
}

static void lagrange3(double x, double w[3]) {
    for (int i = 0; i < 3; i++) {
        w[i] = 1.0;
        for (int j = 0; j < 3; j++) {
            if (i == j) continue;
            w[i] *= (x - G3[j]) / (G3[i] - G3[j]);
        }
    }
}

void upload_shape_tables() {
    static ShapeTables t;  // host copy
    double n8[8];
    const double pts[2] = {-G2, G2};
    int g = 0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) hex8_shape_h(pts[a], pts[b], pts[c], n8, t.h8_dn[g++]);
    for (int i = 0; i < 8; i++)
        hex8_shape_h(HEX_XI[i][0] * G2, HEX_XI[i][1] * G2, HEX_XI[i][2] * G2, n8, t.h8s_dn[i]);
    double s3 = std::sqrt(3.0);
    for (int a = 0; a < 8; a++) {
        double dn[8][3];
        hex8_shape_h(HEX_XI[a][0] * s3, HEX_XI[a][1] * s3, HEX_XI[a][2] * s3, t.h8_ext[a], dn);
    }
    g = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                hex20_shape_h(G3[i], G3[j], G3[k], t.h20_dn[g]);
                t.h20_w[g] = W3[i] * W3[j] * W3[k];
                g++;
            }
    for (int a = 0; a < 20; a++) {
        double lx[3], ly[3], lz[3];
        lagrange3(HEX20_XI[a][0], lx);
        lagrange3(HEX20_XI[a][1], ly);
        lagrange3(HEX20_XI[a][2], lz);
        int idx = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) t.h20_ext[a][idx++] = lx[i] * ly[j] * lz[k];
    }
    AXG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tab), &t, sizeof(t)));
    upload_shape_tables32(&t.h8_dn[0][0][0], &t.h20_dn[0][0][0], t.h20_w);
}

// ------------------------------------------------------------- geometry
// Error record: code in err[0] (first error wins), element index in err[1].
__device__ inline void report(int32_t *err, int32_t code, int64_t el) {
    if (atomicCAS(&err[0], 0, code) == 0) err[1] = (int32_t)el;
}

// J, its inverse and dN/dx at one point: elem.rs:10-29, 180-192.
template <int NN>
__device__ inline int geometry(const double (*dn)[3], const double *x, double *dndx,
                               double &det) {
    double j00 = 0.0, j01 = 0.0, j02 = 0.0, j10 = 0.0, j11 = 0.0, j12 = 0.0, j20 = 0.0,
           j21 = 0.0, j22 = 0.0;
    // J[q][p] += dN_a/dxi_p * x_{a,q}, summed over a in order
#pragma unroll 4
    for (int a = 0; a < NN; a++) {
        double d0 = dn[a][0], d1 = dn[a][1], d2 = dn[a][2];
        double x0 = x[3 * a], x1 = x[3 * a + 1], x2 = x[3 * a + 2];
        j00 += d0 * x0;
        j10 += d0 * x1;
        j20 += d0 * x2;
        j01 += d1 * x0;
        j11 += d1 * x1;
        j21 += d1 * x2;
        j02 += d2 * x0;
        j12 += d2 * x1;
        j22 += d2 * x2;
    }
    det = j00 * (j11 * j22 - j12 * j21) - j01 * (j10 * j22 - j12 * j20) +
          j02 * (j10 * j21 - j11 * j20);
    if (fabs(det) < 1e-18) return AXG_E_JAC_SINGULAR;
    double id = 1.0 / det;
    double b00 = (j11 * j22 - j12 * j21) * id;
    double b01 = (j02 * j21 - j01 * j22) * id;
    double b02 = (j01 * j12 - j02 * j11) * id;
    double b10 = (j12 * j20 - j10 * j22) * id;
    double b11 = (j00 * j22 - j02 * j20) * id;
    double b12 = (j02 * j10 - j00 * j12) * id;
    double b20 = (j10 * j21 - j11 * j20) * id;
    double b21 = (j01 * j20 - j00 * j21) * id;
    double b22 = (j00 * j11 - j01 * j10) * id;
    // dN_a/dx_i = inv[0][i] dN/dxi_0 + inv[1][i] dN/dxi_1 + inv[2][i] dN/dxi_2
#pragma unroll 4
    for (int a = 0; a < NN; a++) {
        double d0 = dn[a][0], d1 = dn[a][1], d2 = dn[a][2];
        dndx[3 * a + 0] = b00 * d0 + b10 * d1 + b20 * d2;
        dndx[3 * a + 1] = b01 * d0 + b11 * d1 + b21 * d2;
        dndx[3 * a + 2] = b02 * d0 + b12 * d1 + b22 * d2;
    }
    return AXG_OK;
}

// d_iso_3d, elem.rs:41-59
__device__ inline bool lame(double e, double nu, double &lam, double &mu, double &diag) {
    if (nu >= 0.5 || nu <= -1.0 || e <= 0.0) return false;
    lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    mu = e / (2.0 * (1.0 + nu));
    diag = lam + 2.0 * mu;
    return true;
}

// The 3x3 block of w * B_a^T D B_b for one point, added to k[9] (row-major
// p, q). Nonzero products of gemm_bt_d_b in ascending k order.
__device__ inline void block_btdb(double ax, double ay, double az, double bx, double by, double bz,
                                  double lam, double mu, double diag, double w, double *k) {
    // tmp[k][3b+q] = D[k][q] * dN_b,q (k < 3) or mu * B[k][3b+q] (k >= 3)
    double t0x = diag * bx, t0y = lam * by, t0z = lam * bz;
    double t1x = lam * bx, t1y = diag * by, t1z = lam * bz;
    double t2x = lam * bx, t2y = lam * by, t2z = diag * bz;
    double t3x = mu * by, t3y = mu * bx;  // row xy
    double t4y = mu * bz, t4z = mu * by;  // row yz
    double t5x = mu * bz, t5z = mu * bx;  // row zx
    double s;
    // p = 0: k in {0, 3, 5} with B = (ax, ay, az)
    s = ax * t0x + ay * t3x + az * t5x;
    k[0] += w * s;
    s = ax * t0y + ay * t3y;
    k[1] += w * s;
    s = ax * t0z + az * t5z;
    k[2] += w * s;
    // p = 1: k in {1, 3, 4} with B = (ay, ax, az)
    s = ay * t1x + ax * t3x;
    k[3] += w * s;
    s = ay * t1y + ax * t3y + az * t4y;
    k[4] += w * s;
    s = ay * t1z + az * t4z;
    k[5] += w * s;
    // p = 2: k in {2, 4, 5} with B = (az, ay, ax)
    s = az * t2x + ax * t5x;
    k[6] += w * s;
    s = az * t2y + ay * t4y;
    k[7] += w * s;
    s = az * t2z + ay * t4z + ax * t5z;
    k[8] += w * s;
}

// ------------------------------------------------------- FP64 Ke kernels
// One workgroup integrates EB elements: phase 1 stages dN/dx and the weight of
// every (element, point) in LDS, phase 2 gives each thread (element, a, b)
// node pairs and accumulates the 3x3 block over the points in registers.
template <int NN, int NG, int EB>
__global__ void __launch_bounds__(256) ke_fp64_kernel(const double *__restrict__ xyz,
                                                      const int32_t *__restrict__ conn,
                                                      const double *__restrict__ E,
                                                      const double *__restrict__ NU, int64_t nelem,
                                                      double *__restrict__ ke,
                                                      double *__restrict__ vol,
                                                      int32_t *__restrict__ err) {
    constexpr int M = 3 * NN;
    __shared__ double s_dndx[EB][NG][3 * NN];
    __shared__ double s_w[EB][NG];
    __shared__ double s_x[EB][3 * NN];
    __shared__ double s_mat[EB][3];
    const int64_t e0 = (int64_t)blockIdx.x * EB;
    const int tid = threadIdx.x;
    // stage coordinates
    for (int t = tid; t < EB * NN; t += blockDim.x) {
        int el = t / NN, a = t % NN;
        int64_t e = e0 + el;
        if (e < nelem) {
            int64_t node = conn[e * NN + a];
            s_x[el][3 * a + 0] = xyz[3 * node + 0];
            s_x[el][3 * a + 1] = xyz[3 * node + 1];
            s_x[el][3 * a + 2] = xyz[3 * node + 2];
        }
    }
    if (tid < EB) {
        int64_t e = e0 + tid;
        if (e < nelem) {
            double lam, mu, diag;
            if (!lame(E[e], NU[e], lam, mu, diag)) {
                report(err, AXG_E_MATERIAL, e);
                lam = mu = diag = 0.0;
            }
            s_mat[tid][0] = lam;
            s_mat[tid][1] = mu;
            s_mat[tid][2] = diag;
        }
    }
    __syncthreads();
    for (int t = tid; t < EB * NG; t += blockDim.x) {
        int el = t / NG, g = t % NG;
        int64_t e = e0 + el;
        if (e >= nelem) continue;
        double det;
        const double(*dn)[3] = (NN == 8) ? c_tab.h8_dn[g] : c_tab.h20_dn[g];
        int st = geometry<NN>(dn, s_x[el], s_dndx[el][g], det);
        if (st != AXG_OK) {
            report(err, st, e);
            det = 0.0;
        } else if (det <= 0.0) {
            report(err, AXG_E_NEG_JACOBIAN, e);
        }
        s_w[el][g] = (NN == 8) ? det : c_tab.h20_w[g] * det;
    }
    __syncthreads();
    // element volume: the weights summed over the points in order (the
    // reference's kef.volume, elem.rs:199-219, quadratic.rs:122-146)
    if (tid < EB && e0 + tid < nelem) {
        double v = 0.0;
        for (int g = 0; g < NG; g++) v += s_w[tid][g];
        vol[e0 + tid] = v;
    }
    for (int t = tid; t < EB * NN * NN; t += blockDim.x) {
        int el = t / (NN * NN), ab = t % (NN * NN);
        int a = ab / NN, b = ab % NN;
        int64_t e = e0 + el;
        if (e >= nelem) continue;
        double lam = s_mat[el][0], mu = s_mat[el][1], diag = s_mat[el][2];
        double k[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int g = 0; g < NG; g++) {
            const double *da = &s_dndx[el][g][3 * a];
            const double *db = &s_dndx[el][g][3 * b];
            block_btdb(da[0], da[1], da[2], db[0], db[1], db[2], lam, mu, diag, s_w[el][g], k);
        }
        double *out = ke + e * (int64_t)(M * M);
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int q = 0; q < 3; q++) out[(3 * a + p) * M + 3 * b + q] = k[3 * p + q];
    }
}

void launch_ke_fp64(axg_mesh *m, hipStream_t s, int32_t *d_err) {
    if (m->nelem == 0) return;
    if (m->nn == 8) {
        constexpr int EB = 16;
        dim3 grid((unsigned)((m->nelem + EB - 1) / EB));
        hipLaunchKernelGGL((ke_fp64_kernel<8, 8, EB>), grid, dim3(256), 0, s, m->xyz.p, m->conn.p,
                           m->e.p, m->nu.p, m->nelem, m->ke.p, m->vol.p, d_err);
    } else {
        constexpr int EB = 2;
        dim3 grid((unsigned)((m->nelem + EB - 1) / EB));
        hipLaunchKernelGGL((ke_fp64_kernel<20, 27, EB>), grid, dim3(256), 0, s, m->xyz.p,
                           m->conn.p, m->e.p, m->nu.p, m->nelem, m->ke.p, m->vol.p, d_err);
    }
    AXG_HIP(hipGetLastError());
}

// ------------------------------------------------------------- gather
// Slot s of K_ff = sum of its contributions in ascending element order
// (csr_from_triplets, linalg.rs:18-24). A slot whose contributions are all
// exact zeros would be absent from the reference (|v| > 0 filter,
// analysis.rs:664 and linalg.rs:11); it is flagged for the host.
__global__ void gather_fp64_kernel(const uint32_t *__restrict__ slotptr,
                                   const uint32_t *__restrict__ contrib,
                                   const double *__restrict__ ke, int64_t nslot,
                                   double *__restrict__ val, uint8_t *__restrict__ slot_nz,
                                   uint32_t *__restrict__ zero_count) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nslot;
         s += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t c0 = slotptr[s], c1 = slotptr[s + 1];
        double v = 0.0;
        bool nz = false;
        // up to 8 contributions in flight, summed in their (ascending element) order
        for (uint32_t c = c0; c < c1; c += 8) {
            uint32_t idx[8];
            double kv[8];
#pragma unroll
            for (int u = 0; u < 8; u++) idx[u] = c + u < c1 ? contrib[c + u] : 0u;
#pragma unroll
            for (int u = 0; u < 8; u++) kv[u] = c + u < c1 ? ke[idx[u]] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (c + u < c1) {
                    v += kv[u];
                    nz |= (kv[u] != 0.0);
                }
        }
        val[s] = v;
        if (slot_nz) slot_nz[s] = nz ? 1 : 0;
        if (zero_count && !nz) atomicAdd(zero_count, 1u);
    }
}

void launch_gather_fp64(axg_mesh *m, hipStream_t s, bool topo, uint32_t *d_zero_count) {
    int64_t nslot = topo ? m->nnz_topo : m->nnz;
    if (nslot == 0) return;
    const uint32_t *sp = topo ? m->slotptr_t.p : m->slotptr_c.p;
    const uint32_t *cp = topo ? m->contrib_t.p : m->contrib_c.p;
    int64_t blocks = std::min<int64_t>((nslot + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(gather_fp64_kernel, dim3((unsigned)blocks), dim3(256), 0, s, sp, cp,
                       m->ke.p, nslot, m->val.p, topo ? m->slot_nz.p : nullptr, d_zero_count);
    AXG_HIP(hipGetLastError());
}

// ------------------------------------------------------- stress recovery
// elem.rs:723-779 (C3D8) and quadratic.rs:148-244 (C3D20), one thread per
// element; nonzero products of sigma_from_b in ascending j order.
template <int NN, int NG>
__global__ void nodal_stress_kernel(const double *__restrict__ xyz,
                                    const int32_t *__restrict__ conn, const double *__restrict__ E,
                                    const double *__restrict__ NU, int64_t nelem,
                                    const double *__restrict__ u, double *__restrict__ out,
                                    int32_t *__restrict__ err) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nelem) return;
    double x[3 * NN], ue[3 * NN];
    for (int a = 0; a < NN; a++) {
        int64_t node = conn[e * NN + a];
        for (int q = 0; q < 3; q++) {
            x[3 * a + q] = xyz[3 * node + q];
            ue[3 * a + q] = u[3 * node + q];
        }
    }
    double lam, mu, diag;
    if (!lame(E[e], NU[e], lam, mu, diag)) {
        report(err, AXG_E_MATERIAL, e);
        return;
    }
    double gsig[NG][6];
    double dndx[3 * NN];
    for (int g = 0; g < NG; g++) {
        double det;
        const double(*dn)[3] = (NN == 8) ? c_tab.h8s_dn[g] : c_tab.h20_dn[g];
        int st = geometry<NN>(dn, x, dndx, det);
        if (st != AXG_OK) {
            report(err, st, e);
            return;
        }
        double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0, e5 = 0.0;
        for (int a = 0; a < NN; a++) {
            double dx = dndx[3 * a], dy = dndx[3 * a + 1], dz = dndx[3 * a + 2];
            double u0 = ue[3 * a], u1 = ue[3 * a + 1], u2 = ue[3 * a + 2];
            e0 += dx * u0;
            e1 += dy * u1;
            e2 += dz * u2;
            e3 += dy * u0;
            e3 += dx * u1;
            e4 += dz * u1;
            e4 += dy * u2;
            e5 += dz * u0;
            e5 += dx * u2;
        }
        gsig[g][0] = diag * e0 + lam * e1 + lam * e2;
        gsig[g][1] = lam * e0 + diag * e1 + lam * e2;
        gsig[g][2] = lam * e0 + lam * e1 + diag * e2;
        gsig[g][3] = mu * e3;
        gsig[g][4] = mu * e4;
        gsig[g][5] = mu * e5;
    }
    double *o = out + e * NN * 6;
    for (int a = 0; a < NN; a++)
        for (int c = 0; c < 6; c++) {
            double v = 0.0;
            for (int g = 0; g < NG; g++)
                v += ((NN == 8) ? c_tab.h8_ext[a][g] : c_tab.h20_ext[a][g]) * gsig[g][c];
            o[a * 6 + c] = v;
        }
}

// Plain mean over incident elements in ascending element order
// (analysis.rs:1136-1153), then von Mises (elem.rs:879-889).
__global__ void nodal_average_kernel(const int64_t *__restrict__ inc_ptr,
                                     const int64_t *__restrict__ inc_slot, int64_t nnode, int nn,
                                     const double *__restrict__ en, double *__restrict__ stress,
                                     double *__restrict__ vm) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnode) return;
    double acc[6] = {0, 0, 0, 0, 0, 0}, cnt = 0.0;
    for (int64_t k = inc_ptr[i]; k < inc_ptr[i + 1]; k++) {
        const double *s = en + inc_slot[k] * 6;  // (element * nn + local) * 6
        for (int c = 0; c < 6; c++) acc[c] += s[c];
        cnt += 1.0;
    }
    double st[6];
    for (int c = 0; c < 6; c++) st[c] = cnt > 0.0 ? acc[c] / cnt : 0.0;
    for (int c = 0; c < 6; c++) stress[i * 6 + c] = st[c];
    if (vm) {
        double a = st[0] - st[1], b = st[1] - st[2], c = st[2] - st[0];
        vm[i] = sqrt(0.5 * (a * a + b * b + c * c) +
                     3.0 * (st[3] * st[3] + st[4] * st[4] + st[5] * st[5]));
    }
}

// --------------------------------------------------------- strain energy
// u_e^T K1_e u_e with K1 the stiffness at unit modulus (same nu):
// sum over points of w * eps^T D1 eps.
template <int NN, int NG>
__global__ void strain_energy_kernel(const double *__restrict__ xyz,
                                     const int32_t *__restrict__ conn,
                                     const double *__restrict__ NU, int64_t nelem,
                                     const double *__restrict__ u, double *__restrict__ energy,
                                     int32_t *__restrict__ err) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nelem) return;
    double x[3 * NN], ue[3 * NN];
    for (int a = 0; a < NN; a++) {
        int64_t node = conn[e * NN + a];
        for (int q = 0; q < 3; q++) {
            x[3 * a + q] = xyz[3 * node + q];
            ue[3 * a + q] = u[3 * node + q];
        }
    }
    double lam, mu, diag;
    if (!lame(1.0, NU[e], lam, mu, diag)) {
        report(err, AXG_E_MATERIAL, e);
        return;
    }
    double dndx[3 * NN];
    double en = 0.0;
    for (int g = 0; g < NG; g++) {
        double det;
        const double(*dn)[3] = (NN == 8) ? c_tab.h8_dn[g] : c_tab.h20_dn[g];
        int st = geometry<NN>(dn, x, dndx, det);
        if (st != AXG_OK) {
            report(err, st, e);
            return;
        }
        double w = (NN == 8) ? det : c_tab.h20_w[g] * det;
        double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0, e5 = 0.0;
        for (int a = 0; a < NN; a++) {
            double dx = dndx[3 * a], dy = dndx[3 * a + 1], dz = dndx[3 * a + 2];
            double u0 = ue[3 * a], u1 = ue[3 * a + 1], u2 = ue[3 * a + 2];
            e0 += dx * u0;
            e1 += dy * u1;
            e2 += dz * u2;
            e3 += dy * u0 + dx * u1;
            e4 += dz * u1 + dy * u2;
            e5 += dz * u0 + dx * u2;
        }
        double s0 = diag * e0 + lam * e1 + lam * e2;
        double s1 = lam * e0 + diag * e1 + lam * e2;
        double s2 = lam * e0 + lam * e1 + diag * e2;
        en += w * (e0 * s0 + e1 * s1 + e2 * s2 + mu * (e3 * e3 + e4 * e4 + e5 * e5));
    }
    energy[e] = en;
}

// Rows of the full (unconstrained) stiffness, from the stored FP64 element
// matrices, one thread per DOF r = 3 node + d. Terms come in the reference's
// triplet order (elements ascending, then the element row's columns) and skip
// exact zeros, as the triplet list does (analysis.rs:651-668):
//   ROW_REACT: out[r] = (sum_t v_t w_j) - f[r]             (analysis.rs:1027-1035)
//   ROW_RHS:   out[ind_of[r]] = f[r] - v_t w_j - ... over prescribed j with
//              w_j != 0, free r only (DofMap::reduce, constraint.rs:138-173)
//   ROW_DIAG:  out[r] = sum of the diagonal triplets of row r
//              (autospc_null_diag, analysis.rs:2164-2228)
enum { ROW_REACT = 0, ROW_RHS = 1, ROW_DIAG = 2 };
template <int NN, int MODE>
__global__ void __launch_bounds__(128) row_apply_kernel(
    int64_t ndof, const int64_t *__restrict__ inc_ptr, const int64_t *__restrict__ inc_slot,
    const int32_t *__restrict__ conn, const double *__restrict__ ke,
    const int32_t *__restrict__ ind_of, const double *__restrict__ w,
    const double *__restrict__ f, double *__restrict__ out) {
    constexpr int M = 3 * NN;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ndof) return;
    if (MODE == ROW_RHS && ind_of[r] < 0) return;
    const int64_t node = r / 3;
    const int d = (int)(r - 3 * node);
    double acc = 0.0;
    if (MODE == ROW_RHS && f[r] != 0.0) acc += f[r];
    for (int64_t k = inc_ptr[node]; k < inc_ptr[node + 1]; k++) {
        const int64_t slot = inc_slot[k];
        const int64_t e = slot / NN;
        const int a = (int)(slot - e * NN);
        const double *kr = ke + e * (int64_t)(M * M) + (int64_t)(3 * a + d) * M;
        if (MODE == ROW_DIAG) {
            acc += kr[3 * a + d];
            continue;
        }
        const int32_t *en = conn + e * NN;
        for (int j = 0; j < M; j++) {
            const double v = kr[j];
            if (v == 0.0) continue;
            const int64_t gj = 3 * (int64_t)en[j / 3] + (j % 3);
            if (MODE == ROW_REACT) {
                acc += v * w[gj];
            } else if (ind_of[gj] < 0 && w[gj] != 0.0) {
                acc -= v * w[gj];
            }
        }
    }
    if (MODE == ROW_REACT)
        out[r] = acc - f[r];
    else if (MODE == ROW_RHS)
        out[ind_of[r]] = acc;
    else
        out[r] = acc;
}

void launch_row_apply(axg_mesh *m, hipStream_t s, int mode, const double *d_w,
                      const double *d_f, double *d_out) {
    if (m->ndof == 0) return;
    dim3 grid((unsigned)((m->ndof + 127) / 128));
#define AXG_RA(NNV, MV)                                                                         \
    hipLaunchKernelGGL((row_apply_kernel<NNV, MV>), grid, dim3(128), 0, s, m->ndof,              \
                       m->inc_ptr.p, m->inc_slot.p, m->conn.p, m->ke.p, m->ind_of.p, d_w, d_f, d_out)
    if (m->nn == 8) {
        if (mode == ROW_REACT) AXG_RA(8, ROW_REACT);
        else if (mode == ROW_RHS) AXG_RA(8, ROW_RHS);
        else AXG_RA(8, ROW_DIAG);
    } else {
        if (mode == ROW_REACT) AXG_RA(20, ROW_REACT);
        else if (mode == ROW_RHS) AXG_RA(20, ROW_RHS);
        else AXG_RA(20, ROW_DIAG);
    }
#undef AXG_RA
    AXG_HIP(hipGetLastError());
}

void launch_strain_energy(axg_mesh *m, hipStream_t s, const double *d_u, double *d_energy,
                          int32_t *d_err) {
    if (m->nelem == 0) return;
    dim3 grid((unsigned)((m->nelem + 127) / 128));
    if (m->nn == 8)
        hipLaunchKernelGGL((strain_energy_kernel<8, 8>), grid, dim3(128), 0, s, m->xyz.p,
                           m->conn.p, m->nu.p, m->nelem, d_u, d_energy, d_err);
    else
        hipLaunchKernelGGL((strain_energy_kernel<20, 27>), grid, dim3(128), 0, s, m->xyz.p,
                           m->conn.p, m->nu.p, m->nelem, d_u, d_energy, d_err);
    AXG_HIP(hipGetLastError());
}
// stress_gp of solid elements (analysis.rs:1145-1151): the extrapolated nodal
// stresses of an element, each divided by nn, summed in node order.
__global__ void element_mean_kernel(const double *__restrict__ en, int64_t nelem, int nn,
                                    double *__restrict__ mean) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nelem * 6) return;
    const int64_t e = t / 6;
    const int c = (int)(t % 6);
    const double div = (double)nn;
    double acc = 0.0;
    for (int a = 0; a < nn; a++) acc += en[(e * nn + a) * 6 + c] / div;
    mean[t] = acc;
}

// Geometric stiffness of one element under its constant mean stress
// (eigen.rs:299-360, hex8_kg / hex20_kg with add_continuum_kg). Only the
// direction-diagonal entries Kg[3a+d][3b+d] are nonzero and they are equal for
// d = 0, 1, 2, so the kernel stores the nn x nn scalars
//   kg_ab = sum over Gauss points with det > 0 of (grad N_a . s . grad N_b) w,
// summed in point order; w = det (2x2x2) or W_i W_j W_k det (3x3x3).
template <int NN, int NG, int EB>
__global__ void __launch_bounds__(256) kg_kernel(const double *__restrict__ xyz,
                                                 const int32_t *__restrict__ conn,
                                                 const double *__restrict__ mean, int64_t nelem,
                                                 double *__restrict__ kg,
                                                 int32_t *__restrict__ err) {
    __shared__ double s_dndx[EB][NG][3 * NN];
    __shared__ double s_w[EB][NG];
    __shared__ double s_x[EB][3 * NN];
    __shared__ double s_sig[EB][6];
    const int64_t e0 = (int64_t)blockIdx.x * EB;
    const int tid = threadIdx.x;
    for (int t = tid; t < EB * NN; t += blockDim.x) {
        int el = t / NN, a = t % NN;
        int64_t e = e0 + el;
        if (e < nelem) {
            int64_t node = conn[e * NN + a];
            s_x[el][3 * a + 0] = xyz[3 * node + 0];
            s_x[el][3 * a + 1] = xyz[3 * node + 1];
            s_x[el][3 * a + 2] = xyz[3 * node + 2];
        }
    }
    for (int t = tid; t < EB * 6; t += blockDim.x) {
        int64_t e = e0 + t / 6;
        if (e < nelem) s_sig[t / 6][t % 6] = mean[e * 6 + t % 6];
    }
    __syncthreads();
    for (int t = tid; t < EB * NG; t += blockDim.x) {
        int el = t / NG, g = t % NG;
        int64_t e = e0 + el;
        if (e >= nelem) continue;
        double det;
        const double(*dn)[3] = (NN == 8) ? c_tab.h8_dn[g] : c_tab.h20_dn[g];
        int st = geometry<NN>(dn, s_x[el], s_dndx[el][g], det);
        if (st != AXG_OK) {
            report(err, st, e);
            det = 0.0;
        }
        // a point with det <= 0 adds nothing (its weight is flagged as skipped)
        s_w[el][g] = det <= 0.0 ? -1.0 : ((NN == 8) ? det : c_tab.h20_w[g] * det);
    }
    __syncthreads();
    for (int t = tid; t < EB * NN * NN; t += blockDim.x) {
        int el = t / (NN * NN), ab = t % (NN * NN);
        int a = ab / NN, b = ab % NN;
        int64_t e = e0 + el;
        if (e >= nelem) continue;
        const double sxx = s_sig[el][0], syy = s_sig[el][1], szz = s_sig[el][2];
        const double sxy = s_sig[el][3], syz = s_sig[el][4], szx = s_sig[el][5];
        double k = 0.0;
        for (int g = 0; g < NG; g++) {
            const double w = s_w[el][g];
            if (w < 0.0) continue;
            const double *ga = &s_dndx[el][g][3 * a];
            const double *gb = &s_dndx[el][g][3 * b];
            const double gtg = ga[0] * (sxx * gb[0] + sxy * gb[1] + szx * gb[2]) +
                               ga[1] * (sxy * gb[0] + syy * gb[1] + syz * gb[2]) +
                               ga[2] * (szx * gb[0] + syz * gb[1] + szz * gb[2]);
            k += gtg * w;
        }
        kg[e * (int64_t)(NN * NN) + ab] = k;
    }
}

// Slot s of Kg_ff on the topological pattern: the contributions of the K
// scatter map whose local row and column directions agree, summed in ascending
// element order (the others are exact zeros, which the reference's triplet
// list omits and which change no sum). sign = -1 stores A = -Kg_ff.
__global__ void gather_kg_kernel(const uint32_t *__restrict__ slotptr,
                                 const uint32_t *__restrict__ contrib,
                                 const double *__restrict__ kg, int nn, int64_t nslot,
                                 double sign, double *__restrict__ val) {
    const uint32_t m = 3u * (uint32_t)nn, mm = m * m;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nslot;
         s += (int64_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (uint32_t c = slotptr[s]; c < slotptr[s + 1]; c++) {
            const uint32_t idx = contrib[c];
            const uint32_t el = idx / mm, ij = idx % mm;
            const uint32_t i = ij / m, j = ij % m;
            if (i % 3 != j % 3) continue;
            v += kg[(size_t)el * nn * nn + (i / 3) * nn + j / 3];
        }
        val[s] = sign * v;
    }
}

// ------------------------------------------------------------- NLGEOM
// Total-Lagrange St. Venant-Kirchhoff hex (nlgeom.rs:127-273), one element per
// workgroup: per Gauss point (the stiffness rule, in order) F = I + sum u (x)
// grad N, E = (F^T F - I) / 2, S = lam tr(E) I + 2 mu E, B_nl (fill_b_nl),
// tmp = D B_nl; each thread owns entries (i, j) of the tangent and adds
// w sum_k B[k][i] tmp[k][j] (gemm_bt_d_b) and then, on equal directions, the
// geometric term (grad N_a . S . grad N_b) w (add_continuum_kg); f_e[j] +=
// (sum_i B[i][j] S[i]) w. Cauchy and Green-Lagrange are weighted and summed in
// point order and divided by the volume (GpAcc::finish). All in the
// reference's operation order (this file is built without FMA contraction).
template <int NN, int NG, int NT>
__global__ void __launch_bounds__(NT) nl_kernel(const double *__restrict__ xyz,
                                                const int32_t *__restrict__ conn,
                                                const double *__restrict__ E,
                                                const double *__restrict__ NU, int64_t nelem,
                                                const double *__restrict__ u,
                                                double *__restrict__ ke,
                                                double *__restrict__ fe_out,
                                                double *__restrict__ cau_out,
                                                double *__restrict__ gl_out,
                                                int32_t *__restrict__ err) {
    constexpr int M = 3 * NN;
    constexpr int KPT = (M * M + NT - 1) / NT;
    __shared__ double s_x[3 * NN], s_u[3 * NN];
    __shared__ double s_dndx[NG][3 * NN];
    __shared__ double s_w[NG];
    __shared__ double s_f[NG][9], s_s[NG][6], s_c[NG][6], s_g[NG][6];
    __shared__ double s_b[6][M], s_t[6][M];
    __shared__ double s_mat[3];
    __shared__ int s_bad;
    const int64_t e = blockIdx.x;
    const int tid = threadIdx.x;
    if (e >= nelem) return;
    for (int t = tid; t < NN; t += NT) {
        const int64_t node = conn[e * NN + t];
        for (int q = 0; q < 3; q++) {
            s_x[3 * t + q] = xyz[3 * node + q];
            s_u[3 * t + q] = u[3 * node + q];
        }
    }
    if (tid == 0) {
        s_bad = 0;
        // lam_mu (nlgeom.rs:117-125) and d_iso_3d's diagonal
        const double ev = E[e], nv = NU[e];
        if (ev <= 0.0 || nv >= 0.5 || nv <= -1.0) {
            report(err, AXG_E_MATERIAL, e);
            s_bad = 1;
            s_mat[0] = s_mat[1] = s_mat[2] = 0.0;
        } else {
            const double lam = ev * nv / ((1.0 + nv) * (1.0 - 2.0 * nv));
            const double mu = ev / (2.0 * (1.0 + nv));
            s_mat[0] = lam;
            s_mat[1] = mu;
            s_mat[2] = lam + 2.0 * mu;
        }
    }
    __syncthreads();
    for (int g = tid; g < NG; g += NT) {
        double det;
        const double(*dn)[3] = (NN == 8) ? c_tab.h8_dn[g] : c_tab.h20_dn[g];
        const double *dx = s_dndx[g];
        int st = geometry<NN>(dn, s_x, s_dndx[g], det);
        if (st != AXG_OK) {
            report(err, st, e);
            s_bad = 1;
            continue;
        }
        if (det <= 0.0) {
            report(err, AXG_E_NEG_JACOBIAN, e);
            s_bad = 1;
            continue;
        }
        s_w[g] = (NN == 8) ? det : c_tab.h20_w[g] * det;
        // deformation_gradient (nlgeom.rs:50-64)
        double f[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
        for (int a = 0; a < NN; a++)
            for (int i = 0; i < 3; i++) {
                const double uu = s_u[3 * a + i];
                for (int j = 0; j < 3; j++) f[i][j] += uu * dx[3 * a + j];
            }
        const double jdet = f[0][0] * (f[1][1] * f[2][2] - f[1][2] * f[2][1]) -
                            f[0][1] * (f[1][0] * f[2][2] - f[1][2] * f[2][0]) +
                            f[0][2] * (f[1][0] * f[2][1] - f[1][1] * f[2][0]);
        if (jdet <= 1e-18) {
            report(err, AXG_E_INVERSION, e);
            s_bad = 1;
            continue;
        }
        // green_lagrange (nlgeom.rs:34-48), voigt_s (nlgeom.rs:22-32)
        double eg[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const double c = f[0][i] * f[0][j] + f[1][i] * f[1][j] + f[2][i] * f[2][j];
                eg[i][j] = 0.5 * (c - (i == j ? 1.0 : 0.0));
            }
        const double lam = s_mat[0], mu = s_mat[1];
        const double tr = eg[0][0] + eg[1][1] + eg[2][2];
        double sv[6];
        sv[0] = lam * tr + 2.0 * mu * eg[0][0];
        sv[1] = lam * tr + 2.0 * mu * eg[1][1];
        sv[2] = lam * tr + 2.0 * mu * eg[2][2];
        sv[3] = 2.0 * mu * eg[0][1];
        sv[4] = 2.0 * mu * eg[1][2];
        sv[5] = 2.0 * mu * eg[2][0];
        // pk2_to_cauchy (nlgeom.rs:72-99)
        const double sm[3][3] = {{sv[0], sv[3], sv[5]}, {sv[3], sv[1], sv[4]}, {sv[5], sv[4], sv[2]}};
        double fs[3][3], cc[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) fs[i][j] = f[i][0] * sm[0][j] + f[i][1] * sm[1][j] + f[i][2] * sm[2][j];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) cc[i][j] = fs[i][0] * f[j][0] + fs[i][1] * f[j][1] + fs[i][2] * f[j][2];
        const double invj = 1.0 / jdet;
        s_c[g][0] = cc[0][0] * invj;
        s_c[g][1] = cc[1][1] * invj;
        s_c[g][2] = cc[2][2] * invj;
        s_c[g][3] = cc[0][1] * invj;
        s_c[g][4] = cc[1][2] * invj;
        s_c[g][5] = cc[2][0] * invj;
        s_g[g][0] = eg[0][0];
        s_g[g][1] = eg[1][1];
        s_g[g][2] = eg[2][2];
        s_g[g][3] = 2.0 * eg[0][1];
        s_g[g][4] = 2.0 * eg[1][2];
        s_g[g][5] = 2.0 * eg[2][0];
        for (int q = 0; q < 9; q++) s_f[g][q] = f[q / 3][q % 3];
        for (int q = 0; q < 6; q++) s_s[g][q] = sv[q];
    }
    __syncthreads();
    if (s_bad) return;  // uniform across the workgroup
    double kacc[KPT];
#pragma unroll
    for (int k = 0; k < KPT; k++) kacc[k] = 0.0;
    double facc = 0.0;
    const double lam = s_mat[0], mu = s_mat[1], dg = s_mat[2];
    for (int g = 0; g < NG; g++) {
        // fill_b_nl (nlgeom.rs:101-115)
        for (int t = tid; t < M; t += NT) {
            const int a = t / 3, i = t % 3;
            const double *gg = &s_dndx[g][3 * a];
            const double *f = s_f[g];
            s_b[0][t] = f[3 * i + 0] * gg[0];
            s_b[1][t] = f[3 * i + 1] * gg[1];
            s_b[2][t] = f[3 * i + 2] * gg[2];
            s_b[3][t] = f[3 * i + 0] * gg[1] + f[3 * i + 1] * gg[0];
            s_b[4][t] = f[3 * i + 1] * gg[2] + f[3 * i + 2] * gg[1];
            s_b[5][t] = f[3 * i + 2] * gg[0] + f[3 * i + 0] * gg[2];
        }
        __syncthreads();
        // tmp = D B (gemm_bt_d_b's first product; D's exact zeros skipped, which
        // leaves every sum unchanged)
        for (int t = tid; t < M; t += NT) {
            const double b0 = s_b[0][t], b1 = s_b[1][t], b2 = s_b[2][t];
            s_t[0][t] = dg * b0 + lam * b1 + lam * b2;
            s_t[1][t] = lam * b0 + dg * b1 + lam * b2;
            s_t[2][t] = lam * b0 + lam * b1 + dg * b2;
            s_t[3][t] = mu * s_b[3][t];
            s_t[4][t] = mu * s_b[4][t];
            s_t[5][t] = mu * s_b[5][t];
        }
        __syncthreads();
        const double w = s_w[g];
        const double sxx = s_s[g][0], syy = s_s[g][1], szz = s_s[g][2];
        const double sxy = s_s[g][3], syz = s_s[g][4], szx = s_s[g][5];
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const int t = tid + k * NT;
            if (t >= M * M) break;
            const int i = t / M, j = t % M;
            double sacc = 0.0;
#pragma unroll
            for (int r = 0; r < 6; r++) sacc += s_b[r][i] * s_t[r][j];
            kacc[k] += w * sacc;
            if (i % 3 == j % 3) {
                const double *ga = &s_dndx[g][3 * (i / 3)];
                const double *gb = &s_dndx[g][3 * (j / 3)];
                const double gtg = ga[0] * (sxx * gb[0] + sxy * gb[1] + szx * gb[2]) +
                                   ga[1] * (sxy * gb[0] + syy * gb[1] + syz * gb[2]) +
                                   ga[2] * (szx * gb[0] + syz * gb[1] + szz * gb[2]);
                kacc[k] += gtg * w;
            }
        }
        if (tid < M) {
            double q = 0.0;
#pragma unroll
            for (int r = 0; r < 6; r++) q += s_b[r][tid] * s_s[g][r];
            facc += q * w;
        }
        __syncthreads();
    }
    double *out = ke + e * (int64_t)(M * M);
#pragma unroll
    for (int k = 0; k < KPT; k++) {
        const int t = tid + k * NT;
        if (t < M * M) out[t] = kacc[k];
    }
    if (tid < M) fe_out[e * M + tid] = facc;
    if (tid < 6) {
        // GpAcc: weighted sums in point order, then / volume
        double ca = 0.0, ga = 0.0, vol = 0.0;
        for (int g = 0; g < NG; g++) {
            ca += s_c[g][tid] * s_w[g];
            ga += s_g[g][tid] * s_w[g];
            vol += s_w[g];
        }
        const double inv = fabs(vol) > 0.0 ? 1.0 / vol : 0.0;
        cau_out[e * 6 + tid] = ca * inv;
        gl_out[e * 6 + tid] = ga * inv;
    }
}

// f_int over all DOFs: the element internal forces summed per node over the
// incident elements in ascending element order (the element loop's
// f_int[gdofs[i]] += fe[i]).
__global__ void nl_fint_kernel(const int64_t *__restrict__ inc_ptr,
                               const int64_t *__restrict__ inc_slot, int64_t nnode,
                               const double *__restrict__ fe, double *__restrict__ fint) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnode * 3) return;
    const int64_t i = t / 3;
    const int d = (int)(t % 3);
    double acc = 0.0;
    for (int64_t k = inc_ptr[i]; k < inc_ptr[i + 1]; k++) acc += fe[inc_slot[k] * 3 + d];
    fint[t] = acc;
}

// hex20_nodal_nl (nlgeom.rs:274-308): Cauchy and Green-Lagrange at the 27
// points, extrapolated to the 20 nodes (quadratic.rs:208-230), one thread per
// element; out_c / out_g: nelem x 20 x 6.
__global__ void nl_nodal20_kernel(const double *__restrict__ xyz, const int32_t *__restrict__ conn,
                                  const double *__restrict__ E, const double *__restrict__ NU,
                                  int64_t nelem, const double *__restrict__ u,
                                  double *__restrict__ out_c, double *__restrict__ out_g,
                                  int32_t *__restrict__ err) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nelem) return;
    double x[60], ue[60], dndx[60];
    for (int a = 0; a < 20; a++) {
        const int64_t node = conn[e * 20 + a];
        for (int q = 0; q < 3; q++) {
            x[3 * a + q] = xyz[3 * node + q];
            ue[3 * a + q] = u[3 * node + q];
        }
    }
    const double ev = E[e], nv = NU[e];
    if (ev <= 0.0 || nv >= 0.5 || nv <= -1.0) {
        report(err, AXG_E_MATERIAL, e);
        return;
    }
    const double lam = ev * nv / ((1.0 + nv) * (1.0 - 2.0 * nv));
    const double mu = ev / (2.0 * (1.0 + nv));
    double oc[20][6], og[20][6];
    for (int a = 0; a < 20; a++)
        for (int c = 0; c < 6; c++) oc[a][c] = og[a][c] = 0.0;
    for (int g = 0; g < 27; g++) {
        double det;
        int st = geometry<20>(c_tab.h20_dn[g], x, dndx, det);
        if (st != AXG_OK) {
            report(err, st, e);
            return;
        }
        if (det <= 0.0) {
            report(err, AXG_E_NEG_JACOBIAN, e);
            return;
        }
        double f[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
        for (int a = 0; a < 20; a++)
            for (int i = 0; i < 3; i++) {
                const double uu = ue[3 * a + i];
                for (int j = 0; j < 3; j++) f[i][j] += uu * dndx[3 * a + j];
            }
        const double jdet = f[0][0] * (f[1][1] * f[2][2] - f[1][2] * f[2][1]) -
                            f[0][1] * (f[1][0] * f[2][2] - f[1][2] * f[2][0]) +
                            f[0][2] * (f[1][0] * f[2][1] - f[1][1] * f[2][0]);
        if (jdet <= 1e-18) {
            report(err, AXG_E_INVERSION, e);
            return;
        }
        double eg[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const double c = f[0][i] * f[0][j] + f[1][i] * f[1][j] + f[2][i] * f[2][j];
                eg[i][j] = 0.5 * (c - (i == j ? 1.0 : 0.0));
            }
        const double tr = eg[0][0] + eg[1][1] + eg[2][2];
        const double sv[6] = {lam * tr + 2.0 * mu * eg[0][0], lam * tr + 2.0 * mu * eg[1][1],
                              lam * tr + 2.0 * mu * eg[2][2], 2.0 * mu * eg[0][1],
                              2.0 * mu * eg[1][2],           2.0 * mu * eg[2][0]};
        const double sm[3][3] = {{sv[0], sv[3], sv[5]}, {sv[3], sv[1], sv[4]}, {sv[5], sv[4], sv[2]}};
        double fs[3][3], cc[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) fs[i][j] = f[i][0] * sm[0][j] + f[i][1] * sm[1][j] + f[i][2] * sm[2][j];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) cc[i][j] = fs[i][0] * f[j][0] + fs[i][1] * f[j][1] + fs[i][2] * f[j][2];
        const double invj = 1.0 / jdet;
        const double gc[6] = {cc[0][0] * invj, cc[1][1] * invj, cc[2][2] * invj,
                              cc[0][1] * invj, cc[1][2] * invj, cc[2][0] * invj};
        const double gg[6] = {eg[0][0], eg[1][1], eg[2][2], 2.0 * eg[0][1], 2.0 * eg[1][2],
                              2.0 * eg[2][0]};
        for (int a = 0; a < 20; a++) {
            const double wv = c_tab.h20_ext[a][g];
            for (int c = 0; c < 6; c++) {
                oc[a][c] += wv * gc[c];
                og[a][c] += wv * gg[c];
            }
        }
    }
    for (int a = 0; a < 20; a++)
        for (int c = 0; c < 6; c++) {
            out_c[(e * 20 + a) * 6 + c] = oc[a][c];
            out_g[(e * 20 + a) * 6 + c] = og[a][c];
        }
}

// every node of an element gets the element's value (nelem x 6 -> nelem x nn x 6)
__global__ void broadcast_nodal_kernel(int64_t nelem, int nn, const double *__restrict__ v,
                                       double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nelem * nn * 6) return;
    const int64_t e = t / (nn * 6);
    out[t] = v[e * 6 + t % 6];
}

// Lumped mass of solid elements (analysis.rs:669-717): density x |volume| of each
// element, 1/nn of it on each translational DOF of each of its nodes, summed per
// node over the incident elements in ascending element order.
__global__ void lumped_mass_kernel(const int64_t *__restrict__ inc_ptr,
                                   const int64_t *__restrict__ inc_slot, int64_t nnode, int nn,
                                   const double *__restrict__ vol, const double *__restrict__ rho,
                                   double *__restrict__ m_full) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnode) return;
    const double w = 1.0 / nn;
    double acc = 0.0;
    for (int64_t k = inc_ptr[i]; k < inc_ptr[i + 1]; k++) {
        const int64_t el = inc_slot[k] / nn;
        const double r = rho[el], v = vol[el];
        if (!(fabs(r) > 0.0) || !(fabs(v) > 0.0)) continue;
        acc += r * fabs(v) * w;
    }
    for (int d = 0; d < 3; d++) m_full[i * 3 + d] = acc;
}

}  // namespace axg

namespace axg {
void launch_lumped_mass(axg_mesh *m, hipStream_t s, const double *d_rho, double *d_m_full) {
    if (m->nnode == 0) return;
    dim3 grid((unsigned)((m->nnode + 255) / 256));
    hipLaunchKernelGGL(lumped_mass_kernel, grid, dim3(256), 0, s, m->inc_ptr.p, m->inc_slot.p,
                       m->nnode, m->nn, m->vol.p, d_rho, d_m_full);
    AXG_HIP(hipGetLastError());
}

void launch_element_mean(axg_mesh *m, hipStream_t s, const double *d_elem_nodal,
                         double *d_mean) {
    if (m->nelem == 0) return;
    const int64_t t = m->nelem * 6;
    hipLaunchKernelGGL(element_mean_kernel, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s,
                       d_elem_nodal, m->nelem, m->nn, d_mean);
    AXG_HIP(hipGetLastError());
}

void launch_kg(axg_mesh *m, hipStream_t s, const double *d_mean, double *d_kg, int32_t *d_err) {
    if (m->nelem == 0) return;
    if (m->nn == 8) {
        constexpr int EB = 16;
        dim3 grid((unsigned)((m->nelem + EB - 1) / EB));
        hipLaunchKernelGGL((kg_kernel<8, 8, EB>), grid, dim3(256), 0, s, m->xyz.p, m->conn.p,
                           d_mean, m->nelem, d_kg, d_err);
    } else {
        constexpr int EB = 2;
        dim3 grid((unsigned)((m->nelem + EB - 1) / EB));
        hipLaunchKernelGGL((kg_kernel<20, 27, EB>), grid, dim3(256), 0, s, m->xyz.p, m->conn.p,
                           d_mean, m->nelem, d_kg, d_err);
    }
    AXG_HIP(hipGetLastError());
}

void launch_gather_kg(axg_mesh *m, hipStream_t s, const double *d_kg, double sign,
                      double *d_val) {
    const int64_t nslot = m->nnz_topo;
    if (nslot == 0) return;
    int64_t blocks = std::min<int64_t>((nslot + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(gather_kg_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                       m->slotptr_t.p, m->contrib_t.p, d_kg, m->nn, nslot, sign, d_val);
    AXG_HIP(hipGetLastError());
}

void launch_nl(axg_mesh *m, hipStream_t s, const double *d_u, double *d_fe, double *d_cau,
               double *d_gl, int32_t *d_err) {
    if (m->nelem == 0) return;
    const dim3 grid((unsigned)m->nelem);
    if (m->nn == 8)
        hipLaunchKernelGGL((nl_kernel<8, 8, 64>), grid, dim3(64), 0, s, m->xyz.p, m->conn.p, m->e.p,
                           m->nu.p, m->nelem, d_u, m->ke.p, d_fe, d_cau, d_gl, d_err);
    else
        hipLaunchKernelGGL((nl_kernel<20, 27, 256>), grid, dim3(256), 0, s, m->xyz.p, m->conn.p,
                           m->e.p, m->nu.p, m->nelem, d_u, m->ke.p, d_fe, d_cau, d_gl, d_err);
    AXG_HIP(hipGetLastError());
}

void launch_nl_fint(axg_mesh *m, hipStream_t s, const double *d_fe, double *d_fint) {
    if (m->nnode == 0) return;
    const int64_t t = m->nnode * 3;
    hipLaunchKernelGGL(nl_fint_kernel, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s,
                       m->inc_ptr.p, m->inc_slot.p, m->nnode, d_fe, d_fint);
    AXG_HIP(hipGetLastError());
}

// Nodal Cauchy / Green-Lagrange of the NLGEOM result (analysis.rs:4745-4799):
// C3D20 from its 27 points, C3D8 the element means on every node; en_c, en_g
// receive the per-element nodal values (nelem x nn x 6).
void launch_nl_nodal(axg_mesh *m, hipStream_t s, const double *d_u, const double *d_cau,
                     const double *d_gl, double *en_c, double *en_g, int32_t *d_err) {
    if (m->nelem == 0) return;
    if (m->nn == 20) {
        hipLaunchKernelGGL(nl_nodal20_kernel, dim3((unsigned)((m->nelem + 63) / 64)), dim3(64), 0, s,
                           m->xyz.p, m->conn.p, m->e.p, m->nu.p, m->nelem, d_u, en_c, en_g, d_err);
    } else {
        const int64_t t = m->nelem * m->nn * 6;
        hipLaunchKernelGGL(broadcast_nodal_kernel, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s,
                           m->nelem, m->nn, d_cau, en_c);
        hipLaunchKernelGGL(broadcast_nodal_kernel, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s,
                           m->nelem, m->nn, d_gl, en_g);
    }
    AXG_HIP(hipGetLastError());
}

void launch_nodal_average(axg_mesh *m, hipStream_t s, const double *en, double *d_stress,
                          double *d_vm) {
    if (m->nnode == 0) return;
    hipLaunchKernelGGL(nodal_average_kernel, dim3((unsigned)((m->nnode + 255) / 256)), dim3(256), 0,
                       s, m->inc_ptr.p, m->inc_slot.p, m->nnode, m->nn, en, d_stress, d_vm);
    AXG_HIP(hipGetLastError());
}

void launch_nodal_stress(axg_mesh *m, hipStream_t s, const double *d_u, double *d_elem_nodal,
                         double *d_stress, double *d_vm, int32_t *d_err) {
    if (m->nelem > 0) {
        dim3 grid((unsigned)((m->nelem + 63) / 64));
        if (m->nn == 8)
            hipLaunchKernelGGL((nodal_stress_kernel<8, 8>), grid, dim3(64), 0, s, m->xyz.p,
                               m->conn.p, m->e.p, m->nu.p, m->nelem, d_u, d_elem_nodal, d_err);
        else
            hipLaunchKernelGGL((nodal_stress_kernel<20, 27>), grid, dim3(64), 0, s, m->xyz.p,
                               m->conn.p, m->e.p, m->nu.p, m->nelem, d_u, d_elem_nodal, d_err);
        AXG_HIP(hipGetLastError());
    }
    if (m->nnode > 0) {
        dim3 grid((unsigned)((m->nnode + 255) / 256));
        hipLaunchKernelGGL(nodal_average_kernel, grid, dim3(256), 0, s, m->inc_ptr.p,
                           m->inc_slot.p, m->nnode, m->nn, d_elem_nodal, d_stress, d_vm);
        AXG_HIP(hipGetLastError());
    }
}
}  // namespace axg
