// FP32 element integration for the mixed-precision operator, and the gather
// of FP32 element matrices into the FP32 BSR-3 blocks.
//
// Ke = sum_g w_g B_g^T D B_g is evaluated as one dense contraction per
// element, Ke = Bs^T Cs, where Bs stacks the strain-displacement matrices of
// all integration points (6 NG x 3 NN) and Cs = blockdiag(w_g D) Bs. That
// product runs on the matrix cores: v_mfma_f32_32x32x2_f32, one wavefront per
// element, 32x32 output tiles (1 tile for C3D8, 2x2 tiles for C3D20), K
// walked two strain rows at a time. Lane l supplies A = Bs^T[row][k] and
// B = Cs[k][col] with row/col = tile offset + l % 32 and k = 2 t + l / 32.
#include <cmath>

#include "axg_internal.h"

namespace axg {

struct ShapeTables32 {
    float h8_dn[8][8][3];
    float h20_dn[27][20][3];
    float h20_w[27];
};
__constant__ ShapeTables32 c_tab32;

void upload_shape_tables32(const double *h8_dn, const double *h20_dn, const double *h20_w) {
    static ShapeTables32 t;
    for (int i = 0; i < 8 * 8 * 3; i++) (&t.h8_dn[0][0][0])[i] = (float)h8_dn[i];
    for (int i = 0; i < 27 * 20 * 3; i++) (&t.h20_dn[0][0][0])[i] = (float)h20_dn[i];
    for (int i = 0; i < 27; i++) t.h20_w[i] = (float)h20_w[i];
    AXG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tab32), &t, sizeof(t)));
}

__device__ inline void report32(int32_t *err, int32_t code, int64_t el) {
    if (atomicCAS(&err[0], 0, code) == 0) err[1] = (int32_t)el;
}

// dN/dx and the weight at one point in FP32 (geometry in FP64 for robustness
// of the inverse on distorted elements, then rounded).
template <int NN>
__device__ inline bool geometry32(const float (*dn)[3], const float *x, float *dndx, float &wdet,
                                  float wgp) {
    double j[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < NN; a++)
        for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++) j[q][p] += (double)dn[a][p] * (double)x[3 * a + q];
    double det = j[0][0] * (j[1][1] * j[2][2] - j[1][2] * j[2][1]) -
                 j[0][1] * (j[1][0] * j[2][2] - j[1][2] * j[2][0]) +
                 j[0][2] * (j[1][0] * j[2][1] - j[1][1] * j[2][0]);
    if (!(det > 0.0)) return false;
    double id = 1.0 / det;
    double b[3][3];
    b[0][0] = (j[1][1] * j[2][2] - j[1][2] * j[2][1]) * id;
    b[0][1] = (j[0][2] * j[2][1] - j[0][1] * j[2][2]) * id;
    b[0][2] = (j[0][1] * j[1][2] - j[0][2] * j[1][1]) * id;
    b[1][0] = (j[1][2] * j[2][0] - j[1][0] * j[2][2]) * id;
    b[1][1] = (j[0][0] * j[2][2] - j[0][2] * j[2][0]) * id;
    b[1][2] = (j[0][2] * j[1][0] - j[0][0] * j[1][2]) * id;
    b[2][0] = (j[1][0] * j[2][1] - j[1][1] * j[2][0]) * id;
    b[2][1] = (j[0][1] * j[2][0] - j[0][0] * j[2][1]) * id;
    b[2][2] = (j[0][0] * j[1][1] - j[0][1] * j[1][0]) * id;
    for (int a = 0; a < NN; a++)
        for (int i = 0; i < 3; i++)
            dndx[3 * a + i] = (float)(b[0][i] * dn[a][0] + b[1][i] * dn[a][1] + b[2][i] * dn[a][2]);
    wdet = (float)(wgp * det);
    return true;
}

// Strain row k (0..5, Voigt xx yy zz xy yz zx) of B for DOF c = 3a + p.
__device__ inline float bval(const float *dndx, int k, int c) {
    int a = c / 3, p = c - 3 * a;
    float dx = dndx[3 * a], dy = dndx[3 * a + 1], dz = dndx[3 * a + 2];
    switch (k) {
        case 0: return p == 0 ? dx : 0.f;
        case 1: return p == 1 ? dy : 0.f;
        case 2: return p == 2 ? dz : 0.f;
        case 3: return p == 0 ? dy : (p == 1 ? dx : 0.f);
        case 4: return p == 1 ? dz : (p == 2 ? dy : 0.f);
        default: return p == 0 ? dz : (p == 2 ? dx : 0.f);
    }
}

// (D B)[k][c] for isotropic D.
__device__ inline float dbval(const float *dndx, int k, int c, float lam, float mu, float diag) {
    if (k < 3) {
        float s = 0.f;
        for (int m = 0; m < 3; m++) s += ((m == k) ? diag : lam) * bval(dndx, m, c);
        return s;
    }
    return mu * bval(dndx, k, c);
}

typedef float floatx16 __attribute__((ext_vector_type(16)));

// One wavefront per element (grid-stride over elements), 4 wavefronts per
// workgroup. LDS per wavefront: coordinates, dN/dx and w of every point.
template <int NN, int NG>
__global__ void __launch_bounds__(256) ke_fp32_mfma_kernel(const double *__restrict__ xyz,
                                                           const int32_t *__restrict__ conn,
                                                           const double *__restrict__ E,
                                                           const double *__restrict__ NU,
                                                           int64_t nelem, float *__restrict__ ke,
                                                           int32_t *__restrict__ err) {
    constexpr int M = 3 * NN;          // 24 or 60
    constexpr int T = (M + 31) / 32;   // output tiles per dimension
    constexpr int KTOT = 6 * NG;       // 48 or 162
    __shared__ float s_x[4][3 * NN];
    __shared__ float s_dndx[4][NG][3 * NN];
    __shared__ float s_w[4][NG];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    const int h = lane >> 5, c32 = lane & 31;
    // the loop bound is uniform over the workgroup, so every wavefront reaches
    // every __syncthreads; wavefronts past the last element idle
    for (int64_t base = (int64_t)blockIdx.x * 4; base < nelem; base += nwaves) {
        const int64_t e = base + wave;
        const bool valid = e < nelem;
        if (valid)
            for (int t = lane; t < 3 * NN; t += 64) {
                int a = t / 3, q = t - 3 * a;
                s_x[wave][t] = (float)xyz[3 * (int64_t)conn[e * NN + a] + q];
            }
        __syncthreads();
        if (valid && lane < NG) {
            const float(*dn)[3] = (NN == 8) ? c_tab32.h8_dn[lane] : c_tab32.h20_dn[lane];
            float wdet;
            float wgp = (NN == 8) ? 1.0f : c_tab32.h20_w[lane];
            if (!geometry32<NN>(dn, s_x[wave], s_dndx[wave][lane], wdet, wgp)) {
                report32(err, AXG_E_NEG_JACOBIAN, e);
                wdet = 0.f;
            }
            s_w[wave][lane] = wdet;
        }
        __syncthreads();
        if (valid) {
            double ed = E[e], nud = NU[e];
            float lam = (float)(ed * nud / ((1.0 + nud) * (1.0 - 2.0 * nud)));
            float mu = (float)(ed / (2.0 * (1.0 + nud)));
            float diag = lam + 2.f * mu;
            floatx16 acc[T][T];
#pragma unroll
            for (int ti = 0; ti < T; ti++)
#pragma unroll
                for (int tj = 0; tj < T; tj++)
#pragma unroll
                    for (int r = 0; r < 16; r++) acc[ti][tj][r] = 0.f;
            for (int t = 0; t < KTOT / 2; t++) {
                int krow = 2 * t + h;  // stacked strain row 0..KTOT-1
                int g = krow / 6, k = krow - 6 * g;
                const float *dd = s_dndx[wave][g];
                float w = s_w[wave][g];
                float av[T], bv[T];
#pragma unroll
                for (int ti = 0; ti < T; ti++) {
                    int cidx = 32 * ti + c32;
                    av[ti] = (cidx < M) ? bval(dd, k, cidx) : 0.f;
                    bv[ti] = (cidx < M) ? w * dbval(dd, k, cidx, lam, mu, diag) : 0.f;
                }
#pragma unroll
                for (int ti = 0; ti < T; ti++)
#pragma unroll
                    for (int tj = 0; tj < T; tj++)
                        acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[ti], bv[tj],
                                                                           acc[ti][tj], 0, 0, 0);
            }
            // output layout of 32x32x2: register r of lane l holds row
            // 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
            float *out = ke + e * (int64_t)(M * M);
#pragma unroll
            for (int ti = 0; ti < T; ti++)
#pragma unroll
                for (int tj = 0; tj < T; tj++)
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        int row = 32 * ti + 8 * (r >> 2) + 4 * h + (r & 3);
                        int col = 32 * tj + c32;
                        if (row < M && col < M) out[row * M + col] = acc[ti][tj][r];
                    }
        }
        __syncthreads();
    }
}

void launch_ke_fp32(axg_mesh *m, hipStream_t s, int32_t *d_err) {
    if (m->nelem == 0) return;
    int64_t blocks = std::min<int64_t>((m->nelem + 3) / 4, (int64_t)m->ctx->cu_count * 16);
    if (m->nn == 8)
        hipLaunchKernelGGL((ke_fp32_mfma_kernel<8, 8>), dim3((unsigned)blocks), dim3(256), 0, s,
                           m->xyz.p, m->conn.p, m->e.p, m->nu.p, m->nelem, m->ke32.p, d_err);
    else
        hipLaunchKernelGGL((ke_fp32_mfma_kernel<20, 27>), dim3((unsigned)blocks), dim3(256), 0, s,
                           m->xyz.p, m->conn.p, m->e.p, m->nu.p, m->nelem, m->ke32.p, d_err);
    AXG_HIP(hipGetLastError());
}

// FP32 sums of the scalar slots' contributions (same scatter map, ascending
// element order, FP32 accumulation) written to their BSR-3 positions.
__global__ void gather_bsr32_kernel(const uint32_t *__restrict__ slotptr,
                                    const uint32_t *__restrict__ contrib,
                                    const float *__restrict__ ke32,
                                    const int32_t *__restrict__ entry_of_slot, int64_t nslot,
                                    float *__restrict__ bval) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nslot;
         s += (int64_t)gridDim.x * blockDim.x) {
        uint32_t c0 = slotptr[s], c1 = slotptr[s + 1];
        float v = 0.f;
        for (uint32_t c = c0; c < c1; c++) v += ke32[contrib[c]];
        bval[entry_of_slot[s]] = v;
    }
}

void launch_gather_bsr32(axg_mesh *m, hipStream_t s, const int32_t *entry_of_slot) {
    int64_t nslot = m->nnz_topo;
    if (nslot == 0) return;
    int64_t blocks = std::min<int64_t>((nslot + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(gather_bsr32_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                       m->slotptr_t.p, m->contrib_t.p, m->ke32.p, entry_of_slot, nslot, m->bval.p);
    AXG_HIP(hipGetLastError());
}

}  // namespace axg
