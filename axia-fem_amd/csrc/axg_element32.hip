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

typedef float floatx16 __attribute__((ext_vector_type(16)));

// Strain row k (Voigt xx yy zz xy yz zx) of B for DOF component p: which
// dN/dx component it holds (0..2) or 3 for an exact zero; 2 bits per (k, p).
__device__ inline int bcomp(int k, int p) {
    // k = 0: (0,3,3)  1: (3,1,3)  2: (3,3,2)  3: (1,0,3)  4: (3,2,1)  5: (2,3,0)
    constexpr unsigned long long tab =
        (0ull << 0) | (3ull << 2) | (3ull << 4) |      // k = 0
        (3ull << 6) | (1ull << 8) | (3ull << 10) |     // k = 1
        (3ull << 12) | (3ull << 14) | (2ull << 16) |   // k = 2
        (1ull << 18) | (0ull << 20) | (3ull << 22) |   // k = 3
        (3ull << 24) | (2ull << 26) | (1ull << 28) |   // k = 4
        (2ull << 30) | (3ull << 32) | (0ull << 34);    // k = 5
    return (int)((tab >> (2 * (3 * k + p))) & 3ull);
}

// One wavefront (one workgroup) per element, grid-stride over elements.
// Phase 1 (lanes < NG): J, its inverse and w det at every point, in FP32 on
// element-local coordinates (node 0 subtracted in FP64). Phase 2 (all lanes):
// dN/dx of every (point, node). Phase 3: Ke = Bs^T Cs on the matrix cores.
template <int NN, int NG>
__global__ void __launch_bounds__(64) ke_fp32_mfma_kernel(const double *__restrict__ xyz,
                                                          const int32_t *__restrict__ conn,
                                                          const double *__restrict__ E,
                                                          const double *__restrict__ NU,
                                                          int64_t nelem, float *__restrict__ ke,
                                                          int32_t *__restrict__ err) {
    constexpr int M = 3 * NN;         // 24 or 60
    constexpr int T = (M + 31) / 32;  // output tiles per dimension
    constexpr int KTOT = 6 * NG;      // 48 or 162
    __shared__ float s_x[3 * NN];
    __shared__ float s_g[NG][10];     // inverse J (row-major) and w det
    __shared__ float s_dndx[NG][3 * NN];
    const int lane = threadIdx.x;
    const int h = lane >> 5, c32 = lane & 31;
    int ca[T], cp[T];
#pragma unroll
    for (int ti = 0; ti < T; ti++) {
        int c = 32 * ti + c32;
        ca[ti] = c < M ? c / 3 : 0;
        cp[ti] = c < M ? c % 3 : 0;
    }
    for (int64_t e = blockIdx.x; e < nelem; e += gridDim.x) {
        const int64_t n0 = conn[e * NN];
        for (int t = lane; t < 3 * NN; t += 64) {
            int a = t / 3, q = t - 3 * a;
            s_x[t] = (float)(xyz[3 * (int64_t)conn[e * NN + a] + q] - xyz[3 * n0 + q]);
        }
        __syncthreads();
        if (lane < NG) {
            const float(*dn)[3] = (NN == 8) ? c_tab32.h8_dn[lane] : c_tab32.h20_dn[lane];
            float j[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
            for (int a = 0; a < NN; a++)
#pragma unroll
                for (int pp = 0; pp < 3; pp++)
#pragma unroll
                    for (int q = 0; q < 3; q++) j[q][pp] += dn[a][pp] * s_x[3 * a + q];
            float det = j[0][0] * (j[1][1] * j[2][2] - j[1][2] * j[2][1]) -
                        j[0][1] * (j[1][0] * j[2][2] - j[1][2] * j[2][0]) +
                        j[0][2] * (j[1][0] * j[2][1] - j[1][1] * j[2][0]);
            if (!(det > 0.f)) {
                report32(err, AXG_E_NEG_JACOBIAN, e);
                det = 1.f;
            }
            float id = 1.f / det;
            float *g = s_g[lane];
            g[0] = (j[1][1] * j[2][2] - j[1][2] * j[2][1]) * id;
            g[1] = (j[0][2] * j[2][1] - j[0][1] * j[2][2]) * id;
            g[2] = (j[0][1] * j[1][2] - j[0][2] * j[1][1]) * id;
            g[3] = (j[1][2] * j[2][0] - j[1][0] * j[2][2]) * id;
            g[4] = (j[0][0] * j[2][2] - j[0][2] * j[2][0]) * id;
            g[5] = (j[0][2] * j[1][0] - j[0][0] * j[1][2]) * id;
            g[6] = (j[1][0] * j[2][1] - j[1][1] * j[2][0]) * id;
            g[7] = (j[0][1] * j[2][0] - j[0][0] * j[2][1]) * id;
            g[8] = (j[0][0] * j[1][1] - j[0][1] * j[1][0]) * id;
            g[9] = ((NN == 8) ? 1.f : c_tab32.h20_w[lane]) * det;
        }
        __syncthreads();
        for (int t = lane; t < NG * NN; t += 64) {
            const int g = t / NN, a = t - NN * (t / NN);
            const float(*dn)[3] = (NN == 8) ? c_tab32.h8_dn[g] : c_tab32.h20_dn[g];
            const float d0 = dn[a][0], d1 = dn[a][1], d2 = dn[a][2];
            const float *iv = s_g[g];
#pragma unroll
            for (int i = 0; i < 3; i++)
                s_dndx[g][3 * a + i] = iv[i] * d0 + iv[3 + i] * d1 + iv[6 + i] * d2;
        }
        __syncthreads();
        const double ed = E[e], nud = NU[e];
        const float lam = (float)(ed * nud / ((1.0 + nud) * (1.0 - 2.0 * nud)));
        const float mu = (float)(ed / (2.0 * (1.0 + nud)));
        const float diag = lam + 2.f * mu;
        floatx16 acc[T][T];
#pragma unroll
        for (int ti = 0; ti < T; ti++)
#pragma unroll
            for (int tj = 0; tj < T; tj++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[ti][tj][r] = 0.f;
        int g = 0, k = h;  // stacked strain row 2 t + h = 6 g + k
        for (int t = 0; t < KTOT / 2; t++) {
            const float w = s_g[g][9];
            float av[T], bv[T];
#pragma unroll
            for (int ti = 0; ti < T; ti++) {
                const float *dd = &s_dndx[g][3 * ca[ti]];
                const int cmp = bcomp(k, cp[ti]);
                const float b = cmp < 3 ? dd[cmp] : 0.f;
                // (D B)[k][c]: normal rows D[k][p] dN_p, shear rows mu B
                const float db = k < 3 ? ((k == cp[ti]) ? diag : lam) * dd[cp[ti]] : mu * b;
                const bool ok = 32 * ti + c32 < M;
                av[ti] = ok ? b : 0.f;
                bv[ti] = ok ? w * db : 0.f;
            }
#pragma unroll
            for (int ti = 0; ti < T; ti++)
#pragma unroll
                for (int tj = 0; tj < T; tj++)
                    acc[ti][tj] =
                        __builtin_amdgcn_mfma_f32_32x32x2f32(av[ti], bv[tj], acc[ti][tj], 0, 0, 0);
            k += 2;
            if (k >= 6) k -= 6, g++;
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
                    const int row = 32 * ti + 8 * (r >> 2) + 4 * h + (r & 3);
                    const int col = 32 * tj + c32;
                    if (row < M && col < M) out[row * M + col] = acc[ti][tj][r];
                }
        __syncthreads();
    }
}

void launch_ke_fp32(axg_mesh *m, hipStream_t s, int32_t *d_err) {
    if (m->nelem == 0) return;
    int64_t blocks = std::min<int64_t>(m->nelem, (int64_t)m->ctx->cu_count * 64);
    if (m->nn == 8)
        hipLaunchKernelGGL((ke_fp32_mfma_kernel<8, 8>), dim3((unsigned)blocks), dim3(64), 0, s,
                           m->xyz.p, m->conn.p, m->e.p, m->nu.p, m->nelem, m->ke32.p, d_err);
    else
        hipLaunchKernelGGL((ke_fp32_mfma_kernel<20, 27>), dim3((unsigned)blocks), dim3(64), 0, s,
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
        const uint32_t c0 = slotptr[s], c1 = slotptr[s + 1];
        float v = 0.f;
        for (uint32_t c = c0; c < c1; c += 8) {
            uint32_t idx[8];
            float kv[8];
#pragma unroll
            for (int u = 0; u < 8; u++) idx[u] = c + u < c1 ? contrib[c + u] : 0u;
#pragma unroll
            for (int u = 0; u < 8; u++) kv[u] = c + u < c1 ? ke32[idx[u]] : 0.f;
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (c + u < c1) v += kv[u];
        }
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
