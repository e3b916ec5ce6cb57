// FP32 assembly of the mixed-precision operator into the BSR-3 planes, in two
// forms: the element integration fused with the block-row assembly (no element
// matrix is stored; the C3D8 default), or the element blocks formed on the
// matrix cores and stored once, then gathered into the block rows (the C3D20
// default, where an element's 20 x 20 blocks are worth forming once).
//
// For isotropic linear elasticity the 3x3 block of B^T D B that couples the
// local nodes a and b of an element is, summed over its Gauss points g,
//   K_ab = lam P + mu P^T + mu tr(P) 1,
//   P    = sum_g s_g^2 grad N_a(g) grad N_b(g)^T,   s_g^2 = w_g det J_g,
// (lam, mu the Lame constants of the element's E, nu). This is B^T D B of the
// element formulation (elem.rs:90-112, 199-219; quadratic.rs:122-146) with the
// zero structure of B and D multiplied out: nine FMAs per (a, b, g) instead of
// the dense 6 x 3NN contraction.
//
// The fused form has two kernels:
//   1. elem_grad32_kernel, lane per (element, Gauss point): J from the
//      element-local coordinates (FP64 differences to the element's first node,
//      rounded to FP32), s_g J^-1, and the element's (lam, mu). For hex8 it
//      stores the eight s_g J^-1 (288 B per element); for hex20 the scaled
//      gradient rows s_g grad N_b (6.7 KB per element).
//   2. row_asm32_kernel, node-centric: a wavefront owns two block rows (one or
//      more slot lanes per block) and builds each from its incident elements
//      in chunks of eight: task (element, b) forms node b's gradient row (hex8:
//      from the element's s_g J^-1 and the reference derivatives), P with the
//      row node's row and K_ab, and the block's slot from a table built with the
//      block layout; then each slot lane adds the chunk's K_ab in ascending
//      element order.
// Every block is summed in the same fixed order (incident elements ascending)
// on every run, and written once, with its column, to the three value planes
// and to the third plane of the LDS-staged SpMM. Constrained DOFs give zero
// entries. Workgroups are mapped XCD-aware: the XCD that runs workgroup b
// (b % 8) takes a contiguous eighth of the rows, so the 16-byte block writes
// of neighbouring rows meet in one L2.
#include <climits>
#include <cmath>

#include "axg_internal.h"

namespace axg {

struct ShapeTables32 {
    float h8_dn[8][8][3];
    float h20_dn[27][20][3];
    float h20_w[27];
};
__constant__ ShapeTables32 c_tab32;

// Tile pairs of P = W W^T and their batches (see BlkTab32), fixed by the
// element type at compile time: pairs (I <= J) in row-major order, cut into
// batches of at most three pairs and 64 blocks (a <= b).
struct BlkPlan {
    int npairs = 0, nbatch = 0;
    int pi[10] = {}, pj[10] = {}, slot[10] = {}, batch[10] = {};
    bool last[10] = {};  // last pair of its batch
    int start[11] = {};  // first storage index of each batch
};
constexpr int plan_nodes(int nn, int T) { return nn - 5 * T < 5 ? nn - 5 * T : 5; }
constexpr int plan_count(int nn, int I, int J) {
    return I < J ? plan_nodes(nn, I) * plan_nodes(nn, J) : plan_nodes(nn, I) * (plan_nodes(nn, I) + 1) / 2;
}
constexpr BlkPlan blk_plan(int nn) {
    BlkPlan p;
    const int rt = (nn + 4) / 5;
    int used = 0, nslot = 0, s = 0;
    for (int I = 0; I < rt; I++)
        for (int J = I; J < rt; J++) {
            const int c = plan_count(nn, I, J);
            if (p.npairs > 0 && (used + c > 64 || nslot == 3)) {
                p.last[p.npairs - 1] = true;
                p.nbatch++;
                used = 0, nslot = 0;
            }
            if (nslot == 0) p.start[p.nbatch] = s;
            p.pi[p.npairs] = I, p.pj[p.npairs] = J, p.slot[p.npairs] = nslot++;
            p.batch[p.npairs] = p.nbatch;
            used += c, s += c;
            p.npairs++;
        }
    p.last[p.npairs - 1] = true;
    p.nbatch++;
    p.start[p.nbatch] = s;
    return p;
}

// Element blocks on the matrix cores (elem_blk32_kernel): the rows of
// W (W[3 a + i][g] = s_g dN_a/dx_i at Gauss point g) are grouped in 16-row
// tiles of five nodes, row 16 T + 3 (a - 5 T) + i, row 15 of a tile zero, so
// that every 3x3 block P_ab of P = W W^T lies in one 16x16 tile pair. The
// upper tile pairs (I <= J) are formed in batches whose blocks (a <= b) fit
// one wavefront; a block's place in that order is its storage index within
// the element's record.
struct BlkTab32 {
    uint32_t ent[8][64];  // lane's block in batch q: 1 | pair slot << 1 | a << 4 | b << 9, 0 = none
    uint8_t idx[20][20];  // storage index of block (a, b), a <= b
};
__constant__ BlkTab32 c_blk8, c_blk20;

static BlkTab32 make_blk_tab(int nn) {
    BlkTab32 t{};
    const BlkPlan p = blk_plan(nn);
    if (p.nbatch > 8) throw Error{AXG_E_INVALID, "FP32 block tables: too many batches"};
    int s = 0, lane = 0;
    for (int pp = 0; pp < p.npairs; pp++) {
        const int q = p.batch[pp], I = p.pi[pp], J = p.pj[pp];
        if (p.slot[pp] == 0) lane = 0;
        for (int a = 5 * I; a < 5 * I + plan_nodes(nn, I); a++)
            for (int b = 5 * J; b < 5 * J + plan_nodes(nn, J); b++) {
                if (b < a) continue;
                t.ent[q][lane++] = 1u | (uint32_t)p.slot[pp] << 1 | (uint32_t)a << 4 | (uint32_t)b << 9;
                t.idx[a][b] = (uint8_t)s++;
            }
    }
    if (s != nn * (nn + 1) / 2 || p.start[p.nbatch] != s)
        throw Error{AXG_E_INVALID, "FP32 block tables: wrong block count"};
    return t;
}

void upload_shape_tables32(const double *h8_dn, const double *h20_dn, const double *h20_w) {
    static ShapeTables32 t;
    for (int i = 0; i < 8 * 8 * 3; i++) (&t.h8_dn[0][0][0])[i] = (float)h8_dn[i];
    for (int i = 0; i < 27 * 20 * 3; i++) (&t.h20_dn[0][0][0])[i] = (float)h20_dn[i];
    for (int i = 0; i < 27; i++) t.h20_w[i] = (float)h20_w[i];
    AXG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tab32), &t, sizeof(t)));
    static BlkTab32 b8, b20;
    b8 = make_blk_tab(8);
    b20 = make_blk_tab(20);
    AXG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_blk8), &b8, sizeof(b8)));
    AXG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_blk20), &b20, sizeof(b20)));
}

__device__ inline void report32(int32_t *err, int32_t code, int64_t el) {
    if (atomicCAS(&err[0], 0, code) == 0) err[1] = (int32_t)el;
}

// LDS written by some lanes of a wavefront and read by others: the wave's LDS
// operations complete in order, the fences keep the compiler from moving them.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// s_g J^-1 at one Gauss point: iv[3 p + i] = s_g (J^-1)[p][i], s_g = sqrt(w_g det J),
// from element-local coordinates x[3 b + q] and the reference derivatives
// dn[3 b + p] of the NN nodes. False (and det J taken as 1) when det J <= 0.
template <int NN>
__device__ inline bool scaled_inverse(const float *dn, const float *x, float wg, float (&iv)[9]) {
    float j[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    for (int b = 0; b < NN; b++) {
        const float d0 = dn[3 * b], d1 = dn[3 * b + 1], d2 = dn[3 * b + 2];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const float xq = x[3 * b + q];
            j[q][0] += d0 * xq;
            j[q][1] += d1 * xq;
            j[q][2] += d2 * xq;
        }
    }
    float det = j[0][0] * (j[1][1] * j[2][2] - j[1][2] * j[2][1]) -
                j[0][1] * (j[1][0] * j[2][2] - j[1][2] * j[2][0]) +
                j[0][2] * (j[1][0] * j[2][1] - j[1][1] * j[2][0]);
    const bool ok = det > 0.f;
    if (!ok) det = 1.f;
    const float id = __builtin_amdgcn_sqrtf(wg * det) * __builtin_amdgcn_rcpf(det);
    iv[0] = (j[1][1] * j[2][2] - j[1][2] * j[2][1]) * id;
    iv[1] = (j[0][2] * j[2][1] - j[0][1] * j[2][2]) * id;
    iv[2] = (j[0][1] * j[1][2] - j[0][2] * j[1][1]) * id;
    iv[3] = (j[1][2] * j[2][0] - j[1][0] * j[2][2]) * id;
    iv[4] = (j[0][0] * j[2][2] - j[0][2] * j[2][0]) * id;
    iv[5] = (j[0][2] * j[1][0] - j[0][0] * j[1][2]) * id;
    iv[6] = (j[1][0] * j[2][1] - j[1][1] * j[2][0]) * id;
    iv[7] = (j[0][1] * j[2][0] - j[0][0] * j[2][1]) * id;
    iv[8] = (j[0][0] * j[1][1] - j[0][1] * j[1][0]) * id;
    return ok;
}

// Scaled gradients of every element: grad[(e NN + b) NGP + 3 g + i] =
// s_g d N_b / d x_i at Gauss point g (NGP = 3 NG rounded up to a multiple of
// 12, the tail zero), and the element's (lam, mu). Each wavefront of a
// workgroup takes its own 64 / NG consecutive elements, one lane per (element,
// Gauss point): J from the element-local coordinates (FP64 differences to the
// element's first node, rounded to FP32) and the reference-element derivative
// table, then s_g J^-1 applied to every node's derivatives. The wavefront's
// output (contiguous in memory) is staged in LDS and written with 16-byte
// stores.
//
// IVO (hex8): instead of the gradients, the element's eight s_g J_g^-1 (72
// floats, 288 B against 768 B), laid out per half h of the Gauss points as
// grad[e 72 + 36 h + 4 k + g % 4] = iv_k at point g = 4 h + g % 4; the block
// rows form the gradients from them.
template <int NN, int NG, int NGP, int WPB, bool IVO = false>
__global__ void __launch_bounds__(64 * WPB)
    elem_grad32_kernel(int64_t nelem, const int32_t *__restrict__ conn,
                       const double *__restrict__ xyz, const double *__restrict__ E,
                       const double *__restrict__ NU, float *__restrict__ grad,
                       float2 *__restrict__ lm, int32_t *__restrict__ err) {
    static_assert(!IVO || (NN == 8 && NG == 8), "inverse output: hex8");
    constexpr int EPW = 64 / NG;
    constexpr int PER = IVO ? 9 * NG : NN * NGP;  // floats per element
    constexpr int OUT = EPW * PER;                // floats of one wavefront's elements
    static_assert(OUT % 4 == 0, "16-byte stores");
    __shared__ float s_dn[NG][3 * NN + 1];
    __shared__ float s_w[NG];
    __shared__ float s_x[WPB][EPW][3 * NN + 1];
    __shared__ float4 s_out[WPB][OUT / 4];
    for (int t = threadIdx.x; t < NG * NN * 3; t += blockDim.x)
        s_dn[t / (3 * NN)][t % (3 * NN)] =
            NN == 8 ? (&c_tab32.h8_dn[0][0][0])[t] : (&c_tab32.h20_dn[0][0][0])[t];
    for (int t = threadIdx.x; t < NG; t += blockDim.x) s_w[t] = NN == 8 ? 1.f : c_tab32.h20_w[t];
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t e0 = ((int64_t)blockIdx.x * WPB + w) * EPW;
    if (e0 >= nelem) return;
    const int ne = (int)(nelem - e0 < EPW ? nelem - e0 : EPW);
    float *so = reinterpret_cast<float *>(s_out[w]);
    for (int t = lane; t < ne * NN; t += 64) {
        const int c = t / NN, b = t - NN * (t / NN);
        const int64_t e = e0 + c;
        const int64_t n0 = conn[e * NN], nd = conn[e * NN + b];
#pragma unroll
        for (int q = 0; q < 3; q++) s_x[w][c][3 * b + q] = (float)(xyz[3 * nd + q] - xyz[3 * n0 + q]);
    }
    if (lane < ne) {
        const double ed = E[e0 + lane], nud = NU[e0 + lane];
        lm[e0 + lane] = make_float2((float)(ed * nud / ((1.0 + nud) * (1.0 - 2.0 * nud))),
                                    (float)(ed / (2.0 * (1.0 + nud))));
    }
    wave_sync();
    const int ce = lane / NG, g = lane - NG * (lane / NG);
    if (ce < ne) {
        float iv[9];
        if (!scaled_inverse<NN>(s_dn[g], s_x[w][ce], s_w[g], iv))
            report32(err, AXG_E_NEG_JACOBIAN, e0 + ce);
        float *out = so + ce * PER;
        if constexpr (IVO) {
#pragma unroll
            for (int k = 0; k < 9; k++) out[36 * (g >> 2) + 4 * k + (g & 3)] = iv[k];
        } else {
        for (int b = 0; b < NN; b++) {
            const float d0 = s_dn[g][3 * b], d1 = s_dn[g][3 * b + 1], d2 = s_dn[g][3 * b + 2];
            out[b * NGP + 3 * g] = iv[0] * d0 + iv[3] * d1 + iv[6] * d2;
            out[b * NGP + 3 * g + 1] = iv[1] * d0 + iv[4] * d1 + iv[7] * d2;
            out[b * NGP + 3 * g + 2] = iv[2] * d0 + iv[5] * d1 + iv[8] * d2;
        }
        if (NGP > 3 * NG && g < NGP - 3 * NG)  // the zero tail
            for (int b = 0; b < NN; b++) out[b * NGP + 3 * NG + g] = 0.f;
        }
    }
    wave_sync();
    float4 *dst = reinterpret_cast<float4 *>(grad + e0 * (int64_t)PER);
    for (int t = lane; t < ne * (PER / 4); t += 64) dst[t] = s_out[w][t];
}

// P += sum over four Gauss points of a(g) b(g)^T; a and b hold the points'
// gradient rows (3 floats each) in three float4.
__device__ inline void p_step(float (&p)[3][3], float4 a0, float4 a1, float4 a2, float4 b0, float4 b1,
                              float4 b2) {
    const float av[12] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w, a2.x, a2.y, a2.z, a2.w};
    const float bv[12] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w, b2.x, b2.y, b2.z, b2.w};
#pragma unroll
    for (int gg = 0; gg < 4; gg++)
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int cc = 0; cc < 3; cc++) p[r][cc] += av[3 * gg + r] * bv[3 * gg + cc];
}

// K_ab = lam P + mu P^T + mu tr(P) 1, row-major.
__device__ inline void k_of_p(const float (&p)[3][3], float2 lm, float (&kv)[9]) {
    const float lam = lm.x, mu = lm.y;
    const float tr = mu * (p[0][0] + p[1][1] + p[2][2]);
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int cc = 0; cc < 3; cc++) kv[3 * r + cc] = lam * p[r][cc] + mu * p[cc][r] + (r == cc ? tr : 0.f);
}

// Per-row LDS of the block-row assembly: CH elements per chunk, SPR slots.
template <int NN, int CH, int SPR, int NQA>
struct RowLds {
    float4 c[CH][NN + 1][2];  // K_ab of the chunk's elements, entries 0-7; b = NN: zero
    float c8[CH][NN + 1];     // and entry 8
    float4 ga[CH][NQA];       // gradient row of the block row's own node in chunk element ce
    uint8_t src[SPR][CH];     // local node b of chunk element ce at slot k, NN for none
};

// Block rows: RPW rows per wavefront (two when every row has at most 32
// blocks), KPL x 64 slot lanes; slot lane (h, k) owns slot k of row h. Per
// chunk of CH incident elements, task (h, ce, b) forms K_ab of row h's
// element ce with its local node b; then every slot lane adds the chunk's
// contributions to its block in ascending element order.
//
// IVG (hex8, chunks of eight elements): the element data are the eight
// s_g J_g^-1 (elem_grad32_kernel's IVO layout). The eight tasks of a chunk
// element copy its 288 B into LDS (over the chunk's block buffer, which is
// written only after they are read); task (h, ce, b) forms its gradient row
// s_g grad N_b(g) = (s_g J_g^-1)^T dN_b(g) from them and node b's reference
// derivatives, one row h at a time, the row node's row reaching the other
// tasks through LDS. Otherwise (hex20) the element data are the gradient rows,
// read by each task from global memory.
template <int NN, int NGP, int CH, int KPL, int RPW, int WPB, int MINW, bool IVG>
__global__ void __launch_bounds__(64 * WPB, MINW)
    row_asm32_kernel(int64_t nb, int64_t per_xcd, int incw, const int32_t *__restrict__ rinc,
                     const float *__restrict__ grad,
                     const float2 *__restrict__ lm, const uint8_t *__restrict__ eslot,
                     const uint8_t *__restrict__ smask, const int64_t *__restrict__ sptr,
                     const int32_t *__restrict__ width,
                     const int32_t *__restrict__ bcol, const int32_t *__restrict__ bloc,
                     int64_t nsl, float *__restrict__ bval, float *__restrict__ lval) {
    constexpr int SPR = 64 * KPL / RPW;
    constexpr int TPR = CH * NN;  // tasks per row and chunk
    constexpr int NIT = (RPW * TPR + 63) / 64;
    constexpr int NQ = NGP / 4;
    using Lds = RowLds<NN, CH, SPR, IVG ? NQ : 1>;
    static_assert(CH % 4 == 0, "src rows are read as words");
    static_assert(!IVG || (NN == 8 && CH == 8 && NGP == 24 && NIT == RPW), "inverse input: hex8 chunks of 8");
    static_assert(!IVG || sizeof(Lds::c) >= CH * 18 * sizeof(float4), "inverse data fit the block buffer");
    __shared__ Lds s_row[WPB][RPW];
    // IVG: reference derivatives, dnt[b][3 g + p] = dN_b/dxi_p at Gauss point g
    __shared__ float4 s_dnt[IVG ? 8 : 1][6];
    if constexpr (IVG) {
        for (int t = threadIdx.x; t < 8 * 24; t += blockDim.x)
            reinterpret_cast<float *>(s_dnt[t / 24])[t % 24] = c_tab32.h8_dn[(t % 24) / 3][t / 24][t % 3];
        __syncthreads();
    }
    const int64_t grp = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    // the wave index is wave-uniform: row metadata comes from scalar loads
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t I0 = (grp * WPB + w) * RPW;
    if (grp >= per_xcd * 8 || I0 >= nb) return;
    // rows of the wavefront: slot count and first slot (wave-uniform loads)
    int rw[RPW];
    int64_t rbase[RPW];
#pragma unroll
    for (int h = 0; h < RPW; h++) {
        const int64_t I = I0 + h < nb ? I0 + h : I0;
        rw[h] = I0 + h < nb ? width[I >> 6] : 0;
        rbase[h] = sptr[I >> 6] + (I & 63);
    }
    // slot lanes: row, slot, position, column, DOF mask (the column, mask and
    // LDS index loads are first used at the end)
    int sh[KPL], sk[KPL];
    bool sv[KPL];
    int64_t spos[KPL];
    int32_t ccol[KPL], cloc[KPL];
    unsigned mask[KPL];
    float acc[KPL][9];
#pragma unroll
    for (int j = 0; j < KPL; j++) {
        const int idx = lane + 64 * j;
        const int h = idx / SPR, k = idx - SPR * (idx / SPR);
        int W = rw[0];
        int64_t base = rbase[0];
#pragma unroll
        for (int hh = 1; hh < RPW; hh++)
            if (h == hh) W = rw[hh], base = rbase[hh];
        sh[j] = h, sk[j] = k, mask[j] = 0u, ccol[j] = cloc[j] = 0;
        sv[j] = h < RPW && k < W;
        // unconditional loads from a clamped position: no branch splits the wait counting
        spos[j] = base + 64 * (int64_t)(sv[j] ? k : 0);
#pragma unroll
        for (int r = 0; r < 9; r++) acc[j][r] = 0.f;
        ccol[j] = bcol[spos[j]];
        cloc[j] = bloc[spos[j]];
        mask[j] = smask[spos[j]];
    }
    const int nch = (incw + CH - 1) / CH;
    auto zero_blocks = [&]() {
        for (int t = lane; t < RPW * CH * 3; t += 64)
            if (t % 3 < 2) s_row[w][t / (CH * 3)].c[(t / 3) % CH][NN][t % 3] = make_float4(0.f, 0.f, 0.f, 0.f);
            else s_row[w][t / (CH * 3)].c8[(t / 3) % CH][NN] = 0.f;
    };
    if constexpr (!IVG) zero_blocks();  // never written again
    for (int c = 0; c < nch; c++) {
        // the tasks' incident pairs first, then their gradient rows
        int32_t sl[NIT];
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int t = lane + 64 * it;
            const int h = t / TPR, ce = (t - TPR * (t / TPR)) / NN;
            const bool ok = t < RPW * TPR && I0 + h < nb && c * CH + ce < incw;
            const int32_t v = rinc[ok ? (I0 + h) * incw + c * CH + ce : 0];
            sl[it] = ok ? v : -1;
        }
        for (int t = lane; t < RPW * CH * SPR / 4; t += 64)
            reinterpret_cast<uint32_t *>(&s_row[w][t / (CH * SPR / 4)].src[0][0])[t % (CH * SPR / 4)] =
                0x01010101u * NN;
        if constexpr (IVG) {
            // task (h = it, ce = lane / 8, b = lane % 8): the element's 18 float4
            // copied by its eight tasks; then per row its gradient rows, the row
            // node's through LDS, and K_ab; the blocks go over the inverse data
            // once every row's are formed
            const int ce = lane >> 3, b = lane & 7;
            unsigned kk[NIT];
            float2 lmv[NIT];
            float kv[NIT][9];
#pragma unroll
            for (int it = 0; it < NIT; it++) {
                const int64_t sc = sl[it] < 0 ? 0 : sl[it];
                const int64_t e = sc / NN;
                const float4 *pv = reinterpret_cast<const float4 *>(grad) + e * 18;
                float4 *ivs = &s_row[w][it].c[0][0][0] + ce * 18;
                ivs[b] = pv[b];
                ivs[b + 8] = pv[b + 8];
                if (b < 2) ivs[b + 16] = pv[b + 16];
                kk[it] = eslot[sc * NN + b];
                lmv[it] = lm[e];
            }
            wave_sync();
#pragma unroll
            for (int it = 0; it < NIT; it++) {
                float4 gb[NQ];
#pragma unroll
                for (int hf = 0; hf < 2; hf++) {  // four Gauss points at a time
                    const float4 q0 = s_dnt[b][3 * hf], q1 = s_dnt[b][3 * hf + 1], q2 = s_dnt[b][3 * hf + 2];
                    const float dd[12] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
                    const float4 *ivs = &s_row[w][it].c[0][0][0] + ce * 18 + 9 * hf;
                    float gv[12];
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        const float4 v0 = ivs[i], v1 = ivs[3 + i], v2 = ivs[6 + i];
                        const float a0[4] = {v0.x, v0.y, v0.z, v0.w}, a1[4] = {v1.x, v1.y, v1.z, v1.w},
                                    a2[4] = {v2.x, v2.y, v2.z, v2.w};
#pragma unroll
                        for (int g4 = 0; g4 < 4; g4++)
                            gv[3 * g4 + i] = a0[g4] * dd[3 * g4] + a1[g4] * dd[3 * g4 + 1] + a2[g4] * dd[3 * g4 + 2];
                    }
#pragma unroll
                    for (int q = 0; q < 3; q++)
                        gb[3 * hf + q] = make_float4(gv[4 * q], gv[4 * q + 1], gv[4 * q + 2], gv[4 * q + 3]);
                }
                Lds &R = s_row[w][it];
                if (sl[it] >= 0 && b == sl[it] % NN) {
#pragma unroll
                    for (int q = 0; q < NQ; q++) R.ga[ce][q] = gb[q];
                }
                wave_sync();
                float p[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll
                for (int q = 0; q < NQ / 3; q++)
                    p_step(p, R.ga[ce][3 * q], R.ga[ce][3 * q + 1], R.ga[ce][3 * q + 2], gb[3 * q], gb[3 * q + 1],
                           gb[3 * q + 2]);
                k_of_p(p, lmv[it], kv[it]);
                // formed here: sunk to the stores below, both rows' operands would stay live
#pragma unroll
                for (int r = 0; r < 9; r++) asm volatile("" : "+v"(kv[it][r]));
            }
            wave_sync();
            zero_blocks();  // over the inverse data, all read
#pragma unroll
            for (int it = 0; it < NIT; it++) {
                if (sl[it] < 0) continue;
                Lds &R = s_row[w][it];
                R.c[ce][b][0] = make_float4(kv[it][0], kv[it][1], kv[it][2], kv[it][3]);
                R.c[ce][b][1] = make_float4(kv[it][4], kv[it][5], kv[it][6], kv[it][7]);
                R.c8[ce][b] = kv[it][8];
                if (kk[it] < (unsigned)SPR) R.src[kk[it]][ce] = (uint8_t)b;
            }
        } else {
        wave_sync();
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int t = lane + 64 * it;
            const int h = t / TPR, u = t - TPR * (t / TPR);
            const int ce = u / NN, b = u - NN * (u / NN);
            if (sl[it] < 0) continue;
            const int64_t e = sl[it] / NN;
            const float4 *pa = reinterpret_cast<const float4 *>(grad + (int64_t)sl[it] * NGP);
            const float4 *pb = reinterpret_cast<const float4 *>(grad + (e * NN + b) * NGP);
            const unsigned k = eslot[(int64_t)sl[it] * NN + b];
            const float2 lmv = lm[e];
            float p[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll
            for (int q = 0; q < NGP / 12; q++)  // four Gauss points per step
                p_step(p, pa[3 * q], pa[3 * q + 1], pa[3 * q + 2], pb[3 * q], pb[3 * q + 1], pb[3 * q + 2]);
            float kv[9];
            k_of_p(p, lmv, kv);
            Lds &R = s_row[w][h];
            R.c[ce][b][0] = make_float4(kv[0], kv[1], kv[2], kv[3]);
            R.c[ce][b][1] = make_float4(kv[4], kv[5], kv[6], kv[7]);
            R.c8[ce][b] = kv[8];
            if (k < (unsigned)SPR) R.src[k][ce] = (uint8_t)b;
        }
        }
        wave_sync();
        // per slot, the chunk's contributions in ascending element order
#pragma unroll
        for (int jj = 0; jj < KPL; jj++) {
            if (!sv[jj]) continue;
            const Lds &R = s_row[w][sh[jj]];
            uint32_t srcw[CH / 4];
#pragma unroll
            for (int q = 0; q < CH / 4; q++)
                srcw[q] = reinterpret_cast<const uint32_t *>(&R.src[sk[jj]][0])[q];
#pragma unroll
            for (int ce = 0; ce < CH; ce++) {
                const unsigned bb = (srcw[ce / 4] >> (8 * (ce % 4))) & 0xffu;
                const float4 v0 = R.c[ce][bb][0], v1 = R.c[ce][bb][1];
                const float vv[9] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w, R.c8[ce][bb]};
#pragma unroll
                for (int r = 0; r < 9; r++) acc[jj][r] += vv[r];
            }
        }
        wave_sync();
    }
#pragma unroll
    for (int jj = 0; jj < KPL; jj++) {
        if (!sv[jj]) continue;
        float v[9];
#pragma unroll
        for (int r = 0; r < 9; r++)
            v[r] = ((mask[jj] >> (r / 3)) & (mask[jj] >> (4 + r % 3)) & 1u) ? acc[jj][r] : 0.f;
        const int64_t pos = spos[jj];
        *reinterpret_cast<float4 *>(bval + 4 * pos) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4 *>(bval + 4 * nsl + 4 * pos) = make_float4(v[4], v[5], v[6], v[7]);
        *reinterpret_cast<float2 *>(bval + 8 * nsl + 2 * pos) = make_float2(v[8], __int_as_float(ccol[jj]));
        *reinterpret_cast<float2 *>(lval + 2 * pos) = make_float2(v[8], __int_as_float(cloc[jj]));
    }
}

typedef float f32x4_t __attribute__((ext_vector_type(4)));

// Element blocks of the FP32 operator on the matrix cores, one element record
// of NN (NN + 1) / 2 blocks (a <= b) per element, three planes (entries 0-3,
// 4-7, 8 of the row-major block) of nelem * NS blocks each. Per wavefront of
// EPI elements:
//   1. lane per (element, Gauss point): s_g J^-1 into LDS (FP32, from the
//      element-local coordinates, as for the gradients);
//   2. per element, every lane forms its operand values: lane l holds
//      W[16 T + l % 16][4 k + l / 16] for every row tile T and Gauss-point step
//      k, which is both the A operand (rows of tile T) and the B operand
//      (columns of tile T) of P = W W^T, so no operand passes through LDS;
//   3. per upper tile pair (I, J): P_IJ = sum_k A_Ik B_kJ with
//      v_mfma_f32_16x16x4f32 (four Gauss points per step), staged in LDS per
//      batch;
//   4. lane per block (a <= b) of the batch: K_ab = lam P_ab + mu P_ab^T +
//      mu tr(P_ab) 1 written at the block's storage index (consecutive lanes,
//      consecutive blocks: 16-byte coalesced stores).
template <int NN, int NG, int EPI, int WPB>
__global__ void __launch_bounds__(64 * WPB)
    elem_blk32_kernel(int64_t nelem, const int32_t *__restrict__ conn,
                      const double *__restrict__ xyz, const double *__restrict__ E,
                      const double *__restrict__ NU, float *__restrict__ kb,
                      int32_t *__restrict__ err) {
    constexpr BlkPlan PL = blk_plan(NN);
    constexpr int RT = (NN + 4) / 5;       // row tiles of W
    constexpr int KS = (NG + 3) / 4;       // Gauss-point steps of the contraction
    constexpr int NS = NN * (NN + 1) / 2;  // blocks per element record
    constexpr int GPL = 64 / EPI;          // lanes per element in phase 1
    static_assert(GPL >= NG, "one lane per Gauss point");
    const BlkTab32 &tab = NN == 8 ? c_blk8 : c_blk20;
    __shared__ float s_dn[NG][3 * NN];
    __shared__ float s_wg[NG];
    __shared__ float s_x[WPB][EPI][3 * NN];
    __shared__ float s_iv[WPB][EPI][NG][9];
    __shared__ float s_pt[WPB][3][16][17];
    __shared__ float2 s_lm[WPB][EPI];
    for (int t = threadIdx.x; t < NG * NN * 3; t += blockDim.x)
        s_dn[t / (3 * NN)][t % (3 * NN)] =
            NN == 8 ? (&c_tab32.h8_dn[0][0][0])[t] : (&c_tab32.h20_dn[0][0][0])[t];
    for (int t = threadIdx.x; t < NG; t += blockDim.x) s_wg[t] = NN == 8 ? 1.f : c_tab32.h20_w[t];
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t e0 = ((int64_t)blockIdx.x * WPB + w) * EPI;
    if (e0 >= nelem) return;
    const int ne = (int)(nelem - e0 < EPI ? nelem - e0 : EPI);
    for (int t = lane; t < ne * NN; t += 64) {
        const int c = t / NN, b = t - NN * (t / NN);
        const int64_t e = e0 + c;
        const int64_t n0 = conn[e * NN], nd = conn[e * NN + b];
#pragma unroll
        for (int q = 0; q < 3; q++) s_x[w][c][3 * b + q] = (float)(xyz[3 * nd + q] - xyz[3 * n0 + q]);
    }
    if (lane < ne) {
        const double ed = E[e0 + lane], nud = NU[e0 + lane];
        s_lm[w][lane] = make_float2((float)(ed * nud / ((1.0 + nud) * (1.0 - 2.0 * nud))),
                                    (float)(ed / (2.0 * (1.0 + nud))));
    }
    wave_sync();
    {
        const int ce = lane / GPL, g = lane - GPL * (lane / GPL);
        if (ce < ne && g < NG) {
            float iv[9];
            if (!scaled_inverse<NN>(s_dn[g], s_x[w][ce], s_wg[g], iv))
                report32(err, AXG_E_NEG_JACOBIAN, e0 + ce);
#pragma unroll
            for (int r = 0; r < 9; r++) s_iv[w][ce][g][r] = iv[r];
        }
    }
    wave_sync();
    const int pos = lane & 15, q4 = lane >> 4, dir = pos % 3;
    float *kA = kb, *kB = kb + 4 * NS * nelem, *kC = kb + 8 * NS * nelem;
    for (int ce = 0; ce < ne; ce++) {
        const int64_t e = e0 + ce;
        float wv[RT][KS];
#pragma unroll
        for (int k = 0; k < KS; k++) {
            const int g = 4 * k + q4;
            const int gg = g < NG ? g : 0;
            const float i0 = s_iv[w][ce][gg][dir], i1 = s_iv[w][ce][gg][3 + dir],
                        i2 = s_iv[w][ce][gg][6 + dir];
#pragma unroll
            for (int T = 0; T < RT; T++) {
                const int a = 5 * T + pos / 3;
                const bool ok = pos < 15 && a < NN && g < NG;
                const int aa = ok ? a : 0;
                const float v = i0 * s_dn[gg][3 * aa] + i1 * s_dn[gg][3 * aa + 1] + i2 * s_dn[gg][3 * aa + 2];
                wv[T][k] = ok ? v : 0.f;
            }
        }
        const float2 lmv = s_lm[w][ce];
#pragma unroll
        for (int pp = 0; pp < PL.npairs; pp++) {
            f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < KS; k++)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(wv[PL.pi[pp]][k], wv[PL.pj[pp]][k], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r++) s_pt[w][PL.slot[pp]][4 * q4 + r][pos] = acc[r];
            if (PL.last[pp]) {
                wave_sync();
                const int q = PL.batch[pp];
                const uint32_t en = tab.ent[q][lane];
                if (en) {
                    const int sl = (en >> 1) & 3, a = (en >> 4) & 31, b = (en >> 9) & 31;
                    const int ra = 3 * (a % 5), cb = 3 * (b % 5);
                    float P[3][3];
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++) P[i][j] = s_pt[w][sl][ra + i][cb + j];
                    const float lam = lmv.x, mu = lmv.y;
                    const float tr = mu * (P[0][0] + P[1][1] + P[2][2]);
                    float kv[9];
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++)
                            kv[3 * i + j] = lam * P[i][j] + mu * P[j][i] + (i == j ? tr : 0.f);
                    const int64_t o = e * NS + PL.start[q] + lane;
                    *reinterpret_cast<float4 *>(kA + 4 * o) = make_float4(kv[0], kv[1], kv[2], kv[3]);
                    *reinterpret_cast<float4 *>(kB + 4 * o) = make_float4(kv[4], kv[5], kv[6], kv[7]);
                    kC[o] = kv[8];
                }
                wave_sync();
            }
        }
    }
}

// Block rows from the element records: the slot lanes of row_asm32_kernel,
// with each task (row h, element ce, local node b) loading its block K_ab
// (a the row node's local index; stored as K_ba^T when a > b) instead of
// integrating it. Summation order and outputs as in row_asm32_kernel.
template <int NN, int CH, int KPL, int RPW, int WPB, int MINW>
__global__ void __launch_bounds__(64 * WPB, MINW)
    row_gather32_kernel(int64_t nb, int64_t per_xcd, int incw, const int32_t *__restrict__ rinc,
                        const float *__restrict__ kb, int64_t nelem,
                        const uint8_t *__restrict__ eslot, const uint8_t *__restrict__ smask,
                        const int64_t *__restrict__ sptr, const int32_t *__restrict__ width,
                        const int32_t *__restrict__ bcol, const int32_t *__restrict__ bloc,
                        int64_t nsl, float *__restrict__ bval, float *__restrict__ lval) {
    constexpr int NS = NN * (NN + 1) / 2;
    constexpr int SPR = 64 * KPL / RPW;
    constexpr int TPR = CH * NN;  // tasks per row and chunk
    constexpr int NIT = (RPW * TPR + 63) / 64;
    using Lds = RowLds<NN, CH, SPR, 1>;
    static_assert(CH % 4 == 0, "src rows are read as words");
    __shared__ Lds s_row[WPB][RPW];
    __shared__ uint8_t s_idx[NN][NN];
    {
        const BlkTab32 &tab = NN == 8 ? c_blk8 : c_blk20;
        for (int t = threadIdx.x; t < NN * NN; t += blockDim.x) {
            const int a = t / NN, b = t - NN * (t / NN);
            s_idx[a][b] = a <= b ? tab.idx[a][b] : tab.idx[b][a];
        }
    }
    __syncthreads();
    const float4 *kA = reinterpret_cast<const float4 *>(kb);
    const float4 *kB = reinterpret_cast<const float4 *>(kb + 4 * NS * nelem);
    const float *kC = kb + 8 * NS * nelem;
    const int64_t grp = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t I0 = (grp * WPB + w) * RPW;
    if (grp >= per_xcd * 8 || I0 >= nb) return;
    int rw[RPW];
    int64_t rbase[RPW];
#pragma unroll
    for (int h = 0; h < RPW; h++) {
        const int64_t I = I0 + h < nb ? I0 + h : I0;
        rw[h] = I0 + h < nb ? width[I >> 6] : 0;
        rbase[h] = sptr[I >> 6] + (I & 63);
    }
    int sh[KPL], sk[KPL];
    bool sv[KPL];
    int64_t spos[KPL];
    int32_t ccol[KPL], cloc[KPL];
    unsigned mask[KPL];
    float acc[KPL][9];
#pragma unroll
    for (int j = 0; j < KPL; j++) {
        const int idx = lane + 64 * j;
        const int h = idx / SPR, k = idx - SPR * (idx / SPR);
        int W = rw[0];
        int64_t base = rbase[0];
#pragma unroll
        for (int hh = 1; hh < RPW; hh++)
            if (h == hh) W = rw[hh], base = rbase[hh];
        sh[j] = h, sk[j] = k;
        sv[j] = h < RPW && k < W;
        spos[j] = base + 64 * (int64_t)(sv[j] ? k : 0);
#pragma unroll
        for (int r = 0; r < 9; r++) acc[j][r] = 0.f;
        ccol[j] = bcol[spos[j]];
        cloc[j] = bloc[spos[j]];
        mask[j] = smask[spos[j]];
    }
    const int nch = (incw + CH - 1) / CH;
    for (int t = lane; t < RPW * CH * 3; t += 64)  // the zero blocks (never written again)
        if (t % 3 < 2) s_row[w][t / (CH * 3)].c[(t / 3) % CH][NN][t % 3] = make_float4(0.f, 0.f, 0.f, 0.f);
        else s_row[w][t / (CH * 3)].c8[(t / 3) % CH][NN] = 0.f;
    for (int c = 0; c < nch; c++) {
        int32_t sl[NIT];
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int t = lane + 64 * it;
            const int h = t / TPR, ce = (t - TPR * (t / TPR)) / NN;
            const bool ok = t < RPW * TPR && I0 + h < nb && c * CH + ce < incw;
            const int32_t v = rinc[ok ? (I0 + h) * incw + c * CH + ce : 0];
            sl[it] = ok ? v : -1;
        }
        for (int t = lane; t < RPW * CH * SPR / 4; t += 64)
            reinterpret_cast<uint32_t *>(&s_row[w][t / (CH * SPR / 4)].src[0][0])[t % (CH * SPR / 4)] =
                0x01010101u * NN;
        float4 va[NIT], vb[NIT];
        float vc[NIT];
        unsigned kk[NIT];
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int t = lane + 64 * it;
            const int b = (t - TPR * (t / TPR)) % NN;
            const int64_t s0 = sl[it] < 0 ? 0 : sl[it];
            const int64_t e = s0 / NN;
            const int a = (int)(s0 - NN * e);
            const int64_t o = e * NS + s_idx[a][b];
            va[it] = kA[o];
            vb[it] = kB[o];
            vc[it] = kC[o];
            kk[it] = eslot[s0 * NN + b];
        }
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int t = lane + 64 * it;
            const int h = t / TPR, u = t - TPR * (t / TPR);
            const int ce = u / NN, b = u - NN * (u / NN);
            if (sl[it] < 0) continue;
            const bool tr = sl[it] % NN > b;  // stored as K_ba
            const float4 A = va[it], B = vb[it];
            Lds &R = s_row[w][h];
            if (tr) {
                R.c[ce][b][0] = make_float4(A.x, A.w, B.z, A.y);
                R.c[ce][b][1] = make_float4(B.x, B.w, A.z, B.y);
            } else {
                R.c[ce][b][0] = A;
                R.c[ce][b][1] = B;
            }
            R.c8[ce][b] = vc[it];
            if (kk[it] < (unsigned)SPR) R.src[kk[it]][ce] = (uint8_t)b;
        }
        wave_sync();
#pragma unroll
        for (int jj = 0; jj < KPL; jj++) {
            if (!sv[jj]) continue;
            const Lds &R = s_row[w][sh[jj]];
            uint32_t srcw[CH / 4];
#pragma unroll
            for (int q = 0; q < CH / 4; q++)
                srcw[q] = reinterpret_cast<const uint32_t *>(&R.src[sk[jj]][0])[q];
#pragma unroll
            for (int ce = 0; ce < CH; ce++) {
                const unsigned bb = (srcw[ce / 4] >> (8 * (ce % 4))) & 0xffu;
                const float4 v0 = R.c[ce][bb][0], v1 = R.c[ce][bb][1];
                const float vv[9] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w, R.c8[ce][bb]};
#pragma unroll
                for (int r = 0; r < 9; r++) acc[jj][r] += vv[r];
            }
        }
        wave_sync();
    }
#pragma unroll
    for (int jj = 0; jj < KPL; jj++) {
        if (!sv[jj]) continue;
        float v[9];
#pragma unroll
        for (int r = 0; r < 9; r++)
            v[r] = ((mask[jj] >> (r / 3)) & (mask[jj] >> (4 + r % 3)) & 1u) ? acc[jj][r] : 0.f;
        const int64_t pos = spos[jj];
        *reinterpret_cast<float4 *>(bval + 4 * pos) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4 *>(bval + 4 * nsl + 4 * pos) = make_float4(v[4], v[5], v[6], v[7]);
        *reinterpret_cast<float2 *>(bval + 8 * nsl + 2 * pos) = make_float2(v[8], __int_as_float(ccol[jj]));
        *reinterpret_cast<float2 *>(lval + 2 * pos) = make_float2(v[8], __int_as_float(cloc[jj]));
    }
}

template <int NN, int NG, int NGP, int WPB, bool IVO = false>
static void launch_grad(axg_mesh *m, hipStream_t s, int32_t *d_err) {
    constexpr int EPW = 64 / NG;
    const int64_t blocks = (m->nelem + EPW * WPB - 1) / (EPW * WPB);
    AXG_REQUIRE(blocks < INT32_MAX, AXG_E_INVALID, "FP32 assembly: too many elements");
    hipLaunchKernelGGL((elem_grad32_kernel<NN, NG, NGP, WPB, IVO>), dim3((unsigned)blocks),
                       dim3(64 * WPB), 0, s, m->nelem, m->conn.p, m->xyz.p, m->e.p, m->nu.p,
                       m->grad32.p, reinterpret_cast<float2 *>(m->lm32.p), d_err);
    AXG_HIP(hipGetLastError());
}

template <int NN, int NGP, int CH, int KPL, int RPW, int WPB, int MINW = 1, bool IVG = NN == 8>
static void launch_rows(axg_mesh *m, hipStream_t s) {
    const int64_t ngrp = (m->nb + WPB * RPW - 1) / (WPB * RPW);
    const int64_t per_xcd = (ngrp + 7) / 8;
    AXG_REQUIRE(8 * per_xcd < INT32_MAX, AXG_E_INVALID, "FP32 assembly: too many block rows");
    hipLaunchKernelGGL((row_asm32_kernel<NN, NGP, CH, KPL, RPW, WPB, MINW, IVG>), dim3((unsigned)(8 * per_xcd)),
                       dim3(64 * WPB), 0, s, m->nb, per_xcd, m->b_incw, m->b_inc.p,
                       m->grad32.p, reinterpret_cast<const float2 *>(m->lm32.p), m->eslot.p,
                       m->bs_mask.p, m->bs_sptr.p, m->bs_width.p, m->bs_col.p,
                       m->bs_loc.p, m->bs_nslots, m->bval.p, m->bval_loc.p);
    AXG_HIP(hipGetLastError());
}

template <int NN, int NG, int EPI, int WPB>
static void launch_blk(axg_mesh *m, hipStream_t s, int32_t *d_err) {
    const int64_t blocks = (m->nelem + EPI * WPB - 1) / (EPI * WPB);
    AXG_REQUIRE(blocks < INT32_MAX, AXG_E_INVALID, "FP32 assembly: too many elements");
    hipLaunchKernelGGL((elem_blk32_kernel<NN, NG, EPI, WPB>), dim3((unsigned)blocks), dim3(64 * WPB),
                       0, s, m->nelem, m->conn.p, m->xyz.p, m->e.p, m->nu.p, m->kb32.p, d_err);
    AXG_HIP(hipGetLastError());
}

template <int NN, int CH, int KPL, int RPW, int WPB, int MINW = 1>
static void launch_gather_rows(axg_mesh *m, hipStream_t s) {
    const int64_t ngrp = (m->nb + WPB * RPW - 1) / (WPB * RPW);
    const int64_t per_xcd = (ngrp + 7) / 8;
    AXG_REQUIRE(8 * per_xcd < INT32_MAX, AXG_E_INVALID, "FP32 assembly: too many block rows");
    hipLaunchKernelGGL((row_gather32_kernel<NN, CH, KPL, RPW, WPB, MINW>), dim3((unsigned)(8 * per_xcd)),
                       dim3(64 * WPB), 0, s, m->nb, per_xcd, m->b_incw, m->b_inc.p, m->kb32.p,
                       m->nelem, m->eslot.p, m->bs_mask.p, m->bs_sptr.p, m->bs_width.p,
                       m->bs_col.p, m->bs_loc.p, m->bs_nslots, m->bval.p, m->bval_loc.p);
    AXG_HIP(hipGetLastError());
}

constexpr int ngp_of(int ng) { return (3 * ng + 11) / 12 * 12; }

int asm32_mode_of(const axg_mesh *m) {
    if (m->asm32_mode >= 0) return m->asm32_mode;
    return m->nn == 8 ? 0 : 1;
}

bool launch_row_asm32(axg_mesh *m, hipStream_t s, int32_t *d_err, hipEvent_t mid) {
    if (m->nelem == 0 || m->nb == 0) return false;
    const int wmax = m->bs_maxw;
    AXG_REQUIRE(wmax <= 255, AXG_E_INVALID,
                "FP32 assembly: a block row couples more than 255 nodes (" + std::to_string(wmax) +
                    ")");
    const int nn = m->nn;
    if (asm32_mode_of(m) == 1) {
        // element blocks on the matrix cores, then the block rows gathered from them
        m->kb32.ensure((size_t)9 * (nn * (nn + 1) / 2) * m->nelem);
        if (nn == 8) {
            launch_blk<8, 8, 8, 4>(m, s, d_err);
            AXG_HIP(hipEventRecord(mid, s));
            if (wmax <= 32) launch_gather_rows<8, 4, 1, 2, 4>(m, s);
            else if (wmax <= 64) launch_gather_rows<8, 8, 1, 1, 4>(m, s);
            else launch_gather_rows<8, 8, 4, 1, 4>(m, s);
        } else {
            launch_blk<20, 27, 2, 4>(m, s, d_err);
            AXG_HIP(hipEventRecord(mid, s));
            if (wmax <= 128) launch_gather_rows<20, 8, 2, 1, 4>(m, s);
            else launch_gather_rows<20, 8, 4, 1, 4>(m, s);
        }
        return true;
    }
    // hex8: the eight s_g J_g^-1 per element (72 floats); hex20: the gradient rows
    m->grad32.ensure((size_t)m->nelem * (nn == 8 ? 72 : nn * ngp_of(27)));
    m->lm32.ensure(2 * (size_t)m->nelem);
    if (nn == 8) {
        launch_grad<8, 8, ngp_of(8), 4, true>(m, s, d_err);
        AXG_HIP(hipEventRecord(mid, s));
        if (wmax <= 32) launch_rows<8, ngp_of(8), 8, 1, 2, 4, 4>(m, s);
        else if (wmax <= 64) launch_rows<8, ngp_of(8), 8, 1, 1, 4>(m, s);
        else launch_rows<8, ngp_of(8), 8, 4, 1, 4>(m, s);
    } else {
        launch_grad<20, 27, ngp_of(27), 1>(m, s, d_err);
        AXG_HIP(hipEventRecord(mid, s));
        if (wmax <= 128) launch_rows<20, ngp_of(27), 8, 2, 1, 4>(m, s);
        else launch_rows<20, ngp_of(27), 8, 4, 1, 4>(m, s);
    }
    return true;
}

// FP32 values of an operator taken from a caller's CSR: each topological
// slot's FP64 value (its contributions summed in order) rounded once and
// stored at its BSR-3 plane position.
__global__ void csr_to_bsr32_kernel(const uint32_t *__restrict__ slotptr,
                                    const uint32_t *__restrict__ contrib,
                                    const double *__restrict__ ke,
                                    const int32_t *__restrict__ entry_of_slot, int64_t nslot,
                                    float *__restrict__ bval) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nslot;
         s += (int64_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (uint32_t c = slotptr[s]; c < slotptr[s + 1]; c++) v += ke[contrib[c]];
        bval[entry_of_slot[s]] = (float)v;
    }
}

void launch_csr_to_bsr32(axg_mesh *m, hipStream_t s) {
    const int64_t nslot = m->nnz_topo;
    if (nslot == 0) return;
    const int64_t blocks = std::min<int64_t>((nslot + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(csr_to_bsr32_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                       m->slotptr_t.p, m->contrib_t.p, m->ke.p, m->entry_of_slot.p, nslot,
                       m->bval.p);
    AXG_HIP(hipGetLastError());
}

}  // namespace axg
