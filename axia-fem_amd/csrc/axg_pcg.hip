// FP64 preconditioned conjugate gradients on gfx950.
//
// The reference's Jacobi PCG (linalg.rs:194-259): x0 = 0, recursive residual,
// stop when ||r|| <= tol ||b||, "not SPD" when |p.Ap| < 1e-30, "not
// converged" when the final ||r|| > 10 tol ||b||, z = r / diag. Each
// iteration is three kernels:
//   spmv_dot   Ap = A p and p.Ap (last workgroup forms alpha)
//   update_xr  x += alpha p, r -= alpha Ap, ||r||^2 and, with Jacobi, z and
//              r.z (last workgroup: stop decision and beta)
//   update_p   p = z + beta p
// (block-Jacobi adds a precond_dot kernel: its z needs the other two DOFs of
// the node). Scalars live in device memory, so batches of iterations run
// without host round trips; a batch is captured once into a hipGraph and
// replayed until the device-side "done" flag is set.
//
// SpMV: L lanes per CSR row, each lane keeping U loads in flight per pass
// (val and column streamed with non-temporal loads, x gathered through the
// caches). (L, U) is picked per matrix by timing the candidates.
#include <cmath>
#include <cstring>

#include <hip/hip_ext.h>

#include "axg_internal.h"
#include "axg_pcg.h"
#include "axg_reduce.h"
#include "axg_mixed.h"

namespace axg {

constexpr int BS = 256;  // threads per workgroup

// ----------------------------------------------------------------- SpMV
// Row sum of one CSR row by the L lanes of a group; valid in lane 0.
template <int L, int U>
__device__ inline double row_dot(const int32_t *__restrict__ colind, const double *__restrict__ val,
                                 const double *__restrict__ x, int32_t k0, int32_t k1, int lane) {
    double s = 0.0;
    for (int32_t base = k0 + lane; base < k1; base += L * U) {
        int32_t c[U];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            int32_t k = base + u * L;
            bool ok = k < k1;
            c[u] = ok ? __builtin_nontemporal_load(&colind[k]) : 0;
            v[u] = ok ? __builtin_nontemporal_load(&val[k]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) s += v[u] * x[c[u]];
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) s += __shfl_down(s, off, L);
    return s;
}

template <int L, int U>
__global__ void __launch_bounds__(BS) spmv_dot_kernel(int64_t n, const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ colind,
                                                     const double *__restrict__ val,
                                                     const double *__restrict__ p,
                                                     double *__restrict__ ap,
                                                     double *__restrict__ part, PcgScalars *sc) {
    if (sc->done) return;
    const int lane = threadIdx.x % L;
    const int64_t group = ((int64_t)blockIdx.x * BS + threadIdx.x) / L;
    const int64_t ngroups = (int64_t)gridDim.x * BS / L;
    double dot[1] = {0.0};
    for (int64_t row = group; row < n; row += ngroups) {
        double s = row_dot<L, U>(colind, val, p, rowptr[row], rowptr[row + 1], lane);
        if (lane == 0) {
            ap[row] = s;
            dot[0] += p[row] * s;
        }
    }
    group_sums<1, BS>(dot);
    if (threadIdx.x == 0) part[blockIdx.x] = dot[0];
}

template <int L, int U>
__global__ void __launch_bounds__(BS) spmv_kernel(int64_t n, const int32_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ colind,
                                                 const double *__restrict__ val,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ y) {
    const int lane = threadIdx.x % L;
    const int64_t group = ((int64_t)blockIdx.x * BS + threadIdx.x) / L;
    const int64_t ngroups = (int64_t)gridDim.x * BS / L;
    for (int64_t row = group; row < n; row += ngroups) {
        double s = row_dot<L, U>(colind, val, x, rowptr[row], rowptr[row + 1], lane);
        if (lane == 0) y[row] = s;
    }
}

// ------------------------------------------------------ sliced ELL
// One thread per row; the 64 rows of a slice are the 64 lanes of a
// wavefront, so entry j of all rows is one coalesced 64-wide load, and the
// row sum needs no cross-lane reduction. U entries per lane are loaded before
// the x gathers are issued.
template <int U>
__device__ inline double sell_row(const int32_t *__restrict__ col, const double *__restrict__ val,
                                  const double *__restrict__ x, int64_t base, int w) {
    // (prefetching the next batch before this batch's gathers measured 3 %
    // slower here: the FP64 loop already keeps enough loads in flight)
    double acc = 0.0;
    int j = 0;
    for (; j + U <= w; j += U) {
        int32_t c[U];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            c[u] = __builtin_nontemporal_load(&col[base + 64 * (int64_t)(j + u)]);
            v[u] = __builtin_nontemporal_load(&val[base + 64 * (int64_t)(j + u)]);
        }
#pragma unroll
        for (int u = 0; u < U; u++) acc += v[u] * x[c[u]];
    }
    for (; j < w; j++) acc += val[base + 64 * (int64_t)j] * x[col[base + 64 * (int64_t)j]];
    return acc;
}

template <int U>
__global__ void __launch_bounds__(BS) sell_spmv_dot_kernel(int64_t n, const int64_t *__restrict__ sptr,
                                                          const int32_t *__restrict__ width,
                                                          const int32_t *__restrict__ col,
                                                          const double *__restrict__ val,
                                                          const double *__restrict__ p,
                                                          double *__restrict__ ap,
                                                          double *__restrict__ part,
                                                          PcgScalars *sc) {
    if (sc->done) return;
    double dot[1] = {0.0};
    const RowBand rb = row_band(n, BS);
    for (int64_t row = rb.first; row < rb.end; row += rb.step) {
        const int64_t sl = row >> 6;
        double s = sell_row<U>(col, val, p, sptr[sl] + (row & 63), width[sl]);
        ap[row] = s;
        dot[0] += p[row] * s;
    }
    group_sums<1, BS>(dot);
    if (threadIdx.x == 0) part[rb.id] = dot[0];
}

template <int U>
__global__ void __launch_bounds__(BS) sell_spmv_kernel(int64_t n, const int64_t *__restrict__ sptr,
                                                      const int32_t *__restrict__ width,
                                                      const int32_t *__restrict__ col,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ y) {
    const RowBand rb = row_band(n, BS);
    for (int64_t row = rb.first; row < rb.end; row += rb.step) {
        const int64_t sl = row >> 6;
        y[row] = sell_row<U>(col, val, x, sptr[sl] + (row & 63), width[sl]);
    }
}

// Build the sliced-ELL columns and the source slot of every entry from the
// CSR pattern (one thread per row).
__global__ void sell_build_kernel(int64_t n, const int32_t *__restrict__ rowptr,
                                  const int32_t *__restrict__ colind,
                                  const int64_t *__restrict__ sptr,
                                  const int32_t *__restrict__ width, int32_t *__restrict__ scol,
                                  int32_t *__restrict__ ssrc) {
    int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const int64_t sl = row >> 6, base = sptr[sl] + (row & 63);
    const int32_t k0 = rowptr[row], len = rowptr[row + 1] - k0, w = width[sl];
    for (int j = 0; j < w; j++) {
        int64_t p = base + 64 * (int64_t)j;
        scol[p] = j < len ? colind[k0 + j] : (int32_t)row;
        ssrc[p] = j < len ? k0 + j : -1;
    }
}

// Rows of the last slice past n: columns pointing at row 0, value 0.
__global__ void sell_tail_kernel(int64_t n, const int64_t *__restrict__ sptr,
                                 const int32_t *__restrict__ width, int32_t *__restrict__ scol,
                                 int32_t *__restrict__ ssrc) {
    const int64_t sl = (n - 1) >> 6;
    const int lane = threadIdx.x;
    if (lane >= 64 || (sl * 64 + lane) < n) return;
    for (int j = 0; j < width[sl]; j++) {
        int64_t p = sptr[sl] + lane + 64 * (int64_t)j;
        scol[p] = 0;
        ssrc[p] = -1;
    }
}

__global__ void sell_fill_kernel(int64_t nslots, const int32_t *__restrict__ ssrc,
                                 const double *__restrict__ val, double *__restrict__ sval) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nslots;
         p += (int64_t)gridDim.x * blockDim.x) {
        int32_t k = ssrc[p];
        sval[p] = k >= 0 ? val[k] : 0.0;
    }
}

// Candidate (L, U) shapes; the index is what the autotuner stores.
struct SpmvShape {
    int L, U;
};
static const SpmvShape kShapes[] = {{8, 4}, {16, 2}, {16, 4}, {32, 1}, {32, 2}, {32, 3},
                                    {64, 1}, {64, 2}, {64, 4}};
static const int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

#define AXG_SHAPE_SWITCH(IDX, MACRO) \
    switch (IDX) {                   \
        case 0: MACRO(8, 4); break;  \
        case 1: MACRO(16, 2); break; \
        case 2: MACRO(16, 4); break; \
        case 3: MACRO(32, 1); break; \
        case 4: MACRO(32, 2); break; \
        case 5: MACRO(32, 3); break; \
        case 6: MACRO(64, 1); break; \
        case 7: MACRO(64, 2); break; \
        default: MACRO(64, 4); break; \
    }

// Grid-stride launches use the same number of workgroups on every CU; 4 per CU
// measured best for the sliced-ELL operators on MI355X (8: -3 %, 6: -8 %).
static int grid_for(axg_ctx *ctx, int64_t work_threads, int per_cu) {
    int64_t g = (work_threads + BS - 1) / BS;
    int64_t cap = (int64_t)ctx->cu_count * per_cu;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

static int default_shape(const CsrView &A) {
    double avg = A.n > 0 ? (double)A.nnz / (double)A.n : 1.0;
    if (avg <= 24) return 0;
    if (avg <= 64) return 2;
    if (avg <= 100) return 5;
    return 8;
}

static void launch_spmv_shape(axg_ctx *ctx, int shape, const CsrView &A, const double *x,
                              double *y) {
    const int L = kShapes[shape].L;
    const int grid = grid_for(ctx, A.n * L, 8);
    hipStream_t s = ctx->stream;
#define AXG_L(LL, UU)                                                                            \
    hipLaunchKernelGGL((spmv_kernel<LL, UU>), dim3(grid), dim3(BS), 0, s, A.n, A.rowptr,         \
                       A.colind, A.val, x, y)
    AXG_SHAPE_SWITCH(shape, AXG_L)
#undef AXG_L
}

// Autotuner: time every candidate on this matrix (a few launches each) and
// keep the fastest. Called once per matrix; small matrices keep the default.
int tune_spmv(axg_ctx *ctx, const CsrView &A) {
    if (A.nnz < 2000000) return default_shape(A);
    hipStream_t s = ctx->stream;
    DevBuf<double> x(A.n), y(A.n);
    AXG_HIP(hipMemsetAsync(x.p, 0, sizeof(double) * A.n, s));
    int best = default_shape(A);
    float best_ms = 1e30f;
    for (int c = 0; c < kNumShapes; c++) {
        launch_spmv_shape(ctx, c, A, x.p, y.p);  // warm
        AXG_HIP(hipEventRecord(ctx->ev0, s));
        for (int r = 0; r < 5; r++) launch_spmv_shape(ctx, c, A, x.p, y.p);
        AXG_HIP(hipEventRecord(ctx->ev1, s));
        AXG_HIP(hipEventSynchronize(ctx->ev1));
        float ms = 0.f;
        AXG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        if (ms < best_ms) best_ms = ms, best = c;
    }
    AXG_HIP(hipGetLastError());
    return best;
}

static int shape_of(axg_ctx *ctx, const CsrView &A) {
    if (A.shape && *A.shape >= 0) return *A.shape;
    int sh = tune_spmv(ctx, A);
    if (A.shape) *A.shape = sh;
    return sh;
}

void spmv_fp64(axg_ctx *ctx, const CsrView &A, const double *x, double *y) {
    if (A.sell_val) {
        const int grid = grid_for(ctx, A.n, 4);
        hipLaunchKernelGGL((sell_spmv_kernel<4>), dim3(grid), dim3(BS), 0, ctx->stream, A.n,
                           A.sell_sptr, A.sell_width, A.sell_col, A.sell_val, x, y);
    } else {
        launch_spmv_shape(ctx, shape_of(ctx, A), A, x, y);
    }
    AXG_HIP(hipGetLastError());
}

void sell_layout(const std::vector<int64_t> &row_len, SellHost &out) {
    const int64_t n = (int64_t)row_len.size(), ns = (n + 63) / 64;
    out.sptr.assign(ns + 1, 0);
    out.width.assign(ns, 0);
    for (int64_t s = 0; s < ns; s++) {
        int64_t w = 0;
        for (int64_t r = 64 * s; r < std::min(n, 64 * s + 64); r++) w = std::max(w, row_len[r]);
        out.width[s] = (int32_t)w;
        out.sptr[s + 1] = out.sptr[s] + 64 * w;
    }
    out.nslots = out.sptr[ns];
}

// Sliced-ELL copy of the mesh's active K_ff: layout once per pattern, values
// refreshed after every assembly.
void ensure_sell(axg_mesh *m) {
    hipStream_t s = m->ctx->stream;
    CsrView A = active_csr(m);
    const int64_t n = A.n;
    if (n == 0) return;
    if (!m->sell_valid) {
        std::vector<int32_t> rp(n + 1);
        AXG_HIP(hipMemcpyAsync(rp.data(), A.rowptr, 4 * (n + 1), hipMemcpyDeviceToHost, s));
        AXG_HIP(hipStreamSynchronize(s));
        std::vector<int64_t> len(n);
        for (int64_t i = 0; i < n; i++) len[i] = rp[i + 1] - rp[i];
        SellHost L;
        sell_layout(len, L);
        AXG_REQUIRE(L.nslots < INT32_MAX, AXG_E_INVALID, "sliced-ELL operator exceeds 2^31 slots");
        m->sell_nslots = L.nslots;
        m->sell_sptr.alloc(L.sptr.size());
        m->sell_width.alloc(L.width.size());
        m->sell_col.alloc(L.nslots);
        m->sell_src.alloc(L.nslots);
        m->sell_val.alloc(L.nslots);
        AXG_HIP(hipMemcpyAsync(m->sell_sptr.p, L.sptr.data(), 8 * L.sptr.size(),
                               hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(m->sell_width.p, L.width.data(), 4 * L.width.size(),
                               hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(sell_build_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           n, A.rowptr, A.colind, m->sell_sptr.p, m->sell_width.p, m->sell_col.p,
                           m->sell_src.p);
        hipLaunchKernelGGL(sell_tail_kernel, dim3(1), dim3(64), 0, s, n, m->sell_sptr.p,
                           m->sell_width.p, m->sell_col.p, m->sell_src.p);
        AXG_HIP(hipGetLastError());
        m->sell_valid = true;
    }
    if (!m->sell_fresh) {
        int64_t blocks = std::min<int64_t>((m->sell_nslots + 255) / 256, 1 << 20);
        hipLaunchKernelGGL(sell_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                           m->sell_nslots, m->sell_src.p, A.val, m->sell_val.p);
        AXG_HIP(hipGetLastError());
        m->sell_fresh = true;
    }
}

// The solver's view of the mesh operator: active CSR plus its sliced-ELL copy.
CsrView solver_csr(axg_mesh *m) {
    ensure_sell(m);
    CsrView A = active_csr(m);
    A.sell_sptr = m->sell_sptr.p;
    A.sell_width = m->sell_width.p;
    A.sell_col = m->sell_col.p;
    A.sell_val = m->sell_val.p;
    return A;
}

// ------------------------------------------------------- vector kernels
// z = M^-1 r for one DOF: Jacobi (r / d, linalg.rs:213-216) or the 3x3
// node-block inverse (block of DOF i, row d).
struct Precond64 {
    const double *dsel;
    const double *binv;
    const int32_t *blk_of_free;
    const int32_t *free_of_blk;
    __device__ inline double apply(const double *r, int64_t i) const {
        if (!binv) return r[i] / dsel[i];
        int32_t bd = blk_of_free[i];
        int32_t b = bd / 3, d = bd % 3;
        const double *B = binv + 9 * (int64_t)b + 3 * d;
        double z = 0.0;
#pragma unroll
        for (int e = 0; e < 3; e++) {
            int32_t f = free_of_blk[3 * b + e];
            if (f >= 0) z += B[e] * r[f];
        }
        return z;
    }
};

template <bool FUSED>
__global__ void __launch_bounds__(BS) update_xr_kernel(int64_t n, double *__restrict__ x,
                                                      double *__restrict__ r,
                                                      const double *__restrict__ p,
                                                      const double *__restrict__ ap,
                                                      double *__restrict__ z,
                                                      const double *__restrict__ dsel,
                                                      double *__restrict__ part, PcgScalars *sc) {
    if (sc->done) return;
    const double alpha = sc->alpha;
    double acc[2] = {0.0, 0.0};  // ||r||^2, r.z
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        x[i] += alpha * p[i];
        double ri = r[i] - alpha * ap[i];
        r[i] = ri;
        acc[0] += ri * ri;
        if (FUSED) {
            double zi = ri / dsel[i];
            z[i] = zi;
            acc[1] += ri * zi;
        }
    }
    group_sums<2, BS>(acc);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = acc[0];
        part[2 * blockIdx.x + 1] = acc[1];
    }
}

// ----------------------------------------------------- finalize kernels
// One workgroup: sum the partials of the kernel just before (fixed order)
// and update the scalars. The scalars a kernel needs are loaded before the
// partials and the stop flag is tested after the sum, so the two reads
// overlap instead of costing one memory round trip each.
__global__ void __launch_bounds__(FIN_TB) fin_pap_kernel(const double *__restrict__ part, int nblk,
                                                    PcgScalars *sc) {
    const int done = sc->done;
    const double rz = sc->rz;
    double pap[1];
    fin_sums<1>(part, nblk, pap);
    if (done || threadIdx.x != 0) return;
    sc->pap = pap[0];
    if (fabs(pap[0]) < 1e-30) {  // linalg.rs:231-233
        sc->status = AXG_E_NOT_SPD;
        sc->done = 1;
    } else {
        sc->alpha = rz / pap[0];
    }
}

// ||r||, iteration count, stop decision (linalg.rs:241-245); with FUSED also
// beta = r.z_new / r.z.
template <bool FUSED>
__global__ void __launch_bounds__(FIN_TB) fin_update_kernel(const double *__restrict__ part, int nblk,
                                                       PcgScalars *sc) {
    const int done = sc->done;
    const double rz = sc->rz, tol_abs = sc->tol_abs;
    const int64_t iters = sc->iters + 1, max_iter = sc->max_iter;
    double tot[2];
    fin_sums<2>(part, nblk, tot);
    if (done || threadIdx.x != 0) return;
    const double res = sqrt(tot[0]);
    sc->residual = res;
    sc->iters = iters;
    if (res <= tol_abs) {
        sc->done = 1;
    } else if (iters >= max_iter) {
        sc->done = 1;
        sc->status = AXG_E_NOT_CONVERGED;
    } else if (FUSED) {
        sc->beta = tot[1] / rz;
        sc->rz = tot[1];
    }
}

__global__ void __launch_bounds__(FIN_TB) fin_rz_kernel(const double *__restrict__ part, int nblk,
                                                   PcgScalars *sc, int init) {
    const int done = sc->done;
    const double rz = sc->rz;
    double tot[1];
    fin_sums<1>(part, nblk, tot);
    if (done || threadIdx.x != 0) return;
    if (!init) sc->beta = tot[0] / rz;
    sc->rz = tot[0];
}

__global__ void __launch_bounds__(FIN_TB) fin_init_kernel(const double *__restrict__ part, int nblk,
                                                     PcgScalars *sc, double tol) {
    double tot[1];
    fin_sums<1>(part, nblk, tot);
    if (threadIdx.x == 0) {
        double bn = sqrt(tot[0]);
        sc->bnorm = bn;
        sc->tol_abs = tol * bn;
        sc->residual = bn;
        if (bn < 1e-30) sc->done = 1;  // linalg.rs:206-214
    }
}

// z = M^-1 r and r.z; the last workgroup forms beta (or, at start, rz).
__global__ void __launch_bounds__(BS) precond_dot_kernel(int64_t n, const double *__restrict__ r,
                                                        double *__restrict__ z, Precond64 M,
                                                        double *__restrict__ part,
                                                        PcgScalars *sc, int init) {
    if (sc->done) return;
    double acc[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        double zi = M.apply(r, i);
        z[i] = zi;
        acc[0] += r[i] * zi;
    }
    group_sums<1, BS>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
    (void)init;
}

__global__ void __launch_bounds__(BS) update_p_kernel(int64_t n, double *__restrict__ p,
                                                     const double *__restrict__ z, PcgScalars *sc,
                                                     int init) {
    if (sc->done) return;
    const double beta = init ? 0.0 : sc->beta;
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
        p[i] = init ? z[i] : z[i] + beta * p[i];
}

// x = 0, r = b, ||b|| (last workgroup: tol_abs and the zero right-hand side).
__global__ void __launch_bounds__(BS) init_kernel(int64_t n, const double *__restrict__ b,
                                                 double *__restrict__ x, double *__restrict__ r,
                                                 double *__restrict__ part, PcgScalars *sc,
                                                 double tol) {
    double acc[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        double bi = b[i];
        x[i] = 0.0;
        r[i] = bi;
        acc[0] += bi * bi;
    }
    group_sums<1, BS>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
    (void)sc, (void)tol;
}

// Per-workgroup partial sums of (b - y)^2.
__global__ void __launch_bounds__(BS) residual_part_kernel(int64_t n, const double *__restrict__ b,
                                                          const double *__restrict__ y,
                                                          double *__restrict__ part) {
    double acc[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        const double d = b[i] - y[i];
        acc[0] += d * d;
    }
    group_sums<1, BS>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// sqrt of the partials' fixed-order sum, into sc->pad[0].
__global__ void __launch_bounds__(FIN_TB) fin_residual_kernel(const double *__restrict__ part,
                                                         int nblk, PcgScalars *sc) {
    double tot[1];
    fin_sums<1>(part, nblk, tot);
    if (threadIdx.x == 0) sc->pad[0] = sqrt(tot[0]);
}

// dsel[i] = first (i, i) entry of row i if |d| > 1e-30, else 1.0 (Csr::diag,
// linalg.rs:52-63, and the divisor rule of linalg.rs:213-216).
__global__ void diag_kernel(int64_t n, const int32_t *__restrict__ rowptr,
                            const int32_t *__restrict__ colind, const double *__restrict__ val,
                            double *__restrict__ dsel, double *__restrict__ diag_raw) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; k++)
        if (colind[k] == i) {
            d = val[k];
            break;
        }
    if (diag_raw) diag_raw[i] = d;
    dsel[i] = fabs(d) > 1e-30 ? d : 1.0;
}

void diag_fp64(axg_ctx *ctx, const CsrView &A, double *dsel, double *diag_raw) {
    if (A.n == 0) return;
    hipLaunchKernelGGL(diag_kernel, dim3((unsigned)((A.n + 255) / 256)), dim3(256), 0, ctx->stream,
                       A.n, A.rowptr, A.colind, A.val, dsel, diag_raw);
    AXG_HIP(hipGetLastError());
}

// ------------------------------------------------------------ host driver
// Buffers of one solve: n-vectors r, p, ap, z; part holds 2 max(grid) + 2
// partials (pcg_part_count); sc one PcgScalars; host the host copy of the
// scalars (pinned, or any host memory when graph is false).
struct PcgBuffers {
    double *r = nullptr, *p = nullptr, *ap = nullptr, *z = nullptr, *part = nullptr;
    PcgScalars *sc = nullptr, *host = nullptr;
    bool graph = true;  // replay captured batches of iterations (false: plain launches)
    int shape = -1;     // CSR SpMV shape (-1: tuned on first use and cached in A.shape)
};

static void pcg_fp64_core(axg_ctx *ctx, const CsrView &A, const Precond64Args &pa, const double *b,
                          double *x, double tol, int64_t max_iter, bool timing, PcgResult &res,
                          const PcgBuffers &bf);

void pcg_fp64(axg_ctx *ctx, const CsrView &A, const Precond64Args &pa, const double *b,
              double *x, double tol, int64_t max_iter, bool timing, PcgResult &res) {
    const int64_t n = A.n;
    Workspace &w = ctx->ws;
    w.r.ensure(n);
    w.p.ensure(n);
    w.ap.ensure(n);
    w.z.ensure(n);
    w.part.ensure(pcg_part_count(ctx));
    w.sc.ensure(1);
    ctx->pinned.ensure(sizeof(PcgScalars) * 2);
    PcgBuffers bf;
    bf.r = w.r.p, bf.p = w.p.p, bf.ap = w.ap.p, bf.z = w.z.p, bf.part = w.part.p;
    bf.sc = w.sc.p;
    bf.host = (PcgScalars *)ctx->pinned.p;
    pcg_fp64_core(ctx, A, pa, b, x, tol, max_iter, timing, res, bf);
}

size_t pcg_part_count(axg_ctx *ctx) { return 2 * (size_t)ctx->cu_count * 8 + 2; }

static size_t round32(size_t v) { return (v + 31) / 32 * 32; }

size_t pcg_noalloc_doubles(axg_ctx *ctx, int64_t n) {
    return round32(5 * (size_t)n) + round32(pcg_part_count(ctx)) +
           round32(sizeof(PcgScalars) / sizeof(double));
}

// linalg::pcg with Jacobi on a device CSR (int32 row pointer and columns)
// into caller buffers: no allocation, no graph, no autotuning (the SpMV shape
// follows the mean row length). scratch (pcg_noalloc_doubles(ctx, n) doubles,
// 256-byte aligned): Jacobi divisor, r, p, ap, z, partials, scalars. Returns
// the FP64 ||b - A x||.
double pcg_fp64_noalloc(axg_ctx *ctx, const CsrView &A, const double *b, double *x, double tol,
                        int64_t max_iter, double *scratch, PcgResult &res) {
    const int64_t n = A.n;
    double *dsel = scratch;
    PcgBuffers bf;
    bf.r = scratch + n, bf.p = scratch + 2 * n, bf.ap = scratch + 3 * n, bf.z = scratch + 4 * n;
    bf.part = scratch + round32(5 * (size_t)n);
    bf.sc = reinterpret_cast<PcgScalars *>(bf.part + round32(pcg_part_count(ctx)));
    PcgScalars host;
    bf.host = &host;
    bf.graph = false;
    bf.shape = default_shape(A);
    diag_fp64(ctx, A, dsel, nullptr);
    Precond64Args pa;
    pa.dsel = dsel;
    pcg_fp64_core(ctx, A, pa, b, x, tol, max_iter, false, res, bf);
    // ||b - A x||: A x into ap, squared differences summed by workgroup, then in
    // fixed order by one workgroup into the scalars' pad slot
    hipStream_t s = ctx->stream;
    launch_spmv_shape(ctx, bf.shape, A, x, bf.ap);
    const int grid_v = grid_for(ctx, n, 4);
    hipLaunchKernelGGL(residual_part_kernel, dim3(grid_v), dim3(BS), 0, s, n, b, bf.ap, bf.part);
    hipLaunchKernelGGL(fin_residual_kernel, dim3(1), dim3(FIN_TB), 0, s, bf.part, grid_v, bf.sc);
    AXG_HIP(hipGetLastError());
    AXG_HIP(hipMemcpyAsync(&host, bf.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    return host.pad[0];
}

static void pcg_fp64_core(axg_ctx *ctx, const CsrView &A, const Precond64Args &pa, const double *b,
                          double *x, double tol, int64_t max_iter, bool timing, PcgResult &res,
                          const PcgBuffers &bf) {
    hipStream_t s = ctx->stream;
    const int64_t n = A.n;
    const bool sell = A.sell_val != nullptr;
    const int shape = sell ? 0 : bf.shape >= 0 ? bf.shape : shape_of(ctx, A);
    const int grid_mv = sell ? grid_for(ctx, n, 4) : grid_for(ctx, n * kShapes[shape].L, 8);
    const int grid_v = grid_for(ctx, n, 4);
    Precond64 M{pa.dsel, pa.binv, pa.blk_of_free, pa.free_of_blk};
    const bool fused = pa.binv == nullptr;

    PcgScalars *h = bf.host;
    std::memset((void *)h, 0, sizeof(PcgScalars));
    h->max_iter = max_iter;
    AXG_HIP(hipMemcpyAsync(bf.sc, h, sizeof(PcgScalars), hipMemcpyHostToDevice, s));
    AXG_HIP(hipEventRecord(ctx->ev0, s));
    hipLaunchKernelGGL(init_kernel, dim3(grid_v), dim3(BS), 0, s, n, b, x, bf.r, bf.part, bf.sc,
                       tol);
    hipLaunchKernelGGL(fin_init_kernel, dim3(1), dim3(FIN_TB), 0, s, bf.part, grid_v, bf.sc, tol);
    hipLaunchKernelGGL(precond_dot_kernel, dim3(grid_v), dim3(BS), 0, s, n, bf.r, bf.z, M,
                       bf.part, bf.sc, 1);
    hipLaunchKernelGGL(fin_rz_kernel, dim3(1), dim3(FIN_TB), 0, s, bf.part, grid_v, bf.sc, 1);
    hipLaunchKernelGGL(update_p_kernel, dim3(grid_v), dim3(BS), 0, s, n, bf.p, bf.z, bf.sc, 1);
    AXG_HIP(hipGetLastError());

    // timing: the events are stamped by the SpMV dispatch itself
    auto one_iter = [&](hipEvent_t e0, hipEvent_t e1) {
        if (sell && e0) {
            hipExtLaunchKernelGGL((sell_spmv_dot_kernel<4>), dim3(grid_mv), dim3(BS), 0, s, e0, e1,
                                  0, n, (const int64_t *)A.sell_sptr, (const int32_t *)A.sell_width,
                                  (const int32_t *)A.sell_col, (const double *)A.sell_val,
                                  (const double *)bf.p, bf.ap, bf.part, bf.sc);
        } else if (sell) {
            hipLaunchKernelGGL((sell_spmv_dot_kernel<4>), dim3(grid_mv), dim3(BS), 0, s, n,
                               A.sell_sptr, A.sell_width, A.sell_col, A.sell_val, bf.p, bf.ap,
                               bf.part, bf.sc);
        } else {
#define AXG_L(LL, UU)                                                                            \
    if (e0)                                                                                      \
        hipExtLaunchKernelGGL((spmv_dot_kernel<LL, UU>), dim3(grid_mv), dim3(BS), 0, s, e0, e1, 0, \
                              n, (const int32_t *)A.rowptr, (const int32_t *)A.colind,           \
                              (const double *)A.val, (const double *)bf.p, bf.ap, bf.part,   \
                              bf.sc);                                                           \
    else                                                                                         \
        hipLaunchKernelGGL((spmv_dot_kernel<LL, UU>), dim3(grid_mv), dim3(BS), 0, s, n, A.rowptr, \
                           A.colind, A.val, bf.p, bf.ap, bf.part, bf.sc)
            AXG_SHAPE_SWITCH(shape, AXG_L)
#undef AXG_L
        }
        hipLaunchKernelGGL(fin_pap_kernel, dim3(1), dim3(FIN_TB), 0, s, bf.part, grid_mv, bf.sc);
        if (fused) {
            hipLaunchKernelGGL((update_xr_kernel<true>), dim3(grid_v), dim3(BS), 0, s, n, x, bf.r,
                               bf.p, bf.ap, bf.z, M.dsel, bf.part, bf.sc);
            hipLaunchKernelGGL((fin_update_kernel<true>), dim3(1), dim3(FIN_TB), 0, s, bf.part,
                               grid_v, bf.sc);
        } else {
            hipLaunchKernelGGL((update_xr_kernel<false>), dim3(grid_v), dim3(BS), 0, s, n, x,
                               bf.r, bf.p, bf.ap, bf.z, M.dsel, bf.part, bf.sc);
            hipLaunchKernelGGL((fin_update_kernel<false>), dim3(1), dim3(FIN_TB), 0, s, bf.part,
                               grid_v, bf.sc);
            hipLaunchKernelGGL(precond_dot_kernel, dim3(grid_v), dim3(BS), 0, s, n, bf.r, bf.z,
                               M, bf.part, bf.sc, 0);
            hipLaunchKernelGGL(fin_rz_kernel, dim3(1), dim3(FIN_TB), 0, s, bf.part, grid_v, bf.sc,
                               0);
        }
        hipLaunchKernelGGL(update_p_kernel, dim3(grid_v), dim3(BS), 0, s, n, bf.p, bf.z, bf.sc,
                           0);
    };

    res.spmv_ms = 0.0;
    res.spmv_launches = 0;
    const int batch = n > 200000 ? 16 : 64;
    hipGraphExec_t gexec = nullptr;
    if (!timing && bf.graph) {
        hipGraph_t graph;
        AXG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < batch; i++) one_iter(nullptr, nullptr);
        AXG_HIP(hipStreamEndCapture(s, &graph));
        AXG_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        AXG_HIP(hipGraphDestroy(graph));
    } else if (timing) {
        while (ctx->timing_events.size() < 2 * (size_t)batch) {
            hipEvent_t e;
            AXG_HIP(hipEventCreate(&e));
            ctx->timing_events.push_back(e);
        }
    }
    int64_t launched = 0;
    for (;;) {
        AXG_HIP(hipMemcpyAsync(h, bf.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
        AXG_HIP(hipStreamSynchronize(s));
        if (h->done || launched >= max_iter) break;
        if (timing) {
            int64_t before = h->iters;
            for (int i = 0; i < batch; i++)
                one_iter(ctx->timing_events[2 * i], ctx->timing_events[2 * i + 1]);
            AXG_HIP(hipMemcpyAsync(h, bf.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
            AXG_HIP(hipStreamSynchronize(s));
            // SpMV launches that did work: the iterations this batch counted
            int64_t did = h->iters - before;
            for (int64_t i = 0; i < did && i < batch; i++) {
                float ms = 0.f;
                AXG_HIP(hipEventElapsedTime(&ms, ctx->timing_events[2 * i],
                                            ctx->timing_events[2 * i + 1]));
                res.spmv_ms += ms;
                res.spmv_launches++;
            }
        } else if (gexec) {
            AXG_HIP(hipGraphLaunch(gexec, s));
        } else {
            for (int i = 0; i < batch; i++) one_iter(nullptr, nullptr);
            AXG_HIP(hipGetLastError());
        }
        launched += batch;
    }
    AXG_HIP(hipEventRecord(ctx->ev1, s));
    AXG_HIP(hipEventSynchronize(ctx->ev1));
    if (gexec) AXG_HIP(hipGraphExecDestroy(gexec));
    float ms = 0.f;
    AXG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    res.solve_ms = ms;
    res.iters = h->iters;
    res.residual = h->residual;
    res.bnorm = h->bnorm;
    res.status = h->status;
    res.spmv_shape_L = sell ? 1 : kShapes[shape].L;
    res.spmv_shape_U = sell ? 4 : kShapes[shape].U;
    if (res.status == AXG_E_NOT_CONVERGED && h->residual <= 10.0 * tol * h->bnorm)
        res.status = AXG_OK;  // linalg.rs:253: only > 10 tol ||b|| is an error
    if (res.status == AXG_OK && !h->done && h->bnorm >= 1e-30) res.status = AXG_E_NOT_CONVERGED;
}

}  // namespace axg
