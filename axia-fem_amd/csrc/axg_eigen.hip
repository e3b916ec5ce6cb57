// Lowest modes of K phi = lambda M phi with a diagonal (lumped) M by inverse
// subspace iteration, after eigen.rs:12-111 (subspace_gen): nvec = nmodes + 4
// vectors from the same seeded generator, at most 40 sweeps of
//   Y = K^-1 M X (one multi-right-hand-side PCG solve for all vectors),
//   M-orthonormalise Y (modified Gram-Schmidt), Kr = Y^T K Y, Kr Q = Q diag(l)
//   (cyclic Jacobi on the host, nvec x nvec), X = Y Q, M-orthonormalise X,
// stopping once every Ritz value moved by at most 1e-8 (1 + |l|) after sweep 5.
// The vectors live on the device (nvec contiguous vectors of n_free); the
// direct solve of the reference becomes the GPU PCG chosen by the options.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "axg_internal.h"
#include "axg_pcg.h"
#include "axg_reduce.h"

namespace axg {

constexpr int EBS = 256;   // workgroup size of the vector kernels
constexpr int NVMAX = 64;  // most vectors one multi-dot handles

// Partials of sum_i a_i w_i b_i (w may be null: weight 1), one per workgroup.
__global__ void __launch_bounds__(EBS) wdot_kernel(int64_t n, const double *__restrict__ a,
                                                   const double *__restrict__ w,
                                                   const double *__restrict__ b,
                                                   double *__restrict__ part) {
    double s[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * EBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * EBS)
        s[0] += w ? a[i] * w[i] * b[i] : a[i] * b[i];
    group_sums<1, EBS>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Partials of y_k . v for every vector k < nvec (y: nvec vectors of n).
template <int NV>
__global__ void __launch_bounds__(EBS) multi_dot_kernel(int64_t n, int nvec,
                                                        const double *__restrict__ y,
                                                        const double *__restrict__ v,
                                                        double *__restrict__ part) {
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) s[k] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * EBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * EBS) {
        const double vi = v[i];
#pragma unroll
        for (int k = 0; k < NV; k++)
            if (k < nvec) s[k] += y[(int64_t)k * n + i] * vi;
    }
    group_sums<NV, EBS>(s);
    if (threadIdx.x == 0)
        for (int k = 0; k < nvec; k++) part[(size_t)blockIdx.x * nvec + k] = s[k];
}

// a -= c b  (c from the host)
__global__ void axpy_kernel(int64_t n, double c, const double *__restrict__ b,
                            double *__restrict__ a) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] -= c * b[i];
}

// y = x / c
__global__ void div_copy_kernel(int64_t n, double c, const double *__restrict__ x,
                                double *__restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i] / c;
}

__global__ void scale_kernel(int64_t n, double c, double *__restrict__ a) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] /= c;
}

// out_k = (m .* x_k) * scale for every vector (scale: a power of two)
__global__ void mass_apply_kernel(int64_t n, int nvec, const double *__restrict__ m,
                                  const double *__restrict__ x, double scale,
                                  double *__restrict__ out) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nvec) return;
    out[t] = (m[t % n] * x[t]) * scale;
}

// x_j = sum_k y_k q[k][j] over the nonzero q (eigen.rs:63-74)
__global__ void combine_kernel(int64_t n, int nvec, const double *__restrict__ y,
                               const double *__restrict__ q, double *__restrict__ x) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int j = 0; j < nvec; j++) {
        double acc = 0.0;
        for (int k = 0; k < nvec; k++) {
            const double c = q[k * nvec + j];
            if (c == 0.0) continue;
            acc += y[(int64_t)k * n + i] * c;
        }
        x[(int64_t)j * n + i] = acc;
    }
}

// Cyclic-by-largest Jacobi rotations of a dense symmetric matrix (row-major),
// eigen.rs:208-266; q receives the eigenvectors as columns.
static void jacobi_eigen(std::vector<double> &a, int n, std::vector<double> &vals,
                         std::vector<double> &q) {
    q.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) q[i * n + i] = 1.0;
    for (int it = 0; it < 40 * std::max(n, 1); it++) {
        int p = 0, r = 1;
        double best = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++)
                if (std::fabs(a[i * n + j]) > best) best = std::fabs(a[i * n + j]), p = i, r = j;
        if (best < 1e-14) break;
        const double app = a[p * n + p], aqq = a[r * n + r], apq = a[p * n + r];
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int i = 0; i < n; i++) {
            if (i == p || i == r) continue;
            const double aip = a[i * n + p], aiq = a[i * n + r];
            a[i * n + p] = a[p * n + i] = c * aip - s * aiq;
            a[i * n + r] = a[r * n + i] = s * aip + c * aiq;
        }
        a[p * n + p] = c * c * app - 2.0 * s * c * apq + s * s * aqq;
        a[r * n + r] = s * s * app + 2.0 * s * c * apq + c * c * aqq;
        a[p * n + r] = a[r * n + p] = 0.0;
        for (int i = 0; i < n; i++) {
            const double vp = q[i * n + p], vq = q[i * n + r];
            q[i * n + p] = c * vp - s * vq;
            q[i * n + r] = s * vp + c * vq;
        }
    }
    vals.resize(n);
    for (int i = 0; i < n; i++) vals[i] = a[i * n + i];
}

// The solves of a subspace sweep: columns R_j (n each) with tolerance tl[j] < 1
// are packed, solved together (multi-right-hand-side FP64 block PCG for the
// Jacobi preconditioner, the mixed solver at the tightest of the tolerances
// for mixed precision, one FP64 PCG per column otherwise) and added to the
// iterate as Y_j += scl[j] x_j.
struct ColumnSolver {
    axg_mesh *m;
    axg_solve_opts o;
    double *P, *D;
    BlockPcg *bp = nullptr;  // the mesh's own block PCG (it outlives this solver)
    int64_t iters = 0;
    double ms = 0.0, res = 0.0, rel = 0.0;
    ColumnSolver(axg_mesh *mesh, const axg_solve_opts &opts, double *pack, double *sol)
        : m(mesh), o(opts), P(pack), D(sol) {
        if (o.precision != AXG_PREC_MIXED) {
            const bool blk3 = o.precond == AXG_PRECOND_BLOCK3;
            BlockPcg *&mine = blk3 ? m->bpcg3 : m->bpcg;
            if (!mine) mine = block_pcg_create(m, o.max_iter, blk3);
            block_pcg_refresh(mine, o.max_iter);
            bp = mine;
        }
    }
    ColumnSolver(const ColumnSolver &) = delete;
    ColumnSolver &operator=(const ColumnSolver &) = delete;

    void solve(const double *R, double *Y, const std::vector<double> &scl,
               const std::vector<double> &tl) {
        hipStream_t s = m->ctx->stream;
        const int64_t n = m->n_free;
        const unsigned gn = (unsigned)((n + 255) / 256);
        iters = 0, ms = 0.0, res = 0.0, rel = 0.0;
        std::vector<int> cols;
        std::vector<double> ctol;
        for (int j = 0; j < (int)tl.size(); j++)
            if (tl[j] < 1.0) {
                AXG_HIP(hipMemcpyAsync(P + (size_t)cols.size() * n, R + (size_t)j * n, 8 * n,
                                       hipMemcpyDeviceToDevice, s));
                cols.push_back(j);
                ctol.push_back(tl[j]);
            }
        const int k = (int)cols.size();
        if (k == 0) return;
        if (bp) {
            PcgResult pr;
            block_pcg_solve(bp, P, k, ctol.data(), D, pr);
            if (pr.status != AXG_OK) throw Error{pr.status, pcg_error(pr)};
            iters = pr.iters, ms = pr.solve_ms, res = pr.residual;
            rel = pr.residual / std::max(pr.bnorm, 1e-300);
        } else if (o.precision == AXG_PREC_MIXED) {
            axg_solve_opts oj = o;
            oj.tol = *std::min_element(ctol.begin(), ctol.end());
            axg_solve_info si;
            const int st = axg_mesh_solve_device(m, P, k, &oj, D, &si);
            if (st != AXG_OK) throw Error{st, axg_last_error()};
            iters = si.iters, ms = si.solve_ms, res = si.residual, rel = si.rel_residual;
        } else {
            for (int c = 0; c < k; c++) {
                axg_solve_opts oj = o;
                oj.tol = ctol[c];
                axg_solve_info si;
                const int st = axg_mesh_solve_device(m, P + (size_t)c * n, 1, &oj, D + (size_t)c * n, &si);
                if (st != AXG_OK) throw Error{st, axg_last_error()};
                iters += si.iters, ms += si.solve_ms;
                if (si.rel_residual >= rel) rel = si.rel_residual, res = si.residual;
            }
        }
        for (int c = 0; c < k; c++)
            hipLaunchKernelGGL(axpy_kernel, dim3(gn), dim3(256), 0, s, n, -scl[cols[c]],
                               D + (size_t)c * n, Y + (size_t)cols[c] * n);
        AXG_HIP(hipGetLastError());
    }
};

}  // namespace axg

using namespace axg;

extern "C" int axg_eigen(axg_mesh *mesh, const double *m_diag, int32_t nmodes,
                         const axg_solve_opts *opts, double *values, double *vectors,
                         int32_t *nfound, int32_t *sweeps, axg_solve_info *info) {
    API_BEGIN
    AXG_REQUIRE(mesh && m_diag && values && nfound, AXG_E_INVALID, "axg_eigen: NULL argument");
    AXG_REQUIRE(mesh->assembled64, AXG_E_INVALID, "mesh not assembled (call axg_assemble)");
    const int64_t n = mesh->n_free;
    AXG_REQUIRE(n > 0 && nmodes > 0, AXG_E_INVALID, "Eigenproblem ohne freie DOF / ohne Moden.");
    const int64_t nm = std::min<int64_t>(nmodes, n);
    const int nvec = (int)std::max(std::min<int64_t>(nm + 4, n), nm);
    AXG_REQUIRE(nvec <= NVMAX, AXG_E_INVALID, "axg_eigen: at most 60 modes");
    axg_ctx *ctx = mesh->ctx;
    AXG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    axg_solve_opts o;
    axg_default_opts(&o);
    if (opts) o = *opts;

    const size_t nv = (size_t)n * nvec;
    DevBuf<double> dm(n), X(nv), Y(nv), R(nv), KY(n), dq((size_t)nvec * nvec);
    const int grid = (int)std::min<int64_t>((n + EBS - 1) / EBS, (int64_t)ctx->cu_count * 4);
    DevBuf<double> part((size_t)grid * nvec);
    std::vector<double> hpart((size_t)grid * nvec);
    const unsigned gn = (unsigned)((n + 255) / 256);
    AXG_HIP(hipMemcpyAsync(dm.p, m_diag, 8 * n, hipMemcpyHostToDevice, s));

    // start block (eigen.rs:24-34)
    {
        std::vector<double> hx(nv);
        uint64_t rng = 1;
        for (int v = 0; v < nvec; v++)
            for (int64_t i = 0; i < n; i++) {
                rng = rng * 6364136223846793005ull + 1ull;
                double val = (double)(rng >> 33) / (double)(1ull << 31) - 1.0;
                hx[(size_t)v * n + i] = m_diag[i] <= 0.0 ? 0.0 : val;
            }
        AXG_HIP(hipMemcpyAsync(X.p, hx.data(), 8 * nv, hipMemcpyHostToDevice, s));
        AXG_HIP(hipStreamSynchronize(s));
    }
    auto wdot = [&](const double *a, const double *w, const double *b) {
        hipLaunchKernelGGL(wdot_kernel, dim3(grid), dim3(EBS), 0, s, n, a, w, b, part.p);
        AXG_HIP(hipMemcpyAsync(hpart.data(), part.p, 8 * grid, hipMemcpyDeviceToHost, s));
        AXG_HIP(hipStreamSynchronize(s));
        double acc = 0.0;
        for (int g = 0; g < grid; g++) acc += hpart[g];
        return acc;
    };
    // eigen.rs:181-206
    auto m_orthonormalize = [&](double *x) {
        for (int a = 0; a < nvec; a++) {
            double *xa = x + (size_t)a * n;
            for (int b = 0; b < a; b++) {
                const double *xb = x + (size_t)b * n;
                const double d = wdot(xa, dm.p, xb);
                hipLaunchKernelGGL(axpy_kernel, dim3(gn), dim3(256), 0, s, n, d, xb, xa);
            }
            const double nrm = std::sqrt(wdot(xa, dm.p, xa));
            if (nrm > 1e-30) hipLaunchKernelGGL(scale_kernel, dim3(gn), dim3(256), 0, s, n, nrm, xa);
        }
        AXG_HIP(hipGetLastError());
    };
    m_orthonormalize(X.p);

    CsrView A = solver_csr(mesh);
    // 1 / Jacobi divisor, for the right-hand-side scale below
    DevBuf<double> dinv(n), P(nv), D(nv);
    {
        std::vector<double> hd(n);
        AXG_HIP(hipMemcpyAsync(hd.data(), mesh->dsel.p, 8 * n, hipMemcpyDeviceToHost, s));
        AXG_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n; i++) hd[i] = 1.0 / hd[i];
        AXG_HIP(hipMemcpyAsync(dinv.p, hd.data(), 8 * n, hipMemcpyHostToDevice, s));
    }
    // The PCG stops with "not positive definite" when |p^T K p| < 1e-30 (an
    // absolute guard, linalg.rs:231-233); with lumped masses of order 1e-10 the
    // curvatures of a converging solve fall below it. Each right-hand side is
    // divided by the power of two nearest sqrt(b^T D^-1 b) (D: the Jacobi
    // divisor) and its solution multiplied back: a power of two changes no
    // rounding, so the iterates are exactly those of the unscaled system.
    auto pow2_scale = [&](double *b) {
        const double beta = wdot(b, dinv.p, b);
        if (!(beta > 0.0) || !std::isfinite(beta)) return 1.0;
        const double sc = std::ldexp(1.0, (int)std::lround(0.5 * std::log2(beta)));
        hipLaunchKernelGGL(scale_kernel, dim3(gn), dim3(256), 0, s, n, sc, b);
        return sc;
    };
    std::vector<double> vals(nvec, 0.0), ev, q, kr((size_t)nvec * nvec);
    std::vector<double> sc(nvec, 1.0), tolj(nvec, o.tol);
    int it = 0;
    int64_t solve_iters = 0, skipped = 0;
    double solve_ms = 0.0, worst_res = 0.0, worst_rel = 0.0;
    ColumnSolver cs(mesh, o, P.p, D.p);
    // Y_j += sc_j K^-1 R_j for every column with tolj < 1, each to its tolerance
    auto solve_columns = [&](double *Rb, const std::vector<double> &scl,
                             const std::vector<double> &tl) {
        cs.solve(Rb, Y.p, scl, tl);
        solve_iters += cs.iters;
        solve_ms += cs.ms;
        if (cs.rel >= worst_rel) worst_rel = cs.rel, worst_res = cs.res;
    };
    for (; it < 40; it++) {
        hipLaunchKernelGGL(mass_apply_kernel, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s,
                           n, nvec, dm.p, X.p, 1.0, R.p);
        // Warm start (after the first sweep): x_j is a Ritz vector with Ritz value
        // l_j, so K^-1 M x_j ~ x_j / l_j. The solve then only corrects that guess:
        // K d_j = M x_j - K (x_j / l_j), stopped at the same absolute residual,
        // tol ||M x_j||, as a solve from zero (so at relative tolerance
        // tol ||M x_j|| / ||M x_j - K x_j / l_j||; none once that reaches 1).
        for (int j = 0; j < nvec; j++) {
            double *rj = R.p + (size_t)j * n, *yj = Y.p + (size_t)j * n;
            const double lam = vals[j];
            if (it > 0 && lam > 0.0 && std::isfinite(lam)) {
                const double nb = std::sqrt(wdot(rj, nullptr, rj));
                hipLaunchKernelGGL(div_copy_kernel, dim3(gn), dim3(256), 0, s, n, lam,
                                   X.p + (size_t)j * n, yj);
                spmv_fp64(ctx, A, yj, KY.p);
                hipLaunchKernelGGL(axpy_kernel, dim3(gn), dim3(256), 0, s, n, 1.0, KY.p, rj);
                const double nbp = std::sqrt(wdot(rj, nullptr, rj));
                tolj[j] = nbp > 0.0 ? o.tol * nb / nbp : HUGE_VAL;
            } else {
                AXG_HIP(hipMemsetAsync(yj, 0, 8 * n, s));
                tolj[j] = o.tol;
            }
            if (tolj[j] >= 1.0) {
                skipped++;
                continue;
            }
            sc[j] = pow2_scale(rj);
        }
        AXG_HIP(hipGetLastError());
        solve_columns(R.p, sc, tolj);
        m_orthonormalize(Y.p);
        // Kr = Y^T K Y, column j from K y_j
        for (int j = 0; j < nvec; j++) {
            spmv_fp64(ctx, A, Y.p + (size_t)j * n, KY.p);
            hipLaunchKernelGGL((multi_dot_kernel<NVMAX>), dim3(grid), dim3(EBS), 0, s, n, nvec, Y.p,
                               KY.p, part.p);
            AXG_HIP(hipMemcpyAsync(hpart.data(), part.p, 8 * (size_t)grid * nvec,
                                   hipMemcpyDeviceToHost, s));
            AXG_HIP(hipStreamSynchronize(s));
            for (int i = 0; i < nvec; i++) {
                double acc = 0.0;
                for (int g = 0; g < grid; g++) acc += hpart[(size_t)g * nvec + i];
                kr[(size_t)i * nvec + j] = acc;
            }
        }
        jacobi_eigen(kr, nvec, ev, q);
        AXG_HIP(hipMemcpyAsync(dq.p, q.data(), 8 * q.size(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(combine_kernel, dim3(gn), dim3(256), 0, s, n, nvec, Y.p, dq.p, X.p);
        AXG_HIP(hipStreamSynchronize(s));
        m_orthonormalize(X.p);
        bool conv = true;
        for (int i = 0; i < nvec; i++) {
            if (std::fabs(ev[i] - vals[i]) > 1e-8 * (1.0 + std::fabs(ev[i]))) conv = false;
            vals[i] = ev[i];
        }
        if (conv && it > 4) {
            it++;
            break;
        }
    }
    if (it > 40) it = 40;
    // ascending, the first nmodes (eigen.rs:88-110)
    std::vector<int> order(nvec);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return vals[a] < vals[b]; });
    std::vector<double> hx;
    if (vectors) {
        hx.resize(nv);
        AXG_HIP(hipMemcpyAsync(hx.data(), X.p, 8 * nv, hipMemcpyDeviceToHost, s));
        AXG_HIP(hipStreamSynchronize(s));
    }
    int32_t k = 0;
    for (int r = 0; r < (int)nm; r++) {
        const int i = order[r];
        if (!std::isfinite(vals[i])) continue;
        values[k] = vals[i];
        if (vectors) std::copy(hx.begin() + (size_t)i * n, hx.begin() + (size_t)(i + 1) * n,
                               vectors + (size_t)k * n);
        k++;
    }
    AXG_REQUIRE(k > 0, AXG_E_INVALID, "Keine Eigenwerte gefunden.");
    *nfound = k;
    if (sweeps) *sweeps = it;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->iters = solve_iters;
        info->outer_iters = it;
        info->nrhs = nvec;
        info->residual = worst_res;
        info->rel_residual = worst_rel;
        info->solve_ms = solve_ms;
        std::snprintf(info->name, sizeof(info->name), "subspace (%d vectors, GPU PCG, %lld warm)", nvec,
                      (long long)skipped);
    }
    API_END
}

// eigen::subspace_ab (eigen.rs:113-179) for *BUCKLE: K phi = lambda A phi with
// A = -Kg_ff from the displacements u_full (analysis.rs:1176-1247). nmodes + 4
// vectors from the seed-7 generator, each scaled to unit length; 30 sweeps of
// Y = K^-1 A X (one multi-right-hand-side solve) with Euclidean Gram-Schmidt;
// then lambda = v.Kv / v.Av per vector (|v.Av| < 1e-30 dropped), sorted by
// |lambda| (stable), the first nmodes returned.
extern "C" int axg_buckle(axg_mesh *mesh, const double *u_full, int32_t nmodes,
                          const axg_solve_opts *opts, double *values, double *vectors,
                          int32_t *nfound, axg_solve_info *info) {
    API_BEGIN
    AXG_REQUIRE(mesh && u_full && values && nfound, AXG_E_INVALID, "axg_buckle: NULL argument");
    AXG_REQUIRE(mesh->assembled64, AXG_E_INVALID, "mesh not assembled (call axg_assemble)");
    const int64_t n = mesh->n_free;
    AXG_REQUIRE(n > 0 && nmodes > 0, AXG_E_INVALID, "Eigenproblem ohne freie DOF.");
    const int64_t nm = std::min<int64_t>(nmodes, n);
    const int nvec = (int)std::max(std::min<int64_t>(nm + 4, n), nm);
    AXG_REQUIRE(nvec <= NVMAX, AXG_E_INVALID, "axg_buckle: at most 60 modes");
    axg_ctx *ctx = mesh->ctx;
    AXG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    axg_solve_opts o;
    axg_default_opts(&o);
    if (opts) o = *opts;

    const size_t nv = (size_t)n * nvec;
    DevBuf<double> du(std::max<int64_t>(mesh->ndof, 1)), aval(std::max<int64_t>(mesh->nnz_topo, 1)),
        X(nv), Y(nv), R(nv), P(nv), D(nv), T(n), dinv(n);
    AXG_HIP(hipMemcpyAsync(du.p, u_full, 8 * mesh->ndof, hipMemcpyHostToDevice, s));
    geometric_stiffness_device(mesh, du.p, -1.0, aval.p, nullptr);
    int ashape = -1;
    CsrView Aop;
    Aop.n = n;
    Aop.nnz = mesh->nnz_topo;
    Aop.rowptr = mesh->rowptr_t.p;
    Aop.colind = mesh->colind_t.p;
    Aop.val = aval.p;
    Aop.shape = &ashape;
    CsrView K = solver_csr(mesh);

    const int grid = (int)std::min<int64_t>((n + EBS - 1) / EBS, (int64_t)ctx->cu_count * 4);
    DevBuf<double> part((size_t)grid);
    std::vector<double> hpart(grid);
    const unsigned gn = (unsigned)((n + 255) / 256);
    auto wdot = [&](const double *a, const double *w, const double *b) {
        hipLaunchKernelGGL(wdot_kernel, dim3(grid), dim3(EBS), 0, s, n, a, w, b, part.p);
        AXG_HIP(hipMemcpyAsync(hpart.data(), part.p, 8 * grid, hipMemcpyDeviceToHost, s));
        AXG_HIP(hipStreamSynchronize(s));
        double acc = 0.0;
        for (int g = 0; g < grid; g++) acc += hpart[g];
        return acc;
    };
    {
        std::vector<double> hd(n);
        AXG_HIP(hipMemcpyAsync(hd.data(), mesh->dsel.p, 8 * n, hipMemcpyDeviceToHost, s));
        AXG_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n; i++) hd[i] = 1.0 / hd[i];
        AXG_HIP(hipMemcpyAsync(dinv.p, hd.data(), 8 * n, hipMemcpyHostToDevice, s));
        // start block (eigen.rs:124-135)
        std::vector<double> hx(nv);
        uint64_t rng = 7;
        for (int v = 0; v < nvec; v++) {
            double *xv = hx.data() + (size_t)v * n;
            for (int64_t i = 0; i < n; i++) {
                rng = rng * 6364136223846793005ull + 1ull;
                xv[i] = (double)(rng >> 33) / (double)(1ull << 31) - 1.0;
            }
            double ss = 0.0;
            for (int64_t i = 0; i < n; i++) ss += xv[i] * xv[i];
            const double nrm = std::max(std::sqrt(ss), 1e-30);
            for (int64_t i = 0; i < n; i++) xv[i] /= nrm;
        }
        AXG_HIP(hipMemcpyAsync(X.p, hx.data(), 8 * nv, hipMemcpyHostToDevice, s));
        AXG_HIP(hipStreamSynchronize(s));
    }
    int64_t solve_iters = 0;
    double solve_ms = 0.0, worst_res = 0.0, worst_rel = 0.0;
    double *x = X.p, *y = Y.p;
    ColumnSolver cs(mesh, o, P.p, D.p);
    std::vector<double> scl(nvec, 1.0), tl(nvec, o.tol);
    int64_t skipped = 0;
    for (int it = 0; it < 30; it++) {
        AXG_HIP(hipMemsetAsync(y, 0, 8 * nv, s));
        for (int k = 0; k < nvec; k++) {
            double *r = R.p + (size_t)k * n, *yk = y + (size_t)k * n;
            const double *xk = x + (size_t)k * n;
            spmv_fp64(ctx, Aop, xk, r);
            tl[k] = o.tol;
            // Warm start (after the first sweep): along x, K^-1 A scales by the
            // Rayleigh quotient mu = x^T A x / x^T K x, so y ~ mu x. The solve
            // corrects that guess, K d = r - mu K x, to the same absolute residual
            // as a solve from zero, tol ||r|| (none once that is already met).
            if (it > 0) {
                spmv_fp64(ctx, K, xk, T.p);
                const double xkx = wdot(xk, nullptr, T.p), mu = wdot(xk, nullptr, r) / xkx;
                if (xkx > 0.0 && std::isfinite(mu) && mu != 0.0) {
                    const double nb = std::sqrt(wdot(r, nullptr, r));
                    hipLaunchKernelGGL(div_copy_kernel, dim3(gn), dim3(256), 0, s, n, 1.0 / mu,
                                       xk, yk);
                    hipLaunchKernelGGL(axpy_kernel, dim3(gn), dim3(256), 0, s, n, mu, T.p, r);
                    const double nbp = std::sqrt(wdot(r, nullptr, r));
                    tl[k] = nbp > 0.0 ? o.tol * nb / nbp : HUGE_VAL;
                }
            }
            if (tl[k] >= 1.0) {
                skipped++;
                continue;
            }
            // power-of-two scale to r^T D^-1 r ~ 1 (the PCG's absolute breakdown
            // guard); exact, and multiplied back onto the solution
            const double beta = wdot(r, dinv.p, r);
            scl[k] = 1.0;
            if (beta > 0.0 && std::isfinite(beta)) {
                scl[k] = std::ldexp(1.0, (int)std::lround(0.5 * std::log2(beta)));
                hipLaunchKernelGGL(scale_kernel, dim3(gn), dim3(256), 0, s, n, scl[k], r);
            }
        }
        cs.solve(R.p, y, scl, tl);
        solve_iters += cs.iters;
        solve_ms += cs.ms;
        if (cs.rel >= worst_rel) worst_rel = cs.rel, worst_res = cs.res;
        // Euclidean Gram-Schmidt (eigen.rs:143-156)
        for (int a = 0; a < nvec; a++) {
            double *ya = y + (size_t)a * n;
            for (int b = 0; b < a; b++) {
                const double *yb = y + (size_t)b * n;
                const double d = wdot(ya, nullptr, yb);
                hipLaunchKernelGGL(axpy_kernel, dim3(gn), dim3(256), 0, s, n, d, yb, ya);
            }
            const double nrm = std::max(std::sqrt(wdot(ya, nullptr, ya)), 1e-30);
            hipLaunchKernelGGL(scale_kernel, dim3(gn), dim3(256), 0, s, n, nrm, ya);
        }
        AXG_HIP(hipGetLastError());
        std::swap(x, y);
    }
    // Rayleigh quotients (eigen.rs:159-173)
    std::vector<std::pair<double, int>> pairs;
    for (int k = 0; k < nvec; k++) {
        const double *v = x + (size_t)k * n;
        spmv_fp64(ctx, K, v, T.p);
        const double num = wdot(v, nullptr, T.p);
        spmv_fp64(ctx, Aop, v, T.p);
        const double den = wdot(v, nullptr, T.p);
        if (std::fabs(den) < 1e-30) continue;
        pairs.push_back({num / den, k});
    }
    std::stable_sort(pairs.begin(), pairs.end(), [](const std::pair<double, int> &a,
                                                    const std::pair<double, int> &b) {
        return std::fabs(a.first) < std::fabs(b.first);
    });
    const int32_t take = (int32_t)std::min<int64_t>(nm, (int64_t)pairs.size());
    for (int32_t k = 0; k < take; k++) {
        values[k] = pairs[k].first;
        if (vectors)
            AXG_HIP(hipMemcpyAsync(vectors + (size_t)k * n, x + (size_t)pairs[k].second * n,
                                   8 * n, hipMemcpyDeviceToHost, s));
    }
    AXG_HIP(hipStreamSynchronize(s));
    *nfound = take;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->iters = solve_iters;
        info->outer_iters = 30;
        info->nrhs = nvec;
        info->residual = worst_res;
        info->rel_residual = worst_rel;
        info->solve_ms = solve_ms;
        std::snprintf(info->name, sizeof(info->name), "buckle (%d vectors, GPU PCG, %lld warm)",
                      nvec, (long long)skipped);
    }
    API_END
}
