// C ABI of libaxia_gpu (include/axia_gpu.h): context, solve_kff drop-in,
// mesh assembly, solves, stress recovery and the compliance adjoint.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <set>

#include "axg_internal.h"
#include "axg_mixed.h"
#include "axg_pcg.h"
#include "axg_triplets.h"

namespace axg {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
void clear_error() { g_err.clear(); }

static std::mutex g_tab_mu;
static std::set<int> g_tab_devices;

static void ensure_tables(int device) {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    if (g_tab_devices.count(device)) return;
    upload_shape_tables();
    g_tab_devices.insert(device);
}

static const char *elem_name(int nn) { return nn == 8 ? "C3D8" : "C3D20"; }

// FemError texts of the reference for the element error codes.
static std::string element_error(int code, int nn, int32_t el, const double *e, const double *nu) {
    switch (code) {
        case AXG_E_NEG_JACOBIAN:  // elem.rs:210, quadratic.rs:138
            return std::string(elem_name(nn)) + ": negative Jakobideterminante. (element index " +
                   std::to_string(el) + ")";
        case AXG_E_JAC_SINGULAR:  // elem.rs:14
            return "Jakobimatrix ist singulär (entartetes Element). (element index " +
                   std::to_string(el) + ")";
        case AXG_E_MATERIAL: {  // elem.rs:43-48
            char buf[160];
            if (nu && (nu[el] >= 0.5 || nu[el] <= -1.0))
                snprintf(buf, sizeof buf, "Ungültige Querkontraktion nu=%g (element index %d)",
                         nu[el], el);
            else
                snprintf(buf, sizeof buf, "Ungültiger E-Modul E=%g (element index %d)",
                         e ? e[el] : 0.0, el);
            return buf;
        }
        default:
            return "element error " + std::to_string(code);
    }
}

std::string pcg_error(const PcgResult &r) {
    if (r.status == AXG_E_NOT_SPD)  // linalg.rs:231-233
        return "CG: unerwartetes Null-Pivot — Matrix nicht positiv definit.";
    if (r.status == AXG_E_NOT_CONVERGED) {  // linalg.rs:253-257
        char buf[160];
        const bool col = r.fail_iters >= 0;  // the failing column of a multi-rhs solve
        snprintf(buf, sizeof buf, "CG konvergierte nicht (Residuum %.3e nach %lld Iterationen).",
                 col ? r.fail_residual : r.residual, (long long)(col ? r.fail_iters : r.iters));
        return buf;
    }
    return "solver error";
}

CsrView active_csr(axg_mesh *m) {
    CsrView A;
    A.n = m->n_free;
    A.nnz = m->nnz;
    A.rowptr = m->compacted ? m->rowptr_c.p : m->rowptr_t.p;
    A.colind = m->compacted ? m->colind_c.p : m->colind_t.p;
    A.val = m->val.p;
    A.shape = &m->spmv_shape;
    return A;
}

static double norm2_host(const double *v, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += v[i] * v[i];
    return std::sqrt(s);
}

// ||b - A x||_2 on the device (FP64), returned to the host.
double true_residual(axg_ctx *ctx, const CsrView &A, const double *d_x, const double *d_b) {
    Workspace &w = ctx->ws;
    w.t0.ensure(A.n);
    spmv_fp64(ctx, A, d_x, w.t0.p);
    std::vector<double> ax(A.n), b(A.n);
    AXG_HIP(hipMemcpyAsync(ax.data(), w.t0.p, sizeof(double) * A.n, hipMemcpyDeviceToHost,
                           ctx->stream));
    AXG_HIP(hipMemcpyAsync(b.data(), d_b, sizeof(double) * A.n, hipMemcpyDeviceToHost,
                           ctx->stream));
    AXG_HIP(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (int64_t i = 0; i < A.n; i++) {
        double d = ax[i] - b[i];
        s += d * d;
    }
    return std::sqrt(s);
}

static double device_norm(axg_ctx *ctx, const double *d_v, int64_t n) {
    std::vector<double> v(n);
    AXG_HIP(hipMemcpyAsync(v.data(), d_v, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    AXG_HIP(hipStreamSynchronize(ctx->stream));
    return norm2_host(v.data(), n);
}

// The block PCG's vectors live in node-block space (3 rows per node with a
// free DOF, padded to whole workgroups, 8 right-hand sides wide) and are
// indexed with 32-bit offsets.
static bool block_pcg_fits(axg_mesh *m) {
    if (m->n_free == 0) return false;
    ensure_block_maps(m);
    return m->nb > 0 && (3 * m->nb + 256) * 8 < INT32_MAX;
}

}  // namespace axg

using namespace axg;

namespace axg {
// Kg_ff on the topological pattern for displacements d_u (full DOF space, on
// the device): extrapolated nodal stresses, the element mean stresses, the
// element Kg at those stresses and the gather, times sign.
void geometric_stiffness_device(axg_mesh *mesh, const double *d_u, double sign, double *d_val,
                                double *d_mean) {
    axg_ctx *ctx = mesh->ctx;
    hipStream_t s = ctx->stream;
    Workspace &w = ctx->ws;
    DevBuf<double> den((size_t)std::max<int64_t>(mesh->nelem, 1) * mesh->nn * 6),
        dst((size_t)std::max<int64_t>(mesh->nnode, 1) * 6), dvm(std::max<int64_t>(mesh->nnode, 1)),
        mean((size_t)std::max<int64_t>(mesh->nelem, 1) * 6),
        kg((size_t)std::max<int64_t>(mesh->nelem, 1) * mesh->nn * mesh->nn);
    AXG_HIP(hipMemsetAsync(w.err.p, 0, 8, s));
    launch_nodal_stress(mesh, s, d_u, den.p, dst.p, dvm.p, w.err.p);
    launch_element_mean(mesh, s, den.p, mean.p);
    launch_kg(mesh, s, mean.p, kg.p, w.err.p);
    launch_gather_kg(mesh, s, kg.p, sign, d_val);
    int32_t err[2];
    AXG_HIP(hipMemcpyAsync(err, w.err.p, 8, hipMemcpyDeviceToHost, s));
    if (d_mean && mesh->nelem > 0)
        AXG_HIP(hipMemcpyAsync(d_mean, mean.p, 8 * (size_t)mesh->nelem * 6,
                               hipMemcpyDeviceToDevice, s));
    AXG_HIP(hipStreamSynchronize(s));
    if (err[0]) throw Error{err[0], element_error(err[0], mesh->nn, err[1], nullptr, nullptr)};
}
}  // namespace axg

// Drop the slots whose contributions were all exact zeros (the reference
// never creates them: analysis.rs:664, linalg.rs:11).
static void compact_pattern(axg_mesh *m) {
    hipStream_t s = m->ctx->stream;
    int64_t nslot = m->nnz_topo, n = m->n_free;
    std::vector<int32_t> rp(n + 1), ci(nslot);
    std::vector<uint32_t> sp(nslot + 1), cb(m->ncontrib_topo);
    std::vector<uint8_t> nz(nslot);
    AXG_HIP(hipMemcpyAsync(rp.data(), m->rowptr_t.p, 4 * (n + 1), hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(ci.data(), m->colind_t.p, 4 * nslot, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(sp.data(), m->slotptr_t.p, 4 * (nslot + 1), hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(cb.data(), m->contrib_t.p, 4 * m->ncontrib_topo,
                           hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(nz.data(), m->slot_nz.p, nslot, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> rp2(n + 1, 0), ci2;
    std::vector<uint32_t> sp2, cb2;
    ci2.reserve(nslot);
    sp2.reserve(nslot + 1);
    cb2.reserve(cb.size());
    for (int64_t i = 0; i < n; i++) {
        for (int32_t k = rp[i]; k < rp[i + 1]; k++) {
            if (!nz[k]) continue;
            ci2.push_back(ci[k]);
            sp2.push_back((uint32_t)cb2.size());
            for (uint32_t c = sp[k]; c < sp[k + 1]; c++) cb2.push_back(cb[c]);
        }
        rp2[i + 1] = (int32_t)ci2.size();
    }
    sp2.push_back((uint32_t)cb2.size());
    auto up = [&](auto &buf, const auto &v) {
        buf.alloc(v.size());
        if (!v.empty())
            AXG_HIP(hipMemcpyAsync(buf.p, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice,
                                   s));
    };
    up(m->rowptr_c, rp2);
    up(m->colind_c, ci2);
    up(m->slotptr_c, sp2);
    up(m->contrib_c, cb2);
    AXG_HIP(hipStreamSynchronize(s));
    m->compacted = true;
    m->nnz = (int64_t)ci2.size();
    m->ncontrib = (int64_t)cb2.size();
}

// After the topological gather of K_ff from mesh->ke: drop the slots whose
// contributions were all exact zeros (re-gathering the compacted pattern), or
// return to the topological pattern when there are none; then the Jacobi
// divisor and the validity flags of the derived operators.
static void finish_gather64(axg_mesh *mesh, uint32_t zero_slots) {
    axg_ctx *ctx = mesh->ctx;
    hipStream_t s = ctx->stream;
    if (zero_slots > 0) {
        compact_pattern(mesh);
        launch_gather_fp64(mesh, s, false, nullptr);
        mesh->sell_valid = false;
    } else if (mesh->compacted) {
        mesh->compacted = false;
        mesh->nnz = mesh->nnz_topo;
        mesh->ncontrib = mesh->ncontrib_topo;
        mesh->sell_valid = false;
    }
    mesh->sell_fresh = false;
    mesh->bsell64_fresh = false;
    CsrView A = active_csr(mesh);
    diag_fp64(ctx, A, mesh->dsel.p, nullptr);
    AXG_HIP(hipStreamSynchronize(s));
    mesh->assembled64 = true;
    mesh->binv_valid = false;
    mesh->assembled32 = false;
}

namespace axg {
// NLGEOM FemError texts (nlgeom.rs:93, 117-125, 240-272)
static std::string nl_error(int code, int nn, int32_t el, double e, double nu) {
    char buf[200];
    switch (code) {
        case AXG_E_INVERSION:
            return "NLGEOM: det(F) ≤ 0 (Element inversion).";
        case AXG_E_NEG_JACOBIAN:
            snprintf(buf, sizeof buf, "NLGEOM %s: negative Jakobideterminante.", elem_name(nn));
            return buf;
        case AXG_E_MATERIAL:
            snprintf(buf, sizeof buf, "Ungültiges Material E=%g nu=%g", e, nu);
            return buf;
        default:
            return element_error(code, nn, el, nullptr, nullptr);
    }
}

// The NLGEOM tangent of every element into mesh->ke, its internal forces (d_fe,
// nelem x 3nn), Cauchy and Green-Lagrange means (nelem x 6), K_T gathered into
// the mesh's K_ff (exact-zero slots dropped as for K) and f_int over all DOFs.
void nl_assemble_device(axg_mesh *mesh, const double *d_u, double *d_fe, double *d_cau,
                        double *d_gl, double *d_fint) {
    axg_ctx *ctx = mesh->ctx;
    hipStream_t s = ctx->stream;
    Workspace &w = ctx->ws;
    AXG_HIP(hipMemsetAsync(w.err.p, 0, 4 * sizeof(int32_t), s));
    AXG_HIP(hipMemsetAsync(w.cnt.p, 0, 4 * sizeof(uint32_t), s));
    launch_nl(mesh, s, d_u, d_fe, d_cau, d_gl, w.err.p);
    launch_gather_fp64(mesh, s, true, w.cnt.p);
    launch_nl_fint(mesh, s, d_fe, d_fint);
    int32_t err[2];
    uint32_t zero_slots = 0;
    AXG_HIP(hipMemcpyAsync(err, w.err.p, 8, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(&zero_slots, w.cnt.p, 4, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    if (err[0] != 0) {
        double e = 0.0, nu = 0.0;
        if (err[1] >= 0 && err[1] < mesh->nelem) {
            AXG_HIP(hipMemcpy(&e, mesh->e.p + err[1], 8, hipMemcpyDeviceToHost));
            AXG_HIP(hipMemcpy(&nu, mesh->nu.p + err[1], 8, hipMemcpyDeviceToHost));
        }
        throw Error{err[0], nl_error(err[0], mesh->nn, err[1], e, nu)};
    }
    finish_gather64(mesh, zero_slots);
}

// Nodal Cauchy and Green-Lagrange of an NLGEOM state (analysis.rs:4745-4799)
// and the element means stress_gp; d_cau / d_gl: the element means of the
// last tangent assembly at the same displacements (C3D8).
void nl_nodal_device(axg_mesh *mesh, const double *d_u, const double *d_cau, const double *d_gl,
                     double *d_stress, double *d_vm, double *d_strain, double *d_gp) {
    axg_ctx *ctx = mesh->ctx;
    hipStream_t s = ctx->stream;
    Workspace &w = ctx->ws;
    const size_t ne = (size_t)std::max<int64_t>(mesh->nelem, 1) * mesh->nn * 6;
    DevBuf<double> enc(ne), eng(ne), dummy((size_t)std::max<int64_t>(mesh->nnode, 1) * 6);
    AXG_HIP(hipMemsetAsync(w.err.p, 0, 8, s));
    launch_nl_nodal(mesh, s, d_u, d_cau, d_gl, enc.p, eng.p, w.err.p);
    launch_nodal_average(mesh, s, enc.p, d_stress, d_vm);
    launch_nodal_average(mesh, s, eng.p, d_strain ? d_strain : dummy.p, nullptr);
    // stress_gp: C3D20 the mean of its nodal Cauchy values / nn, C3D8 the
    // element's volume-averaged Cauchy itself (analysis.rs:4766-4783)
    if (d_gp) {
        if (mesh->nn == 20)
            launch_element_mean(mesh, s, enc.p, d_gp);
        else if (mesh->nelem > 0)
            AXG_HIP(hipMemcpyAsync(d_gp, d_cau, 8 * 6 * mesh->nelem, hipMemcpyDeviceToDevice, s));
    }
    int32_t err[2];
    AXG_HIP(hipMemcpyAsync(err, w.err.p, 8, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    if (err[0]) throw Error{err[0], nl_error(err[0], mesh->nn, err[1], 0.0, 0.0)};
}
}  // namespace axg

extern "C" {

const char *axg_last_error(void) { return g_err.c_str(); }
const char *axg_version(void) { return "axia_gpu 0.1 (gfx950)"; }

void axg_default_opts(axg_solve_opts *o) {
    if (!o) return;
    o->precision = AXG_PREC_FP64;
    o->precond = AXG_PRECOND_JACOBI;
    o->tol = 1e-8;
    o->max_iter = 0;
    o->inner_tol = 1e-3;
    o->max_outer = 30;
    o->timing = 0;
    o->refinement = AXG_REFINE_FCG;
    o->batch_groups = 0;
}

static void ctx_free(axg_ctx *ctx);

int axg_create(int device, axg_ctx **out) {
    API_BEGIN
    AXG_REQUIRE(out, AXG_E_INVALID, "axg_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    AXG_HIP(hipGetDeviceCount(&ndev));
    AXG_REQUIRE(device >= 0 && device < ndev, AXG_E_INVALID,
                "axg_create: device " + std::to_string(device) + " not present");
    AXG_HIP(hipSetDevice(device));
    auto *c = new axg_ctx();
    c->device = device;
    try {
        AXG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        AXG_HIP(hipEventCreate(&c->ev0));
        AXG_HIP(hipEventCreate(&c->ev1));
        AXG_HIP(hipEventCreate(&c->evA));
        hipDeviceProp_t prop;
        AXG_HIP(hipGetDeviceProperties(&prop, device));
        c->cu_count = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        c->arch = prop.gcnArchName;
        ensure_tables(device);
        noalloc_setup(device);
        c->ws.err.ensure(4);
        c->ws.cnt.ensure(4);
    } catch (...) {
        ctx_free(c);
        throw;
    }
    *out = c;
    API_END
}

static void ctx_free(axg_ctx *ctx) {
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto e : ctx->timing_events) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->evA) (void)hipEventDestroy(ctx->evA);
    ctx->ws = Workspace();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

static void ctx_release(axg_ctx *ctx) {
    if (ctx->refs.fetch_sub(1) == 1) ctx_free(ctx);
}

int axg_destroy(axg_ctx *ctx) {
    if (!ctx) return AXG_OK;
    ctx_release(ctx);
    return AXG_OK;
}

void *axg_stream(axg_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

// ------------------------------------------------------------ solve_kff
static void fill_info(axg_solve_info *info, const char *name, int64_t iters, int outer,
                      double residual, double bnorm, double ms, double spmv_ms,
                      int64_t spmv_launches, int nrhs) {
    if (!info) return;
    std::memset(info, 0, sizeof(*info));
    info->iters = iters;
    info->outer_iters = outer;
    info->nrhs = nrhs;
    info->residual = residual;
    info->rel_residual = bnorm > 0 ? residual / bnorm : 0.0;
    info->solve_ms = ms;
    info->spmv_ms = spmv_ms;
    info->spmv_launches = spmv_launches;
    std::snprintf(info->name, sizeof(info->name), "%s", name);
}

static int64_t default_max_iter(const axg_solve_opts &o, int64_t n) {
    return o.max_iter > 0 ? o.max_iter : std::max<int64_t>(4 * n, 200);
}

int axg_csr_from_triplets(axg_ctx *ctx, int64_t n, int64_t ntrip, const int64_t *ti,
                          const int64_t *tj, const double *tv, int64_t *indptr, int64_t *indices,
                          double *data) {
    API_BEGIN
    AXG_REQUIRE(ctx && n >= 0 && ntrip >= 0 && indptr, AXG_E_INVALID,
                "axg_csr_from_triplets: invalid argument");
    AXG_REQUIRE(ntrip == 0 || (ti && tj && tv), AXG_E_INVALID, "axg_csr_from_triplets: NULL triplets");
    AXG_HIP(hipSetDevice(ctx->device));
    DevBuf<int64_t> dti(ntrip), dtj(ntrip);
    DevBuf<double> dtv(ntrip);
    hipStream_t s = ctx->stream;
    if (ntrip) {
        AXG_HIP(hipMemcpyAsync(dti.p, ti, 8 * ntrip, hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(dtj.p, tj, 8 * ntrip, hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(dtv.p, tv, 8 * ntrip, hipMemcpyHostToDevice, s));
    }
    DeviceCsr64 csr;
    csr_from_triplets_device(ctx, n, ntrip, dti.p, dtj.p, dtv.p, csr);
    AXG_HIP(hipMemcpyAsync(indptr, csr.indptr.p, 8 * (n + 1), hipMemcpyDeviceToHost, s));
    if (indices && csr.nnz) {
        AXG_HIP(hipMemcpyAsync(indices, csr.indices.p, 8 * csr.nnz, hipMemcpyDeviceToHost, s));
        AXG_HIP(hipMemcpyAsync(data, csr.data.p, 8 * csr.nnz, hipMemcpyDeviceToHost, s));
    }
    AXG_HIP(hipStreamSynchronize(s));
    API_END
}

static void csr_operator_solve(axg_ctx *ctx, int64_t n, const int32_t *rp, const int32_t *ci,
                               const double *d_val, const double *d_b, double *d_x,
                               const axg_solve_opts &o, axg_solve_info *info);

static void check_solve_opts(const axg_solve_opts &o, const char *what) {
    AXG_REQUIRE(o.precision == AXG_PREC_FP64 || o.precision == AXG_PREC_MIXED, AXG_E_INVALID,
                std::string(what) + ": unknown precision");
    AXG_REQUIRE(o.precond == AXG_PRECOND_JACOBI || o.precond == AXG_PRECOND_BLOCK3, AXG_E_INVALID,
                std::string(what) + ": unknown preconditioner");
    AXG_REQUIRE(o.refinement == AXG_REFINE_FCG || o.refinement == AXG_REFINE_PLAIN ||
                    o.refinement == AXG_REFINE_RELIABLE,
                AXG_E_INVALID,
                std::string(what) + ": unknown refinement");
    AXG_REQUIRE(o.tol > 0.0 && std::isfinite(o.tol), AXG_E_INVALID,
                std::string(what) + ": tol must be positive");
}

// The solve of a device CSR (int32 row pointer and columns; rp, ci: host copies
// of the pattern), through an operator mesh (csr_operator_solve): FP64 + Jacobi
// is the reference's Jacobi PCG (linalg::pcg) on the scalar rows, one right-hand
// side, on the sliced-ELL copy of the rows; block-Jacobi and the mixed FP32 /
// FP64 solve take the operator in node-block form (rows 3i, 3i+1, 3i+2 as block
// row i) and solve it as a mesh's K_ff is.
static void solve_device_csr(axg_ctx *ctx, const CsrView &A, const int32_t *rp, const int32_t *ci,
                             const double *d_b, double *d_x, const axg_solve_opts &o,
                             axg_solve_info *info) {
    const char *what = "GPU PCG (FP64, Jacobi)";
    if (A.n == 0) {
        fill_info(info, what, 0, 1, 0.0, 0.0, 0.0, 0.0, 0, 1);
        return;
    }
    csr_operator_solve(ctx, A.n, rp, ci, A.val, d_b, d_x, o, info);
}

int axg_solve_kff(axg_ctx *ctx, int64_t n, int64_t ntrip, const int64_t *ti, const int64_t *tj,
                  const double *tv, const double *rhs, const axg_solve_opts *opts, double *x_out,
                  axg_solve_info *info) {
    API_BEGIN
    AXG_REQUIRE(ctx && n >= 0 && ntrip >= 0 && rhs && x_out, AXG_E_INVALID,
                "axg_solve_kff: invalid argument");
    AXG_REQUIRE(ntrip == 0 || (ti && tj && tv), AXG_E_INVALID, "axg_solve_kff: NULL triplets");
    axg_solve_opts o;
    axg_default_opts(&o);
    if (opts) o = *opts;
    check_solve_opts(o, "axg_solve_kff");
    AXG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    DevBuf<int64_t> dti(ntrip), dtj(ntrip);
    DevBuf<double> dtv(ntrip);
    if (ntrip) {
        AXG_HIP(hipMemcpyAsync(dti.p, ti, 8 * ntrip, hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(dtj.p, tj, 8 * ntrip, hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(dtv.p, tv, 8 * ntrip, hipMemcpyHostToDevice, s));
    }
    DeviceCsr64 csr;
    csr_from_triplets_device(ctx, n, ntrip, dti.p, dtj.p, dtv.p, csr);
    dti.release(), dtj.release(), dtv.release();
    AXG_REQUIRE(csr.nnz < INT32_MAX, AXG_E_INVALID, "axg_solve_kff: nnz exceeds 2^31");
    CsrView A;
    A.n = n;
    A.nnz = csr.nnz;
    A.rowptr = csr.rowptr32.p;
    A.colind = csr.colind32.p;
    A.val = csr.data.p;
    std::vector<int32_t> rp(n + 1), ci(std::max<int64_t>(csr.nnz, 1));
    AXG_HIP(hipMemcpyAsync(rp.data(), A.rowptr, 4 * (n + 1), hipMemcpyDeviceToHost, s));
    if (csr.nnz)
        AXG_HIP(hipMemcpyAsync(ci.data(), A.colind, 4 * csr.nnz, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    DevBuf<double> db(n), dx(n);
    AXG_HIP(hipMemcpyAsync(db.p, rhs, 8 * n, hipMemcpyHostToDevice, s));
    solve_device_csr(ctx, A, rp.data(), ci.data(), db.p, dx.p, o, info);
    AXG_HIP(hipMemcpyAsync(x_out, dx.p, 8 * n, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    API_END
}

int axg_solve_csr(axg_ctx *ctx, int64_t n, const int64_t *indptr, const int64_t *indices,
                  const double *data, const double *rhs, const axg_solve_opts *opts,
                  double *x_out, axg_solve_info *info) {
    API_BEGIN
    AXG_REQUIRE(ctx && n >= 0 && indptr && rhs && x_out, AXG_E_INVALID,
                "axg_solve_csr: invalid argument");
    axg_solve_opts o;
    axg_default_opts(&o);
    if (opts) o = *opts;
    check_solve_opts(o, "axg_solve_csr");
    AXG_REQUIRE(indptr[0] == 0, AXG_E_INVALID, "axg_solve_csr: indptr[0] must be 0");
    for (int64_t i = 0; i < n; i++)
        AXG_REQUIRE(indptr[i + 1] >= indptr[i], AXG_E_INVALID,
                    "axg_solve_csr: indptr must be non-decreasing");
    int64_t nnz = indptr[n];
    AXG_REQUIRE(nnz >= 0 && nnz < INT32_MAX, AXG_E_INVALID, "axg_solve_csr: nnz out of range");
    AXG_REQUIRE(nnz == 0 || (indices && data), AXG_E_INVALID,
                "axg_solve_csr: NULL indices or data");
    std::vector<int32_t> rp(n + 1), ci(std::max<int64_t>(nnz, 1));
    for (int64_t i = 0; i <= n; i++) rp[i] = (int32_t)indptr[i];
    for (int64_t k = 0; k < nnz; k++) {
        AXG_REQUIRE(indices[k] >= 0 && indices[k] < n, AXG_E_INVALID,
                    "axg_solve_csr: column index out of range");
        ci[k] = (int32_t)indices[k];
    }
    // every argument is checked before the context and the device are touched
    AXG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    DevBuf<int32_t> drp(n + 1), dci(nnz);
    DevBuf<double> dv(nnz), db(n), dx(n);
    AXG_HIP(hipMemcpyAsync(drp.p, rp.data(), 4 * (n + 1), hipMemcpyHostToDevice, s));
    if (nnz) {
        AXG_HIP(hipMemcpyAsync(dci.p, ci.data(), 4 * nnz, hipMemcpyHostToDevice, s));
        AXG_HIP(hipMemcpyAsync(dv.p, data, 8 * nnz, hipMemcpyHostToDevice, s));
    }
    AXG_HIP(hipMemcpyAsync(db.p, rhs, 8 * n, hipMemcpyHostToDevice, s));
    CsrView A;
    A.n = n;
    A.nnz = nnz;
    A.rowptr = drp.p;
    A.colind = dci.p;
    A.val = dv.p;
    solve_device_csr(ctx, A, rp.data(), ci.data(), db.p, dx.p, o, info);
    AXG_HIP(hipMemcpyAsync(x_out, dx.p, 8 * n, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    API_END
}

// ------------------------------------------------------------- assembly
int axg_mesh_create(axg_ctx *ctx, int32_t nn, int64_t nnode, const double *xyz, int64_t nelem,
                    const int32_t *conn, const double *e, const double *nu, const int64_t *ind_of,
                    axg_mesh **out) {
    API_BEGIN
    AXG_REQUIRE(ctx && out && xyz && conn && e && nu && ind_of, AXG_E_INVALID,
                "axg_mesh_create: NULL argument");
    AXG_REQUIRE(nn == 8 || nn == 20, AXG_E_INVALID, "axg_mesh_create: nn must be 8 or 20");
    AXG_REQUIRE(nnode > 0 && nelem >= 0 && nnode < INT32_MAX, AXG_E_INVALID,
                "axg_mesh_create: bad sizes");
    const int m = 3 * nn;
    AXG_REQUIRE((uint64_t)nelem * m * m < (uint64_t)UINT32_MAX, AXG_E_INVALID,
                "axg_mesh_create: nelem * (3 nn)^2 must stay below 2^32");
    *out = nullptr;
    AXG_HIP(hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    int64_t ndof = 3 * nnode, n_free = 0;
    for (int64_t d = 0; d < ndof; d++)
        if (ind_of[d] >= 0) n_free++;
    Pattern pat;
    build_pattern(nn, nnode, nelem, conn, ind_of, n_free, pat);
    auto mesh = std::make_unique<axg_mesh>();
    mesh->ctx = ctx;
    mesh->nn = nn;
    mesh->m = m;
    mesh->nnode = nnode;
    mesh->nelem = nelem;
    mesh->ndof = ndof;
    mesh->n_free = n_free;
    mesh->nnz = mesh->nnz_topo = (int64_t)pat.colind.size();
    mesh->ncontrib = mesh->ncontrib_topo = (int64_t)pat.contrib.size();
    hipStream_t s = ctx->stream;
    auto up = [&](auto &buf, const auto *src, size_t count) {
        buf.alloc(count);
        if (count)
            AXG_HIP(hipMemcpyAsync(buf.p, src, count * sizeof(*src), hipMemcpyHostToDevice, s));
    };
    up(mesh->xyz, xyz, (size_t)nnode * 3);
    up(mesh->conn, conn, (size_t)nelem * nn);
    up(mesh->e, e, (size_t)nelem);
    up(mesh->nu, nu, (size_t)nelem);
    std::vector<int32_t> ind32(ndof);
    for (int64_t d = 0; d < ndof; d++) ind32[d] = (int32_t)ind_of[d];
    up(mesh->ind_of, ind32.data(), (size_t)ndof);
    up(mesh->rowptr_t, pat.rowptr.data(), pat.rowptr.size());
    up(mesh->colind_t, pat.colind.data(), pat.colind.size());
    up(mesh->slotptr_t, pat.slotptr.data(), pat.slotptr.size());
    up(mesh->contrib_t, pat.contrib.data(), pat.contrib.size());
    mesh->h_rowptr_t = std::move(pat.rowptr);
    mesh->h_colind_t = std::move(pat.colind);
    // node -> (element, local) incidence in ascending element order, for stresses
    {
        std::vector<int64_t> ptr(nnode + 1, 0);
        for (int64_t el = 0; el < nelem; el++)
            for (int a = 0; a < nn; a++) ptr[conn[el * nn + a] + 1]++;
        for (int64_t i = 0; i < nnode; i++) ptr[i + 1] += ptr[i];
        std::vector<int64_t> slot(ptr[nnode]), cur(ptr.begin(), ptr.end() - 1);
        for (int64_t el = 0; el < nelem; el++)
            for (int a = 0; a < nn; a++) slot[cur[conn[el * nn + a]]++] = el * nn + a;
        up(mesh->inc_ptr, ptr.data(), ptr.size());
        up(mesh->inc_slot, slot.data(), slot.size());
    }
    mesh->val.alloc(std::max<int64_t>(mesh->nnz_topo, 1));
    mesh->slot_nz.alloc(std::max<int64_t>(mesh->nnz_topo, 1));
    mesh->dsel.alloc(std::max<int64_t>(n_free, 1));
    mesh->ke.alloc((size_t)nelem * m * m);
    mesh->vol.alloc(std::max<int64_t>(nelem, 1));
    AXG_HIP(hipStreamSynchronize(s));
    mesh->symbolic_s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    ctx->refs.fetch_add(1);
    *out = mesh.release();
    API_END
}

int axg_mesh_destroy(axg_mesh *mesh) {
    if (!mesh) return AXG_OK;
    axg_ctx *ctx = mesh->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (mesh->bpcg) block_pcg_destroy(mesh->bpcg);
    if (mesh->bpcg3) block_pcg_destroy(mesh->bpcg3);
    delete mesh;
    ctx_release(ctx);
    return AXG_OK;
}

int axg_mesh_info(axg_mesh *mesh, int64_t *n_free, int64_t *nnz, int64_t *ncontrib,
                  double *symbolic_s) {
    API_BEGIN
    AXG_REQUIRE(mesh, AXG_E_INVALID, "axg_mesh_info: NULL mesh");
    if (n_free) *n_free = mesh->n_free;
    if (nnz) *nnz = mesh->nnz;
    if (ncontrib) *ncontrib = mesh->ncontrib;
    if (symbolic_s) *symbolic_s = mesh->symbolic_s;
    API_END
}

int axg_assembly_times(axg_mesh *mesh, double *ke64_ms, double *gather64_ms, double *ke32_ms,
                       double *gather32_ms) {
    API_BEGIN
    AXG_REQUIRE(mesh, AXG_E_INVALID, "axg_assembly_times: NULL mesh");
    if (ke64_ms) *ke64_ms = mesh->t_ke64;
    if (gather64_ms) *gather64_ms = mesh->t_gather64;
    if (ke32_ms) *ke32_ms = mesh->t_ke32;
    if (gather32_ms) *gather32_ms = mesh->t_gather32;
    API_END
}

int axg_mesh_set_modulus(axg_mesh *mesh, const double *e) {
    API_BEGIN
    AXG_REQUIRE(mesh && e, AXG_E_INVALID, "axg_mesh_set_modulus: NULL argument");
    AXG_HIP(hipSetDevice(mesh->ctx->device));
    AXG_HIP(hipMemcpyAsync(mesh->e.p, e, 8 * mesh->nelem, hipMemcpyHostToDevice,
                           mesh->ctx->stream));
    AXG_HIP(hipStreamSynchronize(mesh->ctx->stream));
    mesh->assembled64 = mesh->assembled32 = false;
    API_END
}

int axg_assemble(axg_mesh *mesh, int32_t precision, double *elapsed_ms) {
    API_BEGIN
    AXG_REQUIRE(mesh, AXG_E_INVALID, "axg_assemble: NULL mesh");
    AXG_REQUIRE(precision == AXG_PREC_FP64 || precision == AXG_PREC_MIXED, AXG_E_INVALID,
                "axg_assemble: unknown precision");
    axg_ctx *ctx = mesh->ctx;
    AXG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    Workspace &w = ctx->ws;
    AXG_HIP(hipMemsetAsync(w.err.p, 0, 4 * sizeof(int32_t), s));
    AXG_HIP(hipMemsetAsync(w.cnt.p, 0, 4 * sizeof(uint32_t), s));
    AXG_HIP(hipEventRecord(ctx->ev0, s));
    launch_ke_fp64(mesh, s, w.err.p);
    AXG_HIP(hipEventRecord(ctx->evA, s));
    launch_gather_fp64(mesh, s, true, w.cnt.p);
    AXG_HIP(hipEventRecord(ctx->ev1, s));
    int32_t err[2];
    uint32_t zero_slots = 0;
    AXG_HIP(hipMemcpyAsync(err, w.err.p, 8, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(&zero_slots, w.cnt.p, 4, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    float ms = 0.f, ms_a = 0.f;
    AXG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    AXG_HIP(hipEventElapsedTime(&ms_a, ctx->ev0, ctx->evA));
    mesh->t_ke64 = ms_a;
    mesh->t_gather64 = ms - ms_a;
    if (err[0] != 0) {
        std::vector<double> e(mesh->nelem), nu(mesh->nelem);
        AXG_HIP(hipMemcpy(e.data(), mesh->e.p, 8 * mesh->nelem, hipMemcpyDeviceToHost));
        AXG_HIP(hipMemcpy(nu.data(), mesh->nu.p, 8 * mesh->nelem, hipMemcpyDeviceToHost));
        throw Error{err[0], element_error(err[0], mesh->nn, err[1], e.data(), nu.data())};
    }
    finish_gather64(mesh, zero_slots);
    if (precision == AXG_PREC_MIXED) {
        double ms32 = 0.0;
        assemble_mixed(mesh, &ms32);
        ms += (float)ms32;
    }
    if (elapsed_ms) *elapsed_ms = ms;
    API_END
}

int axg_csr_export(axg_mesh *mesh, int64_t *indptr, int64_t *indices, double *data) {
    API_BEGIN
    AXG_REQUIRE(mesh && mesh->assembled64, AXG_E_INVALID, "axg_csr_export: mesh not assembled");
    hipStream_t s = mesh->ctx->stream;
    CsrView A = active_csr(mesh);
    std::vector<int32_t> rp(A.n + 1), ci(A.nnz);
    AXG_HIP(hipMemcpyAsync(rp.data(), A.rowptr, 4 * (A.n + 1), hipMemcpyDeviceToHost, s));
    if (A.nnz) AXG_HIP(hipMemcpyAsync(ci.data(), A.colind, 4 * A.nnz, hipMemcpyDeviceToHost, s));
    if (data && A.nnz) AXG_HIP(hipMemcpyAsync(data, A.val, 8 * A.nnz, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    if (indptr)
        for (int64_t i = 0; i <= A.n; i++) indptr[i] = rp[i];
    if (indices)
        for (int64_t k = 0; k < A.nnz; k++) indices[k] = ci[k];
    API_END
}

int axg_scatter_map_export(axg_mesh *mesh, int64_t *slot_ptr, int64_t *contrib) {
    API_BEGIN
    AXG_REQUIRE(mesh, AXG_E_INVALID, "axg_scatter_map_export: NULL mesh");
    hipStream_t s = mesh->ctx->stream;
    const uint32_t *sp = mesh->compacted ? mesh->slotptr_c.p : mesh->slotptr_t.p;
    const uint32_t *cb = mesh->compacted ? mesh->contrib_c.p : mesh->contrib_t.p;
    std::vector<uint32_t> a(mesh->nnz + 1), b(mesh->ncontrib);
    AXG_HIP(hipMemcpyAsync(a.data(), sp, 4 * (mesh->nnz + 1), hipMemcpyDeviceToHost, s));
    if (mesh->ncontrib)
        AXG_HIP(hipMemcpyAsync(b.data(), cb, 4 * mesh->ncontrib, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    if (slot_ptr)
        for (size_t i = 0; i < a.size(); i++) slot_ptr[i] = a[i];
    if (contrib)
        for (size_t i = 0; i < b.size(); i++) contrib[i] = b[i];
    API_END
}

int axg_ke_export(axg_mesh *mesh, int64_t first, int64_t count, double *ke) {
    API_BEGIN
    AXG_REQUIRE(mesh && ke && first >= 0 && count >= 0 && first + count <= mesh->nelem,
                AXG_E_INVALID, "axg_ke_export: invalid range");
    AXG_REQUIRE(mesh->assembled64, AXG_E_INVALID, "axg_ke_export: mesh not assembled");
    size_t mm = (size_t)mesh->m * mesh->m;
    AXG_HIP(hipMemcpy(ke, mesh->ke.p + first * mm, 8 * mm * count, hipMemcpyDeviceToHost));
    API_END
}

int axg_op32_export(axg_mesh *mesh, int64_t *count, int32_t *rows, int32_t *cols, float *vals) {
    API_BEGIN
    AXG_REQUIRE(mesh && count, AXG_E_INVALID, "axg_op32_export: NULL argument");
    AXG_REQUIRE(mesh->assembled32, AXG_E_INVALID, "axg_op32_export: mixed operator not assembled");
    AXG_REQUIRE(!rows == !cols && !rows == !vals, AXG_E_INVALID,
                "axg_op32_export: rows, cols and vals are all given or all NULL");
    const int64_t nb = mesh->nb, nsl = mesh->bs_nslots, ns = (nb + 63) / 64;
    hipStream_t s = mesh->ctx->stream;
    std::vector<int64_t> sptr(ns + 1);
    std::vector<int32_t> width(ns), bcol(nsl), fob(3 * nb);
    std::vector<float> bv(10 * nsl);
    AXG_HIP(hipMemcpyAsync(sptr.data(), mesh->bs_sptr.p, 8 * (ns + 1), hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(width.data(), mesh->bs_width.p, 4 * ns, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(bcol.data(), mesh->bs_col.p, 4 * nsl, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(fob.data(), mesh->free_of_blk.p, 4 * 3 * nb, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(bv.data(), mesh->bval.p, 4 * 10 * nsl, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    const int64_t cap = *count;
    int64_t n = 0;
    for (int64_t I = 0; I < nb; I++) {
        const int64_t base = sptr[I >> 6] + (I & 63);
        for (int64_t k = 0; k < width[I >> 6]; k++) {
            const int64_t pos = base + 64 * k;
            const int32_t J = bcol[pos];
            if (k > 0 && J == (int32_t)I) continue;  // padding slot
            for (int q = 0; q < 9; q++) {
                const int32_t fi = fob[3 * I + q / 3], fj = fob[3 * (int64_t)J + q % 3];
                if (fi < 0 || fj < 0) continue;
                if (rows) {
                    AXG_REQUIRE(n < cap, AXG_E_INVALID, "axg_op32_export: arrays too small");
                    rows[n] = fi;
                    cols[n] = fj;
                    vals[n] = q < 4 ? bv[4 * pos + q] : q < 8 ? bv[4 * nsl + 4 * pos + q - 4]
                                                              : bv[8 * nsl + 2 * pos];
                }
                n++;
            }
        }
    }
    *count = n;
    API_END
}

// ---------------------------------------------------------------- solve
static void mesh_solve_device(axg_mesh *mesh, const double *d_rhs, int32_t nrhs,
                              const axg_solve_opts &o, double *d_x, axg_solve_info *info) {
    AXG_REQUIRE(mesh->assembled64, AXG_E_INVALID, "mesh not assembled (call axg_assemble)");
    axg_ctx *ctx = mesh->ctx;
    CsrView A = solver_csr(mesh);
    const int64_t n = mesh->n_free;
    const int64_t max_iter = default_max_iter(o, n);
    if (o.precision == AXG_PREC_MIXED) {
        if (!mesh->assembled32) {
            double ms = 0.0;
            assemble_mixed(mesh, &ms);
        }
        MixedResult mr;
        solve_mixed(mesh, d_rhs, nrhs, o, d_x, mr);
        if (mr.status != AXG_OK) {
            PcgResult r;
            r.status = mr.status;
            r.residual = mr.residual;
            r.iters = mr.inner_iters;
            throw Error{mr.status, pcg_error(r)};
        }
        char name[64];
        std::snprintf(name, sizeof name, "GPU PCG (FP32 BSR-3 + FP64 %s, %s)",
                      o.refinement == AXG_REFINE_PLAIN      ? "IR"
                      : o.refinement == AXG_REFINE_RELIABLE ? "RU"
                                                            : "FCG",
                      o.precond == AXG_PRECOND_BLOCK3 ? "block-Jacobi" : "Jacobi");
        fill_info(info, name, mr.inner_iters, mr.outer_iters, mr.residual, mr.bnorm, mr.solve_ms,
                  mr.spmv_ms, mr.spmv_launches, nrhs);
        return;
    }
    if (block_pcg_fits(mesh)) {
        // the PCG (the reference's Jacobi, or block-Jacobi), every right-hand
        // side at once on the node-block operator (9 values and one column per
        // 3x3 block)
        const bool blk3 = o.precond == AXG_PRECOND_BLOCK3;
        BlockPcg *&bp = blk3 ? mesh->bpcg3 : mesh->bpcg;
        if (!bp) bp = block_pcg_create(mesh, max_iter, blk3);
        block_pcg_refresh(bp, max_iter);
        std::vector<double> tolr(nrhs, o.tol);
        PcgResult r;
        block_pcg_solve(bp, d_rhs, nrhs, tolr.data(), d_x, r, o.timing != 0);
        if (r.status != AXG_OK) throw Error{r.status, pcg_error(r)};
        double worst = 0.0, worst_b = 0.0;
        for (int32_t k = 0; k < nrhs; k++) {
            const double res = true_residual(ctx, A, d_x + k * n, d_rhs + k * n);
            const double bn = device_norm(ctx, d_rhs + k * n, n);
            if (k == 0 || res / std::max(bn, 1e-300) > worst / std::max(worst_b, 1e-300)) {
                worst = res;
                worst_b = bn;
            }
        }
        fill_info(info, blk3 ? "GPU PCG (FP64 node blocks, block-Jacobi)"
                             : "GPU PCG (FP64 node blocks, Jacobi)",
                  r.iters, 1, worst, worst_b, r.solve_ms, r.spmv_ms, r.spmv_launches, nrhs);
        return;
    }
    Precond64Args pa;
    pa.dsel = mesh->dsel.p;
    if (o.precond == AXG_PRECOND_BLOCK3) {
        ensure_block_inverse64(mesh);
        pa.binv = mesh->binv64.p;
        pa.blk_of_free = mesh->blk_of_free.p;
        pa.free_of_blk = mesh->free_of_blk.p;
    }
    int64_t iters = 0, launches = 0;
    double worst = 0.0, worst_b = 0.0, ms = 0.0, spmv_ms = 0.0;
    for (int32_t k = 0; k < nrhs; k++) {
        PcgResult r;
        pcg_fp64(ctx, A, pa, d_rhs + k * n, d_x + k * n, o.tol, max_iter, o.timing != 0, r);
        if (r.status != AXG_OK) throw Error{r.status, pcg_error(r)};
        double res = true_residual(ctx, A, d_x + k * n, d_rhs + k * n);
        if (k == 0 || res / std::max(r.bnorm, 1e-300) > worst / std::max(worst_b, 1e-300)) {
            worst = res;
            worst_b = r.bnorm;
        }
        iters += r.iters;
        ms += r.solve_ms;
        spmv_ms += r.spmv_ms;
        launches += r.spmv_launches;
    }
    fill_info(info,
              o.precond == AXG_PRECOND_BLOCK3 ? "GPU PCG (FP64, block-Jacobi)"
                                              : "GPU PCG (FP64, Jacobi)",
              iters, 1, worst, worst_b, ms, spmv_ms, launches, nrhs);
}

int axg_mesh_solve(axg_mesh *mesh, const double *rhs, int32_t nrhs, const axg_solve_opts *opts,
                   double *x_out, axg_solve_info *info) {
    API_BEGIN
    AXG_REQUIRE(mesh && rhs && x_out && nrhs >= 1, AXG_E_INVALID, "axg_mesh_solve: invalid argument");
    axg_solve_opts o;
    axg_default_opts(&o);
    if (opts) o = *opts;
    AXG_HIP(hipSetDevice(mesh->ctx->device));
    hipStream_t s = mesh->ctx->stream;
    size_t cnt = (size_t)mesh->n_free * nrhs;
    DevBuf<double> db(cnt), dx(cnt);
    AXG_HIP(hipMemcpyAsync(db.p, rhs, 8 * cnt, hipMemcpyHostToDevice, s));
    mesh_solve_device(mesh, db.p, nrhs, o, dx.p, info);
    AXG_HIP(hipMemcpyAsync(x_out, dx.p, 8 * cnt, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    API_END
}

int axg_mesh_solve_device(axg_mesh *mesh, const double *d_rhs, int32_t nrhs,
                          const axg_solve_opts *opts, double *d_x, axg_solve_info *info) {
    API_BEGIN
    AXG_REQUIRE(mesh && d_rhs && d_x && nrhs >= 1, AXG_E_INVALID,
                "axg_mesh_solve_device: invalid argument");
    axg_solve_opts o;
    axg_default_opts(&o);
    if (opts) o = *opts;
    AXG_HIP(hipSetDevice(mesh->ctx->device));
    mesh_solve_device(mesh, d_rhs, nrhs, o, d_x, info);
    AXG_HIP(hipStreamSynchronize(mesh->ctx->stream));
    API_END
}

// An axg_mesh without elements whose K_ff is a caller's CSR (n rows, int32
// pattern rp / ci on the host, values d_val on the device): block row i holds
// rows 3i .. 3i+2 (the last one padded with absent rows), every stored value is
// the single contribution of its slot (identity scatter map, ke = the values),
// so the block layouts, the 3x3 inverses, the FP32 operator and the solvers of
// a mesh apply unchanged.
static void csr_operator_solve(axg_ctx *ctx, int64_t n, const int32_t *rp, const int32_t *ci,
                               const double *d_val, const double *d_b, double *d_x,
                               const axg_solve_opts &o, axg_solve_info *info) {
    hipStream_t s = ctx->stream;
    const int64_t nnz = rp[n];
    AXG_REQUIRE(n > 0, AXG_E_INVALID, "empty operator");
    auto mesh = std::make_unique<axg_mesh>();
    mesh->ctx = ctx;
    mesh->csr_operator = true;
    mesh->nnode = (n + 2) / 3;
    mesh->nelem = 0;
    mesh->ndof = 3 * mesh->nnode;
    mesh->n_free = n;
    mesh->nnz = mesh->nnz_topo = nnz;
    mesh->ncontrib = mesh->ncontrib_topo = nnz;
    const bool plain = o.precision == AXG_PREC_FP64 && o.precond == AXG_PRECOND_JACOBI;
    std::vector<int32_t> ind(mesh->ndof);
    for (int64_t d = 0; d < mesh->ndof; d++) ind[d] = d < n ? (int32_t)d : -1;
    mesh->h_rowptr_t.assign(rp, rp + n + 1);
    mesh->h_colind_t.assign(ci, ci + nnz);
    for (int64_t i = 0; i < n && mesh->cols_sorted; i++)
        for (int32_t k = rp[i] + 1; k < rp[i + 1]; k++)
            if (ci[k] <= ci[k - 1]) {
                mesh->cols_sorted = false;
                break;
            }
    auto up = [&](auto &buf, const auto *src, size_t count) {
        buf.alloc(count);
        if (count)
            AXG_HIP(hipMemcpyAsync(buf.p, src, count * sizeof(*src), hipMemcpyHostToDevice, s));
    };
    up(mesh->ind_of, ind.data(), ind.size());
    up(mesh->rowptr_t, rp, (size_t)n + 1);
    up(mesh->colind_t, ci, (size_t)nnz);
    mesh->val.alloc(std::max<int64_t>(nnz, 1));
    if (nnz) AXG_HIP(hipMemcpyAsync(mesh->val.p, d_val, 8 * nnz, hipMemcpyDeviceToDevice, s));
    if (!plain) {
        // the node-block layouts gather slot values through the scatter map
        std::vector<uint32_t> ident(nnz + 1);
        for (int64_t k = 0; k <= nnz; k++) ident[k] = (uint32_t)k;
        up(mesh->slotptr_t, ident.data(), ident.size());
        up(mesh->contrib_t, ident.data(), (size_t)nnz);
        mesh->ke.alloc(std::max<int64_t>(nnz, 1));
        if (nnz) AXG_HIP(hipMemcpyAsync(mesh->ke.p, d_val, 8 * nnz, hipMemcpyDeviceToDevice, s));
    }
    mesh->dsel.alloc(n);
    CsrView A = active_csr(mesh.get());
    diag_fp64(ctx, A, mesh->dsel.p, nullptr);
    AXG_HIP(hipStreamSynchronize(s));
    mesh->assembled64 = true;
    ctx->refs.fetch_add(1);
    struct Drop {
        axg_mesh *m;
        ~Drop() { axg_mesh_destroy(m); }
    } drop{mesh.release()};
    if (!plain) {
        mesh_solve_device(drop.m, d_b, 1, o, d_x, info);
        return;
    }
    // linalg::pcg: Jacobi, the scalar rows (sliced-ELL copy), one right-hand side
    CsrView As = solver_csr(drop.m);
    Precond64Args pa;
    pa.dsel = drop.m->dsel.p;
    PcgResult r;
    pcg_fp64(ctx, As, pa, d_b, d_x, o.tol, default_max_iter(o, n), o.timing != 0, r);
    if (r.status != AXG_OK) throw Error{r.status, pcg_error(r)};
    const double res = true_residual(ctx, As, d_x, d_b);
    fill_info(info, "GPU PCG (FP64, Jacobi)", r.iters, 1, res, r.bnorm, r.solve_ms, r.spmv_ms,
              r.spmv_launches, 1);
}

// ------------------------------------------------------ stress, adjoint
int axg_nodal_stress(axg_mesh *mesh, const double *u_full, double *stress, double *vm,
                     double *elem_nodal) {
    API_BEGIN
    AXG_REQUIRE(mesh && u_full && stress, AXG_E_INVALID, "axg_nodal_stress: NULL argument");
    axg_ctx *ctx = mesh->ctx;
    AXG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    Workspace &w = ctx->ws;
    DevBuf<double> du(mesh->ndof), den((size_t)mesh->nelem * mesh->nn * 6),
        dst((size_t)mesh->nnode * 6), dvm(mesh->nnode);
    AXG_HIP(hipMemcpyAsync(du.p, u_full, 8 * mesh->ndof, hipMemcpyHostToDevice, s));
    AXG_HIP(hipMemsetAsync(w.err.p, 0, 8, s));
    launch_nodal_stress(mesh, s, du.p, den.p, dst.p, dvm.p, w.err.p);
    int32_t err[2];
    AXG_HIP(hipMemcpyAsync(err, w.err.p, 8, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(stress, dst.p, 8 * mesh->nnode * 6, hipMemcpyDeviceToHost, s));
    if (vm) AXG_HIP(hipMemcpyAsync(vm, dvm.p, 8 * mesh->nnode, hipMemcpyDeviceToHost, s));
    if (elem_nodal)
        AXG_HIP(hipMemcpyAsync(elem_nodal, den.p, 8 * (size_t)mesh->nelem * mesh->nn * 6,
                               hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    if (err[0]) throw Error{err[0], element_error(err[0], mesh->nn, err[1], nullptr, nullptr)};
    API_END
}


int axg_geometric_stiffness(axg_mesh *mesh, const double *u_full, double *mean, double *kg_val) {
    API_BEGIN
    AXG_REQUIRE(mesh && u_full, AXG_E_INVALID, "axg_geometric_stiffness: NULL argument");
    AXG_REQUIRE(mesh->assembled64, AXG_E_INVALID, "mesh not assembled (call axg_assemble)");
    axg_ctx *ctx = mesh->ctx;
    AXG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    DevBuf<double> du(std::max<int64_t>(mesh->ndof, 1)),
        dval(std::max<int64_t>(mesh->nnz_topo, 1)),
        dmean((size_t)std::max<int64_t>(mesh->nelem, 1) * 6);
    AXG_HIP(hipMemcpyAsync(du.p, u_full, 8 * mesh->ndof, hipMemcpyHostToDevice, s));
    geometric_stiffness_device(mesh, du.p, 1.0, dval.p, dmean.p);
    if (mean && mesh->nelem > 0)
        AXG_HIP(hipMemcpyAsync(mean, dmean.p, 8 * (size_t)mesh->nelem * 6, hipMemcpyDeviceToHost, s));
    if (kg_val && mesh->nnz_topo > 0)
        AXG_HIP(hipMemcpyAsync(kg_val, dval.p, 8 * mesh->nnz_topo, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    API_END
}

int axg_nl_assemble(axg_mesh *mesh, const double *u_full, double *f_int, double *cauchy,
                    double *gl) {
    API_BEGIN
    AXG_REQUIRE(mesh && u_full, AXG_E_INVALID, "axg_nl_assemble: NULL argument");
    axg_ctx *ctx = mesh->ctx;
    AXG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int64_t ne = std::max<int64_t>(mesh->nelem, 1);
    DevBuf<double> du(std::max<int64_t>(mesh->ndof, 1)), fe((size_t)ne * mesh->m), cau(6 * ne),
        glb(6 * ne), fint(std::max<int64_t>(mesh->ndof, 1));
    AXG_HIP(hipMemcpyAsync(du.p, u_full, 8 * mesh->ndof, hipMemcpyHostToDevice, s));
    nl_assemble_device(mesh, du.p, fe.p, cau.p, glb.p, fint.p);
    if (f_int) AXG_HIP(hipMemcpyAsync(f_int, fint.p, 8 * mesh->ndof, hipMemcpyDeviceToHost, s));
    if (cauchy) AXG_HIP(hipMemcpyAsync(cauchy, cau.p, 8 * 6 * mesh->nelem, hipMemcpyDeviceToHost, s));
    if (gl) AXG_HIP(hipMemcpyAsync(gl, glb.p, 8 * 6 * mesh->nelem, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    API_END
}

int axg_compliance_adjoint(axg_mesh *mesh, const double *u_full, const double *f_full, double *c,
                           double *dc_de) {
    API_BEGIN
    AXG_REQUIRE(mesh && u_full && f_full && dc_de, AXG_E_INVALID,
                "axg_compliance_adjoint: NULL argument");
    axg_ctx *ctx = mesh->ctx;
    AXG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    Workspace &w = ctx->ws;
    DevBuf<double> du(mesh->ndof), den(mesh->nelem);
    AXG_HIP(hipMemcpyAsync(du.p, u_full, 8 * mesh->ndof, hipMemcpyHostToDevice, s));
    AXG_HIP(hipMemsetAsync(w.err.p, 0, 8, s));
    launch_strain_energy(mesh, s, du.p, den.p, w.err.p);
    int32_t err[2];
    AXG_HIP(hipMemcpyAsync(err, w.err.p, 8, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipMemcpyAsync(dc_de, den.p, 8 * mesh->nelem, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    if (err[0]) throw Error{err[0], element_error(err[0], mesh->nn, err[1], nullptr, nullptr)};
    for (int64_t el = 0; el < mesh->nelem; el++) dc_de[el] = -dc_de[el];
    if (c) {
        double acc = 0.0;
        for (int64_t d = 0; d < mesh->ndof; d++) acc += f_full[d] * u_full[d];
        *c = acc;
    }
    API_END
}

// Full-space products with the stored FP64 element matrices (last assembly).
// mode 0: reactions (ndof out), 1: reduced rhs (n_free out), 2: diagonal (ndof out)
static void row_apply(axg_mesh *mesh, int mode, const double *w, const double *f, double *out) {
    AXG_REQUIRE(mesh->assembled64, AXG_E_INVALID, "mesh not assembled (call axg_assemble)");
    axg_ctx *ctx = mesh->ctx;
    AXG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int64_t ndof = mesh->ndof, nout = mode == 1 ? mesh->n_free : ndof;
    DevBuf<double> dw(ndof), df(ndof), dout(std::max<int64_t>(nout, 1));
    if (w) AXG_HIP(hipMemcpyAsync(dw.p, w, 8 * ndof, hipMemcpyHostToDevice, s));
    if (f) AXG_HIP(hipMemcpyAsync(df.p, f, 8 * ndof, hipMemcpyHostToDevice, s));
    launch_row_apply(mesh, s, mode, dw.p, df.p, dout.p);
    if (nout) AXG_HIP(hipMemcpyAsync(out, dout.p, 8 * nout, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
}

int axg_reactions(axg_mesh *mesh, const double *u_full, const double *f_full, double *rf) {
    API_BEGIN
    AXG_REQUIRE(mesh && u_full && f_full && rf, AXG_E_INVALID, "axg_reactions: NULL argument");
    row_apply(mesh, 0, u_full, f_full, rf);
    API_END
}

int axg_reduce_rhs(axg_mesh *mesh, const double *f_full, const double *u0, double *rhs) {
    API_BEGIN
    AXG_REQUIRE(mesh && f_full && u0 && rhs, AXG_E_INVALID, "axg_reduce_rhs: NULL argument");
    row_apply(mesh, 1, u0, f_full, rhs);
    API_END
}

int axg_full_diagonal(axg_mesh *mesh, double *diag) {
    API_BEGIN
    AXG_REQUIRE(mesh && diag, AXG_E_INVALID, "axg_full_diagonal: NULL argument");
    row_apply(mesh, 2, nullptr, nullptr, diag);
    API_END
}

int axg_element_volumes(axg_mesh *mesh, double *vol) {
    API_BEGIN
    AXG_REQUIRE(mesh && vol, AXG_E_INVALID, "axg_element_volumes: NULL argument");
    AXG_REQUIRE(mesh->assembled64, AXG_E_INVALID, "mesh not assembled (call axg_assemble)");
    AXG_HIP(hipSetDevice(mesh->ctx->device));
    hipStream_t s = mesh->ctx->stream;
    if (mesh->nelem > 0)
        AXG_HIP(hipMemcpyAsync(vol, mesh->vol.p, 8 * mesh->nelem, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    API_END
}

int axg_lumped_mass(axg_mesh *mesh, const double *density, double *m_full) {
    API_BEGIN
    AXG_REQUIRE(mesh && density && m_full, AXG_E_INVALID, "axg_lumped_mass: NULL argument");
    AXG_REQUIRE(mesh->assembled64, AXG_E_INVALID, "mesh not assembled (call axg_assemble)");
    AXG_HIP(hipSetDevice(mesh->ctx->device));
    hipStream_t s = mesh->ctx->stream;
    DevBuf<double> rho(std::max<int64_t>(mesh->nelem, 1)), dm(std::max<int64_t>(mesh->ndof, 1));
    if (mesh->nelem > 0)
        AXG_HIP(hipMemcpyAsync(rho.p, density, 8 * mesh->nelem, hipMemcpyHostToDevice, s));
    launch_lumped_mass(mesh, s, rho.p, dm.p);
    if (mesh->ndof > 0)
        AXG_HIP(hipMemcpyAsync(m_full, dm.p, 8 * mesh->ndof, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    API_END
}

// Host-only entry points (no device needed).
int axg_pattern_host(int32_t nn, int64_t nnode, int64_t nelem, const int32_t *conn,
                     const int64_t *ind_of, int64_t *n_free, int64_t *nnz, int64_t *ncontrib,
                     int64_t *rowptr, int64_t *colind, int64_t *slot_ptr, int64_t *contrib) {
    API_BEGIN
    AXG_REQUIRE(conn && ind_of && nnz && ncontrib && n_free, AXG_E_INVALID,
                "axg_pattern_host: NULL argument");
    AXG_REQUIRE(nn == 8 || nn == 20, AXG_E_INVALID, "axg_pattern_host: nn must be 8 or 20");
    int64_t nf = 0;
    for (int64_t d = 0; d < 3 * nnode; d++) nf += ind_of[d] >= 0;
    Pattern pat;
    build_pattern(nn, nnode, nelem, conn, ind_of, nf, pat);
    *n_free = nf;
    *nnz = (int64_t)pat.colind.size();
    *ncontrib = (int64_t)pat.contrib.size();
    if (rowptr)
        for (size_t i = 0; i < pat.rowptr.size(); i++) rowptr[i] = pat.rowptr[i];
    if (colind)
        for (size_t i = 0; i < pat.colind.size(); i++) colind[i] = pat.colind[i];
    if (slot_ptr)
        for (size_t i = 0; i < pat.slotptr.size(); i++) slot_ptr[i] = pat.slotptr[i];
    if (contrib)
        for (size_t i = 0; i < pat.contrib.size(); i++) contrib[i] = pat.contrib[i];
    API_END
}

// Face-pressure loads (elem.rs:892-950, quadratic.rs:266-315) accumulated
// into f_full in the given face order (analysis.rs:723-760 visits elements
// in order and, per element, the *DLOAD entries in deck order).
int axg_pressure_loads(int32_t nn, int64_t nnode, const double *xyz, int64_t nface,
                       const int32_t *elem_conn, const int32_t *face, const double *mag,
                       double *f_full) {
    API_BEGIN
    AXG_REQUIRE(nn == 8 || nn == 20, AXG_E_INVALID, "axg_pressure_loads: nn must be 8 or 20");
    AXG_REQUIRE(nface == 0 || (xyz && elem_conn && face && mag && f_full), AXG_E_INVALID,
                "axg_pressure_loads: NULL argument");
    double exyz[60], fe[60];
    for (int64_t k = 0; k < nface; k++) {
        for (int a = 0; a < nn; a++) {
            int32_t node = elem_conn[k * nn + a];
            AXG_REQUIRE(node >= 0 && node < nnode, AXG_E_INVALID,
                        "axg_pressure_loads: node index out of range");
            for (int q = 0; q < 3; q++) exyz[3 * a + q] = xyz[3 * (int64_t)node + q];
        }
        int st = face_pressure_host(nn, exyz, face[k], mag[k], fe);
        AXG_REQUIRE(st == AXG_OK, AXG_E_INVALID,
                    std::string("Ungültige ") + (nn == 8 ? "C3D8" : "C3D20") + "-Fläche P" +
                        std::to_string(face[k]));
        for (int a = 0; a < nn; a++)
            for (int q = 0; q < 3; q++)
                f_full[3 * (int64_t)elem_conn[k * nn + a] + q] += fe[3 * a + q];
    }
    API_END
}

}  // extern "C"
