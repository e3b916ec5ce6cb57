// The FP64 Jacobi PCG of axg_solve_csr on device-resident data for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc): every intermediate lives in one
// fixed array of this file, `scratch`, of which each device holds its own copy
// for the life of the process. Its device address is looked up once per device
// at context creation (axg_create, with the device current) and kept per
// device; host code never passes the array's name to a kernel.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "axg_internal.h"
#include "axg_pcg.h"

namespace axg {

constexpr int64_t kNoallocMaxN = int64_t(1) << 20;        // rows
constexpr size_t kNoallocReserve = 16384;                  // partials + scalars
constexpr size_t kScratchDoubles = 5 * (size_t)kNoallocMaxN + kNoallocReserve;  // 40.1 MiB

static __device__ __align__(256) double scratch[kScratchDoubles];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[64] = {};

// A failed lookup leaves the device without the entry point (it then returns
// AXG_E_DEVICE) instead of failing the context.
void noalloc_setup(int device) {
    if (device < 0 || device >= 64) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    double *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void **>(&ptr), HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();  // clear it for the caller's next check
        return;
    }
    g_scratch_ptr[device] = ptr;
}

static double *scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < 64) ? g_scratch_ptr[device] : nullptr;
}

}  // namespace axg

using namespace axg;

int axg_solve_csr_device_noalloc(axg_ctx *ctx, int64_t n, const int32_t *d_rowptr,
                                 const int32_t *d_colind, const double *d_val, const double *d_rhs,
                                 const axg_solve_opts *opts, double *d_x, axg_solve_info *info) {
    API_BEGIN
    AXG_REQUIRE(ctx && n >= 0 && d_rowptr && d_rhs && d_x, AXG_E_INVALID,
                "axg_solve_csr_device_noalloc: invalid argument");
    AXG_REQUIRE(n <= kNoallocMaxN, AXG_E_INVALID,
                "axg_solve_csr_device_noalloc: n exceeds the fixed scratch (at most 1048576 rows)");
    axg_solve_opts o;
    axg_default_opts(&o);
    if (opts) o = *opts;
    AXG_REQUIRE(o.precision == AXG_PREC_FP64 && o.precond == AXG_PRECOND_JACOBI, AXG_E_INVALID,
                "axg_solve_csr_device_noalloc: FP64 with Jacobi only");
    AXG_REQUIRE(o.tol > 0.0, AXG_E_INVALID, "axg_solve_csr_device_noalloc: tol must be positive");
    AXG_REQUIRE(pcg_noalloc_doubles(ctx, n) <= kScratchDoubles, AXG_E_INVALID,
                "axg_solve_csr_device_noalloc: scratch too small for this device");
    double *buf = scratch_of(ctx->device);
    AXG_REQUIRE(buf != nullptr, AXG_E_DEVICE, "axg_solve_csr_device_noalloc: no scratch on device");
    AXG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    if (info) std::memset(info, 0, sizeof(*info));
    if (info) std::snprintf(info->name, sizeof(info->name), "GPU PCG (FP64, Jacobi)");
    if (info) info->nrhs = 1, info->outer_iters = 1;
    if (n == 0) return AXG_OK;
    int32_t nnz = 0;
    AXG_HIP(hipMemcpyAsync(&nnz, d_rowptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    AXG_REQUIRE(nnz >= 0 && (nnz == 0 || (d_colind && d_val)), AXG_E_INVALID,
                "axg_solve_csr_device_noalloc: bad row pointer or NULL columns / values");
    CsrView A;
    A.n = n;
    A.nnz = nnz;
    A.rowptr = d_rowptr;
    A.colind = d_colind;
    A.val = d_val;
    const int64_t max_iter = o.max_iter > 0 ? o.max_iter : std::max<int64_t>(4 * n, 200);
    PcgResult r;
    const double res = pcg_fp64_noalloc(ctx, A, d_rhs, d_x, o.tol, max_iter, buf, r);
    if (r.status != AXG_OK) throw Error{r.status, pcg_error(r)};
    if (info) {
        info->iters = r.iters;
        info->residual = res;
        info->rel_residual = r.bnorm > 0 ? res / r.bnorm : 0.0;
        info->solve_ms = r.solve_ms;
    }
    API_END
}
