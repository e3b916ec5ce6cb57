// Kernel microbenchmarks behind the C ABI: a STREAM-style read/copy kernel
// that calibrates achievable HBM bandwidth (and the rocprof byte counters),
// and repeated launches of the solver's operator kernels on a mesh.
#include "axg_internal.h"
#include "axg_mixed.h"
#include "axg_pcg.h"

namespace axg {

// y = x (n doubles): reads 8 n bytes, writes 8 n bytes.
__global__ void __launch_bounds__(256) stream_copy_kernel(int64_t n, const double *__restrict__ x,
                                                         double *__restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        y[i] = __builtin_nontemporal_load(&x[i]);
}

void bench_operator_mixed(axg_mesh *m, int nrhs, int reps, double *avg_ms);

}  // namespace axg

using namespace axg;

extern "C" int axg_bench_stream(axg_ctx *ctx, int64_t n, int32_t reps, double *copy_gbs) {
    try {
        clear_error();
        AXG_REQUIRE(ctx && n > 0 && reps > 0 && copy_gbs, AXG_E_INVALID,
                    "axg_bench_stream: invalid argument");
        AXG_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        DevBuf<double> x(n), y(n);
        AXG_HIP(hipMemsetAsync(x.p, 0, 8 * n, s));
        int grid = ctx->cu_count * 16;
        hipLaunchKernelGGL(stream_copy_kernel, dim3(grid), dim3(256), 0, s, n, x.p, y.p);
        AXG_HIP(hipEventRecord(ctx->ev0, s));
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(stream_copy_kernel, dim3(grid), dim3(256), 0, s, n, x.p, y.p);
        AXG_HIP(hipEventRecord(ctx->ev1, s));
        AXG_HIP(hipEventSynchronize(ctx->ev1));
        float ms = 0.f;
        AXG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *copy_gbs = 16.0 * n * reps / (ms / 1e3) / 1e9;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    }
    return AXG_OK;
}

// which: 0 = FP64 SpMV of the solver (sliced ELL), 1 = FP64 CSR SpMV,
// 2 = FP32 block-sliced-ELL SpMM with nrhs right-hand sides.
extern "C" int axg_mesh_spmm_shape(axg_mesh *m, int32_t nrhs, int32_t *shape) {
    try {
        clear_error();
        AXG_REQUIRE(m && shape && nrhs >= 1 && nrhs <= 8, AXG_E_INVALID,
                    "axg_mesh_spmm_shape: invalid argument");
        if (*shape >= 0) {
            AXG_REQUIRE(spmm_shape_ok(*shape, nrhs), AXG_E_INVALID,
                        "axg_mesh_spmm_shape: no such shape for this many right-hand sides");
            m->spmm_shape[nrhs] = *shape;
        }
        *shape = m->spmm_shape[nrhs];
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    }
    return AXG_OK;
}

extern "C" int axg_mesh_asm32_mode(axg_mesh *m, int32_t *mode) {
    try {
        clear_error();
        AXG_REQUIRE(m && mode && *mode <= 1, AXG_E_INVALID, "axg_mesh_asm32_mode: invalid argument");
        if (*mode >= 0) m->asm32_mode = *mode;
        *mode = asm32_mode_of(m);
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    }
    return AXG_OK;
}

extern "C" int axg_bench_operator(axg_mesh *m, int32_t which, int32_t nrhs, int32_t reps,
                                  double *avg_ms) {
    try {
        clear_error();
        AXG_REQUIRE(m && avg_ms && reps > 0, AXG_E_INVALID, "axg_bench_operator: invalid argument");
        AXG_REQUIRE(m->assembled64, AXG_E_INVALID, "axg_bench_operator: mesh not assembled");
        axg_ctx *ctx = m->ctx;
        AXG_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        if (which == 2) {
            bench_operator_mixed(m, nrhs, reps, avg_ms);
            return AXG_OK;
        }
        if (which == 3) {
            AXG_REQUIRE(nrhs == 1 || nrhs == 2 || nrhs == 4 || nrhs == 8, AXG_E_INVALID,
                        "axg_bench_operator: nrhs must be 1, 2, 4 or 8");
            bench_block_spmm(m, nrhs, reps, avg_ms);
            return AXG_OK;
        }
        CsrView A = which == 0 ? solver_csr(m) : active_csr(m);
        DevBuf<double> x(A.n), y(A.n);
        AXG_HIP(hipMemsetAsync(x.p, 0, 8 * A.n, s));
        spmv_fp64(ctx, A, x.p, y.p);
        AXG_HIP(hipEventRecord(ctx->ev0, s));
        for (int r = 0; r < reps; r++) spmv_fp64(ctx, A, x.p, y.p);
        AXG_HIP(hipEventRecord(ctx->ev1, s));
        AXG_HIP(hipEventSynchronize(ctx->ev1));
        float ms = 0.f;
        AXG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *avg_ms = ms / reps;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    }
    return AXG_OK;
}
