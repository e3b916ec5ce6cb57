// linalg::csr_from_triplets (linalg.rs:10-37) on the device:
//   keep (i < n, j < n, |v| > 0) in input order, stable sort by (i, j),
//   sum equal keys left to right starting from 0.0, count entries per row.
// The stable order comes from an LSD radix sort (hipcub), which keeps equal
// keys in input order; the sums walk each run sequentially, so values carry
// the reference's FP64 bits (this file is built with -ffp-contract=off).
#include <hipcub/hipcub.hpp>

#include "axg_internal.h"
#include "axg_triplets.h"

namespace axg {

__global__ void trip_keys_kernel(int64_t ntrip, int64_t n, const int64_t *__restrict__ ti,
                                 const int64_t *__restrict__ tj, const double *__restrict__ tv,
                                 uint64_t *__restrict__ keys, int64_t *__restrict__ idx,
                                 uint8_t *__restrict__ keep) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntrip) return;
    int64_t i = ti[t], j = tj[t];
    bool k = i >= 0 && i < n && j >= 0 && j < n && fabs(tv[t]) > 0.0;
    keep[t] = k ? 1 : 0;
    keys[t] = k ? (uint64_t)i * (uint64_t)n + (uint64_t)j : 0;
    idx[t] = t;
}

__global__ void run_heads_kernel(int64_t m, const uint64_t *__restrict__ keys,
                                 uint32_t *__restrict__ head) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    head[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1u : 0u;
}

// One thread per run: sum its values in sorted (= input) order.
__global__ void run_sum_kernel(int64_t m, int64_t n, const uint64_t *__restrict__ keys,
                               const int64_t *__restrict__ order, const uint32_t *__restrict__ head,
                               const uint32_t *__restrict__ slot, const double *__restrict__ tv,
                               int64_t *__restrict__ indices, double *__restrict__ data,
                               uint32_t *__restrict__ row_count) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m || !head[k]) return;
    uint64_t key = keys[k];
    double v = 0.0;
    int64_t q = k;
    while (q < m && keys[q] == key) {
        v += tv[order[q]];
        q++;
    }
    uint32_t s = slot[k];
    int64_t i = (int64_t)(key / (uint64_t)n), j = (int64_t)(key % (uint64_t)n);
    indices[s] = j;
    data[s] = v;
    atomicAdd(&row_count[i], 1u);
}

__global__ void widen_kernel(int64_t m, const uint32_t *__restrict__ a, int64_t *__restrict__ b) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) b[k] = a[k];
}

__global__ void narrow_kernel(int64_t m, const int64_t *__restrict__ a, int32_t *__restrict__ b) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) b[k] = (int32_t)a[k];
}

void csr_from_triplets_device(axg_ctx *ctx, int64_t n, int64_t ntrip, const int64_t *d_ti,
                              const int64_t *d_tj, const double *d_tv, DeviceCsr64 &out) {
    hipStream_t s = ctx->stream;
    AXG_REQUIRE(ntrip < INT32_MAX, AXG_E_INVALID, "more than 2^31-1 triplets");
    out.n = n;
    out.indptr.ensure(n + 1);
    if (ntrip == 0 || n == 0) {
        AXG_HIP(hipMemsetAsync(out.indptr.p, 0, sizeof(int64_t) * (n + 1), s));
        out.nnz = 0;
        out.indices.ensure(1);
        out.data.ensure(1);
        AXG_HIP(hipStreamSynchronize(s));
        return;
    }
    DevBuf<uint64_t> keys(ntrip), keys2(ntrip);
    DevBuf<int64_t> idx(ntrip), idx2(ntrip);
    DevBuf<uint8_t> keep(ntrip);
    DevBuf<int64_t> nsel(1);
    unsigned g = (unsigned)((ntrip + 255) / 256);
    hipLaunchKernelGGL(trip_keys_kernel, dim3(g), dim3(256), 0, s, ntrip, n, d_ti, d_tj, d_tv,
                       keys.p, idx.p, keep.p);
    AXG_HIP(hipGetLastError());
    // compact kept triplets, preserving input order
    size_t tmp_bytes = 0, tb2 = 0;
    AXG_HIP(hipcub::DeviceSelect::Flagged(nullptr, tmp_bytes, keys.p, keep.p, keys2.p, nsel.p,
                                          (int)ntrip, s));
    AXG_HIP(hipcub::DeviceSelect::Flagged(nullptr, tb2, idx.p, keep.p, idx2.p, nsel.p, (int)ntrip,
                                          s));
    tmp_bytes = std::max(tmp_bytes, tb2);
    DevBuf<char> tmp(tmp_bytes + 16);
    AXG_HIP(hipcub::DeviceSelect::Flagged(tmp.p, tmp_bytes, keys.p, keep.p, keys2.p, nsel.p,
                                          (int)ntrip, s));
    AXG_HIP(hipcub::DeviceSelect::Flagged(tmp.p, tmp_bytes, idx.p, keep.p, idx2.p, nsel.p,
                                          (int)ntrip, s));
    int64_t m = 0;
    AXG_HIP(hipMemcpyAsync(&m, nsel.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    // stable radix sort by key
    int end_bit = 1;
    {
        unsigned __int128 maxkey = (unsigned __int128)n * (unsigned __int128)n;
        while (end_bit < 64 && ((unsigned __int128)1 << end_bit) < maxkey) end_bit++;
    }
    size_t sort_bytes = 0;
    AXG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys2.p, keys.p, idx2.p, idx.p,
                                               (int)m, 0, end_bit, s));
    DevBuf<char> stmp(sort_bytes + 16);
    AXG_HIP(hipcub::DeviceRadixSort::SortPairs(stmp.p, sort_bytes, keys2.p, keys.p, idx2.p, idx.p,
                                               (int)m, 0, end_bit, s));
    // runs of equal keys
    DevBuf<uint32_t> head(m + 1), slot(m + 1);
    if (m > 0) {
        hipLaunchKernelGGL(run_heads_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m,
                           keys.p, head.p);
        AXG_HIP(hipGetLastError());
    }
    size_t scan_bytes = 0;
    AXG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, head.p, slot.p, (int)m, s));
    DevBuf<char> sctmp(scan_bytes + 16);
    AXG_HIP(hipcub::DeviceScan::ExclusiveSum(sctmp.p, scan_bytes, head.p, slot.p, (int)m, s));
    uint32_t last_slot = 0, last_head = 0;
    if (m > 0) {
        AXG_HIP(hipMemcpyAsync(&last_slot, slot.p + (m - 1), 4, hipMemcpyDeviceToHost, s));
        AXG_HIP(hipMemcpyAsync(&last_head, head.p + (m - 1), 4, hipMemcpyDeviceToHost, s));
    }
    AXG_HIP(hipStreamSynchronize(s));
    int64_t nnz = m > 0 ? (int64_t)last_slot + last_head : 0;
    out.nnz = nnz;
    out.indices.ensure(nnz);
    out.data.ensure(nnz);
    DevBuf<uint32_t> rowc(n + 1);
    AXG_HIP(hipMemsetAsync(rowc.p, 0, sizeof(uint32_t) * (n + 1), s));
    if (m > 0) {
        hipLaunchKernelGGL(run_sum_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m,
                           n, keys.p, idx.p, head.p, slot.p, d_tv, out.indices.p, out.data.p,
                           rowc.p);
        AXG_HIP(hipGetLastError());
    }
    // indptr = exclusive scan of row counts (n + 1 entries, last count is 0)
    DevBuf<int64_t> rowc64(n + 1);
    hipLaunchKernelGGL(widen_kernel, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, s,
                       n + 1, rowc.p, rowc64.p);
    AXG_HIP(hipGetLastError());
    size_t b = 0;
    AXG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, rowc64.p, out.indptr.p, (int)(n + 1), s));
    DevBuf<char> t2(b + 16);
    AXG_HIP(hipcub::DeviceScan::ExclusiveSum(t2.p, b, rowc64.p, out.indptr.p, (int)(n + 1), s));
    // 32-bit copies for the SpMV kernels
    AXG_REQUIRE(nnz < INT32_MAX, AXG_E_INVALID, "K_ff has more than 2^31-1 nonzeros");
    out.rowptr32.ensure(n + 1);
    out.colind32.ensure(nnz);
    hipLaunchKernelGGL(narrow_kernel, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, s,
                       n + 1, out.indptr.p, out.rowptr32.p);
    if (nnz > 0)
        hipLaunchKernelGGL(narrow_kernel, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, s,
                           nnz, out.indices.p, out.colind32.p);
    AXG_HIP(hipGetLastError());
    AXG_HIP(hipStreamSynchronize(s));
}

}  // namespace axg
