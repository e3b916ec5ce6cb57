// Device CSR built from triplets (linalg::csr_from_triplets semantics).
#pragma once
#include "axg_internal.h"

namespace axg {
struct DeviceCsr64 {
    int64_t n = 0, nnz = 0;
    DevBuf<int64_t> indptr, indices;
    DevBuf<double> data;
    DevBuf<int32_t> rowptr32, colind32;
};
void csr_from_triplets_device(axg_ctx *ctx, int64_t n, int64_t ntrip, const int64_t *d_ti,
                              const int64_t *d_tj, const double *d_tv, DeviceCsr64 &out);
}  // namespace axg
