// Symbolic phase (host): the K_ff CSR pattern and the deterministic scatter
// map, built once per mesh topology.
//
// The reference forms K_ff by pushing every nonzero Ke entry as a triplet in
// element order (analysis.rs:651-668), dropping prescribed rows/columns
// (DofMap::reduce, constraint.rs:138-173) and stable-sorting by (i, j) with
// duplicates summed left to right (csr_from_triplets, linalg.rs:10-37). So
// the value of slot (i, j) is the sum of the contributions of the elements
// that hold both DOFs, in ascending element order. This file lists, for each
// slot, exactly those contributions in that order; the device gather then
// reproduces the reference sum without sorting triplets or using atomics.
//
// Columns of a row are ascending free indices: DofMap numbers free DOFs in
// ascending DOF order (constraint.rs:99-108) and DOF = 3 * node + d
// (analysis.rs:152-154), so walking the sorted neighbour nodes and d = 0..2
// yields sorted columns.
#include <algorithm>
#include <atomic>
#include <thread>

#include "axg_internal.h"

namespace axg {

template <class F>
static void parallel_for(int64_t n, F f) {
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int64_t nt = std::min<int64_t>(hw > 16 ? 16 : hw, std::max<int64_t>(1, n / 4096));
    if (nt <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    int64_t chunk = (n + nt - 1) / nt;
    for (int64_t t = 0; t < nt; t++) {
        int64_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([=, &f] { f(a, b); });
    }
    for (auto &t : th) t.join();
}

void build_pattern(int nn, int64_t nnode, int64_t nelem, const int32_t *conn,
                   const int64_t *ind_of, int64_t n_free, Pattern &out) {
    const int m = 3 * nn;
    // An element listing one node twice is not a C3D8 / C3D20 the path supports.
    for (int64_t e = 0; e < nelem; e++)
        for (int a = 0; a < nn; a++) {
            int32_t v = conn[e * nn + a];
            if (v < 0 || v >= nnode) throw Error{AXG_E_INVALID, "element node index out of range"};
            for (int b = 0; b < a; b++)
                if (conn[e * nn + b] == v)
                    throw Error{AXG_E_INVALID, "element " + std::to_string(e) + " repeats a node"};
        }
    // node -> incident elements, ascending element index
    std::vector<int64_t> inc_ptr(nnode + 1, 0);
    for (int64_t e = 0; e < nelem; e++)
        for (int a = 0; a < nn; a++) inc_ptr[conn[e * nn + a] + 1]++;
    for (int64_t i = 0; i < nnode; i++) inc_ptr[i + 1] += inc_ptr[i];
    std::vector<int64_t> inc(inc_ptr[nnode]);
    {
        std::vector<int64_t> cur(inc_ptr.begin(), inc_ptr.end() - 1);
        for (int64_t e = 0; e < nelem; e++)
            for (int a = 0; a < nn; a++) inc[cur[conn[e * nn + a]]++] = e;
    }
    // free DOFs per node, and the local offset of DOF d among the node's free DOFs
    std::vector<uint8_t> nfree_node(nnode);
    for (int64_t i = 0; i < nnode; i++) {
        int c = 0;
        for (int d = 0; d < 3; d++) c += ind_of[3 * i + d] >= 0;
        nfree_node[i] = (uint8_t)c;
    }
    // neighbour nodes (sorted, unique) per node
    std::vector<int64_t> nb_ptr(nnode + 1, 0);
    std::vector<int32_t> nb;
    {
        std::vector<int32_t> cnt(nnode, 0);
        parallel_for(nnode, [&](int64_t a0, int64_t a1) {
            std::vector<int32_t> tmp;
            for (int64_t a = a0; a < a1; a++) {
                tmp.clear();
                for (int64_t k = inc_ptr[a]; k < inc_ptr[a + 1]; k++) {
                    int64_t e = inc[k];
                    for (int b = 0; b < nn; b++) tmp.push_back(conn[e * nn + b]);
                }
                std::sort(tmp.begin(), tmp.end());
                cnt[a] = (int32_t)(std::unique(tmp.begin(), tmp.end()) - tmp.begin());
            }
        });
        for (int64_t a = 0; a < nnode; a++) nb_ptr[a + 1] = nb_ptr[a] + cnt[a];
        nb.resize(nb_ptr[nnode]);
        parallel_for(nnode, [&](int64_t a0, int64_t a1) {
            std::vector<int32_t> tmp;
            for (int64_t a = a0; a < a1; a++) {
                tmp.clear();
                for (int64_t k = inc_ptr[a]; k < inc_ptr[a + 1]; k++) {
                    int64_t e = inc[k];
                    for (int b = 0; b < nn; b++) tmp.push_back(conn[e * nn + b]);
                }
                std::sort(tmp.begin(), tmp.end());
                tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
                std::copy(tmp.begin(), tmp.end(), nb.begin() + nb_ptr[a]);
            }
        });
    }
    // row lengths: a row of node a holds the free DOFs of all neighbours of a
    std::vector<int64_t> row_len_node(nnode, 0);
    parallel_for(nnode, [&](int64_t a0, int64_t a1) {
        for (int64_t a = a0; a < a1; a++) {
            int64_t s = 0;
            for (int64_t k = nb_ptr[a]; k < nb_ptr[a + 1]; k++) s += nfree_node[nb[k]];
            row_len_node[a] = s;
        }
    });
    // rows in free-index order == node order, then d
    out.rowptr.assign(n_free + 1, 0);
    std::vector<int64_t> row_of_node(nnode, -1);  // first free row of node
    {
        int64_t r = 0, acc = 0;
        for (int64_t a = 0; a < nnode; a++) {
            for (int d = 0; d < 3; d++) {
                int64_t f = ind_of[3 * a + d];
                if (f < 0) continue;
                if (f != r) throw Error{AXG_E_INVALID, "ind_of must number free DOFs 0..n-1 in DOF order"};
                if (row_of_node[a] < 0) row_of_node[a] = r;
                acc += row_len_node[a];
                if (acc > INT32_MAX) throw Error{AXG_E_INVALID, "K_ff has more than 2^31-1 nonzeros"};
                out.rowptr[r + 1] = (int32_t)acc;
                r++;
            }
        }
        if (r != n_free) throw Error{AXG_E_INVALID, "ind_of does not match n_free"};
    }
    const int64_t nnz = out.rowptr[n_free];
    out.colind.resize(nnz);  // written in full by the count pass
    // per-slot contribution counts, then the contributions themselves; each
    // row's slots belong to the thread of its node, which zeroes its counts and
    // later sets its fill cursors (no serial pass over the slots)
    HostVec<uint32_t> slot_cnt(nnz);
    auto for_rows = [&](bool fill) {
        parallel_for(nnode, [&](int64_t a0, int64_t a1) {
            // per-thread map neighbour node -> offset of its first free DOF in the row
            std::vector<std::pair<int32_t, int32_t>> nbpos;
            for (int64_t a = a0; a < a1; a++) {
                if (row_of_node[a] < 0) continue;
                nbpos.clear();
                int32_t off = 0;
                for (int64_t k = nb_ptr[a]; k < nb_ptr[a + 1]; k++) {
                    nbpos.emplace_back(nb[k], off);
                    off += nfree_node[nb[k]];
                }
                auto offset_of = [&](int32_t b) -> int32_t {
                    auto it = std::lower_bound(
                        nbpos.begin(), nbpos.end(), std::make_pair(b, INT32_MIN));
                    return it->second;
                };
                int64_t r0 = row_of_node[a];
                int pr = 0;
                for (int p = 0; p < 3; p++) {
                    if (ind_of[3 * a + p] < 0) continue;
                    int64_t r = r0 + pr++;
                    int64_t base = out.rowptr[r];
                    if (!fill)
                        std::fill(slot_cnt.begin() + base, slot_cnt.begin() + out.rowptr[r + 1], 0u);
                    else
                        for (int64_t q = base; q < out.rowptr[r + 1]; q++) slot_cnt[q] = out.slotptr[q];
                    if (!fill) {
                        // columns of this row
                        int64_t c = base;
                        for (int64_t k = nb_ptr[a]; k < nb_ptr[a + 1]; k++) {
                            int32_t b = nb[k];
                            for (int q = 0; q < 3; q++)
                                if (ind_of[3 * (int64_t)b + q] >= 0)
                                    out.colind[c++] = (int32_t)ind_of[3 * (int64_t)b + q];
                        }
                    }
                    for (int64_t k = inc_ptr[a]; k < inc_ptr[a + 1]; k++) {
                        int64_t e = inc[k];
                        int la = 0;
                        while (conn[e * nn + la] != a) la++;
                        for (int lb = 0; lb < nn; lb++) {
                            int32_t b = conn[e * nn + lb];
                            int32_t o = offset_of(b);
                            int qq = 0;
                            for (int q = 0; q < 3; q++) {
                                if (ind_of[3 * (int64_t)b + q] < 0) continue;
                                int64_t slot = base + o + qq++;
                                if (!fill) {
                                    slot_cnt[slot]++;
                                } else {
                                    uint64_t idx = (uint64_t)e * m * m +
                                                   (uint64_t)(3 * la + p) * m + (3 * lb + q);
                                    out.contrib[slot_cnt[slot]++] = (uint32_t)idx;
                                }
                            }
                        }
                    }
                }
            }
        });
    };
    for_rows(false);
    // slot pointers: exclusive prefix sum of the counts, one chunk per thread
    out.slotptr.resize(nnz + 1);
    {
        const int nt = host_threads(nnz, 1 << 16);
        std::vector<uint64_t> csum(nt + 1, 0);
        parallel_chunks(nnz, nt, [&](int t, int64_t lo, int64_t hi) {
            uint64_t acc = 0;
            for (int64_t s = lo; s < hi; s++) acc += slot_cnt[s];
            csum[t + 1] = acc;
        });
        for (int t = 0; t < nt; t++) csum[t + 1] += csum[t];
        if (csum[nt] > UINT32_MAX) throw Error{AXG_E_INVALID, "scatter map exceeds 2^32 entries"};
        parallel_chunks(nnz, nt, [&](int t, int64_t lo, int64_t hi) {
            uint64_t acc = csum[t];
            for (int64_t s = lo; s < hi; s++) {
                out.slotptr[s] = (uint32_t)acc;
                acc += slot_cnt[s];
            }
        });
        out.slotptr[nnz] = (uint32_t)csum[nt];
        out.contrib.resize(csum[nt]);  // written in full by the fill pass
    }
    for_rows(true);  // slot_cnt is each row's fill cursor, set from slotptr
}

}  // namespace axg
