// Internal declarations shared by the host and device translation units of
// libaxia_gpu. Nothing here crosses the C ABI (include/axia_gpu.h).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <exception>
#include <new>
#include <cstdint>
#include <cstdio>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/axia_gpu.h"

namespace axg {

// ------------------------------------------------------- host threading
// Threads for n items in chunks of at least min_chunk: at most 16 (and the
// hardware's count), at least 1.
inline int host_threads(int64_t n, int64_t min_chunk) {
    const unsigned hw = std::thread::hardware_concurrency();
    const int64_t cap = hw == 0 ? 1 : (hw < 16 ? (int64_t)hw : 16);
    const int64_t by_size = n / (min_chunk > 0 ? min_chunk : 1);
    return (int)(by_size < 1 ? 1 : (by_size < cap ? by_size : cap));
}

// fn(t, lo, hi) over nt contiguous chunks of [0, n), chunk 0 on the calling
// thread. fn must not throw (report failures through per-chunk flags).
template <class F>
inline void parallel_chunks(int64_t n, int nt, F &&fn) {
    if (nt <= 1) {
        fn(0, (int64_t)0, n);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    for (int t = 1; t < nt; t++) pool.emplace_back([&fn, n, nt, t] { fn(t, n * t / nt, n * (t + 1) / nt); });
    fn(0, (int64_t)0, n / nt);
    for (auto &th : pool) th.join();
}

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
void clear_error();

struct Error {
    int code;
    std::string msg;
};

#define AXG_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t _e = (call);                                                          \
        if (_e != hipSuccess) {                                                          \
            throw ::axg::Error{_e == hipErrorOutOfMemory ? AXG_E_OOM : AXG_E_DEVICE,     \
                               std::string("HIP: ") + hipGetErrorString(_e) + " at " +   \
                                   __FILE__ + ":" + std::to_string(__LINE__)};           \
        }                                                                                \
    } while (0)

#define AXG_REQUIRE(cond, code, msg)                \
    do {                                            \
        if (!(cond)) throw ::axg::Error{(code), (msg)}; \
    } while (0)

// ----------------------------------------------------------- device buffer
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr, o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p, n = o.n;
            o.p = nullptr, o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        AXG_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void ensure(size_t count) {
        if (count > n || p == nullptr) alloc(count);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr, n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Pinned host staging for small scalar read-backs.
struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void ensure(size_t b) {
        if (b <= bytes && p) return;
        if (p) (void)hipHostFree(p);
        AXG_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
        bytes = b;
    }
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
};

// ----------------------------------------------------------- PCG scalars
// One record per right-hand side, resident in device memory. Kernels read
// alpha/beta/done from here so that iterations need no host round trip.
struct PcgScalars {
    double bnorm;     // ||b||_2
    double tol_abs;   // tol * ||b||_2
    double rz;        // r . z
    double alpha, beta;
    double residual;  // ||r||_2 (recursive)
    double pap;
    double t_done_ms; // set by the host for batched solves
    int64_t iters;
    int64_t max_iter;
    int32_t done;     // 1 once converged / failed / max_iter reached
    int32_t status;   // AXG_OK, AXG_E_NOT_SPD, AXG_E_NOT_CONVERGED
    uint32_t ticket[4];
    double pad[1];
};
static_assert(sizeof(PcgScalars) % 16 == 0, "PcgScalars alignment");

// Reusable device workspace of a context (grown on demand).
struct Workspace {
    DevBuf<double> r, p, ap, z, part, t0, t1;
    DevBuf<PcgScalars> sc;
    DevBuf<float> f0, f1, f2, f3, f4, f5;
    DevBuf<int32_t> err;
    DevBuf<uint32_t> cnt;
};
// Allocator whose resize leaves new elements uninitialised (for large host
// arrays that are written in full by parallel passes: no serial zero fill).
template <class T>
struct DefaultInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = DefaultInitAlloc<U>;
    };
    DefaultInitAlloc() = default;
    template <class U>
    DefaultInitAlloc(const DefaultInitAlloc<U> &) {}
    template <class U>
    void construct(U *p) noexcept {
        ::new (static_cast<void *>(p)) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
};
template <class T>
using HostVec = std::vector<T, DefaultInitAlloc<T>>;

}  // namespace axg

namespace axg {
// thread-local text behind axg_last_error
void set_error(const std::string &msg);
void clear_error();
}  // namespace axg

// Body of every extern "C" entry point: errors become status codes + text.
#define API_BEGIN      \
    try {              \
        axg::clear_error();
#define API_END                                   \
    }                                             \
    catch (const axg::Error &e) {                 \
        axg::set_error(e.msg);                    \
        return e.code;                            \
    }                                             \
    catch (const std::bad_alloc &) {              \
        axg::set_error("host allocation failed"); \
        return AXG_E_OOM;                         \
    }                                             \
    catch (const std::exception &e) {             \
        axg::set_error(e.what());                 \
        return AXG_E_INVALID;                     \
    }                                             \
    return AXG_OK;

struct axg_ctx {
    // one reference held by the creator (axg_destroy drops it) and one per live
    // mesh: the context outlives every mesh made on it, whatever the call order
    std::atomic<int> refs{1};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evA = nullptr;
    std::vector<hipEvent_t> timing_events;
    axg::PinnedBuf pinned;
    axg::Workspace ws;
    int cu_count = 256;
    std::string arch;
};

namespace axg {
struct BlockPcg;
}

struct axg_mesh {
    axg_ctx *ctx = nullptr;
    // an operator built from a caller's CSR instead of elements (no elements;
    // three consecutive rows form a block row; each stored value is the one
    // contribution of its own slot, ke holds the values)
    bool csr_operator = false;
    int nn = 8, m = 24;
    int64_t nnode = 0, nelem = 0, ndof = 0, n_free = 0;
    int64_t nnz = 0, ncontrib = 0;          // of the current (possibly compacted) pattern
    int64_t nnz_topo = 0, ncontrib_topo = 0; // of the topological pattern
    double symbolic_s = 0.0;
    double t_ke64 = 0, t_gather64 = 0, t_ke32 = 0, t_gather32 = 0;  // ms, last assembly
    // mesh inputs (device)
    axg::DevBuf<double> xyz, e, nu;
    axg::DevBuf<int32_t> conn, ind_of;
    // topological pattern and scatter map (device)
    axg::DevBuf<int32_t> rowptr_t, colind_t;
    axg::DevBuf<uint32_t> slotptr_t, contrib_t;
    // host copy of rowptr_t / colind_t (written once, at creation), read by the
    // block layout build instead of downloading the pattern again
    axg::HostVec<int32_t> h_rowptr_t, h_colind_t;
    // active pattern (== topological unless exact-zero slots were dropped)
    bool compacted = false;
    axg::DevBuf<int32_t> rowptr_c, colind_c;
    axg::DevBuf<uint32_t> slotptr_c, contrib_c;
    // values
    axg::DevBuf<double> val;       // FP64 K_ff values (nnz)
    axg::DevBuf<double> ke;        // FP64 element matrices (nelem * m * m)
    axg::DevBuf<double> vol;       // element volumes (sum of Gauss weights x det J)
    axg::DevBuf<uint8_t> slot_nz;  // per topological slot: any nonzero contribution
    axg::DevBuf<double> dsel;      // Jacobi divisor (diag or 1.0)
    axg::DevBuf<int64_t> inc_ptr, inc_slot;  // node -> (element * nn + local), ascending element
    bool assembled64 = false;
    // mixed precision operator (BSR-3, FP32)
    bool assembled32 = false;
    int64_t nb = 0;                 // block rows (nodes with a free DOF)
    int64_t nnzb = 0;               // 3x3 blocks
    axg::DevBuf<int32_t> b_node;              // node of each block row
    axg::DevBuf<int32_t> free_of_blk;         // 3*nb: free index or -1
    axg::DevBuf<int32_t> blk_of_free;         // n_free: 3*blockrow + d
    axg::DevBuf<float> bval;                  // block planes, 10 floats per slot
    axg::DevBuf<float> grad32;                // FP32 scaled shape gradients per element
    axg::DevBuf<float> lm32;                  // FP32 (lam, mu) per element
    axg::DevBuf<float> kb32;                  // FP32 element records, blocks a <= b (3 planes)
    int asm32_mode = -1;                      // FP32 assembly: 0 fused rows, 1 MFMA blocks, -1 default
    axg::DevBuf<uint8_t> eslot;               // (e, a, b) -> slot of b's node in a's block row
    axg::DevBuf<int32_t> b_inc;               // b_incw per block row: e * nn + a ascending, -1
    int32_t b_incw = 0;                       // most elements at one block row's node
    axg::DevBuf<uint8_t> bs_mask;             // per slot: free DOFs of row (bits 0-2), column (4-6)
    axg::DevBuf<float> binv32;                // 9*nb block-Jacobi inverse (FP32 copy)
    axg::DevBuf<float> dinv32;                // 3*nb scalar Jacobi inverse
    axg::DevBuf<double> binv64;               // 9*nb FP64 block inverse over free DOFs
    axg::DevBuf<int32_t> entry_of_slot;       // topological slot -> BSR plane position
    bool binv_valid = false;                  // block inverses match the current values
    bool cols_sorted = true;                  // every CSR row's columns strictly increasing
    int spmv_shape = -1;                      // tuned FP64 SpMV shape
    // sliced-ELL copy of the active FP64 K_ff (solver operator)
    bool sell_valid = false;  // layout matches the active pattern
    bool sell_fresh = false;  // values match the last assembly
    int64_t sell_nslots = 0;
    axg::DevBuf<int64_t> sell_sptr;
    axg::DevBuf<int32_t> sell_width, sell_col, sell_src;
    axg::DevBuf<double> sell_val;
    // block-sliced-ELL layout of the FP32 operator (64 block rows per slice)
    int64_t bs_nslots = 0;
    axg::DevBuf<int64_t> bs_sptr;
    axg::DevBuf<int32_t> bs_width, bs_col;
    int32_t bs_maxw = 0;  // widest slice (most blocks in a block row)
    int spmm_shape[9] = {-1, -1, -1, -1, -1, -1, -1, -1, -1};  // tuned BSR SpMM shape per k
    // groups of 256 block rows for the LDS-staged FP32 SpMM: the sorted column
    // blocks of each group (at most a fixed count; the rest are read from
    // global memory), per slot the column's place in its group's list (or the
    // column with the top bit set), and the third value plane with that index
    // in place of the column
    axg::DevBuf<int32_t> grp_ptr, grp_col, bs_loc;
    axg::DevBuf<float> bval_loc;
    // FP64 values in the same block layout (five double2 planes per slot), for
    // the multi-right-hand-side FP64 block PCG; refreshed after every assembly
    bool bsell64_fresh = false;
    axg::DevBuf<double> bval64;
    // the FP64 block PCGs of this mesh, Jacobi and block-Jacobi (each created at
    // the first such solve, destroyed with the mesh)
    axg::BlockPcg *bpcg = nullptr;
    axg::BlockPcg *bpcg3 = nullptr;
};

namespace axg {
// symbolic phase (host)

struct Pattern {
    HostVec<int32_t> rowptr, colind;
    HostVec<uint32_t> slotptr, contrib;
};
void build_pattern(int nn, int64_t nnode, int64_t nelem, const int32_t *conn,
                   const int64_t *ind_of, int64_t n_free, Pattern &out);

// kernels (launch wrappers), axg_element.hip
void launch_ke_fp64(axg_mesh *m, hipStream_t s, int32_t *d_err);
void launch_gather_fp64(axg_mesh *m, hipStream_t s, bool topo, uint32_t *d_zero_count);
// FP32 operator of a mesh; true when it ran as two kernels (element blocks,
// then the block-row gather), with `mid` recorded between them
bool launch_row_asm32(axg_mesh *m, hipStream_t s, int32_t *d_err, hipEvent_t mid);
int asm32_mode_of(const axg_mesh *m);
void launch_csr_to_bsr32(axg_mesh *m, hipStream_t s);
void upload_shape_tables32(const double *h8_dn, const double *h20_dn, const double *h20_w);
void launch_nodal_stress(axg_mesh *m, hipStream_t s, const double *d_u, double *d_elem_nodal,
                         double *d_stress, double *d_vm, int32_t *d_err);
// mode: 0 reactions, 1 reduced right-hand side, 2 full diagonal
void launch_element_mean(axg_mesh *m, hipStream_t s, const double *d_elem_nodal, double *d_mean);
void launch_kg(axg_mesh *m, hipStream_t s, const double *d_mean, double *d_kg, int32_t *d_err);
void launch_gather_kg(axg_mesh *m, hipStream_t s, const double *d_kg, double sign, double *d_val);
void geometric_stiffness_device(axg_mesh *mesh, const double *d_u, double sign, double *d_val,
                                double *d_mean);
void launch_nl(axg_mesh *m, hipStream_t s, const double *d_u, double *d_fe, double *d_cau,
               double *d_gl, int32_t *d_err);
void launch_nl_fint(axg_mesh *m, hipStream_t s, const double *d_fe, double *d_fint);
void launch_nl_nodal(axg_mesh *m, hipStream_t s, const double *d_u, const double *d_cau,
                     const double *d_gl, double *en_c, double *en_g, int32_t *d_err);
void launch_nodal_average(axg_mesh *m, hipStream_t s, const double *en, double *d_stress,
                          double *d_vm);
void nl_assemble_device(axg_mesh *mesh, const double *d_u, double *d_fe, double *d_cau,
                        double *d_gl, double *d_fint);
void nl_nodal_device(axg_mesh *mesh, const double *d_u, const double *d_cau, const double *d_gl,
                     double *d_stress, double *d_vm, double *d_strain, double *d_gp);
void launch_lumped_mass(axg_mesh *m, hipStream_t s, const double *d_rho, double *d_m_full);
void launch_row_apply(axg_mesh *m, hipStream_t s, int mode, const double *d_w,
                      const double *d_f, double *d_out);
void launch_strain_energy(axg_mesh *m, hipStream_t s, const double *d_u, double *d_energy,
                          int32_t *d_err);
void upload_shape_tables();
}  // namespace axg

namespace axg {
// face pressure of one element, host (axg_loads.cpp)
int face_pressure_host(int nn, const double *xyz, int face, double p, double *fe);
}  // namespace axg
