"""axia_gpu — host side of the MI355X linear-static hot path of axia-fem.

Thin Python mirror of the reference's solve boundary over the C ABI
(include/axia_gpu.h); all arithmetic runs in libaxia_gpu.so (HIP, gfx950).

  solve_kff(n, (ti, tj, tv), rhs)   linalg::solve_kff      (linalg.rs:311-404)
  SparseResult                       linalg::SparseResult   (linalg.rs:261-266)
  csr_from_triplets(...)             linalg::csr_from_triplets (linalg.rs:10-37)
  Mesh(...).assemble()/.solve()      element_ke + triplet loop + DofMap::reduce
                                     + solve_kff of solve_linear
                                     (analysis.rs:574-668, 1018-1024)
  Mesh.nodal_stress(u)               stress recovery (analysis.rs:1061-1162)
  Mesh.compliance_adjoint(u, f)      dc/dE per element (new path)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (AxgError, PREC_FP64, PREC_MIXED, PRECOND_BLOCK3, PRECOND_JACOBI, REFINE_FCG,
                   REFINE_PLAIN, REFINE_RELIABLE, check, dptr, f64, i32, i64, iptr, lptr)

__all__ = [
    "AxgError", "Context", "Mesh", "SparseResult", "solve_kff", "solve_csr", "csr_from_triplets",
    "pattern_host", "pressure_loads", "default_context", "PREC_FP64", "PREC_MIXED",
    "PRECOND_JACOBI", "PRECOND_BLOCK3", "REFINE_FCG", "REFINE_PLAIN", "REFINE_RELIABLE",
]


@dataclass
class SparseResult:
    """linalg.rs:261-266: x, solver name, iterations, residual ||K x - b||."""
    x: np.ndarray
    name: str
    iters: int
    residual: float


def _opts(precision=PREC_FP64, precond=PRECOND_JACOBI, tol=1e-8, max_iter=0, inner_tol=1e-3,
          max_outer=30, timing=False, refinement="fcg"):
    o = _lib.SolveOpts()
    _lib.lib().axg_default_opts(C.byref(o))
    o.precision = {"fp64": PREC_FP64, "mixed": PREC_MIXED}.get(precision, precision)
    o.precond = {"jacobi": PRECOND_JACOBI, "block3": PRECOND_BLOCK3}.get(precond, precond)
    o.tol = tol
    o.max_iter = max_iter
    o.inner_tol = inner_tol
    o.max_outer = max_outer
    o.timing = 1 if timing else 0
    o.refinement = {"fcg": REFINE_FCG, "plain": REFINE_PLAIN,
                    "reliable": REFINE_RELIABLE}.get(refinement, refinement)
    return o


def _need(a, size, what):
    """Host-side size check before a buffer crosses the C ABI (the library reads and
    writes exactly the sizes the mesh implies)."""
    if a.size != size:
        raise ValueError(f"{what}: expected {size} values, got {a.size}")
    return a


def _info_dict(info):
    return {
        "iters": int(info.iters), "outer_iters": int(info.outer_iters), "nrhs": int(info.nrhs),
        "residual": float(info.residual), "rel_residual": float(info.rel_residual),
        "solve_ms": float(info.solve_ms), "spmv_ms": float(info.spmv_ms),
        "spmv_launches": int(info.spmv_launches), "name": info.name.decode(),
    }


class Context:
    """One device, one HIP stream (axg_create)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(_lib.lib().axg_create(device, C.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    @property
    def stream(self):
        return _lib.lib().axg_stream(self._h)

    def close(self):
        if self._h:
            _lib.lib().axg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Mesh:
    """All-C3D8 or all-C3D20 isotropic mesh on the device (axg_mesh_create).

    xyz (nnode, 3); conn (nelem, nn) 0-based node indices; e, nu per element;
    ind_of (3 nnode) free index per DOF or -1 (DofMap::build numbering).
    """

    def __init__(self, nn, xyz, conn, e, nu, ind_of, ctx=None):
        self.ctx = ctx or default_context()
        self.nn = int(nn)
        self.xyz = f64(xyz).reshape(-1, 3)
        self.conn = i32(conn).reshape(-1, self.nn)
        self.nnode = self.xyz.shape[0]
        self.nelem = self.conn.shape[0]
        self.e = f64(np.broadcast_to(e, (self.nelem,)))
        self.nu = f64(np.broadcast_to(nu, (self.nelem,)))
        self.ind_of = i64(ind_of)
        _need(self.ind_of, 3 * self.nnode, "ind_of")
        h = C.c_void_p()
        check(_lib.lib().axg_mesh_create(self.ctx.handle, self.nn, self.nnode, dptr(self.xyz),
                                         self.nelem, iptr(self.conn), dptr(self.e), dptr(self.nu),
                                         lptr(self.ind_of), C.byref(h)))
        self._h = h
        self._n_free = self.info()["n_free"]

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().axg_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        nf, nnz, nc = C.c_int64(), C.c_int64(), C.c_int64()
        sym = C.c_double()
        check(_lib.lib().axg_mesh_info(self._h, C.byref(nf), C.byref(nnz), C.byref(nc),
                                       C.byref(sym)))
        return {"n_free": nf.value, "nnz": nnz.value, "ncontrib": nc.value,
                "symbolic_s": sym.value}

    @property
    def n_free(self):
        return self._n_free

    def _u(self, u_full, what="u_full"):
        return _need(f64(u_full).ravel(), 3 * self.nnode, what)

    def set_modulus(self, e):
        self.e = f64(np.broadcast_to(e, (self.nelem,)))
        check(_lib.lib().axg_mesh_set_modulus(self._h, dptr(self.e)))

    def assemble(self, precision="fp64"):
        """Numeric assembly; returns device milliseconds."""
        ms = C.c_double()
        prec = {"fp64": PREC_FP64, "mixed": PREC_MIXED}.get(precision, precision)
        check(_lib.lib().axg_assemble(self._h, prec, C.byref(ms)))
        return ms.value

    def assembly_times(self):
        """Device ms of the last assembly: ke64, gather64, ke32 (the fused FP32 integration
        and block assembly), gather32 (0: no separate FP32 gather)."""
        v = [C.c_double() for _ in range(4)]
        check(_lib.lib().axg_assembly_times(self._h, *[C.byref(x) for x in v]))
        return {k: x.value for k, x in zip(("ke64", "gather64", "ke32", "gather32"), v)}

    def csr(self):
        inf = self.info()
        n, nnz = inf["n_free"], inf["nnz"]
        indptr = np.zeros(n + 1, np.int64)
        indices = np.zeros(max(nnz, 1), np.int64)
        data = np.zeros(max(nnz, 1), np.float64)
        check(_lib.lib().axg_csr_export(self._h, lptr(indptr), lptr(indices), dptr(data)))
        return indptr, indices[:nnz], data[:nnz]

    def scatter_map(self):
        inf = self.info()
        sp = np.zeros(inf["nnz"] + 1, np.int64)
        cb = np.zeros(max(inf["ncontrib"], 1), np.int64)
        check(_lib.lib().axg_scatter_map_export(self._h, lptr(sp), lptr(cb)))
        return sp, cb[:inf["ncontrib"]]

    def ke(self, first=0, count=None):
        count = self.nelem - first if count is None else count
        m = 3 * self.nn
        out = np.zeros(count * m * m)
        check(_lib.lib().axg_ke_export(self._h, first, count, dptr(out)))
        return out.reshape(count, m, m)

    def op32(self):
        """The FP32 inner operator of the mixed solve as a scipy CSR matrix over the
        free DOFs (after assemble("mixed"))."""
        import scipy.sparse as sp
        lib = _lib.lib()
        cnt = C.c_int64(0)
        check(lib.axg_op32_export(self._h, C.byref(cnt), None, None, None))
        n = cnt.value
        rows, cols = np.zeros(n, np.int32), np.zeros(n, np.int32)
        vals = np.zeros(n, np.float32)
        ip = C.POINTER(C.c_int32)
        check(lib.axg_op32_export(self._h, C.byref(cnt), rows.ctypes.data_as(ip),
                                  cols.ctypes.data_as(ip),
                                  vals.ctypes.data_as(C.POINTER(C.c_float))))
        return sp.csr_matrix((vals, (rows, cols)), shape=(self.n_free, self.n_free))

    def solve(self, rhs, precision="fp64", precond="jacobi", tol=1e-8, max_iter=0,
              inner_tol=1e-3, max_outer=30, timing=False, refinement="fcg"):
        """Solve K_ff x = rhs; rhs (n_free,) or (nrhs, n_free). Returns (x, info)."""
        b = f64(rhs)
        single = b.ndim == 1
        b2 = b.reshape(1, -1) if single else b
        if b2.ndim != 2 or b2.shape[1] != self.n_free or b2.shape[0] < 1:
            raise ValueError(f"rhs: expected (n_free,) or (nrhs, n_free) with n_free = "
                             f"{self.n_free}, got shape {b.shape}")
        nrhs = b2.shape[0]
        x = np.zeros_like(b2)
        info = _lib.SolveInfo()
        o = _opts(precision, precond, tol, max_iter, inner_tol, max_outer, timing, refinement)
        check(_lib.lib().axg_mesh_solve(self._h, dptr(b2), nrhs, C.byref(o), dptr(x),
                                        C.byref(info)))
        return (x[0] if single else x), _info_dict(info)

    def solve_device(self, rhs_ptr, nrhs, x_ptr, precision="fp64", precond="jacobi", tol=1e-8,
                     max_iter=0, inner_tol=1e-3, max_outer=30, timing=False, refinement="fcg"):
        """Device-pointer solve (inputs already in HBM); returns info."""
        info = _lib.SolveInfo()
        o = _opts(precision, precond, tol, max_iter, inner_tol, max_outer, timing, refinement)
        check(_lib.lib().axg_mesh_solve_device(self._h, C.c_void_p(rhs_ptr), nrhs, C.byref(o),
                                               C.c_void_p(x_ptr), C.byref(info)))
        return _info_dict(info)

    def bench_operator(self, which, nrhs=1, reps=20):
        """Average ms of one operator launch: 0 FP64 solver SpMV, 1 FP64 CSR, 2 FP32 SpMM."""
        ms = C.c_double()
        check(_lib.lib().axg_bench_operator(self._h, which, nrhs, reps, C.byref(ms)))
        return ms.value

    def spmm_shape(self, nrhs, shape=-1):
        """FP32 SpMM variant for nrhs right-hand sides: pin it with shape >= 0; returns the
        active one (-1 before the first mixed solve tunes it)."""
        v = C.c_int32(shape)
        check(_lib.lib().axg_mesh_spmm_shape(self._h, nrhs, C.byref(v)))
        return v.value

    def asm32_mode(self, mode=-1):
        """FP32 operator assembly form: 0 fused integration + block rows, 1 element blocks on
        the matrix cores + block-row gather; pin it with mode >= 0; returns the active one."""
        v = C.c_int32(mode)
        check(_lib.lib().axg_mesh_asm32_mode(self._h, C.byref(v)))
        return v.value

    def nodal_stress(self, u_full):
        u = self._u(u_full)
        stress = np.zeros((self.nnode, 6))
        vm = np.zeros(self.nnode)
        en = np.zeros((self.nelem, self.nn, 6))
        check(_lib.lib().axg_nodal_stress(self._h, dptr(u), dptr(stress), dptr(vm), dptr(en)))
        return stress, vm, en

    def reactions(self, u_full, f_full):
        """rf = K_full u_full - f_full over all DOFs (analysis.rs:1027-1035)."""
        rf = np.zeros(3 * self.nnode)
        u, f = self._u(u_full), self._u(f_full, "f_full")
        check(_lib.lib().axg_reactions(self._h, dptr(u), dptr(f), dptr(rf)))
        return rf

    def reduce_rhs(self, f_full, u0):
        """DofMap::reduce of the load vector with prescribed values u0 (constraint.rs:138-173)."""
        rhs = np.zeros(max(self.n_free, 1))
        f, u = self._u(f_full, "f_full"), self._u(u0, "u0")
        check(_lib.lib().axg_reduce_rhs(self._h, dptr(f), dptr(u), dptr(rhs)))
        return rhs[:self.n_free]

    def full_diagonal(self):
        """Diagonal of the full stiffness over all DOFs, summed in element order."""
        d = np.zeros(3 * self.nnode)
        check(_lib.lib().axg_full_diagonal(self._h, dptr(d)))
        return d

    def element_volumes(self):
        """Element volumes of the last FP64 assembly (Gauss weight x det J, point order)."""
        v = np.zeros(self.nelem)
        check(_lib.lib().axg_element_volumes(self._h, dptr(v)))
        return v

    def lumped_mass(self, density):
        """Lumped mass over all DOFs: density_e |vol_e| / nn per node DOF, element order."""
        rho = f64(np.broadcast_to(np.asarray(density, np.float64), (self.nelem,))).copy()
        m = np.zeros(3 * self.nnode)
        check(_lib.lib().axg_lumped_mass(self._h, dptr(rho), dptr(m)))
        return m

    def eigen(self, m_diag, nmodes, precision="fp64", precond="jacobi", tol=1e-10, max_iter=0,
              inner_tol=1e-3, vectors=True):
        """Lowest nmodes of K_ff phi = lambda diag(m_diag) phi by subspace iteration
        (eigen.rs:12-111). Returns (values, vectors (k, n_free) or None, sweeps, info)."""
        m = f64(m_diag).ravel()
        if m.size != self.n_free:
            raise ValueError("m_diag must have n_free entries")
        vals = np.zeros(nmodes)
        vecs = np.zeros((nmodes, self.n_free)) if vectors else None
        nf, sw = C.c_int32(), C.c_int32()
        info = _lib.SolveInfo()
        o = _opts(precision, precond, tol, max_iter, inner_tol)
        check(_lib.lib().axg_eigen(self._h, dptr(m), nmodes, C.byref(o), dptr(vals),
                                   dptr(vecs) if vectors else None, C.byref(nf), C.byref(sw),
                                   C.byref(info)))
        k = nf.value
        return vals[:k], (vecs[:k] if vectors else None), sw.value, _info_dict(info)

    def geometric_stiffness(self, u_full):
        """(element mean stresses (nelem, 6), Kg_ff values on the topological pattern)."""
        u = self._u(u_full)
        mean = np.zeros((self.nelem, 6))
        nnz_t = self._nnz_topo()
        kg = np.zeros(max(nnz_t, 1))
        check(_lib.lib().axg_geometric_stiffness(self._h, dptr(u), dptr(mean), dptr(kg)))
        return mean, kg[:nnz_t]

    def _nnz_topo(self):
        if getattr(self, "_pattern", None) is None:
            self._pattern = pattern_host(self.nn, self.nnode, self.conn, self.ind_of)
        return len(self._pattern[1])

    def topological_pattern(self):
        """(rowptr, colind) of K_ff before exact-zero slots are dropped."""
        self._nnz_topo()
        return self._pattern[0], self._pattern[1]

    def buckle(self, u_full, nmodes, precision="fp64", precond="jacobi", tol=1e-10, max_iter=0,
               inner_tol=1e-3, vectors=True):
        """Buckling factors of K_ff phi = lambda (-Kg_ff) phi for the displacements u_full
        (eigen.rs:113-179). Returns (values, vectors (k, n_free) or None, info)."""
        u = self._u(u_full)
        vals = np.zeros(nmodes)
        vecs = np.zeros((nmodes, self.n_free)) if vectors else None
        nf = C.c_int32()
        info = _lib.SolveInfo()
        o = _opts(precision, precond, tol, max_iter, inner_tol)
        check(_lib.lib().axg_buckle(self._h, dptr(u), nmodes, C.byref(o), dptr(vals),
                                    dptr(vecs) if vectors else None, C.byref(nf),
                                    C.byref(info)))
        k = nf.value
        return vals[:k], (vecs[:k] if vectors else None), _info_dict(info)

    def nl_assemble(self, u_full):
        """NLGEOM tangent (into this mesh's K_ff) and internal force at u_full: returns
        (f_int (3 nnode), element mean Cauchy (nelem, 6), element mean Green-Lagrange)."""
        u = self._u(u_full)
        fint = np.zeros(3 * self.nnode)
        cau = np.zeros((self.nelem, 6))
        gl = np.zeros((self.nelem, 6))
        check(_lib.lib().axg_nl_assemble(self._h, dptr(u), dptr(fint), dptr(cau), dptr(gl)))
        return fint, cau, gl

    def compliance_adjoint(self, u_full, f_full):
        """c = f^T u and dc/dE_e = -u_e^T K1_e u_e (K1: Ke at unit modulus)."""
        u = self._u(u_full)
        f = self._u(f_full, "f_full")
        c = C.c_double()
        dc = np.zeros(self.nelem)
        check(_lib.lib().axg_compliance_adjoint(self._h, dptr(u), dptr(f), C.byref(c), dptr(dc)))
        return c.value, dc


def solve_batch(meshes, rhs_list, tol=1e-8, max_iter=0, ctx=None, groups=0):
    """Solve all meshes' load cases in one batched call.

    meshes: assembled Mesh objects on one context; rhs_list[m]: (k_m, n_free_m)
    array (k_m <= 8). groups: 1 advances every mesh in one launch sequence, G > 1
    splits the meshes by size into G groups solved one after another, smallest
    first (at most 16), 0 lets the library choose (one group per 6 meshes). Returns (xs, iters,
    latency_ms, info): xs[m] has the shape of rhs_list[m]; iters and latency_ms
    list every (mesh, load case) solve in order, latency measured from the start
    of the batch to that solve's convergence on the device clock.
    """
    ctx = ctx or meshes[0].ctx
    nm = len(meshes)
    if len(rhs_list) != nm:
        raise ValueError("one right-hand-side block per mesh")
    bs = [f64(np.atleast_2d(b)) for b in rhs_list]
    for m, b in zip(meshes, bs):
        if b.ndim != 2 or b.shape[1] != m.n_free or not 1 <= b.shape[0] <= 8:
            raise ValueError(f"rhs block of shape {b.shape} for a mesh with n_free = {m.n_free}"
                             " (1 to 8 load cases)")
    xs = [np.zeros_like(b) for b in bs]
    nrhs = np.array([b.shape[0] for b in bs], np.int32)
    H = (C.c_void_p * nm)(*[m._h for m in meshes])
    R = (_lib._dp * nm)(*[dptr(b) for b in bs])
    X = (_lib._dp * nm)(*[dptr(x) for x in xs])
    total = int(nrhs.sum())
    iters = np.zeros(total, np.int64)
    lat = np.zeros(total)
    info = _lib.SolveInfo()
    o = _opts(PREC_FP64, PRECOND_JACOBI, tol, max_iter)
    o.batch_groups = int(groups)
    check(_lib.lib().axg_solve_batch(ctx.handle, nm, H, R, iptr(nrhs), C.byref(o), X, lptr(iters),
                                     dptr(lat), C.byref(info)))
    return xs, iters, lat, _info_dict(info)


def bench_stream(n=1 << 27, reps=10, ctx=None):
    """STREAM-style copy bandwidth (GB/s, read + write) of n doubles."""
    ctx = ctx or default_context()
    gbs = C.c_double()
    check(_lib.lib().axg_bench_stream(ctx.handle, n, reps, C.byref(gbs)))
    return gbs.value


def csr_from_triplets(n, ti, tj, tv, ctx=None):
    """linalg::csr_from_triplets on the device: (indptr, indices, data)."""
    ctx = ctx or default_context()
    ti, tj, tv = i64(ti), i64(tj), f64(tv)
    indptr = np.zeros(n + 1, np.int64)
    L = _lib.lib()
    check(L.axg_csr_from_triplets(ctx.handle, n, len(ti), lptr(ti), lptr(tj), dptr(tv),
                                  lptr(indptr), None, None))
    nnz = int(indptr[n])
    indices = np.zeros(max(nnz, 1), np.int64)
    data = np.zeros(max(nnz, 1), np.float64)
    check(L.axg_csr_from_triplets(ctx.handle, n, len(ti), lptr(ti), lptr(tj), dptr(tv),
                                  lptr(indptr), lptr(indices), dptr(data)))
    return indptr, indices[:nnz], data[:nnz]


def solve_kff(n, trips, rhs, ctx=None, tol=1e-8, max_iter=0, timing=False, precision="fp64",
              precond="jacobi", inner_tol=1e-3, refinement="fcg"):
    """linalg::solve_kff drop-in: trips = (ti, tj, tv) arrays; returns SparseResult.

    precision "fp64" with precond "jacobi" is the reference's linalg::pcg on the scalar
    rows; "block3" (3x3 block-Jacobi over rows 3i..3i+2) and "mixed" (FP32 inner PCG under
    FP64 refinement) solve the operator in node-block form."""
    ctx = ctx or default_context()
    ti, tj, tv = (i64(trips[0]), i64(trips[1]), f64(trips[2]))
    if not len(ti) == len(tj) == len(tv):
        raise ValueError("triplet arrays differ in length")
    b = _need(f64(rhs), n, "rhs")
    x = np.zeros(n)
    info = _lib.SolveInfo()
    o = _opts(precision, precond, tol, max_iter, inner_tol, timing=timing, refinement=refinement)
    check(_lib.lib().axg_solve_kff(ctx.handle, n, len(ti), lptr(ti), lptr(tj), dptr(tv), dptr(b),
                                   C.byref(o), dptr(x), C.byref(info)))
    d = _info_dict(info)
    return SparseResult(x, d["name"], d["iters"], d["residual"])


def solve_csr(indptr, indices, data, rhs, ctx=None, tol=1e-8, max_iter=0, timing=False,
              precision="fp64", precond="jacobi", inner_tol=1e-3, refinement="fcg"):
    """Solve K x = rhs for a host CSR matrix (axg_solve_csr). Returns (x, info)."""
    ctx = ctx or default_context()
    indptr, indices, data = i64(indptr), i64(indices), f64(data)
    n = len(indptr) - 1
    if n < 0 or len(indices) != len(data) or len(indices) < int(indptr[-1]):
        raise ValueError("inconsistent CSR arrays")
    b = _need(f64(rhs), n, "rhs")
    x = np.zeros(n)
    info = _lib.SolveInfo()
    o = _opts(precision, precond, tol, max_iter, inner_tol, timing=timing, refinement=refinement)
    check(_lib.lib().axg_solve_csr(ctx.handle, n, lptr(indptr), lptr(indices), dptr(data),
                                   dptr(b), C.byref(o), dptr(x), C.byref(info)))
    return x, _info_dict(info)


def pattern_host(nn, nnode, conn, ind_of):
    """Symbolic phase on the host: (rowptr, colind, slot_ptr, contrib)."""
    conn = i32(conn)
    ind_of = i64(ind_of)
    nelem = conn.size // nn
    nf, nnz, nc = C.c_int64(), C.c_int64(), C.c_int64()
    L = _lib.lib()
    check(L.axg_pattern_host(nn, nnode, nelem, iptr(conn), lptr(ind_of), C.byref(nf), C.byref(nnz),
                             C.byref(nc), None, None, None, None))
    rp = np.zeros(nf.value + 1, np.int64)
    ci = np.zeros(max(nnz.value, 1), np.int64)
    sp = np.zeros(nnz.value + 1, np.int64)
    cb = np.zeros(max(nc.value, 1), np.int64)
    check(L.axg_pattern_host(nn, nnode, nelem, iptr(conn), lptr(ind_of), C.byref(nf), C.byref(nnz),
                             C.byref(nc), lptr(rp), lptr(ci), lptr(sp), lptr(cb)))
    return rp, ci[:nnz.value], sp, cb[:nc.value]


def pressure_loads(nn, xyz, elem_conn, faces, mags, f_full):
    """Accumulate face-pressure loads into f_full (in place) in the given order."""
    xyz = f64(xyz)
    ec = i32(elem_conn)
    fa = i32(faces)
    mg = f64(mags)
    assert f_full.dtype == np.float64 and f_full.flags.c_contiguous
    check(_lib.lib().axg_pressure_loads(nn, xyz.size // 3, dptr(xyz), len(fa), iptr(ec), iptr(fa),
                                        dptr(mg), dptr(f_full)))
    return f_full
