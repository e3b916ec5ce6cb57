"""GPU parity of the hot path against the oracle and the committed goldens.

Bars: CSR pattern, scatter map and FP64 values bit-exact; Ke bit-exact;
displacements within 1e-8 relative of the direct FP64 solve; stresses from the
same displacements bit-exact (von Mises within 1e-14), and within 1e-8 from
the GPU displacements. Every test calls through the C ABI (axia_gpu).
"""
import os
import re
import sys

import numpy as np
import pytest

import axia_gpu as G
from axia_gpu import meshgen as MG
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
DECKS = ["cube_tension", "c3d20_patch", "patch_c3d8", "le10_c3d20", "le10_c3d20_fine"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def deck(name):
    m = O.parse_inp(open(os.path.join(GOLD, name + ".inp")).read())
    return m, O.deck_setup(m)


def mesh_of(ctx, s):
    return G.Mesh(s.nn, s.xyz, s.conn, s.e, s.nu, s.ind_of, ctx=ctx)


def sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", DECKS)
def test_assembly_bitexact_decks(ctx, name):
    say("case: assembly", name)
    g = np.load(os.path.join(GOLD, name + ".npz"))
    m, s = deck(name)
    mesh = mesh_of(ctx, s)
    mesh.assemble("fp64")
    indptr, indices, data = mesh.csr()
    assert int(g["n_free"]) == len(indptr) - 1
    assert sha(indptr.astype(np.int64)) == str(g["sha_indptr"])
    assert sha(indices.astype(np.int64)) == str(g["sha_indices"])
    assert sha(data.astype(np.float64)) == str(g["sha_data"]), "FP64 K_ff values not bit-exact"
    # scatter map: oracle Ke summed through the exported map gives the same bits
    sp, cb = mesh.scatter_map()
    kes = np.stack([O.hex_stiffness(s.xyz[s.conn[e]], s.e[e], s.nu[e])[0]
                    for e in range(len(s.conn))]).ravel()
    vals = np.array([sum(kes[cb[sp[k]:sp[k + 1]]].tolist(), 0.0) for k in range(len(indices))])
    assert np.array_equal(vals, data)


@pytest.mark.parametrize("nn,seed", [(8, 1), (20, 2)])
def test_ke_bitexact_jittered(ctx, nn, seed):
    say("case: Ke", nn, seed)
    if nn == 8:
        xyz, conn = MG.brick_hex8(3, 2, 2)
    else:
        xyz, conn = MG.brick_hex20(2, 2, 1)
    rng = np.random.default_rng(seed)
    xyz = xyz + rng.uniform(-0.1, 0.1, xyz.shape)
    ne = len(conn)
    e = rng.uniform(1e3, 3e5, ne)
    nu = rng.uniform(0.0, 0.45, ne)
    ind = MG.free_map(len(xyz), [0, 1, 2])
    mesh = G.Mesh(nn, xyz, conn, e, nu, ind, ctx=ctx)
    mesh.assemble("fp64")
    ke = mesh.ke()
    for el in range(ne):
        ref, _ = O.hex_stiffness(xyz[conn[el]], e[el], nu[el])
        assert np.array_equal(ke[el], ref), f"element {el}"


@pytest.mark.parametrize("name", DECKS + ["cantilever_20x4x4", "cantilever_20x9x9"])
def test_solve_and_stress_vs_oracle(ctx, name):
    say("case: solve", name)
    g = np.load(os.path.join(GOLD, name + ".npz"))
    if name.startswith("cantilever"):
        shape = tuple(int(v) for v in name.split("_")[1].split("x"))
        p = MG.cantilever(*shape)
        mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
        nn, nnode = 8, p.nnode
    else:
        m, s = deck(name)
        mesh = mesh_of(ctx, s)
        nn, nnode = s.nn, len(s.xyz)
    mesh.assemble("fp64")
    rhs = g["rhs"]
    x, info = mesh.solve(rhs, tol=1e-12)
    err = rel(x, g["x"])
    say(f"  n={len(rhs)} iters={info['iters']} rel.err={err:.2e} {info['name']}")
    assert err < 1e-8
    assert info["rel_residual"] < 1e-11
    # stresses from the oracle's displacements: same FP64 bits
    st, vm, en = mesh.nodal_stress(g["u_full"])
    assert np.array_equal(st, g["stress"]), f"max diff {np.abs(st - g['stress']).max()}"
    assert np.allclose(vm, g["von_mises"], rtol=1e-14, atol=0)
    if "elem_nodal" in g:
        assert np.array_equal(en, g["elem_nodal"])
    # stresses from the GPU displacements
    u = np.zeros(3 * nnode)
    free = g["ind_of"] >= 0
    u[free] = x[g["ind_of"][free]]
    st2, _, _ = mesh.nodal_stress(u)
    assert rel(st2, g["stress"]) < 1e-8


def test_patch_tests_through_gpu(ctx):
    """lib.rs:429-462 and 751-774: the reference's acceptance values."""
    for name, ux_ref, sxx_ref, tol_u, tol_s in (("cube_tension", 0.01, 210.0, 1e-6, 1e-3),
                                                ("c3d20_patch", -0.01, -210.0, 1e-5, 1.0)):
        say("case: patch", name)
        m, s = deck(name)
        mesh = mesh_of(ctx, s)
        mesh.assemble("fp64")
        oi, oj, ov, rhs = O.reduce(s.ind_of, s.u0, s.n_free, *s.trips, s.f_full)
        x, _ = mesh.solve(rhs, tol=1e-12)
        u = np.zeros(s.ndof)
        u[s.ind_of >= 0] = x[s.ind_of[s.ind_of >= 0]]
        ids = [m.node_index(i) for i in (2, 3, 6, 7)]
        ux = np.mean([u[3 * i] for i in ids])
        assert abs(ux - ux_ref) < tol_u
        st, _, _ = mesh.nodal_stress(u)
        assert abs(st[:, 0].mean() - sxx_ref) < tol_s


@pytest.mark.parametrize("name,target", [("le10_c3d20", -5.51), ("le10_c3d20_fine", -5.50)])
def test_nafems_le10_through_gpu(ctx, name, target):
    """nafems.rs:50-90 and examples/nafems/README.md:12-13."""
    say("case: LE10", name)
    m, s = deck(name)
    mesh = mesh_of(ctx, s)
    mesh.assemble("fp64")
    oi, oj, ov, rhs = O.reduce(s.ind_of, s.u0, s.n_free, *s.trips, s.f_full)
    x, _ = mesh.solve(rhs, tol=1e-12)
    u = np.zeros(s.ndof)
    u[s.ind_of >= 0] = x[s.ind_of[s.ind_of >= 0]]
    st, _, _ = mesh.nodal_stress(u)
    syy = st[m.node_index(m.nsets["D"][0]), 1]
    say(f"  sigma_yy(D) = {syy:.4f}")
    assert round(syy, 2) == target
    assert abs((syy + 5.38) / 5.38) < (0.05 if name == "le10_c3d20" else 0.03)


def test_solve_kff_dropin(ctx):
    """linalg::solve_kff on the reduced triplets of a deck (analysis.rs:1019-1020)."""
    say("case: solve_kff")
    m, s = deck("le10_c3d20")
    oi, oj, ov, rhs = O.reduce(s.ind_of, s.u0, s.n_free, *s.trips, s.f_full)
    g = np.load(os.path.join(GOLD, "le10_c3d20.npz"))
    r = G.solve_kff(s.n_free, (oi, oj, ov), rhs, ctx=ctx, tol=1e-12)
    assert isinstance(r, G.SparseResult)
    assert r.name.startswith("GPU PCG")
    assert rel(r.x, g["x"]) < 1e-8
    assert r.iters > 0 and r.residual < 1e-10 * np.linalg.norm(rhs)


@pytest.mark.parametrize("precision,precond", [("fp64", "block3"), ("mixed", "jacobi"),
                                               ("mixed", "block3")])
def test_solve_kff_dropin_options(ctx, precision, precond):
    """The drop-in honours precision and preconditioner: LE10's reduced system (partly
    constrained nodes, so row triples straddle nodes) against the oracle's direct solve."""
    say("case: solve_kff", precision, precond)
    m, s = deck("le10_c3d20")
    oi, oj, ov, rhs = O.reduce(s.ind_of, s.u0, s.n_free, *s.trips, s.f_full)
    g = np.load(os.path.join(GOLD, "le10_c3d20.npz"))
    r = G.solve_kff(s.n_free, (oi, oj, ov), rhs, ctx=ctx, tol=1e-12, precision=precision,
                    precond=precond)
    say(f"  {r.name}: iters {r.iters}, err {rel(r.x, g['x']):.2e}")
    assert ("FP32" in r.name) == (precision == "mixed")
    assert ("block-Jacobi" in r.name) == (precond == "block3")
    assert rel(r.x, g["x"]) < 1e-8
    assert r.residual <= 1e-11 * np.linalg.norm(rhs)


@pytest.mark.parametrize("order", ["sorted", "shuffled"])
@pytest.mark.parametrize("precision,precond", [("fp64", "jacobi"), ("fp64", "block3"),
                                               ("mixed", "block3")])
def test_solve_csr_generic_spd(ctx, precision, precond, order):
    """A caller's CSR without node structure (n = 301, not a multiple of 3, rows of
    different patterns) through axg_solve_csr in each mode, against LAPACK; with the
    columns of every row in ascending order, and shuffled (the 3x3 block inverses then
    scan whole rows instead of searching them)."""
    say("case: solve_csr", precision, precond, order)
    rng = np.random.default_rng(12)
    n = 301
    a = np.zeros((n, n))
    for off in (1, 2, 5, 17):
        v = rng.uniform(-1.0, 1.0, n - off) * (rng.random(n - off) < 0.7)
        a += np.diag(v, off) + np.diag(v, -off)
    a += np.diag(np.abs(a).sum(1) + rng.uniform(0.5, 2.0, n))
    import scipy.sparse as sps
    csr = sps.csr_matrix(a)
    b = rng.standard_normal(n)
    indices, data = csr.indices.copy(), csr.data.copy()
    if order == "shuffled":
        for i in range(n):
            lo, hi = csr.indptr[i], csr.indptr[i + 1]
            perm = lo + rng.permutation(hi - lo)
            indices[lo:hi], data[lo:hi] = csr.indices[perm], csr.data[perm]
        assert any(np.any(np.diff(indices[csr.indptr[i]:csr.indptr[i + 1]]) < 0) for i in range(n))
    x, info = G.solve_csr(csr.indptr, indices, data, b, ctx=ctx, tol=1e-12,
                          precision=precision, precond=precond)
    ref = np.linalg.solve(a, b)
    say(f"  {info['name']}: iters {info['iters']}, err {rel(x, ref):.2e}")
    assert rel(x, ref) < 1e-10
    assert info["rel_residual"] <= 1e-11


def test_csr_from_triplets_device_bitexact(ctx):
    say("case: csr_from_triplets")
    rng = np.random.default_rng(3)
    n, t = 300, 20000
    ti = rng.integers(-2, n + 2, t)
    tj = rng.integers(-2, n + 2, t)
    tv = rng.standard_normal(t)
    tv[rng.random(t) < 0.1] = 0.0
    tv[rng.random(t) < 0.02] = -0.0
    ref = O.csr_from_triplets(n, ti, tj, tv)
    indptr, indices, data = G.csr_from_triplets(n, ti, tj, tv, ctx=ctx)
    assert np.array_equal(indptr, ref.indptr)
    assert np.array_equal(indices, ref.indices)
    assert np.array_equal(data, ref.data)


def test_block_jacobi_fp64(ctx):
    say("case: block-Jacobi")
    g = np.load(os.path.join(GOLD, "le10_c3d20.npz"))
    m, s = deck("le10_c3d20")
    mesh = mesh_of(ctx, s)
    mesh.assemble("fp64")
    x, info = mesh.solve(g["rhs"], precond="block3", tol=1e-12)
    assert rel(x, g["x"]) < 1e-8
    assert "block" in info["name"]


def test_error_paths(ctx):
    say("case: errors")
    xyz, conn = MG.brick_hex8(2, 1, 1)
    ind = MG.free_map(len(xyz), [0, 3])
    bad = conn.copy()
    bad[1] = bad[1][[4, 5, 6, 7, 0, 1, 2, 3]]  # mirrored element: det < 0
    mesh = G.Mesh(8, xyz, bad, 2e5, 0.3, ind, ctx=ctx)
    with pytest.raises(G.AxgError) as ei:
        mesh.assemble()
    assert ei.value.code == 3 and "negative Jakobideterminante" in str(ei.value)
    mesh = G.Mesh(8, xyz, conn, 2e5, 0.5, ind, ctx=ctx)
    with pytest.raises(G.AxgError) as ei:
        mesh.assemble()
    assert ei.value.code == 4 and "Querkontraktion" in str(ei.value)
    # not converged: too few iterations (linalg.rs:253-257)
    g = np.load(os.path.join(GOLD, "cantilever_20x4x4.npz"))
    p = MG.cantilever(20, 4, 4)
    mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
    mesh.assemble()
    with pytest.raises(G.AxgError) as ei:
        mesh.solve(g["rhs"], tol=1e-10, max_iter=5)
    assert ei.value.code == 6 and "CG konvergierte nicht" in str(ei.value)
    # zero pivot (linalg.rs:231-233): p.Ap = 0 on [[0,1],[1,0]]
    with pytest.raises(G.AxgError) as ei:
        G.solve_kff(2, ([0, 1], [1, 0], [1.0, 1.0]), np.array([1.0, 0.0]), ctx=ctx)
    assert ei.value.code == 5 and "Null-Pivot" in str(ei.value)
    # zero right-hand side: x = 0, zero iterations (linalg.rs:206-214)
    r = G.solve_kff(2, ([0, 1], [0, 1], [2.0, 3.0]), np.zeros(2), ctx=ctx)
    assert r.iters == 0 and np.all(r.x == 0.0)


def test_brick_properties_40(ctx):
    """Size-independent checks on a 40^3 lognormal brick: FP64 residual on the host,
    energy identity u^T K u = f^T u, symmetric pattern."""
    say("case: brick 40^3")
    import scipy.sparse as sp
    p = MG.brick_lognormal(40, seed=1)
    mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
    ms = mesh.assemble("fp64")
    indptr, indices, data = mesh.csr()
    n = len(indptr) - 1
    K = sp.csr_matrix((data, indices, indptr), shape=(n, n))
    assert (abs(K - K.T) > 1e-9 * abs(K).max()).nnz == 0
    b = p.rhs()
    x, info = mesh.solve(b, tol=1e-10)
    r = b - K @ x
    assert np.linalg.norm(r) / np.linalg.norm(b) < 2e-10
    assert abs(x @ (K @ x) - b @ x) / abs(b @ x) < 1e-9
    say(f"  assembly {ms:.2f} ms, iters {info['iters']}, solve {info['solve_ms']:.1f} ms")


# ------------------------------------------------------------ mixed precision
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("nn,seed,offset,dup", [(8, 4, 0.0, 1), (20, 5, 0.0, 1), (8, 6, 1e5, 1),
                                                (20, 7, 1e5, 1), (8, 8, 0.0, 3), (20, 9, 0.0, 3)])
def test_mixed_op32_vs_fp64(ctx, nn, seed, offset, dup, mode):
    """The FP32 operator agrees with the FP64 K_ff entry by entry to FP32 accuracy (relative
    to its row's largest entry), in both assembly forms (0: integration fused with the
    block-row assembly; 1: element blocks on the matrix cores, then the block-row gather),
    also far from the origin (coordinates are taken relative to each element's first node)
    and with every element listed three times (24 elements at an interior node: several
    element chunks per block row); the constrained DOFs are absent from it."""
    say("case: op32", nn, seed, offset, dup, "mode", mode)
    if nn == 8:
        xyz, conn = MG.brick_hex8(3, 3, 2)
    else:
        xyz, conn = MG.brick_hex20(2, 2, 2)
    rng = np.random.default_rng(seed)
    xyz = xyz + rng.uniform(-0.1, 0.1, xyz.shape) + offset
    conn = np.concatenate([conn] * dup)
    ne = len(conn)
    e = rng.uniform(1e3, 3e5, ne)
    mesh = G.Mesh(nn, xyz, conn, e, 0.3, MG.free_map(len(xyz), [0, 1]), ctx=ctx)
    assert mesh.asm32_mode(mode) == mode
    check_op32(mesh)
    mesh.close()


def check_op32(mesh):
    """The FP32 operator entry by entry against the FP64 K_ff (relative to the row's largest
    entry), its pattern covering K_ff's, and its symmetry."""
    import scipy.sparse as sp
    mesh.assemble("mixed")
    indptr, indices, data = mesh.csr()
    n = len(indptr) - 1
    k64 = sp.csr_matrix((data, indices, indptr), shape=(n, n)).toarray()
    k32 = mesh.op32().toarray().astype(np.float64)
    rowmax = np.abs(k64).max(axis=1, keepdims=True)
    err = np.abs(k32 - k64) / rowmax
    say(f"  n={n}, worst entry error / row max {err.max():.2e}")
    assert err.max() < 3e-6
    # every coupling of K_ff is stored; nothing outside the block pattern
    assert np.all((k32 != 0) | (k64 == 0))
    assert np.abs(k32 - k32.T).max() < 3e-6 * rowmax.max()


def fan_hex8(nsec, layers):
    """nsec kite-shaped hexahedra around the z axis per layer: every node on the axis is shared
    by 2 nsec elements (nsec in the end layers) and couples with 6 nsec + 3 nodes."""
    per = 1 + 2 * nsec
    pts = []
    for z in range(layers + 1):
        pts.append((0.0, 0.0, float(z)))
        for i in range(nsec):
            t = 2 * np.pi * i / nsec
            pts.append((np.cos(t), np.sin(t), float(z)))
        for i in range(nsec):
            t = 2 * np.pi * (i + 0.5) / nsec
            pts.append((1.3 * np.cos(t), 1.3 * np.sin(t), float(z)))
    conn = []
    for z in range(layers):
        for i in range(nsec):
            q = [0, 1 + i, 1 + nsec + i, 1 + (i + 1) % nsec]
            conn.append([z * per + v for v in q] + [(z + 1) * per + v for v in q])
    return np.array(pts), np.array(conn, np.int64)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("nsec", [6, 16])
def test_mixed_op32_wide_block_rows(ctx, nsec, mode):
    """Block rows wider than the two-rows-per-wavefront form (the axis nodes of a fan of
    hexahedra: 39 and 99 blocks, 12 and 32 incident elements), partially constrained nodes (the
    outer top nodes keep only y and z free): the FP32 operator against the FP64 K_ff, both
    assembly forms."""
    say("case: op32 fan", nsec, "mode", mode)
    xyz, conn = fan_hex8(nsec, 2)
    per = 1 + 2 * nsec
    ind = np.full(3 * len(xyz), -1, np.int64)
    pres = np.zeros(3 * len(xyz), bool)
    pres[:3 * per] = True                                    # bottom layer clamped
    pres[3 * (2 * per + 1 + nsec + np.arange(nsec))] = True  # outer top nodes: x prescribed
    ind[~pres] = np.arange(int((~pres).sum()))
    rng = np.random.default_rng(nsec)
    mesh = G.Mesh(8, xyz, conn, rng.uniform(1e3, 3e5, len(conn)), 0.3, ind, ctx=ctx)
    mesh.asm32_mode(mode)
    check_op32(mesh)
    mesh.close()


@pytest.mark.parametrize("name,precond", [("le10_c3d20", "block3"), ("le10_c3d20_fine", "jacobi"),
                                          ("cantilever_20x9x9", "block3")])
def test_mixed_solve_vs_oracle(ctx, name, precond):
    say("case: mixed", name, precond)
    g = np.load(os.path.join(GOLD, name + ".npz"))
    if name.startswith("cantilever"):
        p = MG.cantilever(20, 9, 9)
        mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
    else:
        m, s = deck(name)
        mesh = mesh_of(ctx, s)
    mesh.assemble("mixed")
    x, info = mesh.solve(g["rhs"], precision="mixed", precond=precond, tol=1e-12)
    err = rel(x, g["x"])
    say(f"  inner={info['iters']} sweeps={info['outer_iters']} rel.err={err:.2e} {info['name']}")
    assert err < 1e-8
    assert info["rel_residual"] <= 1e-12


def test_mixed_multi_rhs_matches_fp64(ctx):
    say("case: mixed multi-rhs")
    p = MG.brick_lognormal(12, seed=3, nloads=4)
    mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
    mesh.assemble("mixed")
    B = p.rhs_all()
    X64, _ = mesh.solve(B, precision="fp64", tol=1e-13)
    X, info = mesh.solve(B, precision="mixed", precond="block3", tol=1e-12)
    assert info["nrhs"] == 4
    for k in range(4):
        assert rel(X[k], X64[k]) < 1e-9


@pytest.mark.parametrize("refinement", ["fcg", "plain", "reliable"])
def test_mixed_refinement_modes_match_fp64(ctx, refinement):
    """Flexible-CG, classical refinement and reliable updates all reach the FP64 solution."""
    say("case: refinement", refinement)
    p = MG.brick_lognormal(10, seed=8, nloads=2)
    mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
    mesh.assemble("mixed")
    B = p.rhs_all()
    X64, _ = mesh.solve(B, precision="fp64", precond="block3", tol=1e-13)
    X, info = mesh.solve(B, precision="mixed", precond="block3", tol=1e-12, refinement=refinement)
    assert {"fcg": "FCG", "plain": "IR", "reliable": "RU"}[refinement] in info["name"]
    for k in range(2):
        assert rel(X[k], X64[k]) < 1e-9
    assert info["rel_residual"] <= 1e-12


def test_mixed_every_spmm_shape():
    """Every FP32 block SpMM variant the autotuner can pick (pinned one after the other)
    gives a mixed solve within 1e-9 of the FP64 solve, at 4 and 1 right-hand sides."""
    say("case: every SpMM shape")
    c = G.Context(0)
    p = MG.brick_lognormal(11, seed=9, nloads=4)
    B = p.rhs_all()
    ref = None
    for shape in range(17):
        mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=c)
        mesh.assemble("mixed")
        if ref is None:
            ref, _ = mesh.solve(B, precision="fp64", tol=1e-13)
        assert mesh.spmm_shape(4, shape) == shape and mesh.spmm_shape(1, shape) == shape
        X, info = mesh.solve(B, precision="mixed", precond="block3", tol=1e-12)
        x1, _ = mesh.solve(B[2], precision="mixed", precond="jacobi", tol=1e-12)
        err = max(rel(X[k], ref[k]) for k in range(4))
        say(f"  shape {shape}: inner {info['iters']}, err {err:.2e}, single {rel(x1, ref[2]):.2e}")
        assert err < 1e-9 and rel(x1, ref[2]) < 1e-9
        mesh.close()
    c.close()


@pytest.mark.parametrize("shape", [13, 15])
def test_mixed_lds_shapes_wide_groups(ctx, shape):
    """C3D20 rows couple up to 81 nodes: on a 24 x 24 x 2 brick a group of 256 block rows
    has up to 2,080 distinct column blocks, more than the LDS-staged SpMM stages (1,360), so
    the rest come from global memory (the column rides in the third plane); mixed solves
    with LDS shapes pinned (4 and 1 right-hand sides) match the FP64 solve."""
    say("case: LDS shapes, wide groups", shape)
    xyz, conn = MG.brick_hex20(24, 24, 2)
    rng = np.random.default_rng(shape)
    ind = MG.free_map(len(xyz), np.nonzero(xyz[:, 2] == 0.0)[0])
    mesh = G.Mesh(20, xyz, conn, rng.uniform(5e4, 2.5e5, len(conn)), 0.3, ind, ctx=ctx)
    mesh.assemble("mixed")
    n = mesh.n_free
    B = rng.standard_normal((4, n))
    ref, _ = mesh.solve(B, precision="fp64", tol=1e-13)
    assert mesh.spmm_shape(4, shape) == shape and mesh.spmm_shape(1, shape) == shape
    X, info = mesh.solve(B, precision="mixed", precond="block3", tol=1e-11)
    err = max(rel(X[k], ref[k]) for k in range(4))
    x1, _ = mesh.solve(B[1], precision="mixed", precond="block3", tol=1e-11)
    say(f"  n={n}, inner {info['iters']}, err {err:.2e}, single {rel(x1, ref[1]):.2e}")
    assert err < 1e-8 and rel(x1, ref[1]) < 1e-8
    mesh.close()


def test_mixed_c3d20_stiff_beam_floor():
    """A slender jittered C3D20 beam: its FP64 residual has a round-off floor near 1e-10
    relative. The mixed solve reaches 1e-9 (displacements within 1e-8 of FP64), and an
    unreachable 1e-14 stops on stagnation after a few sweeps instead of running to the
    iteration cap."""
    say("case: c3d20 beam floor")
    c = G.Context(0)
    p = MG.brick_c3d20_jitter(32, 6, 6, seed=3)
    mesh = G.Mesh(20, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=c)
    mesh.assemble("mixed")
    b = p.rhs()
    x64, _ = mesh.solve(b, precision="fp64", precond="block3", tol=1e-13)
    x, info = mesh.solve(b, precision="mixed", precond="block3", tol=1e-9)
    assert rel(x, x64) < 1e-8
    with pytest.raises(G.AxgError) as e:
        mesh.solve(b, precision="mixed", precond="block3", tol=1e-14)
    m = re.search(r"nach (\d+) Iterationen", str(e.value))
    say(f"  stopped after {m.group(1) if m else '?'} inner iterations (n = {p.n_free})")
    assert m and int(m.group(1)) < 4 * p.n_free
    mesh.close()
    c.close()


def test_mixed_after_fp64_multi_rhs_on_fresh_context():
    """A multi-RHS FP64 solve before the first mixed solve of a context leaves stale
    FP64 bits where the FP32 search directions are allocated; the inner PCG must not
    read them (its first direction is z itself)."""
    say("case: mixed after fp64 multi-rhs, fresh context")
    c = G.Context(0)
    p = MG.brick_lognormal(12, seed=3, nloads=4)
    mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=c)
    mesh.assemble("mixed")
    B = p.rhs_all()
    X64, _ = mesh.solve(B, precision="fp64", tol=1e-13)
    for k in range(4):
        x, info = mesh.solve(B[k], precision="mixed", precond="block3", tol=1e-12)
        assert rel(x, X64[k]) < 1e-9
    X, _ = mesh.solve(B, precision="mixed", precond="jacobi", tol=1e-12)
    for k in range(4):
        assert rel(X[k], X64[k]) < 1e-9
    # the context may be released before its meshes: the mesh keeps it alive
    c.close()
    x, _ = mesh.solve(B[0], precision="fp64", tol=1e-13)
    assert rel(x, X64[0]) < 1e-12
    mesh.close()


# ------------------------------------------------------------------ batched
@pytest.mark.parametrize("groups", [1, 2, 4])
def test_batched_multi_mesh_matches_single_solves(ctx, groups):
    """Uneven meshes, several load cases each, in one batched call (one launch sequence, or
    size groups solved smallest first); every solve matches the single-mesh FP64 PCG and
    the reference's stopping rule, in the caller's order."""
    say("case: batch, groups", groups)
    probs = [MG.batch_meshes(count=1, lo=m, hi=m, nloads=k, seed=m)[0]
             for m, k in ((4, 4), (7, 2), (5, 3), (9, 1))]
    meshes = []
    for p in probs:
        me = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
        me.assemble("fp64")
        meshes.append(me)
    rhs = [p.rhs_all() for p in probs]
    xs, iters, lat, info = G.solve_batch(meshes, rhs, tol=1e-12, groups=groups)
    assert f"{groups} groups" in info["name"]
    q = 0
    for p, me, B, X in zip(probs, meshes, rhs, xs):
        for k in range(B.shape[0]):
            x1, inf1 = me.solve(B[k], tol=1e-12)
            assert rel(X[k], x1) < 1e-9
            assert abs(int(iters[q]) - inf1["iters"]) <= 2
            assert 0.0 < lat[q] <= info["solve_ms"] * 1.05 + 1.0
            q += 1
    assert info["rel_residual"] <= 1e-12 * 10
