"""The mesh's FP64 Jacobi PCG on the node-block operator (linalg::pcg, linalg.rs:194-259, every
right-hand side at once) against the CSR Jacobi PCG of the solve_kff drop-in on the same K_ff,
the oracle's direct solve (golden x), and itself across right-hand-side widths, re-assembly and
iteration limits."""
import os
import sys

import numpy as np
import pytest

import axia_gpu as G
from axia_gpu import meshgen as MG

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def csr_solve(mesh, rhs, ctx, tol):
    """The same K_ff through axg_solve_kff (triplets -> CSR -> CSR Jacobi PCG)."""
    ip, ix, dv = mesh.csr()
    n = len(ip) - 1
    rows = np.repeat(np.arange(n), np.diff(ip))
    return G.solve_kff(n, (rows, ix, dv), rhs, ctx=ctx, tol=tol)


def test_matches_the_csr_pcg_and_the_golden(ctx):
    say("case: node-block vs CSR Jacobi PCG, cantilever 20x4x4")
    g = np.load(os.path.join(GOLD, "cantilever_20x4x4.npz"))
    p = MG.cantilever(20, 4, 4)
    mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
    mesh.assemble()
    x, info = mesh.solve(g["rhs"], tol=1e-10)
    assert "node blocks" in info["name"]
    ref = csr_solve(mesh, g["rhs"], ctx, 1e-10)
    say("  iters", info["iters"], ref.iters, "rel vs CSR", rel(x, ref.x), "vs golden", rel(x, g["x"]))
    # the same iteration in exact arithmetic: counts within the rounding of the stop test
    assert abs(info["iters"] - ref.iters) <= 2
    assert rel(x, ref.x) < 1e-8
    assert rel(x, g["x"]) < 1e-8
    assert info["rel_residual"] < 1e-9


@pytest.mark.parametrize("nn", [8, 20])
def test_block_jacobi_on_the_node_blocks(ctx, nn):
    """Block-Jacobi (the 3x3 node inverses applied through the workgroup's residual slab in
    LDS, 255-row workgroups of whole nodes) on the node-block operator: the solution of the
    Jacobi solve and of the direct reference, for one and for several right-hand sides."""
    say("case: block-Jacobi", nn)
    if nn == 8:
        p = MG.brick_lognormal(11, seed=5)
        mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
        b = p.rhs()
    else:
        xyz, conn = MG.brick_hex20(6, 3, 3)
        ind = MG.free_map(len(xyz), np.nonzero(xyz[:, 0] == 0.0)[0])
        mesh = G.Mesh(20, xyz, conn, np.full(len(conn), 70000.0), np.full(len(conn), 0.3), ind,
                      ctx=ctx)
        b = np.random.default_rng(1).standard_normal(int((ind >= 0).sum()))
    mesh.assemble()
    xj, ij = mesh.solve(b, tol=1e-12)
    xb, ib = mesh.solve(b, precond="block3", tol=1e-12)
    ref = csr_solve(mesh, b, ctx, 1e-12)
    say("  iters jacobi", ij["iters"], "block", ib["iters"], "rel", rel(xb, xj), rel(xb, ref.x))
    assert "node blocks, block-Jacobi" in ib["name"]
    assert rel(xb, xj) < 1e-9 and rel(xb, ref.x) < 1e-9
    B = np.stack([b, 0.5 * b + 1.0, -b])
    XB, info = mesh.solve(B, precond="block3", tol=1e-12)
    for k in range(3):
        assert rel(XB[k], mesh.solve(B[k], precond="block3", tol=1e-12)[0]) < 1e-9


@pytest.mark.parametrize("nrhs", [2, 3, 6, 11])
def test_every_width_matches_single_solves(ctx, nrhs):
    """nrhs columns go through the 8 / 4 / 2 / 1 wide kernels with masked padding columns;
    each column is the single right-hand-side solve."""
    say("case: widths", nrhs)
    p = MG.brick_lognormal(10, seed=4)
    mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
    mesh.assemble()
    rng = np.random.default_rng(nrhs)
    B = rng.standard_normal((nrhs, mesh.n_free))
    B[1] = 0.0  # a zero column: x = 0 without iterations (linalg.rs:206-214)
    X, info = mesh.solve(B, tol=1e-10)
    total = 0
    for j in range(nrhs):
        xj, ij = mesh.solve(B[j], tol=1e-10)
        total += ij["iters"]
        if j == 1:
            assert np.all(X[j] == 0.0) and ij["iters"] == 0
        else:
            assert rel(X[j], xj) < 1e-8, j
    assert abs(info["iters"] - total) <= 2 * nrhs


def test_refreshes_after_reassembly_and_limits(ctx):
    """The cached solver re-reads the operator after an NLGEOM re-assembly (new values, same
    pattern) and re-captures its graphs when max_iter changes."""
    say("case: refresh")
    xyz, conn = MG.brick_hex8(6, 3, 3)
    ind = MG.free_map(len(xyz), np.nonzero(xyz[:, 0] == 0.0)[0])
    mesh = G.Mesh(8, xyz, conn, np.full(len(conn), 1000.0), np.full(len(conn), 0.3), ind, ctx=ctx)
    mesh.assemble()
    rng = np.random.default_rng(2)
    b = rng.standard_normal(mesh.n_free)
    x0, _ = mesh.solve(b, tol=1e-11)
    u = rng.uniform(-0.05, 0.05, 3 * len(xyz))
    u[ind < 0] = 0.0
    mesh.nl_assemble(u)  # K_T replaces K
    x1, i1 = mesh.solve(b, tol=1e-11)
    ref = csr_solve(mesh, b, ctx, 1e-11)
    say("  rel vs CSR after re-assembly", rel(x1, ref.x), "change", rel(x1, x0))
    assert rel(x1, ref.x) < 1e-8 and rel(x1, x0) > 1e-3
    with pytest.raises(G.AxgError) as ei:
        mesh.solve(b, tol=1e-11, max_iter=3)
    assert ei.value.code == 6 and "nach 3 Iterationen" in str(ei.value)
    x2, i2 = mesh.solve(b, tol=1e-11)
    assert np.array_equal(x2, x1) and i2["iters"] == i1["iters"]


def _partly_fixed(nx, ny, nz, seed):
    """A brick whose bottom nodes keep some DOFs: z fixed on the whole bottom face, x on its
    x = 0 edge, y at one corner, so block rows hold 1, 2 or 3 free DOFs."""
    xyz, conn = MG.brick_hex8(nx, ny, nz)
    ind = np.zeros(3 * len(xyz), np.int64)
    bot = xyz[:, 2] == 0.0
    ind[3 * np.nonzero(bot)[0] + 2] = -1
    ind[3 * np.nonzero(bot & (xyz[:, 0] == 0.0))[0]] = -1
    ind[3 * np.nonzero(bot & (xyz[:, 0] == 0.0) & (xyz[:, 1] == 0.0))[0] + 1] = -1
    free = ind >= 0
    ind[free] = np.arange(int(free.sum()))
    rng = np.random.default_rng(seed)
    e = rng.uniform(800.0, 1200.0, len(conn))
    return xyz, conn, e, ind


def test_batch_and_block_solves_with_partly_fixed_nodes(ctx):
    """Block rows with prescribed DOFs (zero rows and columns, divisor 1) inside the batch's
    concatenated node-block operator and the single-mesh block PCG, against the CSR PCG."""
    say("case: partly fixed nodes")
    meshes, rhs = [], []
    for k, (shape, seed) in enumerate((((5, 4, 3), 1), ((3, 3, 6), 2), ((7, 2, 2), 3))):
        xyz, conn, e, ind = _partly_fixed(*shape, seed)
        me = G.Mesh(8, xyz, conn, e, np.full(len(conn), 0.3), ind, ctx=ctx)
        me.assemble("fp64")
        meshes.append(me)
        rng = np.random.default_rng(10 + k)
        rhs.append(rng.standard_normal((k + 1, me.n_free)))
    xs, iters, lat, info = G.solve_batch(meshes, rhs, tol=1e-11)
    assert "node blocks" in info["name"]
    q = 0
    for me, B, X in zip(meshes, rhs, xs):
        for k in range(B.shape[0]):
            ref = csr_solve(me, B[k], ctx, 1e-11)
            x1, i1 = me.solve(B[k], tol=1e-11)
            say("  mesh", me.n_free, "rhs", k, "iters", int(iters[q]), i1["iters"], ref.iters)
            assert rel(X[k], ref.x) < 1e-8 and rel(x1, ref.x) < 1e-8
            assert abs(int(iters[q]) - ref.iters) <= 2 and abs(i1["iters"] - ref.iters) <= 2
            q += 1
