"""Compliance adjoint on the GPU: against the FP64 CPU adjoint (oracle Ke),
central finite differences (step 1e-4 on 64 random elements) and
torch.autograd.gradcheck through the autograd binding."""
import sys

import numpy as np
import pytest

import axia_gpu as G
from axia_gpu import meshgen as MG
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def simp_problem(nx, ny, nz, seed=0):
    return MG.simp_brick(nx, ny, nz, seed=seed)


def solve_c(mesh, p, e, tol=1e-13):
    mesh.set_modulus(e)
    mesh.assemble("fp64")
    x, _ = mesh.solve(p.rhs(), precond="block3", tol=tol)
    u = p.expand(x)
    return float(p.f_full[0] @ u), u


def test_adjoint_vs_cpu_fp64_adjoint(ctx):
    say("case: adjoint vs CPU")
    p = simp_problem(10, 5, 5, seed=1)
    mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
    c, u = solve_c(mesh, p, p.e)
    c2, dc_de = mesh.compliance_adjoint(u, p.f_full[0])
    assert abs(c2 - c) <= 1e-14 * abs(c)
    ref = np.empty(p.nelem)
    for el in range(p.nelem):
        nodes = p.conn[el]
        ke1, _ = O.hex_stiffness(p.xyz[nodes], 1.0, p.nu[el])
        ue = u.reshape(-1, 3)[nodes].ravel()
        ref[el] = -ue @ ke1 @ ue
    assert np.max(np.abs(dc_de - ref)) <= 1e-11 * np.max(np.abs(ref))
    # energy identity: sum_e E_e u_e^T K1_e u_e = u^T K u = c
    assert abs(-np.dot(p.e, dc_de) - c) <= 1e-9 * abs(c)


def test_adjoint_vs_central_differences(ctx):
    say("case: adjoint vs FD")
    p = simp_problem(12, 6, 6, seed=2)
    rho = p.meta["rho"]
    mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
    c0, u = solve_c(mesh, p, p.e)
    _, dc_de = mesh.compliance_adjoint(u, p.f_full[0])
    dc_drho = dc_de * 3.0 * rho ** 2  # E = 1e-3 + rho^3
    rng = np.random.default_rng(5)
    pick = rng.choice(p.nelem, 64, replace=False)
    h = 1e-4
    worst = 0.0
    for el in pick:
        rp, rm = rho.copy(), rho.copy()
        rp[el] += h
        rm[el] -= h
        cp, _ = solve_c(mesh, p, 1e-3 + rp ** 3)
        cm, _ = solve_c(mesh, p, 1e-3 + rm ** 3)
        fd = (cp - cm) / (2 * h)
        worst = max(worst, abs(fd - dc_drho[el]) / abs(dc_drho[el]))
    say(f"  worst relative FD deviation {worst:.2e}")
    assert worst < 1e-5


def test_autograd_gradcheck(ctx):
    say("case: gradcheck")
    import torch
    from axia_gpu.autograd import compliance, simp_compliance
    p = simp_problem(3, 2, 2, seed=3)
    mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
    e = torch.tensor(p.e, dtype=torch.float64, requires_grad=True)
    f = p.f_full[0]
    assert torch.autograd.gradcheck(lambda t: compliance(t, mesh, f, tol=1e-14), (e,), eps=1e-6,
                                    atol=1e-8, rtol=1e-5)
    rho = torch.tensor(p.meta["rho"], dtype=torch.float64, requires_grad=True)
    c = simp_compliance(rho, mesh, f, tol=1e-14)
    c.backward()
    g = rho.grad.numpy()
    assert np.all(g < 0)  # stiffer material never increases compliance
    assert torch.autograd.gradcheck(lambda t: simp_compliance(t, mesh, f, tol=1e-14), (rho,),
                                    eps=1e-6, atol=1e-8, rtol=1e-5)
