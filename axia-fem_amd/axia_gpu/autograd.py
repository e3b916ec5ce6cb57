"""torch.autograd binding of the compliance adjoint over the C ABI.

forward:  E (per-element modulus, float64 tensor) -> assemble K(E) on the
          device, solve K u = f (FP64 PCG), return c = f^T u; keeps u.
backward: dc/dE_e = -u_e^T K1_e u_e (self-adjoint system: the adjoint
          solution is -u, so no second solve), from axg_compliance_adjoint on
          the kept u. K1_e depends on the geometry and nu only, so the mesh's
          modulus may change between forward and backward. A forward without
          a backward (no gradient needed) never runs the adjoint kernel.

The modulus enters K linearly (K = sum_e E_e K1_e with K1_e the element
stiffness at unit modulus), so any parametrisation E(rho) composes through
ordinary torch operations, e.g. SIMP E = emin E0 + rho^p E0.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Mesh


class Compliance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, e, mesh: Mesh, f_full, tol=1e-12, precond="block3"):
        e_np = e.detach().to("cpu", torch.float64).numpy()
        mesh.set_modulus(e_np)
        mesh.assemble("fp64")
        free = mesh.ind_of >= 0
        b = np.zeros(int(free.sum()))
        b[mesh.ind_of[free]] = f_full[free]
        x, _ = mesh.solve(b, precision="fp64", precond=precond, tol=tol)
        u = np.zeros(3 * mesh.nnode)
        u[free] = x[mesh.ind_of[free]]
        c = float(np.dot(f_full, u))
        ctx.mesh = mesh
        ctx.u = u
        ctx.f_full = f_full
        return torch.tensor(c, dtype=torch.float64, device=e.device)

    @staticmethod
    def backward(ctx, grad_c):
        _, dc_de = ctx.mesh.compliance_adjoint(ctx.u, ctx.f_full)
        g = torch.from_numpy(dc_de).to(device=grad_c.device, dtype=torch.float64)
        return grad_c * g, None, None, None, None


def compliance(e, mesh, f_full, tol=1e-12, precond="block3"):
    """c(E) with gradient dc/dE (see Compliance)."""
    return Compliance.apply(e, mesh, np.ascontiguousarray(f_full, dtype=np.float64), tol, precond)


def simp_compliance(rho, mesh, f_full, e0=1.0, emin=1e-3, p=3.0, tol=1e-12):
    """SIMP: E_e = emin E0 + rho_e^p E0, c = f^T u, differentiable in rho."""
    e = emin * e0 + rho ** p * e0
    return compliance(e, mesh, f_full, tol)
