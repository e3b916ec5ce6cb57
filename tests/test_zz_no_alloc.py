"""The no-allocation entry point (axg_solve_csr_device_noalloc) against axg_solve_csr
(FP64, Jacobi) on the same systems, through the C ABI: two calls in a row, the second
with a smaller system, then the size limit. Device buffers come from torch (plumbing)."""
import ctypes as C
import sys

import numpy as np
import pytest

import axia_gpu as G
from axia_gpu import _lib
from axia_gpu import meshgen as MG

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def csr_of(shape, ctx):
    p = MG.cantilever(*shape)
    mesh = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
    mesh.assemble("fp64")
    indptr, indices, data = mesh.csr()
    mesh.close()
    return indptr, indices, data, p.rhs()


def noalloc(ctx, indptr, indices, data, b, tol):
    import torch
    dev = torch.device("cuda:0")
    rp = torch.tensor(indptr.astype(np.int32), device=dev)
    ci = torch.tensor(indices.astype(np.int32), device=dev)
    va = torch.tensor(data, dtype=torch.float64, device=dev)
    bd = torch.tensor(b, dtype=torch.float64, device=dev)
    xd = torch.full_like(bd, float("nan"))  # stale contents must not matter
    torch.cuda.synchronize()
    o = _lib.SolveOpts()
    _lib.lib().axg_default_opts(C.byref(o))
    o.tol = tol
    info = _lib.SolveInfo()
    st = _lib.lib().axg_solve_csr_device_noalloc(ctx.handle, len(b), rp.data_ptr(), ci.data_ptr(),
                                                 va.data_ptr(), bd.data_ptr(), C.byref(o),
                                                 xd.data_ptr(), C.byref(info))
    _lib.check(st)
    return xd.cpu().numpy(), info


def test_noalloc_matches_solve_csr():
    ctx = G.Context(0)
    for shape in ((20, 9, 9), (12, 4, 4)):  # the second call on a smaller system
        say("case: no-alloc solve, cantilever", shape)
        indptr, indices, data, b = csr_of(shape, ctx)
        x_ref, inf_ref = G.solve_csr(indptr, indices, data, b, ctx=ctx, tol=1e-10)
        x, info = noalloc(ctx, indptr, indices, data, b, 1e-10)
        err = float(np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref))
        say(f"  n={len(b)} iters {info.iters} (solve_csr {inf_ref['iters']}), rel.diff {err:.2e}, "
            f"rel.res {info.rel_residual:.2e}")
        assert err < 1e-8
        assert abs(info.iters - inf_ref["iters"]) <= 2
        assert info.rel_residual <= 1e-9 and info.name.decode() == "GPU PCG (FP64, Jacobi)"
        # the host residual of the returned solution
        import scipy.sparse as sp
        K = sp.csr_matrix((data, indices, indptr), shape=(len(b), len(b)))
        assert np.linalg.norm(b - K @ x) / np.linalg.norm(b) <= 1e-9
    ctx.close()


def test_noalloc_rejects_what_it_cannot_do():
    ctx = G.Context(0)
    L = _lib.lib()
    o = _lib.SolveOpts()
    L.axg_default_opts(C.byref(o))
    info = _lib.SolveInfo()
    dummy = C.c_void_p(8)  # never dereferenced: the checks come first
    st = L.axg_solve_csr_device_noalloc(ctx.handle, (1 << 20) + 1, dummy, dummy, dummy, dummy,
                                        C.byref(o), dummy, C.byref(info))
    assert st == _lib.AXG_E_INVALID and "at most 1048576" in L.axg_last_error().decode()
    o.precision = _lib.PREC_MIXED
    st = L.axg_solve_csr_device_noalloc(ctx.handle, 10, dummy, dummy, dummy, dummy, C.byref(o),
                                        dummy, C.byref(info))
    assert st == _lib.AXG_E_INVALID and "FP64 with Jacobi only" in L.axg_last_error().decode()
    ctx.close()
