"""Benchmark of the linear-static hot path on one MI355X.

Workload (default): C3D8 brick 80x80x80 (512,000 elements, 531,441 nodes,
1,574,640 free DOFs), lognormal E (median 210000, sigma 0.5, seed 0),
nu = 0.3, bottom face clamped, four load cases on the top face (1 MPa
pressure P2, x and y shear of the same total, a seeded random normal load).

One step = numeric assembly (FP64 CSR for the true residual + FP32 MFMA
integration into the FP32 BSR-3 operator) followed by the mixed-precision
multi-right-hand-side PCG (FP32 inner, FP64 refinement) of all load cases to
||f - K u|| <= tol ||f||. Inputs are resident in HBM when the timed region
starts. After the timed steps the plain FP64 GPU PCG (Jacobi, the reference's
linalg::pcg) solves the same load cases for comparison, and both are checked
against an FP64 solution converged to 1e-13.

Prints one JSON line (see the keys below).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "axia-fem_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E, GB/s
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 (vector and matrix)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n_sample=40, pcg_iters=60, asm_n=28):
    """The oracle (C restatement of the reference's CPU path) on this host."""
    import numpy as np

    from axia_gpu import meshgen as MG
    from oracle import oracle as O

    L = O.lib()
    cores = os.cpu_count() or 1
    # B1: element integration + triplets + csr_from_triplets, one thread (as the reference)
    p = MG.brick_lognormal(asm_n, seed=1)
    t0 = time.perf_counter()
    ti, tj, tv = O.assemble_triplets(8, p.xyz, p.conn, p.e, p.nu)
    sel = (p.ind_of[ti] >= 0) & (p.ind_of[tj] >= 0)
    O.csr_from_triplets(p.n_free, p.ind_of[ti[sel]], p.ind_of[tj[sel]], tv[sel])
    t_asm = time.perf_counter() - t0
    asm_eps = p.nelem / t_asm
    # B3 / B4: the reference Jacobi PCG, bounded number of iterations
    q = MG.brick_lognormal(n_sample, seed=2)
    ti, tj, tv = O.assemble_triplets(8, q.xyz, q.conn, q.e, q.nu)
    sel = (q.ind_of[ti] >= 0) & (q.ind_of[tj] >= 0)
    csr = O.csr_from_triplets(q.n_free, q.ind_of[ti[sel]], q.ind_of[tj[sel]], tv[sel])
    b = q.rhs()
    x = np.zeros(q.n_free)
    it = np.zeros(1, np.int64)
    res = np.zeros(1)
    out = {}
    for threads in (1, cores):
        L.or_set_threads(threads)
        t0 = time.perf_counter()
        L.or_pcg_omp(q.n_free, O._l(csr.indptr), O._l(csr.indices), O._d(csr.data), O._d(b),
                     1e-30, 10 ** 9, pcg_iters, O._d(x), O._l(it), O._d(res))
        dt = time.perf_counter() - t0
        out[threads] = q.n_free * int(it[0]) / dt
    L.or_set_threads(0)
    return {
        "value": out[cores], "unit": "DOF·iter/s", "cores": cores, "kind": "port",
        "sample": (f"oracle linalg::pcg restatement (Jacobi, FP64 CSR, OpenMP) on a C3D8 "
                   f"{n_sample}^3 brick, n={q.n_free}, nnz={csr.nnz}, {pcg_iters} iterations; "
                   f"assembly sample {asm_n}^3 single-thread"),
        "pcg_dof_iter_per_s_1thread": out[1],
        "assembly_elements_per_s_1thread": asm_eps,
    }


def kernels_only(a, G, mesh, prob, n, nnz):
    """Repeated launches of the dominant kernels (input for rocprofv3 runs)."""
    out = {"workload": prob.name, "n_free": n, "nnz": nnz, "reps": a.reps}
    nstream = 1 << 27  # 1 GiB per array
    out["stream_copy_gbs"] = G.bench_stream(nstream, a.reps, ctx=mesh.ctx)
    log(f"stream copy {out['stream_copy_gbs']:.0f} GB/s")
    mesh.assemble("mixed")
    t = mesh.assembly_times()
    out["assembly_ms"] = t
    nb, nnzb = n // 3, nnz // 9
    for which, name, nbytes in ((0, "fp64_sell_spmv", 12 * nnz + 16 * n),
                                (1, "fp64_csr_spmv", 12 * nnz + 4 * (n + 1) + 16 * n)):
        ms = mesh.bench_operator(which, 1, a.reps)
        out[name] = {"ms": ms, "gbs": nbytes / (ms / 1e3) / 1e9}
        log(f"{name}: {ms:.4f} ms {out[name]['gbs']:.0f} GB/s")
    for k in (1, 4, 8):
        ms = mesh.bench_operator(2, k, a.reps)
        nbytes = 40 * nnzb + 2 * 4 * 3 * nb * k
        out[f"fp32_bsell_spmm_k{k}"] = {"ms": ms, "gbs": nbytes / (ms / 1e3) / 1e9}
        log(f"fp32 spmm k={k}: {ms:.4f} ms {out[f'fp32_bsell_spmm_k{k}']['gbs']:.0f} GB/s")
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=80, help="brick edge (elements)")
    ap.add_argument("--nrhs", type=int, default=4)
    ap.add_argument("--tol", type=float, default=1e-11)
    ap.add_argument("--precond", default="block3")
    ap.add_argument("--inner-tol", type=float, default=1e-4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--kernels", action="store_true",
                    help="only the operator kernels (repeated launches), for profiling")
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU "
              "sharding of this path is a design (see DESIGN.md), not a measured run.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import axia_gpu as G
    from axia_gpu import meshgen as MG

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    prob = MG.brick_lognormal(a.n, seed=0, nloads=a.nrhs)
    ctx = G.Context(0)
    mesh = G.Mesh(8, prob.xyz, prob.conn, prob.e, prob.nu, prob.ind_of, ctx=ctx)
    info = mesh.info()
    n, nnz = info["n_free"], info["nnz"]
    log(f"mesh {prob.name}: n={n} nnz={nnz} contributions={info['ncontrib']} "
        f"symbolic {info['symbolic_s']:.2f}s setup {time.perf_counter() - t0:.1f}s")
    if a.kernels:
        kernels_only(a, G, mesh, prob, n, nnz)
        return
    B = torch.tensor(prob.rhs_all(), dtype=torch.float64, device=dev).contiguous()
    X = torch.zeros_like(B)
    Xref = torch.zeros_like(B)
    k = B.shape[0]

    # FP64 reference (converged to 1e-13, untimed)
    mesh.assemble("mixed")
    ref = mesh.solve_device(B.data_ptr(), k, Xref.data_ptr(), precision="fp64", tol=1e-13)
    log(f"reference FP64 PCG: iters {ref['iters']} rel.res {ref['rel_residual']:.2e} "
        f"{ref['solve_ms'] / 1e3:.2f}s")

    def step(timing):
        ms_asm = mesh.assemble("mixed")
        t = mesh.assembly_times()
        inf = mesh.solve_device(B.data_ptr(), k, X.data_ptr(), precision="mixed",
                                precond=a.precond, tol=a.tol, inner_tol=a.inner_tol,
                                timing=timing)
        return ms_asm, t, inf

    for _ in range(a.warmup):
        step(False)
        log("warmup step done")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    steps = []
    for s_ in range(a.steps):
        steps.append(step(True))
        log(f"step {s_}: solve {steps[-1][2]['solve_ms']:.1f} ms inner {steps[-1][2]['iters']} "
            f"sweeps {steps[-1][2]['outer_iters']}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = 1e3 * wall / a.steps

    def rel_err(Xa):
        d = (Xa - Xref).double()
        return float((d.norm(dim=1) / Xref.norm(dim=1)).max().item())

    err_mixed = rel_err(X)
    inner_iters = float(np.mean([s[2]["iters"] for s in steps]))
    solve_ms = float(np.mean([s[2]["solve_ms"] for s in steps]))
    # DOF.iter/s of the inner PCG, aggregated over right-hand sides: each inner
    # iteration advances all k right-hand sides by one CG step over n DOFs
    # (inner_iters: per sweep the largest count over right-hand sides, summed)
    mixed_dof_iter_s = n * inner_iters * k / (solve_ms / 1e3)
    spmv_ms = float(np.mean([s[2]["spmv_ms"] / max(1, s[2]["spmv_launches"]) for s in steps]))
    asm = {kk: float(np.mean([s[1][kk] for s in steps])) for kk in steps[0][1]}
    nelem = prob.nelem

    # FP64 plain method on the same load cases
    fp64 = None
    if not a.no_fp64:
        X64 = torch.zeros_like(B)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        mesh.assemble("fp64")
        r64 = mesh.solve_device(B.data_ptr(), k, X64.data_ptr(), precision="fp64", tol=a.tol,
                                timing=True)
        torch.cuda.synchronize()
        wall64 = time.perf_counter() - t1
        fp64 = {
            "iters_total": r64["iters"], "solve_ms": r64["solve_ms"], "wall_ms": 1e3 * wall64,
            "dof_iter_per_s": n * r64["iters"] / (r64["solve_ms"] / 1e3),
            "spmv_ms_avg": r64["spmv_ms"] / max(1, r64["spmv_launches"]),
            "rel_err_vs_ref": rel_err(X64), "rel_residual": r64["rel_residual"],
        }
        fp64["spmv_gbs"] = (12 * nnz + 4 * (n + 1) + 16 * n) / (fp64["spmv_ms_avg"] / 1e3) / 1e9
        mesh.assemble("mixed")
        log(f"fp64: {fp64['solve_ms']:.0f} ms, iters {r64['iters']}, err {fp64['rel_err_vs_ref']:.2e}")

    # roofline of the dominant kernel: FP32 BSR-3 SpMM with k right-hand sides
    nb = n // 3
    nnzb = nnz // 9  # every free node keeps all 3 DOFs in this workload
    spmm_bytes = 36 * nnzb + 4 * nnzb + 4 * (nb + 1) + 2 * 4 * 3 * nb * k
    achieved = spmm_bytes / (spmv_ms / 1e3) / 1e9
    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    # summary of this workload (profiles/hbm_traffic_brick80.json), when it applies
    traffic = None
    try:
        prof = json.load(open(os.path.join(ROOT, "profiles", "hbm_traffic_brick80.json")))
        if a.n == 80:
            traffic = prof["bytes_per_launch"].get(f"axg::bsell_spmm_kernel<{k}, 1, false>")
    except (OSError, ValueError, KeyError):
        traffic = None
    flops_el = 8 * (72 * 24 + 14 * 24 * 24)
    ke32_tflops = flops_el * nelem / (asm["ke32"] / 1e3) / 1e12
    out = {
        "metric": "DOF·iter/s PCG + elements/s assembly, mixed vs FP64 GPU at 1e-8 rel. error",
        "value": mixed_dof_iter_s,
        "unit": "DOF·iter/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner PCG + f64 refinement",
        "data": "synthetic",
        "config": {
            "workload": f"C3D8 brick {a.n}^3, lognormal E, bottom clamped, {k} top-face load "
                        f"cases, mixed-precision multi-RHS PCG to rel. residual {a.tol:g}",
            "elements": nelem, "n_free": n, "nnz": nnz, "nrhs": k, "precond": a.precond,
            "inner_tol": a.inner_tol,
        },
        "mixed": {
            "solve_ms": solve_ms, "inner_iters": inner_iters,
            "sweeps": float(np.mean([s[2]["outer_iters"] for s in steps])),
            "rel_err_vs_fp64_ref": err_mixed,
            "rel_residual": float(max(s[2]["rel_residual"] for s in steps)),
            "time_to_solution_per_rhs_ms": solve_ms / k,
        },
        "fp64": fp64,
        "assembly": {
            "ke64_ms": asm["ke64"], "gather64_ms": asm["gather64"], "ke32_ms": asm["ke32"],
            "gather32_ms": asm["gather32"],
            "elements_per_s_fp64": nelem / ((asm["ke64"] + asm["gather64"]) / 1e3),
            "elements_per_s_fp32": nelem / ((asm["ke32"] + asm["gather32"]) / 1e3),
            "ke32_mfma_tflops": ke32_tflops, "ke32_mfma_frac": ke32_tflops / FP32_PEAK_TFLOPS,
        },
        "speedup_mixed_vs_fp64_time_to_solution": (fp64["solve_ms"] / solve_ms) if fp64 else None,
        "roofline": {
            "bound": "hbm", "kernel": f"bsell_spmm_kernel<K={k}> (FP32 block-sliced-ELL SpMM)",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "bytes_per_launch": spmm_bytes, "avg_launch_ms": spmv_ms,
        },
        "reference_fp64_iters": ref["iters"],
    }
    if not a.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
