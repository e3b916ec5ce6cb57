"""Benchmark of the linear-static hot path on one MI355X.

Workloads (--workload; each prints one JSON line):

brick (default)  C3D8 brick 80x80x80 (512,000 elements, 531,441 nodes,
    1,574,640 free DOFs), lognormal E (median 210000, sigma 0.5, seed 0),
    nu = 0.3, bottom face clamped, four load cases on the top face (1 MPa
    pressure P2, x and y shear of the same total, a seeded random normal
    load). One step = numeric assembly (FP64 CSR for the true residual +
    the fused FP32 integration and block-row assembly of the FP32 BSR-3
    operator) followed by the
    mixed-precision multi-right-hand-side PCG (FP32 inner, FP64 refinement)
    of all load cases to ||f - K u|| <= tol ||f||. After the timed steps two
    FP64 GPU solves of the same load cases to the same tolerance are timed:
    fp64_plain, the drop-in axg_solve_csr on the exported K_ff (the reference's
    linalg::pcg: Jacobi, scalar CSR rows, one load case at a time), and
    fp64_best, the library's fastest FP64 Jacobi PCG (all load cases in one
    pass on the FP64 3x3 node-block operator). All three are checked against
    an FP64 solution converged to 1e-13. Every solve is timed on the solver's
    production path (iterations replayed as hipGraphs, no per-launch events);
    one more instrumented solve of each leg times every SpMM launch with HIP
    events for the roofline and the operator bandwidths.
cantilever  C3D8 cantilever 20x9x9 (2,100 nodes, 6,000 free DOFs), E 210000,
    nu 0.3, x = 0 clamped, -1000 N in z over the tip face; FP64 Jacobi PCG to
    1e-10 against the oracle's Cholesky solution.
c3d20  C3D20 brick 96x24x24 (55,296 elements, 237,025 nodes, 705,600 free
    DOFs), jittered interior corners (+-0.15 h), x = 0 clamped, tip shear;
    FP64 vs FP32+FP64-refinement integration and solve, Jacobi vs 3x3
    block-Jacobi.
batch  48 C3D8 cubes m^3, m ~ U{8..48}, 4 top-face load cases each, in one
    batched call (size groups of graph-replayed PCG, smallest first); p50/p99
    per-solve latency and aggregate throughput, against every mesh in one
    launch sequence and against the meshes solved one after another.
simp  C3D8 brick 120x60x60 (432,000 elements), SIMP modulus, point load;
    compliance, dc/drho through the adjoint kernel, central differences
    (step 1e-4) on 64 random elements, and the FP64 CPU adjoint of the oracle.
freq  *FREQUENCY: the 10 lowest modes of a steel C3D8 brick 48x48x48 (110,592
    elements, 345,744 free DOFs; E 210000, nu 0.3, rho 7.85e-9, bottom face
    clamped) by the reference's subspace iteration (14 vectors, at most 40
    sweeps) over the multi-right-hand-side FP64 block PCG (warm-started from
    the Ritz vectors). One step = assembly, lumped mass and the eigen solve.
nlgeom  *STEP, NLGEOM: a C3D8 cantilever 200x20x20 (80,000 elements, 264,600
    free DOFs; 10 x 1 x 1, E 1000, nu 0.3, x = 0 clamped) under a tip load of
    0.5 in -z (about 20 % of the length in linear theory) in 4 increments: the
    reference's continuum Newton driver with the St. Venant-Kirchhoff tangent
    re-assembled on the device at every iteration and Jacobi PCG solves on the
    FP64 node-block operator.
    One step = one whole analysis.

The default (brick) line also carries `other_configs`: the cantilever, C3D20,
batch and SIMP workloads run once each after a warm-up, without their CPU legs
(--no-configs skips them).

Inputs are resident in HBM when a timed region starts.
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
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
METRIC = "DOF·iter/s PCG + elements/s assembly, mixed vs FP64 GPU at 1e-8 rel. error"
WORKLOADS = ("brick", "cantilever", "c3d20", "batch", "simp", "freq", "nlgeom", "deck",
             "buckle")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """Threads the CPU baseline uses: OMP_NUM_THREADS when set, else the host's cores."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return os.cpu_count() or 1


# --------------------------------------------------------------- CPU baseline
def _oracle_csr(O, p):
    ti, tj, tv = O.assemble_triplets(p.nn, p.xyz, p.conn, p.e, p.nu)
    sel = (p.ind_of[ti] >= 0) & (p.ind_of[tj] >= 0)
    return O.csr_from_triplets(p.n_free, p.ind_of[ti[sel]], p.ind_of[tj[sel]], tv[sel])


def _oracle_pcg_rate(O, csr, b, iters, threads, min_s=1.0):
    """DOF.iter/s of the oracle's linalg::pcg restatement: runs of a bounded iteration count
    (short of convergence), repeated until min_s seconds have passed."""
    import numpy as np
    L = O.lib()
    x = np.zeros(csr.n)
    it = np.zeros(1, np.int64)
    res = np.zeros(1)
    L.or_set_threads(threads)
    done = 0
    t0 = time.perf_counter()
    while True:
        st = L.or_pcg_omp(csr.n, O._l(csr.indptr), O._l(csr.indices), O._d(csr.data), O._d(b),
                          1e-30, 10 ** 9, iters, O._d(x), O._l(it), O._d(res))
        if st != 0:
            raise RuntimeError(f"oracle PCG sample failed (status {st}); lower the iteration count")
        done += int(it[0])
        dt = time.perf_counter() - t0
        if dt >= min_s:
            break
    L.or_set_threads(0)
    return csr.n * done / dt


def host_cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_brick(indptr, indices, data, b, prob, layers=8, tol=1e-13, iters_1=20):
    """The oracle (C restatement of the reference's CPU path) on this host, on the benchmarked
    brick itself: linalg::pcg on the exported K_ff (bit-equal to the oracle's own assembly of
    it) solving load case 0 to ||r|| <= tol ||b|| on all host threads (the rate and the time to
    solution; the solution is the oracle's answer the GPU results are checked against), the
    same PCG single-threaded for a bounded number of iterations, and the reference's
    single-threaded element loop + triplets + csr_from_triplets on the bottom `layers` element
    layers of the same brick (same elements, same moduli). Returns (summary, oracle x)."""
    import numpy as np
    from axia_gpu import meshgen as MG
    from oracle import oracle as O

    cores = host_threads()
    n = len(indptr) - 1
    csr = O.Csr(n, indptr, indices, data)
    L = O.lib()
    xo = np.zeros(n)
    it = np.zeros(1, np.int64)
    res = np.zeros(1)
    L.or_set_threads(cores)
    t0 = time.perf_counter()
    st = L.or_pcg_omp(n, O._l(csr.indptr), O._l(csr.indices), O._d(csr.data), O._d(b), tol,
                      max(4 * n, 200), 0, O._d(xo), O._l(it), O._d(res))
    t_solve = time.perf_counter() - t0
    L.or_set_threads(0)
    if st != 0:
        raise RuntimeError(f"oracle PCG failed (status {st})")
    rc = n * int(it[0]) / t_solve
    r1 = _oracle_pcg_rate(O, csr, b, iters_1, 1, min_s=3.0)
    nx = round(prob.nelem ** (1.0 / 3.0))
    xyz, conn = MG.brick_hex8(nx, nx, layers)
    ne = len(conn)
    q = MG.Problem("slab", 8, xyz, conn, prob.e[:ne], prob.nu[:ne],
                   MG.free_map(len(xyz), [i for i in range(len(xyz)) if xyz[i, 2] == 0.0]))
    t0 = time.perf_counter()
    _oracle_csr(O, q)
    asm_eps = ne / (time.perf_counter() - t0)
    return {
        "value": rc, "unit": "DOF·iter/s", "cores": cores, "kind": "port",
        "sample": (f"oracle linalg::pcg restatement (Jacobi, FP64 CSR, OpenMP) on the "
                   f"benchmarked K_ff itself (n={n}, nnz={csr.nnz}), load case 0 solved to "
                   f"{tol:g}; single-thread {iters_1}-iteration runs; assembly on the brick's "
                   f"bottom {layers} element layers ({ne} elements), single-thread"),
        "cpu_model": host_cpu_model(),
        "oracle_solve_s": t_solve, "oracle_iters": int(it[0]),
        "oracle_rel_residual": float(res[0]) / float(np.linalg.norm(b)),
        "pcg_dof_iter_per_s_1thread": r1,
        "assembly_elements_per_s_1thread": asm_eps,
    }, xo


# ------------------------------------------------------------------ helpers
def base_line(a, value, ms_per_step, dtype, config, unit="DOF·iter/s"):
    return {
        "metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "config": config,
    }


def spmm_bytes(nb, nnzb, k):
    """Algorithmic bytes of one FP32 BSR-3 SpMM launch: 36 B values + 4 B column per block,
    the row pointer, and x read + y written once (12 B per block row and column)."""
    return 36 * nnzb + 4 * nnzb + 4 * (nb + 1) + 2 * 4 * 3 * nb * k


def sell64_bytes(n, nnz, k=1):
    """Algorithmic bytes of one FP64 SpMV: 8 B value + 4 B column per entry, x read and y
    written once per row and right-hand side."""
    return 12 * nnz + 16 * n * k


def bsell64_bytes(nb, nnzb, k):
    """Algorithmic bytes of one FP64 block SpMM launch: 72 B values + 4 B column per 3x3
    block, x read and y written once per block row and right-hand side (24 B each)."""
    return 76 * nnzb + 2 * 8 * 3 * nb * k


def spmm_kernel_name(k, shape):
    """The kernel behind an FP32 SpMM shape of the autotuner (axg_mixed.hip's shape table),
    as rocprofv3 names it (the solve's launches carry the p.Ap partials: DOT = true)."""
    if shape >= 13:
        return f"axg::bsell_spmm_lds_kernel<{k}, {shape - 12}, true>"
    u = {0: 1, 1: 2, 2: 3, 3: 4, 4: 1, 5: 2, 6: 3, 7: 4, 8: 1, 9: 2, 10: 3, 11: 6, 12: 8}[shape]
    return f"axg::bsell_spmm_kernel<{k}, {u}, true, {str(4 <= shape <= 7).lower()}, " \
           f"{str(8 <= shape <= 10).lower()}>"


def traffic_for(kernel_key, workload_key):
    try:
        prof = json.load(open(os.path.join(ROOT, "profiles", "hbm_traffic_brick80.json")))
        if prof.get("workload_key") not in (None, workload_key):
            return None
        return prof["bytes_per_launch"].get(kernel_key)
    except (OSError, ValueError, KeyError):
        return None


def roofline(kernel, nbytes, launch_ms, traffic=None):
    achieved = nbytes / (launch_ms / 1e3) / 1e9 if launch_ms > 0 else None
    return {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic, "bytes_per_launch": nbytes, "avg_launch_ms": launch_ms}


def avg_launch(info):
    return info["spmv_ms"] / max(1, info["spmv_launches"])


def asm32_forms(a, mesh, nelem, reps=None):
    """Device time of both FP32 assembly forms on a mesh (the form in use is restored):
    0 integration fused with the block-row assembly, 1 element blocks on MFMA + gather."""
    import numpy as np
    keep = mesh.asm32_mode()
    out = {}
    for mode in (0, 1):
        mesh.asm32_mode(mode)
        mesh.assemble("mixed")
        ts = []
        for _ in range(reps or max(a.reps, 3)):
            mesh.assemble("mixed")
            ts.append(mesh.assembly_times())
        ke = float(np.mean([t["ke32"] for t in ts]))
        ga = float(np.mean([t["gather32"] for t in ts]))
        out[("fused_rows", "mfma_blocks")[mode]] = {
            "element_ms": ke, "gather_ms": ga, "elements_per_s": nelem / ((ke + ga) / 1e3)}
        log(f"FP32 assembly form {mode}: {ke:.3f} + {ga:.3f} ms")
    mesh.asm32_mode(keep)
    mesh.assemble("mixed")
    return out


# ----------------------------------------------------------- brick (default)
def kernels_only(a, G, mesh, prob, n, nnz):
    """Repeated launches of the dominant kernels (input for rocprofv3 runs)."""
    import numpy as np
    out = {"workload": prob.name, "n_free": n, "nnz": nnz, "reps": a.reps}
    nstream = 1 << 27  # 1 GiB per array
    out["stream_copy_gbs"] = G.bench_stream(nstream, a.reps, ctx=mesh.ctx)
    log(f"stream copy {out['stream_copy_gbs']:.0f} GB/s")
    times = []
    for _ in range(a.reps):
        mesh.assemble("mixed")
        times.append(mesh.assembly_times())
    out["assembly_ms"] = {k: float(np.mean([t[k] for t in times])) for k in times[0]}
    log("assembly (ms): " + ", ".join(f"{k} {v:.3f}" for k, v in out["assembly_ms"].items()))
    nb, nnzb = n // 3, nnz // 9
    for which, name, nbytes in ((0, "fp64_sell_spmv", sell64_bytes(n, nnz)),
                                (1, "fp64_csr_spmv", sell64_bytes(n, nnz) + 4 * (n + 1))):
        ms = mesh.bench_operator(which, 1, a.reps)
        out[name] = {"ms": ms, "gbs": nbytes / (ms / 1e3) / 1e9}
        log(f"{name}: {ms:.4f} ms {out[name]['gbs']:.0f} GB/s")
    for k in (1, 4, 8):
        ms = mesh.bench_operator(2, k, a.reps)
        out[f"fp32_bsell_spmm_k{k}"] = {"ms": ms, "gbs": spmm_bytes(nb, nnzb, k) / (ms / 1e3) / 1e9}
        log(f"fp32 spmm k={k}: {ms:.4f} ms {out[f'fp32_bsell_spmm_k{k}']['gbs']:.0f} GB/s")
    print(json.dumps(out))


def run_brick(a, G, MG, torch, np):
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    prob = MG.brick_lognormal(a.n, seed=0, nloads=a.nrhs)
    ctx = G.Context(0)
    mesh = G.Mesh(8, prob.xyz, prob.conn, prob.e, prob.nu, prob.ind_of, ctx=ctx)
    mesh.asm32_mode(a.asm32_mode)
    info = mesh.info()
    n, nnz = info["n_free"], info["nnz"]
    log(f"mesh {prob.name}: n={n} nnz={nnz} contributions={info['ncontrib']} "
        f"symbolic {info['symbolic_s']:.2f}s setup {time.perf_counter() - t0:.1f}s")
    if a.kernels:
        kernels_only(a, G, mesh, prob, n, nnz)
        return None
    B = torch.tensor(prob.rhs_all(), dtype=torch.float64, device=dev).contiguous()
    X = torch.zeros_like(B)
    Xref = torch.zeros_like(B)
    k = B.shape[0]

    mesh.assemble("mixed")
    ref = mesh.solve_device(B.data_ptr(), k, Xref.data_ptr(), precision="fp64",
                            precond="block3", tol=1e-13)
    log(f"reference FP64 PCG: iters {ref['iters']} rel.res {ref['rel_residual']:.2e} "
        f"{ref['solve_ms'] / 1e3:.2f}s")

    def step(timing):
        ms_asm = mesh.assemble("mixed")
        t = mesh.assembly_times()
        inf = mesh.solve_device(B.data_ptr(), k, X.data_ptr(), precision="mixed",
                                precond=a.precond, tol=a.tol, inner_tol=a.inner_tol,
                                timing=timing, refinement=a.refinement)
        return ms_asm, t, inf

    for _ in range(a.warmup):
        step(False)
        log("warmup step done")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    steps = []
    # the timed steps run the solver's production path (hipGraph replay, no per-launch
    # events); one instrumented step afterwards times every SpMM launch for the roofline
    for s_ in range(a.steps):
        steps.append(step(False))
        log(f"step {s_}: solve {steps[-1][2]['solve_ms']:.1f} ms inner {steps[-1][2]['iters']} "
            f"sweeps {steps[-1][2]['outer_iters']}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = 1e3 * wall / a.steps
    inst = step(True)
    log(f"instrumented step: solve {inst[2]['solve_ms']:.1f} ms")

    def rel_err(Xa):
        d = (Xa - Xref).double()
        return float((d.norm(dim=1) / Xref.norm(dim=1)).max().item())

    err_mixed = rel_err(X)
    inner_iters = float(np.mean([s[2]["iters"] for s in steps]))
    solve_ms = float(np.mean([s[2]["solve_ms"] for s in steps]))
    # each inner iteration advances all k right-hand sides by one CG step over n DOFs
    # (inner_iters: per sweep the largest count over right-hand sides, summed)
    mixed_dof_iter_s = n * inner_iters * k / (solve_ms / 1e3)
    spmv_ms = avg_launch(inst[2])
    asm = {kk: float(np.mean([s[1][kk] for s in steps])) for kk in steps[0][1]}
    nelem = prob.nelem
    forms = asm32_forms(a, mesh, nelem)

    fp64 = plain = None
    indptr = indices = data = None
    if not a.no_fp64:
        X64 = torch.zeros_like(B)
        mesh.assemble("fp64")
        # warm (tuners, graphs), then timed on the production path, then instrumented
        mesh.solve_device(B.data_ptr(), k, X64.data_ptr(), precision="fp64", tol=a.tol)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        r64 = mesh.solve_device(B.data_ptr(), k, X64.data_ptr(), precision="fp64", tol=a.tol)
        torch.cuda.synchronize()
        wall64 = time.perf_counter() - t1
        r64i = mesh.solve_device(B.data_ptr(), k, X64.data_ptr(), precision="fp64", tol=a.tol,
                                 timing=True)
        fp64 = {
            "solver": r64["name"],
            "iters_total": r64["iters"], "solve_ms": r64["solve_ms"], "wall_ms": 1e3 * wall64,
            "solve_ms_instrumented": r64i["solve_ms"],
            "dof_iter_per_s": n * r64["iters"] / (r64["solve_ms"] / 1e3),
            "spmv_ms_avg": avg_launch(r64i),
            "rel_err_vs_ref": rel_err(X64), "rel_residual": r64["rel_residual"],
        }
        # the node-block operator takes all k right-hand sides in one pass
        op_bytes = (bsell64_bytes(n // 3, nnz // 9, k) if "node blocks" in r64["name"]
                    else sell64_bytes(n, nnz))
        fp64["spmv_gbs"] = op_bytes / (fp64["spmv_ms_avg"] / 1e3) / 1e9
        fp64["spmv_frac"] = fp64["spmv_gbs"] / HBM_PEAK_GBS
        log(f"fp64 best: {fp64['solve_ms']:.0f} ms, iters {r64['iters']}, "
            f"err {fp64['rel_err_vs_ref']:.2e}")
        # the plain FP64 GPU PCG: the drop-in on the exported K_ff, one load case at a time
        indptr, indices, data = mesh.csr()
        Bh = B.cpu().numpy()
        Xp = np.zeros_like(Bh)
        pl = []
        for j in range(k):
            Xp[j], inf = G.solve_csr(indptr, indices, data, Bh[j], ctx=ctx, tol=a.tol)
            pl.append(inf)
        # load case 0 again with every SpMV launch timed
        _, pli = G.solve_csr(indptr, indices, data, Bh[0], ctx=ctx, tol=a.tol, timing=True)
        Xpt = torch.tensor(Xp, device=dev)
        plain = {
            "solver": pl[0]["name"], "calls": k,
            "iters_total": int(sum(r["iters"] for r in pl)),
            "solve_ms": float(sum(r["solve_ms"] for r in pl)),
            "solve_ms_instrumented_case0": pli["solve_ms"],
            "spmv_ms_avg": avg_launch(pli),
            "rel_err_vs_ref": rel_err(Xpt),
            "rel_residual": float(max(r["rel_residual"] for r in pl)),
        }
        plain["dof_iter_per_s"] = n * plain["iters_total"] / (plain["solve_ms"] / 1e3)
        plain["spmv_gbs"] = sell64_bytes(n, nnz) / (plain["spmv_ms_avg"] / 1e3) / 1e9
        plain["spmv_frac"] = plain["spmv_gbs"] / HBM_PEAK_GBS
        log(f"fp64 plain: {plain['solve_ms']:.0f} ms, iters {plain['iters_total']}, "
            f"err {plain['rel_err_vs_ref']:.2e}")
        mesh.assemble("mixed")

    nb, nnzb = n // 3, nnz // 9  # every free node keeps all 3 DOFs in this workload
    # the reference's dense element count (B^T D B per Gauss point)
    flops_el = 8 * (72 * 24 + 14 * 24 * 24)
    asm32_ms = asm["ke32"] + asm["gather32"]
    asm32_tflops = flops_el * nelem / (asm32_ms / 1e3) / 1e12
    asm64_tflops = flops_el * nelem / ((asm["ke64"] + asm["gather64"]) / 1e3) / 1e12
    # FP32 assembly against HBM: the operator it must write (three value planes and the
    # LDS-staged SpMM's third plane, 48 B per block) and the mesh it must read
    asm32_bytes = 48 * nnzb + 24 * prob.nnode + (4 * 8 + 16) * nelem
    out = base_line(a, mixed_dof_iter_s, ms_per_step, "f32 inner PCG + f64 refinement", {
        "workload": f"C3D8 brick {a.n}^3, lognormal E, bottom clamped, {k} top-face load "
                    f"cases, mixed-precision multi-RHS PCG to rel. residual {a.tol:g}",
        "elements": nelem, "n_free": n, "nnz": nnz, "nrhs": k, "precond": a.precond,
        "inner_tol": a.inner_tol, "refinement": a.refinement,
    })
    out.update({
        "mixed": {
            "solve_ms": solve_ms, "inner_iters": inner_iters,
            "sweeps": float(np.mean([s[2]["outer_iters"] for s in steps])),
            "rel_err_vs_fp64_ref": err_mixed,
            "rel_residual": float(max(s[2]["rel_residual"] for s in steps)),
            "time_to_solution_per_rhs_ms": solve_ms / k,
            "solve_ms_instrumented": inst[2]["solve_ms"],
        },
        "fp64_best": fp64,
        "fp64_plain": plain,
        "assembly": {
            "ke64_ms": asm["ke64"], "gather64_ms": asm["gather64"], "ke32_ms": asm["ke32"],
            "gather32_ms": asm["gather32"],
            "elements_per_s_fp64": nelem / ((asm["ke64"] + asm["gather64"]) / 1e3),
            "elements_per_s_fp32": nelem / (asm32_ms / 1e3),
            "assembly_fp32_bytes": asm32_bytes,
            "assembly_fp32_hbm_frac": asm32_bytes / (asm32_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
            "assembly_fp32_roofline_frac": asm32_tflops / FP32_PEAK_TFLOPS,
            "assembly_fp64_roofline_frac": asm64_tflops / FP64_PEAK_TFLOPS,
            "fp32_forms": forms,
        },
        "speedup_mixed_vs_fp64_plain": (plain["solve_ms"] / solve_ms) if plain else None,
        "speedup_mixed_vs_fp64_best": (fp64["solve_ms"] / solve_ms) if fp64 else None,
        "roofline": roofline(spmm_kernel_name(k, mesh.spmm_shape(k)) +
                             " (FP32 block-sliced-ELL SpMM)", spmm_bytes(nb, nnzb, k), spmv_ms,
                             traffic_for(spmm_kernel_name(k, mesh.spmm_shape(k)), f"brick{a.n}")),
        "reference_fp64_iters": ref["iters"],
    })
    if not a.no_cpu:
        log("cpu baseline ...")
        if indptr is None:
            indptr, indices, data = mesh.csr()
        out["cpu_baseline"], xo = cpu_baseline_brick(indptr, indices, data, prob.rhs(0), prob)

        def err_vs_oracle(v):
            return float(np.linalg.norm(v - xo) / np.linalg.norm(xo))

        # load case 0 of every GPU solution against the oracle's own solve of it
        out["mixed"]["rel_err_vs_oracle_case0"] = err_vs_oracle(X[0].cpu().numpy())
        if plain:
            out["fp64_plain"]["rel_err_vs_oracle_case0"] = err_vs_oracle(Xp[0])
            out["fp64_best"]["rel_err_vs_oracle_case0"] = err_vs_oracle(X64[0].cpu().numpy())
        out["cpu_baseline"]["time_to_solution_vs_mixed_per_rhs"] = (
            1e3 * out["cpu_baseline"]["oracle_solve_s"] / out["mixed"]["time_to_solution_per_rhs_ms"])
        log(f"oracle: {out['cpu_baseline']['oracle_iters']} iterations "
            f"{out['cpu_baseline']['oracle_solve_s']:.1f} s; mixed vs oracle "
            f"{out['mixed']['rel_err_vs_oracle_case0']:.2e}")
    return out


# --------------------------------------------------------------- cantilever
def run_cantilever(a, G, MG, torch, np):
    prob = MG.cantilever(20, 9, 9)
    mesh = G.Mesh(8, prob.xyz, prob.conn, prob.e, prob.nu, prob.ind_of)
    n, nnz = prob.n_free, mesh.info()["nnz"]
    b = torch.tensor(prob.rhs(), dtype=torch.float64, device="cuda:0")
    x = torch.zeros_like(b)
    tol = 1e-10

    def step(timing):
        ms = mesh.assemble("fp64")
        return ms, mesh.solve_device(b.data_ptr(), 1, x.data_ptr(), precision="fp64",
                                     precond="jacobi", tol=tol, timing=timing)

    for _ in range(a.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = [step(False) for _ in range(a.steps)]
    torch.cuda.synchronize()
    ms_per_step = 1e3 * (time.perf_counter() - t0) / a.steps
    iters = res[-1][1]["iters"]
    solve_ms = float(np.mean([r[1]["solve_ms"] for r in res]))
    asm_ms = float(np.mean([r[0] for r in res]))
    out = base_line(a, n * iters / (solve_ms / 1e3), ms_per_step, "f64", {
        "workload": "C3D8 cantilever 20x9x9 (2,100 nodes, 6,000 free DOFs), E 210000, nu 0.3, "
                    "x=0 clamped, -1000 N z on the tip face; FP64 Jacobi PCG to 1e-10",
        "elements": prob.nelem, "n_free": n, "nnz": nnz, "nrhs": 1, "precond": "jacobi"})
    spmv = avg_launch(step(True)[1])  # one instrumented solve: every SpMV launch timed
    out.update({
        "fp64": {"iters": iters, "solve_ms": solve_ms, "assembly_ms": asm_ms,
                 "elements_per_s": prob.nelem / (asm_ms / 1e3),
                 "rel_residual": res[-1][1]["rel_residual"]},
        "roofline": (roofline("bsell64_spmm_kernel<K=1> (FP64 node-block sliced-ELL SpMV)",
                              bsell64_bytes(n // 3, nnz // 9, 1), spmv)
                     if "node blocks" in res[-1][1]["name"] else
                     roofline("sell_spmv_dot_kernel (FP64 sliced-ELL SpMV)", sell64_bytes(n, nnz),
                              spmv)),
    })
    if not a.no_cpu:
        from oracle import oracle as O
        t1 = time.perf_counter()
        csr = _oracle_csr(O, prob)
        xc, name, _, _ = O.solve_kff(n, csr, prob.rhs(), "lapack")
        t_chol = time.perf_counter() - t1
        xg = x.cpu().numpy()
        cores = host_threads()
        out["cpu_baseline"] = {
            "value": _oracle_pcg_rate(O, csr, prob.rhs(), 100, cores), "unit": "DOF·iter/s",
            "cores": cores, "kind": "port",
            "sample": "oracle linalg::pcg restatement, runs of 100 Jacobi iterations on this problem; "
                      "oracle assembly + dense LAPACK Cholesky (faer stand-in) timed beside it",
            "cholesky_assembly_and_solve_s": t_chol,
            "gpu_rel_err_vs_cholesky": float(np.linalg.norm(xg - xc) / np.linalg.norm(xc)),
        }
    return out


# -------------------------------------------------------------------- C3D20
def run_c3d20(a, G, MG, torch, np):
    t0 = time.perf_counter()
    prob = MG.brick_c3d20_jitter(96, 24, 24, seed=0)
    mesh = G.Mesh(20, prob.xyz, prob.conn, prob.e, prob.nu, prob.ind_of)
    mesh.asm32_mode(a.asm32_mode)
    info = mesh.info()
    n, nnz = info["n_free"], info["nnz"]
    log(f"mesh {prob.name}: n={n} nnz={nnz} setup {time.perf_counter() - t0:.1f}s")
    forms = asm32_forms(a, mesh, prob.nelem)
    if a.kernels:
        print(json.dumps({"workload": prob.name, "n_free": n, "nnz": nnz, "reps": a.reps,
                          "assembly_fp32_forms": forms}))
        return None
    b = torch.tensor(prob.rhs(), dtype=torch.float64, device="cuda:0")
    x = torch.zeros_like(b)
    xref = torch.zeros_like(b)
    # the FP64 residual of this slender beam bottoms out near 1e-10 relative (round-off
    # of f - K u), so the solves stop at 1e-9; displacements are compared at 1e-8.
    # Its first inner solve gains only ~3x whatever the inner tolerance (the FP32
    # operator's rounding, amplified by the beam's condition number, limits it), so
    # each inner solve stops at 1e-2 (a sweep of 1e-1 ... 1e-4 measured this fastest)
    tol = max(a.tol, 1e-9)
    inner_tol = 1e-2
    mesh.assemble("mixed")
    ref = mesh.solve_device(b.data_ptr(), 1, xref.data_ptr(), precision="fp64", precond="block3",
                            tol=1e-13)
    log(f"reference FP64 block-Jacobi: {ref['iters']} iterations {ref['solve_ms']:.0f} ms")

    def err(v):
        return float((v - xref).norm().item() / xref.norm().item())

    def step(timing, precond="block3"):
        ms = mesh.assemble("mixed")
        inf = mesh.solve_device(b.data_ptr(), 1, x.data_ptr(), precision="mixed", precond=precond,
                                tol=tol, inner_tol=inner_tol, timing=timing)
        return ms, mesh.assembly_times(), inf

    for _ in range(a.warmup):
        step(False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # timed on the production path (graph replay); one instrumented step for the SpMM
    res = [step(False) for _ in range(a.steps)]
    torch.cuda.synchronize()
    ms_per_step = 1e3 * (time.perf_counter() - t1) / a.steps
    e_mixed = err(x)
    solve_ms = float(np.mean([r[2]["solve_ms"] for r in res]))
    inner = float(np.mean([r[2]["iters"] for r in res]))
    asm = {kk: float(np.mean([r[1][kk] for r in res])) for kk in res[0][1]}
    spmm = avg_launch(step(True)[2])
    variants = {"mixed_block3": {"solve_ms": solve_ms, "iters": inner,
                                 "sweeps": res[-1][2]["outer_iters"], "rel_err": e_mixed}}
    _, _, rj = step(False, "jacobi")
    variants["mixed_jacobi"] = {"solve_ms": rj["solve_ms"], "iters": rj["iters"],
                                "sweeps": rj["outer_iters"], "rel_err": err(x)}
    mesh.assemble("fp64")
    for pc in ("block3", "jacobi"):
        mesh.solve_device(b.data_ptr(), 1, x.data_ptr(), precision="fp64", precond=pc, tol=tol)
        r = mesh.solve_device(b.data_ptr(), 1, x.data_ptr(), precision="fp64", precond=pc, tol=tol)
        variants[f"fp64_{pc}"] = {"solve_ms": r["solve_ms"], "iters": r["iters"], "rel_err": err(x)}
        if pc == "block3":
            ri = mesh.solve_device(b.data_ptr(), 1, x.data_ptr(), precision="fp64", precond=pc,
                                   tol=tol, timing=True)
            variants[f"fp64_{pc}"]["spmv_ms_avg"] = avg_launch(ri)
        log(f"fp64 {pc}: {r['iters']} iterations {r['solve_ms']:.0f} ms")
    v64 = variants["fp64_block3"]
    # C3D20's nodes all keep 3 DOFs here (the clamped face is fully fixed): n / 3 block rows
    v64["spmv_gbs"] = bsell64_bytes(n // 3, nnz // 9, 1) / (v64["spmv_ms_avg"] / 1e3) / 1e9
    ne = prob.nelem
    out = base_line(a, n * inner / (solve_ms / 1e3), ms_per_step, "f32 inner PCG + f64 refinement", {
        "workload": "C3D20 brick 96x24x24, jittered interior corners +-0.15 h (seed 0), x=0 "
                    f"clamped, tip shear 1 N/node; mixed block-Jacobi PCG to {tol:g}",
        "elements": ne, "n_free": n, "nnz": nnz, "nrhs": 1, "precond": "block3",
        "inner_tol": inner_tol})
    out.update({
        "variants": variants,
        "speedup_mixed_vs_fp64_block3": v64["solve_ms"] / solve_ms,
        "assembly": {
            "ke64_ms": asm["ke64"], "gather64_ms": asm["gather64"], "ke32_ms": asm["ke32"],
            "gather32_ms": asm["gather32"],
            "elements_per_s_fp64": ne / ((asm["ke64"] + asm["gather64"]) / 1e3),
            "elements_per_s_fp32": ne / ((asm["ke32"] + asm["gather32"]) / 1e3),
            "fp32_forms": forms,
        },
        "roofline": roofline("bsell_spmm_kernel<K=1> (FP32 block-sliced-ELL SpMV)",
                             spmm_bytes(n // 3, nnz // 9, 1), spmm),
    })
    if not a.no_cpu:
        from oracle import oracle as O
        cores = host_threads()
        q = MG.brick_c3d20_jitter(12, 6, 6, seed=1)
        t2 = time.perf_counter()
        csr = _oracle_csr(O, q)
        asm_eps = q.nelem / (time.perf_counter() - t2)
        out["cpu_baseline"] = {
            "value": _oracle_pcg_rate(O, csr, q.rhs(), 200, cores), "unit": "DOF·iter/s",
            "cores": cores, "kind": "port",
            "sample": f"oracle assembly (single thread) and linalg::pcg (200 Jacobi iterations) "
                      f"on a jittered C3D20 12x6x6 brick, n={q.n_free}, nnz={csr.nnz}",
            "assembly_elements_per_s_1thread": asm_eps,
        }
    return out


# -------------------------------------------------------------------- batch
def run_batch(a, G, MG, torch, np):
    t0 = time.perf_counter()
    probs = MG.batch_meshes(count=48, lo=8, hi=48, nloads=4, seed=0)
    ctx = G.Context(0)
    meshes = []
    for p in probs:
        m = G.Mesh(8, p.xyz, p.conn, p.e, p.nu, p.ind_of, ctx=ctx)
        m.assemble("fp64")
        meshes.append(m)
    rhs = [p.rhs_all() for p in probs]
    ns = np.array([p.n_free for p in probs])
    nnzs = np.array([m.info()["nnz"] for m in meshes])
    log(f"48 meshes: {ns.sum()} DOFs, {nnzs.sum()} nonzeros, setup {time.perf_counter() - t0:.1f}s")
    tol = 1e-10
    refs = [m.solve(b, precision="fp64", precond="block3", tol=1e-13)[0] for m, b in
            zip(meshes, rhs)]
    for _ in range(a.warmup):
        G.solve_batch(meshes, rhs, tol=tol, groups=a.batch_groups)
    t1 = time.perf_counter()
    res = [G.solve_batch(meshes, rhs, tol=tol, groups=a.batch_groups) for _ in range(a.steps)]
    ms_per_step = 1e3 * (time.perf_counter() - t1) / a.steps
    xs, iters, lat, info = res[-1]
    # every mesh in one launch sequence (lockstep), for comparison and for the roofline of
    # the SpMM over the whole concatenated operator
    ls = [G.solve_batch(meshes, rhs, tol=tol, groups=1) for _ in range(max(1, a.steps // 2))]
    ls_lat = np.concatenate([r[2] for r in ls])
    ls_info = ls[-1][3]
    lockstep = {"solve_ms": float(np.mean([r[3]["solve_ms"] for r in ls])),
                "latency_p50_ms": float(np.percentile(ls_lat, 50)),
                "latency_p99_ms": float(np.percentile(ls_lat, 99))}
    log(f"batch: {info['name']}: {np.mean([r[3]['solve_ms'] for r in res]):.0f} ms, p50 "
        f"{np.percentile(np.concatenate([r[2] for r in res]), 50):.0f} ms; lockstep "
        f"{lockstep['solve_ms']:.0f} ms, p50 {lockstep['latency_p50_ms']:.0f} ms")
    err = max(float(np.linalg.norm(X[k] - R[k]) / np.linalg.norm(R[k]))
              for X, R in zip(xs, refs) for k in range(X.shape[0]))
    dof_iters = float(sum(ns[m] * iters[4 * m + k] for m in range(48) for k in range(4)))
    solve_ms = float(np.mean([r[3]["solve_ms"] for r in res]))
    lat_all = np.concatenate([r[2] for r in res])
    # the same solves one mesh at a time (each mesh's 4 load cases in one call)
    t2 = time.perf_counter()
    seq_ms = 0.0
    for m, b in zip(meshes, rhs):
        _, inf = m.solve(b, precision="fp64", tol=tol)
        seq_ms += inf["solve_ms"]
    seq_wall = 1e3 * (time.perf_counter() - t2)
    # concatenated node-block operator, 4 right-hand sides per row (every free node of these
    # bottom-clamped cubes keeps its 3 DOFs: n / 3 block rows, nnz / 9 blocks)
    nbytes = int(sum(bsell64_bytes(n // 3, z // 9, 4) for n, z in zip(ns, nnzs)))
    out = base_line(a, dof_iters / (solve_ms / 1e3), ms_per_step, "f64", {
        "workload": "48 C3D8 cubes m^3, m ~ U{8..48} (seed 0), E ~ U(5e4, 2.5e5), 4 Gaussian "
                    f"top-face load cases of 1 kN each, one batched Jacobi PCG to {tol:g}",
        "meshes": 48, "solves": 192, "n_free_total": int(ns.sum()), "nnz_total": int(nnzs.sum()),
        "precond": "jacobi"})
    out.update({
        "batch": {"solve_ms": solve_ms, "solves_per_s": 192 / (solve_ms / 1e3),
                  "latency_p50_ms": float(np.percentile(lat_all, 50)),
                  "latency_p99_ms": float(np.percentile(lat_all, 99)),
                  "latency_max_ms": float(lat_all.max()),
                  "iters_min": int(iters.min()), "iters_max": int(iters.max()),
                  "rel_err_vs_fp64_ref_max": err, "rel_residual_max": info["rel_residual"]},
        "solver": info["name"],
        "lockstep": lockstep,
        "sequential": {"solve_ms_sum": seq_ms, "wall_ms": seq_wall,
                       "speedup_batch_vs_sequential": seq_ms / solve_ms},
        "roofline": roofline("bsell64_spmm_kernel<K=4> (FP64 concatenated node-block SpMM, "
                             "every mesh in one launch)", nbytes, avg_launch(ls_info)),
    })
    if not a.no_cpu:
        from oracle import oracle as O
        cores = host_threads()
        i = int(np.argsort(ns)[len(ns) // 2])
        csr = _oracle_csr(O, probs[i])
        out["cpu_baseline"] = {
            "value": _oracle_pcg_rate(O, csr, rhs[i][0], 200, cores), "unit": "DOF·iter/s",
            "cores": cores, "kind": "port",
            "sample": f"oracle linalg::pcg restatement, 200 Jacobi iterations on the median mesh "
                      f"of the batch (n={csr.n}, nnz={csr.nnz})",
        }
    return out


# --------------------------------------------------------------------- freq
def run_freq(a, G, MG, torch, np):
    import scipy.sparse as sp
    N, nm, rho, tol = a.freq_n, a.nmodes, 7.85e-9, a.eig_tol
    xyz, conn = MG.brick_hex8(N, N, N)
    ind = MG.free_map(len(xyz), np.nonzero(xyz[:, 2] == 0.0)[0])
    ne = len(conn)
    p = MG.Problem(f"brick_c3d8_{N}x{N}x{N}_freq", 8, xyz, conn, np.full(ne, 210000.0),
                   np.full(ne, 0.3), ind)
    ctx = G.Context(0)
    m = G.Mesh(8, xyz, conn, p.e, p.nu, ind, ctx=ctx)
    m.assemble("fp64")
    n, nnz = m.n_free, m.info()["nnz"]
    log(f"freq: n={n} nnz={nnz}")
    if a.kernels:
        out = {"workload": p.name, "n_free": n, "nnz": nnz, "reps": a.reps}
        for k in (1, 2, 4, 8):
            ms = m.bench_operator(3, k, a.reps)
            out[f"fp64_block_spmm_k{k}"] = {
                "ms": ms, "gbs": bsell64_bytes(n // 3, nnz // 9, k) / (ms / 1e3) / 1e9}
            log(f"fp64 block spmm k={k}: {ms:.4f} ms")
        print(json.dumps(out))
        return None

    def step(precision="fp64", vectors=True):
        m.assemble("mixed" if precision == "mixed" else "fp64")
        md = m.lumped_mass(rho)[ind >= 0]
        return md, m.eigen(md, nm, precision=precision, tol=tol, vectors=vectors)

    for _ in range(a.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        md, (vals, vecs, sweeps, info) = step()
    ms_per_step = 1e3 * (time.perf_counter() - t0) / a.steps
    freqs = np.sqrt(vals) / (2 * np.pi)
    log(f"freq: {ms_per_step:.0f} ms/step, sweeps {sweeps}, {info['iters']} column iterations, "
        f"f = {freqs[:4]}")
    # size-independent checks: M-orthonormal modes solving K phi = lambda M phi
    ip, ix, dv = m.csr()
    K = sp.csr_matrix((dv, ix, ip), shape=(n, n))
    gram = (vecs * md) @ vecs.T
    orth = float(np.abs(gram - np.eye(len(vals))).max())
    resid = [float(np.linalg.norm(K @ v - lam * md * v) / np.linalg.norm(K @ v))
             for lam, v in zip(vals, vecs)]
    # the same analysis with the FP32-operator solver
    _, (vmx, _, sw_mx, info_mx) = step("mixed", vectors=False)
    mixed_dev = float(np.abs(vmx / vals - 1.0).max())
    value = n * info["iters"] / (info["solve_ms"] / 1e3)
    launch_ms = m.bench_operator(3, 8, a.reps)
    nb, nnzb = n // 3, nnz // 9  # every node of this brick has 3 free DOFs or none
    out = base_line(a, value, ms_per_step, "f64", {
        "workload": f"*FREQUENCY, {nm} lowest modes of a steel C3D8 brick {N}^3 (n={n}), "
                    f"bottom clamped, lumped mass, subspace iteration over the FP64 block PCG "
                    f"to {tol:g}",
        "n_free": n, "nnz": nnz, "nmodes": nm, "precond": "jacobi"})
    out.update({
        "eigen": {"time_to_solution_ms": ms_per_step, "sweeps": sweeps,
                  "column_iterations": info["iters"], "solve_ms": info["solve_ms"],
                  "frequencies_hz": [float(f) for f in freqs],
                  "m_orthonormality_max_dev": orth, "rel_eigen_residual_max": max(resid),
                  "mixed_solve_ms": info_mx["solve_ms"], "mixed_sweeps": sw_mx,
                  "mixed_rel_value_dev_max": mixed_dev, "solver": info["name"]},
        "roofline": roofline("bsell64_spmm_kernel<K=8> (FP64 block-sliced-ELL SpMM of the "
                             "block PCG)", bsell64_bytes(nb, nnzb, 8), launch_ms),
    })
    if not a.no_cpu:
        from oracle import oracle as O
        cores = host_threads()
        csr = _oracle_csr(O, p)
        b = np.asarray(md, dtype=np.float64) * 1e6
        rate = _oracle_pcg_rate(O, csr, b, 60, cores)
        out["cpu_baseline"] = {
            "value": rate, "unit": "DOF·iter/s", "cores": cores, "kind": "port",
            "sample": f"oracle linalg::pcg restatement (Jacobi, FP64 CSR, OpenMP), 60 iterations "
                      f"on this brick's K_ff (n={csr.n}, nnz={csr.nnz}); the eigen solve's "
                      f"{info['iters']} column iterations at this rate take "
                      f"{info['iters'] * n / rate:.1f} s",
        }
    return out


def run_buckle(a, G, MG, torch, np):
    """*BUCKLE on a clamped C3D8 column under an end load: the preload solve, the geometric
    stiffness on the device, and the reference's subspace iteration (subspace_ab) over the FP64
    block PCG; the lowest factor against Euler's fixed-free column, pi^2 E I / (4 L^2)."""
    from axia_gpu import deck as D
    nx, ny, nz = a.buckle_shape
    E, P = 210000.0, -1.0
    text = MG.buckling_column_deck(nx, ny, nz, p_total=P, e=E, nbuckle=a.nbuckle)
    dk = D.Deck(text)
    n = dk.n_free
    log(f"buckle: {dk.nelem} elements, n={n}")
    for _ in range(a.warmup):
        dk.buckle(tol=a.eig_tol)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res = dk.buckle(tol=a.eig_tol)
    ms_per_step = 1e3 * (time.perf_counter() - t0) / a.steps
    inertia = min(ny * nz ** 3, nz * ny ** 3) / 12.0
    euler = np.pi ** 2 * E * inertia / (4.0 * nx ** 2) / abs(P)
    log(f"buckle: {ms_per_step:.0f} ms per analysis, factors {res.factors}, Euler {euler:.1f}, "
        f"{res.iters} column iterations in {res.solve_ms:.0f} ms")
    arr = dk.arrays()
    m = G.Mesh(8, arr["xyz"], arr["conn"], arr["e"], arr["nu"], arr["ind_of"])
    m.assemble("fp64")
    nnz = m.info()["nnz"]
    launch_ms = m.bench_operator(3, 8, a.reps)
    m.close()
    out = base_line(a, n * res.iters / (res.solve_ms / 1e3), ms_per_step, "f64", {
        "workload": f"*BUCKLE, {a.nbuckle} factors of a C3D8 column {nx}x{ny}x{nz} (n={n}), "
                    f"x = 0 clamped, end load {P:g} in x; subspace iteration over the FP64 "
                    f"block PCG to {a.eig_tol:g}",
        "n_free": n, "nnz": nnz, "elements": dk.nelem, "nbuckle": a.nbuckle, "precond": "jacobi"})
    out.update({
        "buckle": {"time_to_solution_ms": ms_per_step, "factors": [float(f) for f in res.factors],
                   "euler_fixed_free": euler,
                   "lowest_over_euler": float(abs(res.factors[0]) / euler),
                   "column_iterations": res.iters, "solve_ms": res.solve_ms,
                   "solver": res.solver},
        "roofline": roofline("bsell64_spmm_kernel<K=8> (FP64 block-sliced-ELL SpMM of the "
                             "block PCG)", bsell64_bytes(n // 3, nnz // 9, 8), launch_ms),
    })
    return out


# ------------------------------------------------------------------- nlgeom
def nl_flops(nn, ngp):
    """Algorithmic FP64 flops of one NLGEOM element tangent + internal force: the dense
    B^T D B of gemm_bt_d_b (72 nd + 14 nd^2 per point, as for Ke), the geometric term on the
    nd^2 / 3 equal-direction entries (20 flops each) and B^T S (12 nd) per point."""
    nd = 3 * nn
    return ngp * (72 * nd + 14 * nd * nd + 20 * nd * nd // 3 + 12 * nd)


def run_nlgeom(a, G, MG, torch, np):
    from axia_gpu import deck as D
    nx, ny, nz = a.nl_shape
    text = MG.nlgeom_cantilever_deck(nx, ny, nz, dims=(10.0, 1.0, 1.0), p_total=-0.5, e=1000.0,
                                     nu=0.3, ninc=4)
    dk = D.Deck(text)
    arr = dk.arrays()
    n = dk.n_free
    log(f"nlgeom: {dk.nelem} elements, n={n}")
    for _ in range(a.warmup):
        dk.nlgeom(tol=a.nl_tol, precond=a.nl_precond)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res = dk.nlgeom(tol=a.nl_tol, precond=a.nl_precond)
    ms_per_step = 1e3 * (time.perf_counter() - t0) / a.steps
    log(f"nlgeom: {ms_per_step:.0f} ms per analysis, {res.ninc} increments, {res.pcg_iters} PCG "
        f"iterations in {res.solve_ms:.0f} ms, {res.solver}")
    # the tangent re-assembly alone (tangent + gather + f_int), at the converged state
    m = G.Mesh(8, arr["xyz"], arr["conn"], arr["e"], arr["nu"], arr["ind_of"])
    m.nl_assemble(res.u_full)
    reps = 10
    t1 = time.perf_counter()
    for _ in range(reps):
        m.nl_assemble(res.u_full)
    asm_ms = 1e3 * (time.perf_counter() - t1) / reps
    nnz = m.info()["nnz"]
    block_op = "node blocks" in res.solver  # every free node keeps its 3 DOFs here
    spmv_ms = m.bench_operator(3 if block_op else 0, 1, a.reps)
    xyz = arr["xyz"]
    tip = np.nonzero(xyz[:, 0] == 10.0)[0]
    fixed = np.nonzero(xyz[:, 0] == 0.0)[0]
    value = n * res.pcg_iters / (res.solve_ms / 1e3)
    # the result files (frd.rs / dat.rs text), formatted on the host's threads
    t3 = time.perf_counter()
    frd = dk.write_frd(res.u_full, res.stress, res.rf, res.strain)
    dat = dk.write_dat(res.u_full, res.rf, res.stress_gp)
    write_ms = 1e3 * (time.perf_counter() - t3)
    write_mb = (len(frd) + len(dat)) / 1e6
    out = base_line(a, value, ms_per_step, "f64", {
        "workload": f"*STEP, NLGEOM C3D8 cantilever {nx}x{ny}x{nz} (10 x 1 x 1, n={n}), tip "
                    f"load 0.5, 4 increments, Newton with the device tangent and "
                    f"{'Jacobi' if a.nl_precond == 'jacobi' else 'block-Jacobi'} PCG to {a.nl_tol:g}",
        "n_free": n, "nnz": nnz, "elements": dk.nelem, "precond": a.nl_precond})
    asm_eps = dk.nelem / (asm_ms / 1e3)
    out.update({
        "nlgeom": {"time_per_analysis_ms": ms_per_step, "increments": res.ninc,
                   "newton_iters_last_increment": res.iters, "pcg_iterations": res.pcg_iters,
                   "pcg_solve_ms": res.solve_ms, "tip_deflection": float(res.u_full[3 * tip + 2].mean()),
                   "residual": res.residual,
                   "reaction_z_sum": float(res.rf[3 * fixed + 2].sum()),
                   "tangent_reassembly_ms": asm_ms, "tangent_reassembly_elements_per_s": asm_eps,
                   "tangent_fp64_tflops": nl_flops(8, 8) * asm_eps / 1e12,
                   "frd_dat_write_ms": write_ms, "frd_dat_mb": write_mb,
                   "solver": res.solver},
        "roofline": (roofline("bsell64_spmm_kernel<K=1> (FP64 node-block SpMV of the tangent)",
                              bsell64_bytes(n // 3, nnz // 9, 1), spmv_ms) if block_op else
                     roofline("sell_spmv_kernel (FP64 SpMV of the tangent in the PCG)",
                              sell64_bytes(n, nnz), spmv_ms)),
    })
    if not a.no_cpu:
        from oracle import oracle as O
        cores = host_threads()
        # the oracle's element tangent (single thread) on a sample of elements
        u3 = res.u_full.reshape(-1, 3)
        conn, e, nu = arr["conn"], arr["e"], arr["nu"]
        sample = min(2000, dk.nelem)
        t2 = time.perf_counter()
        for el in range(sample):
            O.hex_nl(xyz[conn[el]], u3[conn[el]].ravel(), e[el], nu[el])
        cpu_eps = sample / (time.perf_counter() - t2)
        p = MG.Problem("nlgeom_tangent_sample", 8, xyz, conn, e, nu, arr["ind_of"])
        csr = _oracle_csr(O, p)
        rate = _oracle_pcg_rate(O, csr, np.ones(csr.n), 60, cores)
        out["cpu_baseline"] = {
            "value": rate, "unit": "DOF·iter/s", "cores": cores, "kind": "port",
            "sample": f"oracle linalg::pcg restatement (Jacobi, FP64 CSR, OpenMP), 60 iterations "
                      f"on this cantilever's K_ff (n={csr.n}, nnz={csr.nnz}); the oracle's "
                      f"NLGEOM element tangent single-thread on {sample} elements",
            "tangent_elements_per_s_1thread": cpu_eps,
        }
    return out


# --------------------------------------------------------------------- SIMP
def cpu_adjoint(O, p, u, threads, chunk=20000):
    """FP64 CPU adjoint: dc/dE_e = -u_e^T K1_e u_e with the oracle's element matrices."""
    import numpy as np
    L = O.lib()
    ne = p.nelem
    out = np.empty(ne)
    U = u.reshape(-1, 3)
    ones = np.ones(chunk)
    for s in range(0, ne, chunk):
        c = min(chunk, ne - s)
        conn = np.ascontiguousarray(p.conn[s:s + c], np.int32)
        ke = np.empty(c * 576)
        L.or_element_ke_batch(8, p.nnode, O._d(p.xyz), c, O._i(conn),
                              O._d(ones), O._d(np.ascontiguousarray(p.nu[s:s + c])), O._d(ke),
                              threads)
        ue = U[conn].reshape(c, 24)
        out[s:s + c] = -np.einsum("ei,eij,ej->e", ue, ke.reshape(c, 24, 24), ue)
    return out


def run_deck(a, G, MG, torch, np):
    """A deck end to end: parse + setup (the native INP reader), device assembly + autospc +
    FP64 Jacobi PCG to 1e-8 (linalg::pcg's defaults) + reconstruction + reactions + stresses,
    then the DAT and FRD files; the deck writes the lognormal brick the way the reference needs a
    per-element modulus written: one element set, material and section per element."""
    from axia_gpu import deck as D
    n = a.deck_n
    t0 = time.perf_counter()
    text, prob = MG.lognormal_brick_deck(n, seed=0)
    mb = len(text) / 1e6
    log(f"deck: {mb:.1f} MB, {prob.nelem} elements, generated in {time.perf_counter() - t0:.1f} s")

    def step():
        t1 = time.perf_counter()
        dk = D.Deck(text)
        t2 = time.perf_counter()
        res = dk.solve()
        t3 = time.perf_counter()
        frd = dk.write_frd(res.u_full, res.stress, res.rf)
        dat = dk.write_dat(res.u_full, res.rf, res.stress_gp, res.stress)
        t4 = time.perf_counter()
        return dk, res, (t2 - t1, t3 - t2, t4 - t3), len(frd) + len(dat)

    for _ in range(a.warmup):
        step()
    t5 = time.perf_counter()
    runs = [step() for _ in range(a.steps)]
    ms_per_step = 1e3 * (time.perf_counter() - t5) / a.steps
    dk, res, _, out_bytes = runs[-1]
    parse_s, solve_s, write_s = (float(np.mean([r[2][k] for r in runs])) for k in range(3))
    log(f"deck: parse {parse_s:.2f} s, solve {solve_s:.2f} s ({res.iters} it., PCG "
        f"{res.solve_ms:.0f} ms), files {write_s:.2f} s")
    arr = dk.arrays()
    nf = dk.n_free
    bot = np.nonzero(arr["xyz"][:, 2] == 0.0)[0]
    reaction_z = float(res.rf[3 * bot + 2].sum())
    # the deck's operator on a mesh of the same arrays: the block SpMV the PCG runs
    m = G.Mesh(8, arr["xyz"], arr["conn"], arr["e"], arr["nu"], arr["ind_of"])
    m.assemble("fp64")
    nnz = m.info()["nnz"]
    spmv_ms = m.bench_operator(3, 1, a.reps)
    out = base_line(a, nf * res.iters / (res.solve_ms / 1e3), ms_per_step, "f64", {
        "workload": f"INP deck of the C3D8 brick {n}^3 with lognormal E, one *ELSET, *MATERIAL "
                    f"and *SOLID SECTION per element ({mb:.0f} MB), bottom clamped, P2 = 1 on the "
                    "top layer: parse, FP64 Jacobi PCG to 1e-8, reactions, stresses, DAT + FRD",
        "deck_mb": mb, "elements": prob.nelem, "n_free": nf, "nnz": nnz, "precond": "jacobi"})
    out.update({
        "deck": {"time_per_deck_ms": ms_per_step, "parse_and_setup_s": parse_s,
                 "parse_mb_per_s": mb / parse_s, "solve_s": solve_s, "pcg_ms": res.solve_ms,
                 "pcg_iterations": res.iters, "files_s": write_s, "files_mb": out_bytes / 1e6,
                 "reaction_z_sum": reaction_z, "applied_load": float(n * n),
                 "solver": res.solver},
        "roofline": roofline("bsell64_spmm_kernel<K=1> (FP64 node-block SpMV of the deck's K_ff)",
                             bsell64_bytes(nf // 3, nnz // 9, 1), spmv_ms),
    })
    m.close()
    if not a.no_cpu:
        from oracle import oracle as O
        # the oracle's reader and setup (a restatement of inp.rs / analysis.rs:548-946 in
        # Python, one thread) on the same kind of deck, 16^3
        small, _ = MG.lognormal_brick_deck(16, seed=0)
        t6 = time.perf_counter()
        O.deck_setup(O.parse_inp(small))
        t_small = time.perf_counter() - t6
        out["cpu_baseline"] = {
            "value": len(small) / 1e6 / t_small, "unit": "MB/s (deck parse + setup)", "cores": 1,
            "kind": "port",
            "sample": f"oracle parse_inp + deck_setup on the same deck form at 16^3 "
                      f"({len(small) / 1e6:.1f} MB); the native reader did "
                      f"{mb / parse_s:.1f} MB/s on the {n}^3 deck"}
    return out


def run_simp(a, G, MG, torch, np):
    t0 = time.perf_counter()
    prob = MG.simp_brick(120, 60, 60, seed=0)
    rho = prob.meta["rho"]
    mesh = G.Mesh(8, prob.xyz, prob.conn, prob.e, prob.nu, prob.ind_of)
    n = prob.n_free
    nnz = mesh.info()["nnz"]
    f = prob.f_full[0]
    log(f"mesh {prob.name}: n={n} nnz={nnz} setup {time.perf_counter() - t0:.1f}s")
    tol = 1e-12

    def forward(e):
        mesh.set_modulus(e)
        ms_asm = mesh.assemble("fp64")
        x, inf = mesh.solve(prob.rhs(), precision="fp64", precond="block3", tol=tol)
        return ms_asm, prob.expand(x), inf

    def step():
        ms_asm, u, inf = forward(prob.e)
        t = time.perf_counter()
        c, dc_de = mesh.compliance_adjoint(u, f)
        adj_s = time.perf_counter() - t
        return ms_asm, inf, c, dc_de, u, adj_s

    for _ in range(a.warmup):
        step()
    t1 = time.perf_counter()
    res = [step() for _ in range(a.steps)]
    ms_per_step = 1e3 * (time.perf_counter() - t1) / a.steps
    ms_asm, inf, c0, dc_de, u, adj_s = res[-1]
    dc_drho = dc_de * 3.0 * rho ** 2
    # central differences, step 1e-4 in rho, on random elements. c(rho +- h) - c(rho) is
    # formed as f . du with K(rho +- h) du = -(E(rho +- h) - E(rho)) K1_e u_e (an exact
    # identity), so the difference of two nearly equal compliances is never taken.
    rng = np.random.default_rng(7)
    pick = rng.choice(prob.nelem, a.fd_elements, replace=False)
    h = 1e-4
    b_free = prob.rhs()
    fd = []
    t_fd = time.perf_counter()
    for el in pick:
        dofs = (3 * prob.conn[el][:, None] + np.arange(3)).ravel()
        k1 = mesh.ke(int(el), 1)[0] / prob.e[el]
        ku = k1 @ u[dofs]
        dc = []
        for sgn in (1.0, -1.0):
            r = rho.copy()
            r[el] += sgn * h
            e = 1e-3 + r ** 3
            rhs_full = np.zeros(3 * prob.nnode)
            rhs_full[dofs] = -(e[el] - prob.e[el]) * ku
            free = prob.ind_of >= 0
            rhs = np.zeros(n)
            rhs[prob.ind_of[free]] = rhs_full[free]
            mesh.set_modulus(e)
            mesh.assemble("mixed")
            du, _ = mesh.solve(rhs, precision="mixed", precond="block3", tol=1e-12)
            dc.append(float(b_free @ du))
        fd.append((dc[0] - dc[1]) / (2 * h))
        mesh.set_modulus(prob.e)
        mesh.assemble("fp64")
    t_fd = time.perf_counter() - t_fd
    fd = np.array(fd)
    fd_dev = np.abs(fd - dc_drho[pick]) / np.abs(dc_drho[pick])
    mesh.set_modulus(prob.e)
    solve_ms = float(np.mean([r[1]["solve_ms"] for r in res]))
    iters = res[-1][1]["iters"]
    # adjoint kernel: one pass over the elements reading u_e (8 nodes x 3) and writing one value
    out = base_line(a, n * iters / (solve_ms / 1e3), ms_per_step, "f64", {
        "workload": "SIMP C3D8 brick 120x60x60, E = 1e-3 + rho^3, rho ~ U(0.3, 1) (seed 0), "
                    "left face clamped, unit point load; compliance and dc/drho via the adjoint",
        "elements": prob.nelem, "n_free": n, "nnz": nnz, "precond": "block3", "tol": tol})
    out.update({
        "compliance": c0,
        "forward": {"assembly_ms": ms_asm, "solve_ms": solve_ms, "iters": iters},
        "adjoint": {"host_call_ms": 1e3 * adj_s,
                    "elements_per_s_host_call": prob.nelem / adj_s},
        "fd_check": {"elements": len(pick), "step": h, "max_rel_dev": float(fd_dev.max()),
                     "median_rel_dev": float(np.median(fd_dev)), "seconds": t_fd},
    })
    if not a.no_cpu:
        from oracle import oracle as O
        cores = host_threads()
        t2 = time.perf_counter()
        ref = cpu_adjoint(O, prob, u, cores)
        dt = time.perf_counter() - t2
        out["cpu_baseline"] = {
            "value": prob.nelem / dt, "unit": "elements/s (FP64 adjoint)", "cores": cores,
            "kind": "port",
            "sample": "oracle element matrices at unit modulus and u_e^T K1_e u_e over all "
                      f"{prob.nelem} elements",
            "gpu_vs_cpu_adjoint_max_rel_dev": float(np.max(np.abs(dc_de - ref)) /
                                                    np.max(np.abs(ref))),
        }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="brick", choices=WORKLOADS)
    ap.add_argument("--n", type=int, default=80, help="brick edge (elements)")
    ap.add_argument("--nrhs", type=int, default=4)
    ap.add_argument("--tol", type=float, default=1e-11)
    ap.add_argument("--precond", default="block3")
    ap.add_argument("--inner-tol", type=float, default=1e-3)
    ap.add_argument("--refinement", default="fcg", choices=("fcg", "plain", "reliable"))
    ap.add_argument("--batch-groups", type=int, default=0,
                    help="batch workload: size groups solved smallest first (0: library default)")
    ap.add_argument("--asm32-mode", type=int, default=-1, choices=(-1, 0, 1),
                    help="FP32 assembly form: 0 fused rows, 1 MFMA element blocks + gather, "
                         "-1 the library's default for the element type")
    ap.add_argument("--fd-elements", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="brick: skip the summaries of the other benchmark configurations")
    ap.add_argument("--kernels", action="store_true",
                    help="only the operator kernels (repeated launches), for profiling")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--freq-n", type=int, default=48, help="freq brick edge (elements)")
    ap.add_argument("--nmodes", type=int, default=10)
    ap.add_argument("--eig-tol", type=float, default=1e-10)
    ap.add_argument("--nl-shape", type=int, nargs=3, default=[200, 20, 20])
    ap.add_argument("--nl-tol", type=float, default=1e-10)
    ap.add_argument("--nl-precond", default="jacobi", choices=("jacobi", "block3"))
    ap.add_argument("--deck-n", type=int, default=80, help="deck brick edge (elements)")
    ap.add_argument("--buckle-shape", type=int, nargs=3, default=[320, 16, 16])
    ap.add_argument("--nbuckle", type=int, default=4)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU "
              "sharding of this path is a design, not a measured run.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import axia_gpu as G
    from axia_gpu import meshgen as MG

    run = {"brick": run_brick, "cantilever": run_cantilever, "c3d20": run_c3d20,
           "batch": run_batch, "simp": run_simp, "freq": run_freq,
           "nlgeom": run_nlgeom, "deck": run_deck, "buckle": run_buckle}[a.workload]
    out = run(a, G, MG, torch, np)
    if out is not None and a.workload == "brick" and not a.kernels and not a.no_configs:
        out["other_configs"] = other_configs(a, G, MG, torch, np)
    if out is not None:
        print(json.dumps(out))


def other_configs(a, G, MG, torch, np):
    """The other benchmark configurations, each once after a warm-up (no CPU legs, 8
    finite-difference elements for the adjoint), summarised beside the headline line:
    the cantilever (FP64 vs the oracle's Cholesky is the parity test; here the GPU solve),
    the jittered C3D20 brick, the 48-mesh batch and the SIMP adjoint."""
    keep_out = {"metric", "n_gpus", "steps", "warmup", "higher_is_better", "scaling",
                "vs_baseline", "data"}
    res = {}
    for name, fn in (("cantilever_20x9x9", run_cantilever), ("c3d20_96x24x24", run_c3d20),
                     ("batch48", run_batch), ("simp_120x60x60", run_simp)):
        b = argparse.Namespace(**vars(a))
        b.steps, b.warmup, b.no_cpu, b.fd_elements, b.reps = 1, 1, True, 8, 3
        t0 = time.perf_counter()
        try:
            o = fn(b, G, MG, torch, np)
            res[name] = {k: v for k, v in o.items() if k not in keep_out}
        except Exception as e:  # a failed companion must not hide the headline line
            res[name] = {"error": f"{type(e).__name__}: {e}"}
        res[name]["wall_s"] = time.perf_counter() - t0
        log(f"{name}: {res[name]['wall_s']:.1f} s")
    return res


if __name__ == "__main__":
    main()
