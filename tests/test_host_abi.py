"""Host-side checks of the C ABI library (no GPU needed).

The library must load and export every entry point include/axia_gpu.h
declares; the host-only entry points (symbolic phase, face-pressure loads)
are compared with the oracle here.
"""
import os
import re

import numpy as np
import pytest

import axia_gpu as G
from axia_gpu import _lib
from axia_gpu import meshgen as MG
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "axia_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(axg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), f"{s} not exported"
        assert s in _lib.SIGNATURES, f"{s} has no ctypes signature"
    assert L.axg_version().decode().startswith("axia_gpu")


def test_default_opts_match_reference_pcg():
    import ctypes as C
    o = _lib.SolveOpts()
    _lib.lib().axg_default_opts(C.byref(o))
    assert o.tol == 1e-8 and o.max_iter == 0 and o.precision == 0 and o.precond == 0


def oracle_csr(p):
    ti, tj, tv = O.assemble_triplets(p.nn, p.xyz, p.conn, p.e, p.nu)
    sel = (p.ind_of[ti] >= 0) & (p.ind_of[tj] >= 0)
    return O.csr_from_triplets(p.n_free, p.ind_of[ti[sel]], p.ind_of[tj[sel]], tv[sel])


@pytest.mark.parametrize("maker", [
    lambda: MG.cantilever(5, 2, 2),
    lambda: MG.brick_lognormal(3, seed=4),
    lambda: MG.brick_c3d20_jitter(3, 2, 2, seed=1),
])
def test_symbolic_pattern_and_scatter_map_vs_oracle(maker):
    """Pattern bit-exact, and oracle Ke summed through the map gives the oracle's K_ff bits."""
    p = maker()
    csr = oracle_csr(p)
    rp, ci, sp, cb = G.pattern_host(p.nn, p.nnode, p.conn, p.ind_of)
    assert np.array_equal(rp, csr.indptr)
    assert np.array_equal(ci, csr.indices)
    kes = np.stack([O.hex_stiffness(p.xyz[p.conn[e]], p.e[e], p.nu[e])[0]
                    for e in range(p.nelem)]).ravel()
    vals = np.array([sum(kes[cb[sp[k]:sp[k + 1]]].tolist(), 0.0) for k in range(len(ci))])
    assert np.array_equal(vals, csr.data)
    # contributions of one slot come in ascending element order
    m2 = (3 * p.nn) ** 2
    for k in range(0, len(ci), max(1, len(ci) // 200)):
        els = cb[sp[k]:sp[k + 1]] // m2
        assert np.all(np.diff(els) > 0)


def test_symbolic_pattern_on_host_threads():
    """A mesh large enough for the symbolic phase to split its rows over host threads
    (C3D8 28^3, 24,389 nodes): the pattern equals the oracle's csr_from_triplets pattern,
    and every slot's contributions are its (element, i, j) entries in ascending element
    order, each naming the slot's row and column DOFs."""
    p = MG.brick_lognormal(28, seed=2)
    csr = oracle_csr(p)
    rp, ci, sp, cb = G.pattern_host(p.nn, p.nnode, p.conn, p.ind_of)
    assert np.array_equal(rp, csr.indptr) and np.array_equal(ci, csr.indices)
    m = 3 * p.nn
    el, rest = cb // (m * m), cb % (m * m)
    li, lj = rest // m, rest % m
    dof_i = 3 * p.conn[el, li // 3].astype(np.int64) + li % 3
    dof_j = 3 * p.conn[el, lj // 3].astype(np.int64) + lj % 3
    slot = np.repeat(np.arange(len(ci)), np.diff(sp))
    row = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    assert np.array_equal(p.ind_of[dof_i], row[slot])
    assert np.array_equal(p.ind_of[dof_j], ci[slot])
    first = np.ones(len(cb), bool)
    first[sp[:-1]] = False  # the first contribution of each slot has no predecessor
    assert np.all(np.diff(el)[first[1:]] > 0)


def test_symbolic_partial_constraints():
    """Nodes with only some DOFs prescribed (LE10-style YS,2,2) keep their other rows."""
    xyz, conn = MG.brick_hex8(3, 2, 2)
    pres = np.zeros(3 * len(xyz), bool)
    pres[3 * np.nonzero(xyz[:, 1] == 0)[0] + 1] = True
    pres[3 * np.nonzero(xyz[:, 0] == 0)[0] + 0] = True
    pres[[2, 5]] = True
    ind = np.full(3 * len(xyz), -1, np.int64)
    ind[~pres] = np.arange(int((~pres).sum()))
    p = MG.Problem("partial", 8, xyz, conn, np.full(len(conn), 1e5), np.full(len(conn), 0.3), ind)
    csr = oracle_csr(p)
    rp, ci, sp, cb = G.pattern_host(8, len(xyz), conn, ind)
    assert np.array_equal(rp, csr.indptr) and np.array_equal(ci, csr.indices)


def test_symbolic_rejects_bad_meshes():
    xyz, conn = MG.brick_hex8(1, 1, 1)
    bad = conn.copy()
    bad[0, 1] = bad[0, 0]
    with pytest.raises(G.AxgError) as e:
        G.pattern_host(8, len(xyz), bad, MG.free_map(len(xyz), [0]))
    assert "repeats a node" in str(e.value)
    bad = conn.copy()
    bad[0, 3] = 99
    with pytest.raises(G.AxgError):
        G.pattern_host(8, len(xyz), bad, MG.free_map(len(xyz), [0]))


@pytest.mark.parametrize("nn", [8, 20])
def test_pressure_loads_bitexact_vs_oracle(nn):
    rng = np.random.default_rng(nn)
    xyz, conn = MG.brick_hex8(2, 2, 1) if nn == 8 else MG.brick_hex20(2, 1, 1)
    xyz = xyz + rng.uniform(-0.1, 0.1, xyz.shape)
    faces = rng.integers(1, 7, len(conn)).astype(np.int32)
    mags = rng.uniform(-2, 2, len(conn))
    f = np.zeros(3 * len(xyz))
    G.pressure_loads(nn, xyz, conn, faces, mags, f)
    ref = np.zeros(3 * len(xyz))
    for e in range(len(conn)):
        fe = O.hex_face_pressure(xyz[conn[e]], faces[e], mags[e])
        for a in range(nn):
            for d in range(3):
                ref[3 * conn[e][a] + d] += fe[3 * a + d]
    assert np.array_equal(f, ref)
    with pytest.raises(G.AxgError):
        G.pressure_loads(nn, xyz, conn[:1], np.array([7], np.int32), np.array([1.0]), f)


def test_meshgen_counts_of_the_benchmark_shapes():
    """Exact node / element / DOF counts of the benchmark shapes."""
    xyz, conn = MG.brick_hex20(96, 24, 24)
    assert len(xyz) == 237025 and len(conn) == 55296
    ind = MG.free_map(len(xyz), np.nonzero(xyz[:, 0] == 0)[0])
    assert int((ind >= 0).sum()) == 705600
    xyz, conn = MG.brick_hex8(80, 80, 80)
    assert len(xyz) == 531441
    ind = MG.free_map(len(xyz), np.nonzero(xyz[:, 2] == 0)[0])
    assert int((ind >= 0).sum()) == 1574640
    p = MG.cantilever(20, 9, 9)
    assert p.nnode == 2100 and p.n_free == 6000
    assert abs(p.f_full.__getitem__(0).sum() + 1000.0) < 1e-9


# ------------------------------------------------- struct layouts across the ABI
RUST_SIZES = {"i32": (4, 4), "i64": (8, 8), "f64": (8, 8), "[c_char; 64]": (64, 1)}
RUST_BINDING_DOC = "INTEGRATION.md"


def rust_struct_fields(name):
    """Fields of a #[repr(C)] struct of the published Rust binding (the integration guide
    at the repository root, whose code blocks a Rust crate would copy)."""
    text = open(os.path.join(ROOT, RUST_BINDING_DOC)).read()
    body = re.search(r"pub struct " + name + r" \{(.*?)\n\}", text, re.S).group(1)
    return re.findall(r"pub (\w+): ([^,]+),", body)


def repr_c_layout(fields):
    """Offsets and size by the C layout rules #[repr(C)] follows."""
    off, align, out = 0, 1, {}
    for f, t in fields:
        size, al = RUST_SIZES[t.strip()]
        off = (off + al - 1) // al * al
        out[f] = off
        off += size
        align = max(align, al)
    return out, (off + align - 1) // align * align


def c_layout(struct, fields):
    """sizeof / offsetof of the C header's struct, from a program compiled with gcc."""
    import subprocess
    import tempfile
    lines = [f'printf("%zu\\n", offsetof({struct}, {f}));' for f in fields]
    src = ("#include <stddef.h>\n#include <stdio.h>\n#include \"axia_gpu.h\"\nint main(void) {\n"
           + "\n".join(lines) + f'\nprintf("%zu\\n", sizeof({struct}));\nreturn 0;\n}}\n')
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), c, "-o", exe],
                       check=True)
        vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                                check=True).stdout.split()]
    return dict(zip(fields, vals[:-1])), vals[-1]


@pytest.mark.parametrize("c_name,rust_name,py", [("axg_solve_opts", "AxgSolveOpts", "SolveOpts"),
                                                 ("axg_solve_info", "AxgSolveInfo", "SolveInfo")])
def test_struct_layout_matches_rust_and_ctypes_mirrors(c_name, rust_name, py):
    """The C header's structs, the Rust #[repr(C)] mirror a maintainer adds and the ctypes
    mirror of the Python binding agree field for field in order, offset and size."""
    rust = rust_struct_fields(rust_name)
    names = [f for f, _ in rust]
    cls = getattr(_lib, py)
    assert names == [f for f, _ in cls._fields_]
    c_off, c_size = c_layout(c_name, names)
    r_off, r_size = repr_c_layout(rust)
    assert c_off == r_off and c_size == r_size
    import ctypes
    assert {f: getattr(cls, f).offset for f in names} == c_off
    assert ctypes.sizeof(cls) == c_size


# ------------------------------------------------- argument checks before the device
def test_solve_csr_rejects_malformed_csr_before_any_device_work():
    """axg_solve_csr validates the host CSR and the options before it touches the device:
    the handle below is never dereferenced on these paths."""
    import ctypes as C
    L = _lib.lib()
    fake = C.c_void_p(1)
    o = _lib.SolveOpts()
    L.axg_default_opts(C.byref(o))
    x = np.zeros(3)
    rhs = np.ones(3)
    info = _lib.SolveInfo()

    def call(indptr, indices, data, opts=o):
        ip = np.asarray(indptr, np.int64)
        ix = None if indices is None else np.asarray(indices, np.int64)
        dv = None if data is None else np.asarray(data, np.float64)
        return L.axg_solve_csr(fake, 3, _lib.lptr(ip), _lib.lptr(ix) if ix is not None else None,
                               _lib.dptr(dv) if dv is not None else None, _lib.dptr(rhs),
                               C.byref(opts), _lib.dptr(x), C.byref(info))

    cases = [([1, 1, 2, 3], [0, 1, 2], [1.0, 1.0, 1.0], "indptr[0]"),
             ([0, 2, 1, 3], [0, 1, 2], [1.0, 1.0, 1.0], "non-decreasing"),
             ([0, 1, 2, 3], None, None, "NULL indices"),
             ([0, 1, 2, 3], [0, 1, 5], [1.0, 1.0, 1.0], "column index")]
    for ip, ix, dv, msg in cases:
        assert call(ip, ix, dv) == _lib.AXG_E_INVALID
        assert msg in L.axg_last_error().decode()
    bad = _lib.SolveOpts()
    L.axg_default_opts(C.byref(bad))
    bad.precision = 7
    assert call([0, 1, 2, 3], [0, 1, 2], [1.0, 1.0, 1.0], bad) == _lib.AXG_E_INVALID
    assert "unknown precision" in L.axg_last_error().decode()
    bad.precision, bad.precond = 0, 3
    assert call([0, 1, 2, 3], [0, 1, 2], [1.0, 1.0, 1.0], bad) == _lib.AXG_E_INVALID
    assert "unknown preconditioner" in L.axg_last_error().decode()


def test_solve_batch_refuses_other_precisions_and_preconditioners():
    """axg_solve_batch runs the FP64 Jacobi PCG only: a caller asking for the mixed form or
    block-Jacobi gets AXG_E_INVALID before any mesh or device access (the handles below
    are never dereferenced on these paths)."""
    import ctypes as C
    L = _lib.lib()
    fake = C.c_void_p(1)
    meshes = (C.c_void_p * 1)(fake)
    buf = np.zeros(3)
    ptrs = (C.POINTER(C.c_double) * 1)(_lib.dptr(buf))
    nrhs = (C.c_int32 * 1)(1)
    for field, value, msg in (("precision", 1, "precision must be AXG_PREC_FP64"),
                              ("precond", 1, "precond must be AXG_PRECOND_JACOBI")):
        o = _lib.SolveOpts()
        L.axg_default_opts(C.byref(o))
        setattr(o, field, value)
        st = L.axg_solve_batch(fake, 1, meshes, ptrs, nrhs, C.byref(o), ptrs, None, None, None)
        assert st == _lib.AXG_E_INVALID
        assert msg in L.axg_last_error().decode()


def test_nlgeom_refuses_mixed_precision():
    """axg_deck_nlgeom re-assembles the FP64 tangent every Newton iteration, so it has no FP32
    operator: precision = mixed is refused before the device is touched."""
    import ctypes as C
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from eigen_decks import nl_cantilever_deck
    from axia_gpu import deck as D
    L = _lib.lib()
    dk = D.Deck(nl_cantilever_deck(ninc=2))
    o = _lib.SolveOpts()
    L.axg_default_opts(C.byref(o))
    o.precision = 1
    buf = np.zeros(3 * dk.nnode)
    st = L.axg_deck_nlgeom(C.c_void_p(1), dk._h, C.byref(o), _lib.dptr(buf), None, None, None,
                           None, None, None, None, None)
    assert st == _lib.AXG_E_INVALID
    assert "precision must be AXG_PREC_FP64" in L.axg_last_error().decode()


def test_python_binding_checks_sizes_before_the_call():
    """Mesh methods raise ValueError for arrays of the wrong length instead of letting the
    library read or write past them (the checks run before any library call)."""
    m = G.Mesh.__new__(G.Mesh)
    m._h, m.nn, m.nnode, m.nelem, m._n_free = None, 8, 10, 2, 27
    with pytest.raises(ValueError):
        m.solve(np.zeros(26))
    with pytest.raises(ValueError):
        m.solve(np.zeros((2, 28)))
    for call in (lambda: m.nodal_stress(np.zeros(29)),
                 lambda: m.reactions(np.zeros(30), np.zeros(31)),
                 lambda: m.reduce_rhs(np.zeros(30), np.zeros(29)),
                 lambda: m.compliance_adjoint(np.zeros(30), np.zeros(3)),
                 lambda: m.nl_assemble(np.zeros(3)),
                 lambda: m.geometric_stiffness(np.zeros(31))):
        with pytest.raises(ValueError):
            call()
    with pytest.raises(ValueError):
        G.solve_kff(3, ([0], [0], [1.0]), np.ones(4), ctx=object())
    with pytest.raises(ValueError):
        G.solve_batch([m], [np.zeros((1, 26))], ctx=object())
