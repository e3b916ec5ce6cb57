"""CPU FP64 oracle for the axia linear-static hot path — TEST INFRASTRUCTURE.

Python side of the oracle: a ctypes binding of ``liboracle`` (the C
restatement of the reference's element, CSR and PCG routines, built from
``oracle/axia_oracle.c``), a restatement of the reference's CalculiX INP
subset reader, and of the driver ``solve_linear`` for all-Hex8 / all-Hex20
isotropic decks.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module; it is the checker, never the product.

Reference files followed (solver/src/...):
  inp.rs:11-309, 383-1061, 2256-2330   deck subset, deferred sets
  model.rs:1234-1309, 1391-1407        compact (sorting), elset rebinding,
                                       material lookup
  analysis.rs:548-668, 723-760, 926-1035, 1061-1162, 2164-2355, 2709-2764
  constraint.rs:17-218                  DofMap (prescribed DOFs only)
  linalg.rs:10-37, 194-259, 305-404     CSR, PCG, solve_kff dispatch
"""
from __future__ import annotations

import ctypes as C
import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libaxia_oracle.so")
_lib = None

ERRORS = {
    1: "invalid input",
    2: "Jakobimatrix ist singulär (entartetes Element).",
    3: "negative Jakobideterminante.",
    4: "Ungültige Materialkonstanten",
    5: "CG: unerwartetes Null-Pivot — Matrix nicht positiv definit.",
    6: "CG konvergierte nicht",
    7: "Steifigkeitsmatrix ist indefinit.",
    8: "Steifigkeitsmatrix ist singulär — Randbedingungen unzureichend "
       "(Starrkörperbewegung oder entartete Elemente).",
    9: "out of memory",
    10: "NLGEOM: det(F) ≤ 0 (Element inversion).",
}


class OracleError(RuntimeError):
    def __init__(self, code, msg=None):
        self.code = code
        super().__init__(msg or ERRORS.get(code, f"oracle error {code}"))


def build():
    """Compile liboracle with the committed Makefile (gcc, no GPU needed)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        lp = C.POINTER(C.c_int64)
        ip = C.POINTER(C.c_int32)
        i64 = C.c_int64
        L.or_d_iso_3d.argtypes = [C.c_double, C.c_double, dp]
        L.or_hex8_shape.argtypes = [C.c_double] * 3 + [dp, dp]
        L.or_hex20_shape.argtypes = [C.c_double] * 3 + [dp, dp]
        L.or_hex_dndx.argtypes = [C.c_int, dp, C.c_double, C.c_double, C.c_double, dp, dp, dp]
        L.or_hex_stiffness.argtypes = [C.c_int, dp, dp, dp, dp]
        L.or_hex_face_pressure.argtypes = [C.c_int, dp, C.c_int, C.c_double, dp]
        L.or_hex_nodal_stress.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, dp]
        L.or_hex_kg.argtypes = [C.c_int, dp, dp, dp]
        L.or_hex_nl.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, dp, dp, dp, dp, dp]
        L.or_hex20_nodal_nl.argtypes = [dp, dp, C.c_double, C.c_double, dp, dp]
        L.or_von_mises.argtypes = [dp]
        L.or_von_mises.restype = C.c_double
        L.or_assemble_triplets.argtypes = [C.c_int, i64, dp, i64, ip, dp, dp, lp, lp, dp, lp, lp]
        L.or_reduce.argtypes = [i64, lp, dp, i64, i64, lp, lp, dp, dp, lp, lp, dp, lp, dp]
        L.or_csr_from_triplets.argtypes = [i64, i64, lp, lp, dp, lp, lp, dp]
        L.or_csr_matvec.argtypes = [i64, lp, lp, dp, dp, dp]
        L.or_pcg.argtypes = [i64, lp, lp, dp, dp, C.c_double, i64, dp, lp, dp]
        L.or_chol_solve.argtypes = [i64, dp, dp, dp]
        L.or_element_ke_batch.argtypes = [C.c_int, i64, dp, i64, ip, dp, dp, dp, C.c_int]
        L.or_pcg_omp.argtypes = [i64, lp, lp, dp, dp, C.c_double, i64, i64, dp, lp, dp]
        L.or_omp_threads.restype = C.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _l(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _check(st):
    if st != 0:
        raise OracleError(st)


# ----------------------------------------------------------------- elements
def d_iso_3d(e, nu):
    d = np.zeros(36)
    _check(lib().or_d_iso_3d(e, nu, _d(d)))
    return d.reshape(6, 6)


def hex_shape(nn, xi, eta, zeta):
    n = np.zeros(nn)
    dn = np.zeros(3 * nn)
    f = lib().or_hex8_shape if nn == 8 else lib().or_hex20_shape
    f(xi, eta, zeta, _d(n), _d(dn))
    return n, dn.reshape(nn, 3)


def hex_dndx(nn, xyz, xi, eta, zeta):
    xyz = _f64(xyz)
    dndx = np.zeros(3 * nn)
    det = np.zeros(1)
    n = np.zeros(nn)
    _check(lib().or_hex_dndx(nn, _d(xyz), xi, eta, zeta, _d(dndx), _d(det), _d(n)))
    return dndx.reshape(nn, 3), float(det[0]), n


def hex_stiffness(xyz, e, nu, d=None):
    xyz = _f64(xyz)
    nn = xyz.shape[0]
    if d is None:
        d = d_iso_3d(e, nu)
    d = _f64(d)
    ke = np.zeros((3 * nn) ** 2)
    vol = np.zeros(1)
    _check(lib().or_hex_stiffness(nn, _d(xyz), _d(d), _d(ke), _d(vol)))
    return ke.reshape(3 * nn, 3 * nn), float(vol[0])


def hex_face_pressure(xyz, face, p):
    xyz = _f64(xyz)
    nn = xyz.shape[0]
    fe = np.zeros(3 * nn)
    _check(lib().or_hex_face_pressure(nn, _d(xyz), int(face), float(p), _d(fe)))
    return fe


def hex_nodal_stress(xyz, ue, e, nu):
    xyz = _f64(xyz)
    ue = _f64(ue)
    nn = xyz.shape[0]
    out = np.zeros(nn * 6)
    _check(lib().or_hex_nodal_stress(nn, _d(xyz), _d(ue), float(e), float(nu), _d(out)))
    return out.reshape(nn, 6)


def hex_kg(xyz, stress):
    """eigen.rs:299-360 (hex8_kg / hex20_kg, add_continuum_kg): geometric stiffness of one
    element under a constant stress (Voigt xx yy zz xy yz zx)."""
    xyz = _f64(xyz)
    nn = xyz.shape[0]
    kg = np.zeros((3 * nn) ** 2)
    _check(lib().or_hex_kg(nn, _d(xyz), _d(_f64(stress)), _d(kg)))
    return kg.reshape(3 * nn, 3 * nn)


def von_mises(s):
    s = _f64(s)
    return lib().or_von_mises(_d(s))


# ----------------------------------------------------------------- assembly
def assemble_triplets(nn, xyz, conn, e, nu):
    """Full-space triplets of every element, analysis.rs:574-668."""
    xyz = _f64(xyz).reshape(-1, 3)
    conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, nn)
    nelem = conn.shape[0]
    e = _f64(np.broadcast_to(e, (nelem,)))
    nu = _f64(np.broadcast_to(nu, (nelem,)))
    cap = nelem * (3 * nn) ** 2
    ti = np.empty(cap, np.int64)
    tj = np.empty(cap, np.int64)
    tv = np.empty(cap, np.float64)
    nt = np.zeros(1, np.int64)
    bad = np.full(1, -1, np.int64)
    st = lib().or_assemble_triplets(nn, xyz.shape[0], _d(xyz), nelem, _i(conn), _d(e), _d(nu),
                                    _l(ti), _l(tj), _d(tv), _l(nt), _l(bad))
    if st:
        raise OracleError(st, f"{ERRORS.get(st)} (element index {int(bad[0])})")
    k = int(nt[0])
    return ti[:k], tj[:k], tv[:k]


def reduce(ind_of, u0, n_free, ti, tj, tv, f_full):
    ind_of = np.ascontiguousarray(ind_of, np.int64)
    u0 = _f64(u0)
    f_full = _f64(f_full)
    ti = np.ascontiguousarray(ti, np.int64)
    tj = np.ascontiguousarray(tj, np.int64)
    tv = _f64(tv)
    nt = len(ti)
    oi = np.empty(nt, np.int64)
    oj = np.empty(nt, np.int64)
    ov = np.empty(nt, np.float64)
    nout = np.zeros(1, np.int64)
    rhs = np.zeros(n_free)
    _check(lib().or_reduce(len(ind_of), _l(ind_of), _d(u0), n_free, nt, _l(ti), _l(tj), _d(tv),
                           _d(f_full), _l(oi), _l(oj), _d(ov), _l(nout), _d(rhs)))
    k = int(nout[0])
    return oi[:k], oj[:k], ov[:k], rhs


@dataclass
class Csr:
    n: int
    indptr: np.ndarray
    indices: np.ndarray
    data: np.ndarray

    @property
    def nnz(self):
        return int(self.indptr[-1])

    def matvec(self, x):
        y = np.zeros(self.n)
        x = _f64(x)
        lib().or_csr_matvec(self.n, _l(self.indptr), _l(self.indices), _d(self.data), _d(x), _d(y))
        return y

    def to_dense(self):
        a = np.zeros((self.n, self.n))
        for i in range(self.n):
            s, t = self.indptr[i], self.indptr[i + 1]
            a[i, self.indices[s:t]] = self.data[s:t]
        return a

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.data, self.indices, self.indptr), shape=(self.n, self.n))


def csr_from_triplets(n, ti, tj, tv):
    """linalg.rs:10-37 (exact-zero filter, stable (i, j) sort, left-to-right sums)."""
    ti = np.ascontiguousarray(ti, np.int64)
    tj = np.ascontiguousarray(tj, np.int64)
    tv = _f64(tv)
    indptr = np.zeros(n + 1, np.int64)
    L = lib()
    _check(L.or_csr_from_triplets(n, len(ti), _l(ti), _l(tj), _d(tv), _l(indptr), None, None))
    nnz = int(indptr[n])
    indices = np.empty(max(nnz, 1), np.int64)
    data = np.empty(max(nnz, 1), np.float64)
    _check(L.or_csr_from_triplets(n, len(ti), _l(ti), _l(tj), _d(tv), _l(indptr), _l(indices),
                                  _d(data)))
    return Csr(n, indptr, indices[:nnz], data[:nnz])


def pcg(csr, b, tol=1e-8, max_iter=None):
    """linalg.rs:194-259. Returns (x, iters, residual)."""
    n = csr.n
    if max_iter is None:
        max_iter = max(4 * n, 200)
    b = _f64(b)
    x = np.zeros(n)
    it = np.zeros(1, np.int64)
    res = np.zeros(1)
    st = lib().or_pcg(n, _l(csr.indptr), _l(csr.indices), _d(csr.data), _d(b), tol, max_iter,
                      _d(x), _l(it), _d(res))
    if st == 6:
        raise OracleError(st, f"CG konvergierte nicht (Residuum {res[0]:.3e} nach {it[0]} "
                              "Iterationen).")
    _check(st)
    return x, int(it[0]), float(res[0])


def chol_solve(a, b):
    """linalg.rs:78-119 (dense lower Cholesky)."""
    a = np.array(a, dtype=np.float64, order="C", copy=True)
    n = a.shape[0]
    b = _f64(b)
    x = np.zeros(n)
    _check(lib().or_chol_solve(n, _d(a), _d(b), _d(x)))
    return x


def residual_of(csr, x, b):
    """linalg.rs:268-278."""
    d = csr.matvec(x) - b
    return float(math.sqrt(float(np.sum(d * d))))


# ---------------------------------------------------------------- INP reader
def _parse_f64(s):
    """inp.rs:100-141: Fortran floats (1.d0, 1.0D-3, 1., .5, 1.0+3)."""
    raw = s.strip()
    if not raw:
        return None
    for cand in (raw.rstrip("."), raw):
        try:
            return float(cand)
        except ValueError:
            pass
    b = raw.replace("d", "e").replace("D", "e").rstrip(".")
    try:
        return float(b)
    except ValueError:
        pass
    for i in range(1, len(b)):
        if b[i] in "+-" and (b[i - 1].isdigit() or b[i - 1] == "."):
            try:
                return float(b[:i] + "e" + b[i:])
            except ValueError:
                return None
    return None


def _parse_i32(s):
    t = s.strip()
    if re.fullmatch(r"[+-]?\d+", t):
        return int(t)
    f = _parse_f64(t)
    if f is None:
        raise ValueError(f"Keine ganze Zahl: {s}")
    return int(round(f))


def _round_half_away(x):
    """Rust's f64::round: halfway cases away from zero."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def _is_int(s):
    return re.fullmatch(r"[+-]?\d+", s.strip()) is not None


def _tokens(line):
    return [t for t in re.split(r"[,\s]+", line) if t]


def _is_kw(line):
    t = line.strip()
    return t.startswith("*") and not t.startswith("**")


def _keyword(line):
    t = line.strip().lstrip("﻿").upper()
    parts = t.split(",")
    kw = parts[0].strip()
    params = {}
    for p in parts[1:]:
        p = p.strip()
        if not p:
            continue
        if "=" in p:
            k, v = p.split("=", 1)
            params[k.strip()] = v.strip()
        else:
            params[p] = ""
    kw = {"*SOLIDSECTION": "*SOLID SECTION", "*ENDSTEP": "*END STEP", "*ELFILE": "*EL FILE",
          "*NODEFILE": "*NODE FILE", "*NODEOUTPUT": "*NODE OUTPUT",
          "*ELOUTPUT": "*EL OUTPUT"}.get(kw, kw)
    return kw, params


@dataclass
class Model:
    node_ids: list = field(default_factory=list)
    coords: list = field(default_factory=list)
    elements: list = field(default_factory=list)  # (id, kind_nn, nodes(list of ids), elset)
    nsets: dict = field(default_factory=dict)
    elsets: dict = field(default_factory=dict)
    materials: dict = field(default_factory=dict)  # name -> (E, nu)
    elset_material: dict = field(default_factory=dict)
    bcs: list = field(default_factory=list)  # (node id, dof0, value)
    cloads: list = field(default_factory=list)  # (node id, dof0, mag)
    dloads: list = field(default_factory=list)  # (elem id, face, mag)
    id_to_index: dict = field(default_factory=dict)
    deferred: list = field(default_factory=list)
    densities: dict = field(default_factory=dict)  # material name -> density (*DENSITY)
    nmodes: int = 0  # *FREQUENCY n (0: static step)
    nbuckle: int = 0  # *BUCKLE n (0: no buckling step)
    heading: str = ""  # *HEADING
    output: dict = field(default_factory=lambda: {"u": True, "s": True, "rf": True, "e": False,
                                                  "nt": False})  # *NODE FILE / *EL FILE
    nlgeom: bool = False  # *STEP, NLGEOM
    max_inc: int = 100  # *STEP, INC=
    ninc: int = 1  # increments of the static step (*STATIC: round(period / dt0), capped)

    def node_index(self, nid):
        return self.id_to_index[nid]

    def expand_nset(self, name):
        if _is_int(name):
            return [int(name)]
        up = name.upper()
        if up in self.nsets:
            return list(self.nsets[up])
        if up in ("NALL", "ALL"):
            return list(self.node_ids)
        raise ValueError(f"Unbekanntes Knotenset {name}")

    def material_name_for(self, elset):
        """model.rs:1391-1407: the section's material, else the only one, else STEEL."""
        m = self.elset_material.get(elset)
        if m is not None and m in self.materials:
            return m
        if len(self.materials) == 1:
            return next(iter(self.materials))
        if "STEEL" in self.materials:
            return "STEEL"
        raise ValueError(f"Kein Material für ELSET={elset}")

    def material_for(self, elset):
        return self.materials[self.material_name_for(elset)]

    def density_for(self, elset):
        return self.densities.get(self.material_name_for(elset), 0.0)

    # -- arrays for the hot path
    def mesh_arrays(self):
        """(nn, xyz[nnode,3], conn[nelem,nn] 0-based, E[nelem], nu[nelem])."""
        kinds = {el[1] for el in self.elements}
        if len(kinds) != 1 or next(iter(kinds)) not in (8, 20):
            raise ValueError("only all-C3D8 or all-C3D20 decks are on this path")
        nn = next(iter(kinds))
        xyz = np.asarray(self.coords, dtype=np.float64).reshape(-1, 3)
        conn = np.array([[self.id_to_index[i] for i in el[2]] for el in self.elements],
                        dtype=np.int32).reshape(-1, nn)
        mats = [self.material_for(el[3]) for el in self.elements]
        e = np.array([m[0] for m in mats], dtype=np.float64)
        nu = np.array([m[1] for m in mats], dtype=np.float64)
        return nn, xyz, conn, e, nu


_ELEM_NN = {"C3D8": 8, "C3D20": 20}


def parse_inp(text):
    """Restates inp.rs:383-1061 for the keyword subset on this path."""
    m = Model()
    lines = text.splitlines()
    n = len(lines)
    i = 0
    cur_mat = None

    def collect(j):
        toks = []
        while j < n:
            raw = lines[j]
            if raw.strip().startswith("**") or not raw.strip():
                j += 1
                continue
            if _is_kw(raw):
                break
            toks.extend(_tokens(raw))
            j += 1
        return toks, j

    def gen_ids(toks):
        ids = []
        k = 0
        while k + 1 < len(toks):
            a, b = _parse_i32(toks[k]), _parse_i32(toks[k + 1])
            if k + 2 < len(toks):
                s = _parse_i32(toks[k + 2]) or 1
                k += 3
            else:
                s = 1
                k += 2
            v = a
            if s > 0:
                while v <= b:
                    ids.append(v)
                    v += s
            else:
                while v >= b:
                    ids.append(v)
                    v += s
        return ids

    while i < n:
        raw = lines[i]
        if raw.strip().startswith("**") or not raw.strip() or not _is_kw(raw):
            i += 1
            continue
        kw, params = _keyword(raw)
        if kw == "*NODE":
            nset = params.get("NSET")
            i += 1
            added = []
            while i < n:
                raw = lines[i]
                if raw.strip().startswith("**") or not raw.strip():
                    i += 1
                    continue
                if _is_kw(raw):
                    break
                t = _tokens(raw)
                i += 1
                if not t:
                    continue
                nid = _parse_i32(t[0])
                xyz = [(_parse_f64(t[k]) if len(t) > k else 0.0) or 0.0 for k in (1, 2, 3)]
                m.node_ids.append(nid)
                m.coords.append(xyz)
                added.append(nid)
            if nset:
                m.nsets.setdefault(nset, []).extend(added)
        elif kw == "*ELEMENT":
            typ = params.get("TYPE")
            if typ not in _ELEM_NN:
                raise ValueError(f"Nicht unterstützter Elementtyp {typ} (this path: C3D8, C3D20).")
            nn = _ELEM_NN[typ]
            elset = params.get("ELSET", "EALL")
            toks, i = collect(i + 1)
            k = 0
            added = []
            while k + nn < len(toks):
                eid = _parse_i32(toks[k])
                nodes = [_parse_i32(t) for t in toks[k + 1:k + 1 + nn]]
                k += 1 + nn
                m.elements.append((eid, nn, nodes, elset))
                added.append(eid)
            m.elsets.setdefault(elset, []).extend(added)
        elif kw in ("*NSET", "*ELSET"):
            key = kw[1:]
            name = params.get(key)
            if name is None:
                raise ValueError(f"{kw} ohne {key}=")
            toks, i = collect(i + 1)
            if "GENERATE" in params:
                ids = gen_ids(toks)
            else:
                ids = []
                for t in toks:
                    try:
                        ids.append(_parse_i32(t))
                    except ValueError:
                        m.deferred.append(("NESTED_" + key, name, t.upper()))
            (m.nsets if key == "NSET" else m.elsets).setdefault(name, []).extend(ids)
        elif kw == "*MATERIAL":
            cur_mat = params.get("NAME", "MATERIAL-1")
            m.materials.setdefault(cur_mat, (0.0, 0.0))
            i += 1
        elif kw == "*ELASTIC":
            toks, i = collect(i + 1)
            if len(toks) < 2:
                raise ValueError("*ELASTIC benötigt E und nu")
            e = _parse_f64(toks[0])
            nu = _parse_f64(toks[1])
            if abs(e) <= 1e-18:
                cands = [v for v in (_parse_f64(t) for t in toks) if v is not None and abs(v) > 1e-18]
                e = cands[0] if cands else 1.0
            nu = min(max(nu, -0.49), 0.49)
            m.materials[cur_mat or "MATERIAL-1"] = (e, nu)
        elif kw == "*DENSITY":
            toks, i = collect(i + 1)
            if not toks:
                raise ValueError("*DENSITY ohne Wert")
            rho = _parse_f64(toks[0])
            if rho is None:
                raise ValueError(f"Keine Zahl: {toks[0]}")
            name = cur_mat or "MATERIAL-1"
            m.materials.setdefault(name, (0.0, 0.0))
            m.densities[name] = rho
        elif kw == "*FREQUENCY":
            toks, i = collect(i + 1)
            nm = 10
            if toks:
                try:
                    nm = _parse_i32(toks[0])
                except ValueError:
                    nm = 10
            m.nmodes = max(nm, 1)
        elif kw == "*HEADING":
            # inp.rs:410-417
            i += 1
            while i < n and (not lines[i].strip() or lines[i].strip().startswith("**")):
                i += 1
            if i < n and not _is_kw(lines[i]):
                m.heading = lines[i].strip()
                i += 1
        elif kw in ("*NODE FILE", "*NODE OUTPUT"):
            # inp.rs:1989-2004
            toks, i = collect(i + 1)
            m.output["u"] = m.output["rf"] = False
            for t in toks:
                t = t.strip().upper()
                if t == "U":
                    m.output["u"] = True
                elif t == "RF":
                    m.output["rf"] = True
                elif t == "NT":
                    m.output["nt"] = True
            if not m.output["u"] and not m.output["rf"]:
                m.output["u"] = True
        elif kw in ("*EL FILE", "*ELEMENT OUTPUT"):
            # inp.rs:2006-2021
            toks, i = collect(i + 1)
            m.output["s"] = m.output["e"] = False
            for t in toks:
                t = t.strip().upper()
                if t == "S":
                    m.output["s"] = True
                elif t == "E":
                    m.output["e"] = True
            if not m.output["s"] and not m.output["e"]:
                m.output["s"] = True
        elif kw == "*STEP":
            # inp.rs:953-982: NLGEOM (any value) switches the static step to NLGEOM
            if "NLGEOM" in params:
                m.nlgeom = True
                m.ninc = 1
            if "INC" in params:
                try:
                    v = _parse_i32(params["INC"])
                    if v > 0:
                        m.max_inc = v
                except ValueError:
                    pass
            i += 1
        elif kw == "*STATIC":
            # inp.rs:983-1061 without RIKS: initial increment and period
            toks, i = collect(i + 1)
            def pos(k):
                v = _parse_f64(toks[k]) if len(toks) > k else None
                return v if v is not None and v > 0.0 else 1.0
            dt0, period = pos(0), pos(1)
            n_from_time = max(int(_round_half_away(period / dt0)), 1)
            m.ninc = max(min(n_from_time, max(m.max_inc, 1)), 1)
        elif kw == "*BUCKLE":
            # inp.rs:1072-1081: n factors, 1 when absent or unreadable
            toks, i = collect(i + 1)
            nb = 1
            if toks:
                try:
                    nb = _parse_i32(toks[0])
                except ValueError:
                    nb = 1
            m.nbuckle = max(nb, 1)
        elif kw == "*SOLID SECTION":
            elset = params.get("ELSET", "EALL")
            if "MATERIAL" in params:
                m.elset_material[elset] = params["MATERIAL"]
            _, i = collect(i + 1)
        elif kw == "*BOUNDARY":
            i += 1
            while i < n:
                raw = lines[i]
                if raw.strip().startswith("**") or not raw.strip():
                    i += 1
                    continue
                if _is_kw(raw):
                    break
                card = _boundary_card(_tokens(raw))
                if card:
                    _push_boundary(m, *card)
                i += 1
        elif kw == "*CLOAD":
            toks, i = collect(i + 1)
            k = 0
            while k + 2 < len(toks):
                name = toks[k].upper()
                dof = _parse_i32(toks[k + 1])
                mag = _parse_f64(toks[k + 2])
                k += 3
                if dof < 1 or dof > 6:
                    continue
                if _is_int(name):
                    m.cloads.append((int(name), dof - 1, mag))
                elif m.nsets.get(name):
                    for nid in m.nsets[name]:
                        m.cloads.append((nid, dof - 1, mag))
                else:
                    m.deferred.append(("CL", name, dof - 1, mag))
        elif kw == "*DLOAD":
            toks, i = collect(i + 1)
            k = 0
            while k + 1 < len(toks):
                name = toks[k].upper()
                typ = toks[k + 1].upper()
                k += 2
                mt = re.fullmatch(r"P(\d+)", typ)
                if not mt:
                    raise ValueError(f"DLOAD-Typ {typ} is not on this path")
                face = int(mt.group(1))
                mag = _parse_f64(toks[k])
                k += 1
                if _is_int(name):
                    m.dloads.append((int(name), face, mag))
                elif m.elsets.get(name):
                    for eid in m.elsets[name]:
                        m.dloads.append((eid, face, mag))
                else:
                    m.deferred.append(("DL", name, face, mag))
        else:
            i += 1
    _compact(m)
    _expand_deferred(m)
    return m


def _boundary_card(toks):
    """inp.rs:204-230."""
    if len(toks) < 2:
        return None
    name = toks[0].upper()
    d1 = _parse_i32(toks[1])

    def dof_id(s):
        if not _is_int(s):
            return None
        v = int(s)
        return v if (1 <= v <= 6 or v == 11) else None

    if len(toks) == 2:
        return name, d1, d1, 0.0
    if len(toks) == 3:
        d2 = dof_id(toks[2])
        if d2 is not None:
            return name, d1, d2, 0.0
        return name, d1, d1, _parse_f64(toks[2]) or 0.0
    d2 = dof_id(toks[2])
    return name, d1, d1 if d2 is None else d2, _parse_f64(toks[3]) or 0.0


def _push_boundary(m, name, first, last, val):
    """inp.rs:232-268."""
    first = max(first, 1)
    last = max(last, first)
    if _is_int(name):
        ids = [int(name)]
    elif m.nsets.get(name):
        ids = list(m.nsets[name])
    else:
        for d in range(first, last + 1):
            if 1 <= d <= 6:
                m.deferred.append(("BC", name, d - 1, val))
        return
    for nid in ids:
        for d in range(first, last + 1):
            if 1 <= d <= 6:
                m.bcs.append((nid, d - 1, val))


def _compact(m):
    """model.rs:1234-1309: sort nodes by id (first wins), elements by id, rebind elsets."""
    pairs = sorted(zip(m.node_ids, m.coords), key=lambda p: p[0])
    ids, coords, seen = [], [], set()
    for nid, c in pairs:
        if nid in seen:
            continue
        seen.add(nid)
        ids.append(nid)
        coords.append(c)
    m.node_ids, m.coords = ids, coords
    m.id_to_index = {nid: k for k, nid in enumerate(ids)}
    m.elements.sort(key=lambda e: e[0])
    m.nsets["NALL"] = list(ids)
    m.elsets["EALL"] = [e[0] for e in m.elements]
    sectioned = set(m.elset_material)
    if sectioned:
        names = sorted(sectioned, key=lambda a: (0 if a in m.elset_material else 1, a))
        id_to_set = {}
        for name in names:
            for eid in m.elsets.get(name, []):
                id_to_set.setdefault(eid, name)
        m.elements = [(e[0], e[1], e[2], e[3] if e[3] in sectioned else id_to_set.get(e[0], e[3]))
                      for e in m.elements]


def _expand_deferred(m):
    """inp.rs:2256-2330 (the subset: nested sets, BC, CLOAD, DLOAD)."""
    for item in m.deferred:
        if item[0] == "NESTED_NSET":
            m.nsets.setdefault(item[1], []).extend(m.expand_nset(item[2]))
        elif item[0] == "NESTED_ELSET":
            m.elsets.setdefault(item[1], []).extend(m.elsets.get(item[2], []))
    for item in m.deferred:
        if item[0] == "BC":
            for nid in m.expand_nset(item[1]):
                m.bcs.append((nid, item[2], item[3]))
        elif item[0] == "CL":
            for nid in m.expand_nset(item[1]):
                m.cloads.append((nid, item[2], item[3]))
        elif item[0] == "DL":
            for eid in m.elsets.get(item[1], []):
                m.dloads.append((eid, item[2], item[3]))
    m.deferred = []


# ------------------------------------------------------ linear static driver
@dataclass
class LinearSetup:
    """Everything solve_linear builds before the solve (full DOF space)."""
    nn: int
    xyz: np.ndarray
    conn: np.ndarray
    e: np.ndarray
    nu: np.ndarray
    ndof: int
    f_full: np.ndarray
    prescribed: dict
    ind_of: np.ndarray
    u0: np.ndarray
    n_free: int
    trips: tuple  # full-space triplets (ti, tj, tv)


def linear_setup(nn, xyz, conn, e, nu, bcs, cloads, dloads, autospc=True):
    """analysis.rs:548-946 for an all-hex isotropic model.

    bcs: iterable of (node index, dof0, value) in deck order (last wins)
    cloads: iterable of (node index, dof0, mag)
    dloads: iterable of (element index, face, mag) in deck order
    """
    xyz = _f64(xyz).reshape(-1, 3)
    conn = np.ascontiguousarray(conn, np.int32).reshape(-1, nn)
    nnode = xyz.shape[0]
    nelem = conn.shape[0]
    ndof = 3 * nnode
    e = _f64(np.broadcast_to(e, (nelem,)))
    nu = _f64(np.broadcast_to(nu, (nelem,)))
    prescribed = {}
    for ni, d, v in bcs:  # analysis.rs:555-563, last BC wins
        if d < 3:
            prescribed[3 * ni + d] = float(v)
    # pin_unused_dofs (analysis.rs:2230-2355): isolated nodes
    attached = np.zeros(nnode, bool)
    attached[conn.ravel()] = True
    for ni in np.nonzero(~attached)[0]:
        for d in range(3):
            prescribed.setdefault(3 * int(ni) + d, 0.0)
    # element loop: pressure loads in element order, then dload order
    f_full = np.zeros(ndof)
    by_elem = {}
    for k, (el, face, mag) in enumerate(dloads):
        by_elem.setdefault(int(el), []).append((face, mag))
    for el in sorted(by_elem):
        nodes = conn[el]
        for face, mag in by_elem[el]:
            fe = hex_face_pressure(xyz[nodes], face, mag)
            for a in range(nn):  # scatter_fe, analysis.rs:1546-1553
                for d in range(3):
                    f_full[3 * nodes[a] + d] += fe[3 * a + d]
    ti, tj, tv = assemble_triplets(nn, xyz, conn, e, nu)
    # add_cloads (analysis.rs:2709-2764): per (node, dof) sums, then added once
    acc = {}
    for ni, d, mag in cloads:
        if d < 3:
            acc[(ni, d)] = acc.get((ni, d), 0.0) + mag * 1.0
    for (ni, d), mag in acc.items():
        f_full[3 * ni + d] += mag
    # autospc_null_diag (analysis.rs:2164-2228)
    if autospc:
        diag = np.zeros(ndof)
        sel = ti == tj
        for i, v in zip(ti[sel], tv[sel]):
            diag[i] += v
        max_d = float(np.max(np.abs(diag))) if ndof else 0.0
        if max_d >= 1e-30:
            tol = max_d * 1e-14
            for i in np.nonzero(np.abs(diag) <= tol)[0]:
                prescribed.setdefault(int(i), 0.0)
    # DofMap::build (constraint.rs:26-136) without MPCs
    u0 = np.zeros(ndof)
    is_ind = np.ones(ndof, bool)
    for d, v in prescribed.items():
        if d < ndof:
            u0[d] = v
            is_ind[d] = False
    ind_of = np.full(ndof, -1, np.int64)
    n_free = int(is_ind.sum())
    ind_of[is_ind] = np.arange(n_free)
    if n_free == 0 and not prescribed:
        raise ValueError("Modell ohne freie Freiheitsgrade.")
    return LinearSetup(nn, xyz, conn, e, nu, ndof, f_full, prescribed, ind_of, u0, n_free,
                       (ti, tj, tv))


@dataclass
class LinearResult:
    setup: LinearSetup
    csr: Csr
    rhs: np.ndarray
    x: np.ndarray
    u_full: np.ndarray
    rf_full: np.ndarray
    stress: np.ndarray  # nodal averaged, nnode x 6
    elem_nodal_stress: np.ndarray  # nelem x nn x 6
    von_mises: np.ndarray
    solver: str
    iters: int
    residual: float


def solve_kff(n, csr, rhs, solver="auto", tol=1e-8):
    """linalg.rs:311-404 with faer replaced by the dense / LAPACK stand-ins.

    solver: "cholesky" (reference chol_solve, n <= 900 as in the reference),
    "lapack" (numpy Cholesky, faer stand-in for n up to ~2e4), "pcg"
    (reference linalg::pcg), "auto" (cholesky if n <= 900 else lapack).
    """
    if solver == "auto":
        solver = "cholesky" if n <= 900 else "lapack"
    if solver == "cholesky":
        x = chol_solve(csr.to_dense(), rhs)
        return x, "Cholesky", 1, residual_of(csr, x, rhs)
    if solver == "lapack":
        a = csr.to_dense()
        lchol = np.linalg.cholesky(a)
        y = np.linalg.solve(lchol, rhs)
        x = np.linalg.solve(lchol.T, y)
        return x, "dense LAPACK Cholesky (faer stand-in)", 1, residual_of(csr, x, rhs)
    if solver == "superlu":
        import scipy.sparse.linalg as sla
        x = sla.spsolve(csr.to_scipy().tocsc(), rhs)
        return x, "scipy SuperLU (faer stand-in)", 1, residual_of(csr, x, rhs)
    if solver == "pcg":
        x, it, res = pcg(csr, rhs, tol, max(4 * n, 200))
        return x, "PCG", it, res
    raise ValueError(solver)


def linear_solve(setup: LinearSetup, solver="auto", tol=1e-8):
    """analysis.rs:1018-1162: reduce, solve_kff, reconstruct, reactions, stresses."""
    s = setup
    ti, tj, tv = s.trips
    oi, oj, ov, rhs = reduce(s.ind_of, s.u0, s.n_free, ti, tj, tv, s.f_full)
    csr = csr_from_triplets(s.n_free, oi, oj, ov)
    x, name, iters, res = solve_kff(s.n_free, csr, rhs, solver, tol)
    u_full = s.u0.copy()
    free = s.ind_of >= 0
    u_full[free] = u_full[free] + 1.0 * x[s.ind_of[free]]
    ku = np.zeros(s.ndof)
    np.add.at(ku, ti, tv * u_full[tj])  # same order as the sequential loop per index
    rf = ku - s.f_full
    stress, sn_all, vm = nodal_stress(s.nn, s.xyz, s.conn, s.e, s.nu, u_full)
    return LinearResult(s, csr, rhs, x, u_full, rf, stress, sn_all, vm, name, iters, res)


def nodal_stress(nn, xyz, conn, e, nu, u_full):
    """analysis.rs:1061-1162: per-element nodal stress, plain mean per node."""
    nnode = xyz.shape[0]
    acc = np.zeros((nnode, 6))
    cnt = np.zeros(nnode)
    nelem = conn.shape[0]
    sn_all = np.zeros((nelem, nn, 6))
    u3 = u_full.reshape(-1, 3)
    for el in range(nelem):
        nodes = conn[el]
        ue = u3[nodes].ravel()
        sn = hex_nodal_stress(xyz[nodes], ue, e[el], nu[el])
        sn_all[el] = sn
        for a in range(nn):
            acc[nodes[a]] += sn[a]
            cnt[nodes[a]] += 1.0
    stress = np.zeros((nnode, 6))
    has = cnt > 0
    stress[has] = acc[has] / cnt[has, None]
    vm = np.array([von_mises(s) for s in stress])
    return stress, sn_all, vm


def deck_setup(model: Model):
    """Model (from parse_inp) -> LinearSetup, mapping ids to indices."""
    nn, xyz, conn, e, nu = model.mesh_arrays()
    bcs = [(model.node_index(nid), d, v) for nid, d, v in model.bcs]
    cl = [(model.node_index(nid), d, mag) for nid, d, mag in model.cloads]
    eidx = {el[0]: k for k, el in enumerate(model.elements)}
    dl = [(eidx[eid], face, mag) for eid, face, mag in model.dloads]
    return linear_setup(nn, xyz, conn, e, nu, bcs, cl, dl)


def solve_deck(text, solver="auto", tol=1e-8):
    """solve_native for the decks on this path: parse -> setup -> solve."""
    m = parse_inp(text)
    return m, linear_solve(deck_setup(m), solver, tol)


# ------------------------------------------------------------- eigenvalues
def lumped_mass(nn, xyz, conn, density):
    """analysis.rs:669-717 for solid elements: mass = density |volume| of each element (the
    volume summed with the stiffness' Gauss loop), 1/nn of it on each translational DOF of each
    node, elements in order."""
    xyz = _f64(xyz).reshape(-1, 3)
    conn = np.ascontiguousarray(conn, np.int32).reshape(-1, nn)
    m_full = np.zeros(3 * xyz.shape[0])
    w = 1.0 / nn
    for el in range(conn.shape[0]):
        rho = float(density[el])
        if not abs(rho) > 0.0:
            continue
        _, vol = hex_stiffness(xyz[conn[el]], 1.0, 0.0)
        if not abs(vol) > 0.0:
            continue
        mnode = rho * abs(vol) * w
        for a in conn[el]:
            for d in range(3):
                m_full[3 * a + d] += mnode
    return m_full


def lcg_start_vectors(n, nvec, m_diag):
    """eigen.rs:24-34: the seeded start block (LCG, zero where the mass is not positive)."""
    x = np.zeros((nvec, n))
    rng = 1
    mask = (1 << 64) - 1
    for v in range(nvec):
        for i in range(n):
            rng = (rng * 6364136223846793005 + 1) & mask
            x[v, i] = ((rng >> 33) / float(1 << 31)) - 1.0
            if m_diag[i] <= 0.0:
                x[v, i] = 0.0
    return x


def m_orthonormalize(x, m):
    """eigen.rs:181-206: modified Gram-Schmidt in the diagonal M inner product."""
    for a in range(x.shape[0]):
        for b in range(a):
            dot = float(np.sum(x[a] * m * x[b]))
            x[a] -= dot * x[b]
        s = math.sqrt(float(np.sum(x[a] * m * x[a])))
        if s > 1e-30:
            x[a] /= s
    return x


def jacobi_eigen(a):
    """eigen.rs:208-266: cyclic-by-largest Jacobi rotations of a dense symmetric matrix.
    Returns (eigenvalues, eigenvectors as columns)."""
    a = np.array(a, dtype=np.float64, copy=True)
    n = a.shape[0]
    v = np.eye(n)
    for _ in range(40 * max(n, 1)):
        best, p, q = 0.0, 0, 1
        for i in range(n):
            for j in range(i + 1, n):
                if abs(a[i, j]) > best:
                    best, p, q = abs(a[i, j]), i, j
        if best < 1e-14:
            break
        app, aqq, apq = a[p, p], a[q, q], a[p, q]
        tau = (aqq - app) / (2.0 * apq)
        t = (1.0 if tau >= 0.0 else -1.0) / (abs(tau) + math.sqrt(1.0 + tau * tau))
        c = 1.0 / math.sqrt(1.0 + t * t)
        s_ = t * c
        for i in range(n):
            if i != p and i != q:
                aip, aiq = a[i, p], a[i, q]
                a[i, p] = a[p, i] = c * aip - s_ * aiq
                a[i, q] = a[q, i] = s_ * aip + c * aiq
        a[p, p] = c * c * app - 2.0 * s_ * c * apq + s_ * s_ * aqq
        a[q, q] = s_ * s_ * app + 2.0 * s_ * c * apq + c * c * aqq
        a[p, q] = a[q, p] = 0.0
        vp, vq = v[:, p].copy(), v[:, q].copy()
        v[:, p] = c * vp - s_ * vq
        v[:, q] = s_ * vp + c * vq
    return np.diag(a).copy(), v


def subspace_gen(csr, m_diag, nmodes, solve=None):
    """eigen.rs:12-111: lowest nmodes of K phi = lambda M phi (diagonal M) by inverse subspace
    iteration: nvec = nmodes + 4 vectors, 40 sweeps at most, Rayleigh-Ritz with Jacobi,
    stop when every Ritz value moved by at most 1e-8 (1 + |lambda|) after sweep 5.
    solve(B) returns K^-1 B for a block B (rows = vectors); the reference calls solve_kff
    (faer LLT) here, the default is a dense Cholesky (small n)."""
    n = csr.n
    if n == 0 or nmodes == 0:
        raise ValueError("Eigenproblem ohne freie DOF / ohne Moden.")
    nvec = min(max(min(nmodes, n) + 4, min(nmodes, n)), n)
    m = _f64(m_diag)
    if solve is None:
        lchol = np.linalg.cholesky(csr.to_dense())

        def solve(B):
            y = np.linalg.solve(lchol, B.T)
            return np.linalg.solve(lchol.T, y).T
    x = m_orthonormalize(lcg_start_vectors(n, nvec, m), m)
    values = np.zeros(nvec)
    k = csr.to_scipy()
    for it in range(40):
        y = m_orthonormalize(solve(x * m[None, :]), m)
        kr = y @ (k @ y.T)
        evals, evecs = jacobi_eigen(kr)
        nx = m_orthonormalize(evecs.T @ y, m)
        conv = bool(np.all(np.abs(evals - values) <= 1e-8 * (1.0 + np.abs(evals))))
        values = evals
        x = nx
        if conv and it > 4:
            break
    order = np.argsort(values, kind="stable")[:min(nmodes, nvec)]
    keep = [i for i in order if np.isfinite(values[i])]
    if not keep:
        raise ValueError("Keine Eigenwerte gefunden.")
    return values[keep], x[keep], it + 1


def frequency_deck(text):
    """*FREQUENCY for an all-hex deck (analysis.rs:955-1000): K_ff, lumped M restricted to
    the free DOFs, subspace iteration, f = sqrt(lambda) / (2 pi)."""
    model = parse_inp(text)
    s = deck_setup(model)
    dens = np.array([model.density_for(el[3]) for el in model.elements])
    m_full = lumped_mass(s.nn, s.xyz, s.conn, dens)
    m_ind = np.zeros(s.n_free)
    free = s.ind_of >= 0
    np.add.at(m_ind, s.ind_of[free], 1.0 * 1.0 * m_full[free])
    if np.all(m_ind <= 0.0):
        raise ValueError("*FREQUENCY: *DENSITY fehlt oder Masse ist null.")
    ti, tj, tv = s.trips
    oi, oj, ov, _ = reduce(s.ind_of, s.u0, s.n_free, ti, tj, tv, s.f_full)
    csr = csr_from_triplets(s.n_free, oi, oj, ov)
    vals, vecs, sweeps = subspace_gen(csr, m_ind, max(model.nmodes, 1))
    freqs = np.array([math.sqrt(v) / (2.0 * math.pi) if v > 0.0 else 0.0 for v in vals])
    return model, s, csr, m_ind, vals, vecs, freqs, sweeps


# ------------------------------------------------------------------ buckling
def element_mean_stress(sn_all):
    """analysis.rs:1145-1151: stress_gp of an element = sum over its nodes of the
    extrapolated nodal stress / nn, in node order."""
    nelem, nn, _ = sn_all.shape
    mean = np.zeros((nelem, 6))
    for a in range(nn):
        mean += sn_all[:, a, :] / float(nn)
    return mean


def kg_triplets(nn, xyz, conn, mean):
    """analysis.rs:1176-1247: per-element Kg (hex8_kg / hex20_kg at the element's mean
    stress), nonzero entries in element order then row-major, full DOF space."""
    m = 3 * nn
    ti, tj, tv = [], [], []
    for el in range(conn.shape[0]):
        kg = hex_kg(xyz[conn[el]], mean[el])
        g = (3 * conn[el][:, None] + np.arange(3)[None, :]).ravel()
        ii, jj = np.nonzero(np.abs(kg) > 0.0)
        ti.append(g[ii])
        tj.append(g[jj])
        tv.append(kg[ii, jj])
    if not ti:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0)
    return (np.concatenate(ti).astype(np.int64), np.concatenate(tj).astype(np.int64),
            np.concatenate(tv))


def subspace_ab(k_csr, a_csr, nmodes, solve=None):
    """eigen.rs:113-179: K phi = lambda A phi by 30 sweeps of inverse subspace iteration
    with Euclidean Gram-Schmidt (start block: the LCG from seed 7, each vector scaled to
    unit length), then Rayleigh quotients v.Kv / v.Av (|v.Av| < 1e-30 dropped), sorted
    by |lambda| (stable). solve(B) returns K^-1 B (rows = vectors); default: dense
    Cholesky (small n)."""
    n = k_csr.n
    if n == 0 or nmodes == 0:
        raise ValueError("Eigenproblem ohne freie DOF.")
    nvec = min(max(min(nmodes, n) + 4, min(nmodes, n)), n)
    if solve is None:
        lchol = np.linalg.cholesky(k_csr.to_dense())

        def solve(B):
            y = np.linalg.solve(lchol, B.T)
            return np.linalg.solve(lchol.T, y).T
    x = np.zeros((nvec, n))
    rng = 7
    mask = (1 << 64) - 1
    for v in range(nvec):
        for i in range(n):
            rng = (rng * 6364136223846793005 + 1) & mask
            x[v, i] = ((rng >> 33) / float(1 << 31)) - 1.0
        x[v] /= max(math.sqrt(float(np.sum(x[v] * x[v]))), 1e-30)
    a = a_csr.to_scipy()
    k = k_csr.to_scipy()
    for _ in range(30):
        y = solve(np.asarray((a @ x.T).T))
        for p in range(nvec):
            for q in range(p):
                y[p] -= float(y[p] @ y[q]) * y[q]
            y[p] /= max(math.sqrt(float(y[p] @ y[p])), 1e-30)
        x = y
    pairs = []
    for v in x:
        num = float(v @ (k @ v))
        den = float(v @ (a @ v))
        if abs(den) < 1e-30:
            continue
        pairs.append((num / den, v))
    pairs.sort(key=lambda t: abs(t[0]))
    take = min(nmodes, len(pairs))
    return np.array([p[0] for p in pairs[:take]]), np.array([p[1] for p in pairs[:take]])


def buckle_deck(text):
    """*BUCKLE for an all-hex deck (analysis.rs:1176-1260): the static solution of the
    step's loads, element mean stresses, Kg_ff with A = -Kg_ff, subspace_ab on (K_ff, A).
    Returns (model, linear result, kg csr (unnegated), factors, vectors)."""
    model = parse_inp(text)
    s = deck_setup(model)
    res = linear_solve(s)
    mean = element_mean_stress(res.elem_nodal_stress)
    ti, tj, tv = kg_triplets(s.nn, s.xyz, s.conn, mean)
    oi, oj, ov, _ = reduce(s.ind_of, s.u0, s.n_free, ti, tj, tv, s.f_full)
    kg = csr_from_triplets(s.n_free, oi, oj, ov)
    a = Csr(kg.n, kg.indptr, kg.indices, -kg.data)
    vals, vecs = subspace_ab(res.csr, a, max(model.nbuckle, 1))
    return model, res, kg, vals, vecs


# ------------------------------------------------------------------- NLGEOM
def hex_nl(xyz0, ue, e, nu):
    """nlgeom.rs:127-273 (hex8_nl / hex20_nl): St. Venant-Kirchhoff tangent, internal force,
    volume-averaged Cauchy and Green-Lagrange, reference volume of one hex."""
    xyz0 = _f64(xyz0)
    nn = xyz0.shape[0]
    ke = np.zeros((3 * nn) ** 2)
    fe = np.zeros(3 * nn)
    cau = np.zeros(6)
    gl = np.zeros(6)
    vol = np.zeros(1)
    _check(lib().or_hex_nl(nn, _d(xyz0), _d(_f64(ue)), float(e), float(nu), _d(ke), _d(fe),
                           _d(cau), _d(gl), _d(vol)))
    return ke.reshape(3 * nn, 3 * nn), fe, cau, gl, float(vol[0])


def hex20_nodal_nl(xyz0, ue, e, nu):
    """nlgeom.rs:274-308: nodal Cauchy and Green-Lagrange of a hex20 (20 x 6 each)."""
    cs = np.zeros(120)
    gs = np.zeros(120)
    _check(lib().or_hex20_nodal_nl(_d(_f64(xyz0)), _d(_f64(ue)), float(e), float(nu), _d(cs),
                                   _d(gs)))
    return cs.reshape(20, 6), gs.reshape(20, 6)


@dataclass
class NlResult:
    u_full: np.ndarray
    rf_full: np.ndarray  # f_int - f_ext over all DOFs
    stress: np.ndarray  # nodal Cauchy (nnode x 6)
    strain: np.ndarray  # nodal Green-Lagrange (nnode x 6)
    von_mises: np.ndarray
    stress_gp: np.ndarray  # element means (nelem x 6)
    iters: int
    ninc: int
    residual: float
    n_free: int
    ind_of: np.ndarray


def nl_assemble(nn, xyz, conn, e, nu, u_full):
    """The element loop of solve_continuum_newton (analysis.rs:4474-4545): per element the
    NL tangent and internal force; f_int summed in element order, tangent triplets (|v| > 0,
    row-major) in the full DOF space, element Cauchy / Green-Lagrange means."""
    nelem = conn.shape[0]
    m = 3 * nn
    f_int = np.zeros(3 * xyz.shape[0])
    u3 = u_full.reshape(-1, 3)
    ti, tj, tv = [], [], []
    cau = np.zeros((nelem, 6))
    gl = np.zeros((nelem, 6))
    for el in range(nelem):
        nodes = conn[el]
        ke, fe, c, g, _ = hex_nl(xyz[nodes], u3[nodes].ravel(), e[el], nu[el])
        gd = (3 * nodes[:, None] + np.arange(3)[None, :]).ravel()
        for i in range(m):
            f_int[gd[i]] += fe[i]
        ii, jj = np.nonzero(np.abs(ke) > 0.0)
        ti.append(gd[ii])
        tj.append(gd[jj])
        tv.append(ke[ii, jj])
        cau[el] = c
        gl[el] = g
    return (f_int, (np.concatenate(ti).astype(np.int64), np.concatenate(tj).astype(np.int64),
                    np.concatenate(tv)), cau, gl)


def nl_fext(nn, xyz, conn, cloads, dloads, u_full=None):
    """assemble_fext_u (analysis.rs:3222-3300) for CLOAD and face pressure: nodal loads summed
    per (node, dof) and added first, then each element's pressure faces (on the deformed
    coordinates when u_full is given) in element order."""
    f = np.zeros(3 * xyz.shape[0])
    acc = {}
    for ni, d, mag in cloads:
        if d < 3:
            acc[(ni, d)] = acc.get((ni, d), 0.0) + mag * 1.0
    for (ni, d), mag in acc.items():
        f[3 * ni + d] += mag
    by_elem = {}
    for el, face, mag in dloads:
        by_elem.setdefault(int(el), []).append((face, mag))
    for el in sorted(by_elem):
        nodes = conn[el]
        x = xyz[nodes].copy()
        if u_full is not None:
            x = x + u_full.reshape(-1, 3)[nodes]
        for face, mag in by_elem[el]:
            fe = hex_face_pressure(x, face, mag)
            for a in range(nn):
                for d in range(3):
                    f[3 * nodes[a] + d] += fe[3 * a + d]
    return f


def _norm_seq(v):
    s = 0.0
    for x in v:
        s += x * x
    return math.sqrt(s)


def nlgeom_deck(text, newton_tol=1e-8, max_newton=25, solve=None):
    """solve_continuum_newton (analysis.rs:4313-4860) for an all-hex St. Venant-Kirchhoff deck
    without contact, pretension, rigid bodies, plasticity or amplitudes: increments of
    1 / ninc (grown by 1.5x after a success, halved on failure), full Newton with the
    consistent tangent and follower pressure, prescribed values ramped with the load factor,
    convergence at ||r_free|| < newton_tol (1 + ||f_inc||). solve(csr, rhs) defaults to a
    dense LAPACK solve (small decks)."""
    model = parse_inp(text)
    if not model.nlgeom:
        raise ValueError("the deck has no *STEP, NLGEOM")
    nn, xyz, conn, e, nu = model.mesh_arrays()
    nnode, nelem, ndof = xyz.shape[0], conn.shape[0], 3 * xyz.shape[0]
    cloads = [(model.node_index(nid), d, mag) for nid, d, mag in model.cloads]
    eidx = {el[0]: k for k, el in enumerate(model.elements)}
    dloads = [(eidx[eid], face, mag) for eid, face, mag in model.dloads]
    prescribed = {}
    for nid, d, v in model.bcs:
        if d < 3:
            prescribed[3 * model.node_index(nid) + d] = float(v)
    attached = np.zeros(nnode, bool)
    attached[conn.ravel()] = True
    for ni in np.nonzero(~attached)[0]:
        for d in range(3):
            prescribed.setdefault(3 * int(ni) + d, 0.0)
    is_ind = np.ones(ndof, bool)
    u_target = np.zeros(ndof)
    for d, v in prescribed.items():
        is_ind[d] = False
        u_target[d] = v
    ind_of = np.full(ndof, -1, np.int64)
    n_free = int(is_ind.sum())
    ind_of[is_ind] = np.arange(n_free)
    if n_free == 0:
        raise ValueError("NLGEOM Kontinuum: keine freien DOF.")
    if solve is None:
        def solve(csr, rhs):
            return np.linalg.solve(csr.to_dense(), rhs)
    f_tgt = nl_fext(nn, xyz, conn, cloads, dloads)
    f_0 = np.zeros(ndof)
    ninc = max(model.ninc, 1)
    u_full = np.zeros(ndof)
    lam, dt, ninc_done, retries = 0.0, 1.0 / ninc, 0, 0
    f_now = f_0.copy()
    iters, residual = 0, 0.0
    last_fint = np.zeros(ndof)
    last_cau = np.zeros((nelem, 6))
    last_gl = np.zeros((nelem, 6))
    pres = sorted(prescribed)
    while lam < 1.0 - 1e-14:
        lam_new = min(lam + dt, 1.0)
        u_work = u_full.copy()
        for d in pres:
            u_work[d] = (1.0 - lam_new) * 0.0 + lam_new * u_target[d]
        f_inc = (1.0 - lam_new) * f_0 + lam_new * f_tgt
        last_du = np.zeros(ndof)
        have_du, inc_ok, inc_err = False, False, None
        it = 0
        while it < max_newton:
            iters = it + 1
            if dloads:
                fd = nl_fext(nn, xyz, conn, cloads, dloads, u_work)
                f_inc = (1.0 - lam_new) * f_0 + lam_new * fd
            try:
                f_int, trips, cau, gl = nl_assemble(nn, xyz, conn, e, nu, u_work)
            except OracleError as ex:
                inc_err = str(ex)
                if ex.code == 10 and have_du:
                    u_work -= 0.5 * last_du
                    last_du *= 0.5
                    for d in pres:
                        u_work[d] = (1.0 - lam_new) * 0.0 + lam_new * u_target[d]
                    if _norm_seq(last_du) > 1e-18:
                        inc_err = None
                        it += 1
                        continue
                break
            last_fint = f_int
            last_cau, last_gl = cau, gl
            r = f_inc - f_int
            ti, tj, tv = trips
            oi, oj, ov, rhs = reduce(ind_of, np.zeros(ndof), n_free, ti, tj, tv, r)
            residual = _norm_seq(rhs)
            fref = _norm_seq(f_inc)
            if residual < newton_tol * (1.0 + fref):
                inc_ok = True
                break
            if it + 1 == max_newton:
                inc_err = f"Newton konvergierte nicht (r={residual:.3e} nach {iters} Iterationen)."
                break
            x = solve(csr_from_triplets(n_free, oi, oj, ov), rhs)
            du_full = np.zeros(ndof)
            du_full[is_ind] = x[ind_of[is_ind]]
            u_work = u_work + du_full
            for d in pres:
                u_work[d] = (1.0 - lam_new) * 0.0 + lam_new * u_target[d]
            last_du = du_full
            have_du = True
            if _norm_seq(x) < 1e-14:
                inc_ok = True
                break
            it += 1
        if inc_ok:
            u_full = u_work
            lam = lam_new
            f_now = f_inc
            ninc_done += 1
            retries = 0
            remain = 1.0 - lam
            if remain > 1e-14:
                dt = min(dt * 1.5, 1.0 / ninc, remain)
        else:
            retries += 1
            dt *= 0.5
            if dt < 1e-6 or retries > 16:
                raise ValueError(f"NLGEOM Inkrement bei λ={lam:.4f} fehlgeschlagen ({inc_err}).")
    rf = last_fint - f_now
    # nodal Cauchy / Green-Lagrange (analysis.rs:4745-4799): C3D20 from its 27 points,
    # C3D8 the element mean on every node; plain mean per node in element order
    accs = np.zeros((nnode, 6))
    gacc = np.zeros((nnode, 6))
    cnt = np.zeros(nnode)
    stress_gp = np.zeros((nelem, 6))
    u3 = u_full.reshape(-1, 3)
    for el in range(nelem):
        nodes = conn[el]
        if nn == 20:
            cs, gs = hex20_nodal_nl(xyz[nodes], u3[nodes].ravel(), e[el], nu[el])
            mean = np.zeros(6)
            for a in range(nn):
                mean += cs[a] / float(nn)
        else:
            cs = np.repeat(last_cau[el][None, :], nn, 0)
            gs = np.repeat(last_gl[el][None, :], nn, 0)
            mean = last_cau[el]
        for a in range(nn):
            accs[nodes[a]] += cs[a]
            gacc[nodes[a]] += gs[a]
            cnt[nodes[a]] += 1.0
        stress_gp[el] = mean
    stress = np.zeros((nnode, 6))
    strain = np.zeros((nnode, 6))
    has = cnt > 0
    stress[has] = accs[has] / cnt[has, None]
    strain[has] = gacc[has] / cnt[has, None]
    vm = np.array([von_mises(sv) for sv in stress])
    return NlResult(u_full, rf, stress, strain, vm, stress_gp, iters, ninc_done, residual,
                    n_free, ind_of)


# ------------------------------------------------------------------ writers
def rust_exp(v, prec, width):
    """Rust's "{:W.PE}" for an f64: the exponent as a plain integer (no '+', no padding)."""
    if math.isnan(v):
        body = "NaN"
    elif math.isinf(v):
        body = "inf" if v > 0 else "-inf"
    else:
        mant, ex = format(v, f".{prec}E").split("E")
        body = f"{mant}E{int(ex)}"
    return body.rjust(width)


def _powi(a, b):
    """f64::powi with a variable exponent: binary exponentiation, one reciprocal."""
    recip = b < 0
    r = 1.0
    while True:
        if b & 1:
            r *= a
        b = int(b / 2)
        if b == 0:
            break
        a *= a
    return 1.0 / r if recip else r


def frd_e12(v):
    """frd.rs:7-40: Fortran ES12.5."""
    if not math.isfinite(v) or v == 0.0:
        return " 0.00000E+00"
    sign = "-" if math.copysign(1.0, v) < 0 else " "
    av = abs(v)
    ex = math.floor(math.log10(av))
    p10 = _powi(10.0, ex)
    mant = av / p10 if p10 != 0.0 else math.inf  # IEEE: a subnormal over 10^-324 (= 0)
    if mant >= 10.0:
        mant /= 10.0
        ex += 1
    elif mant < 1.0:
        mant *= 10.0
        ex -= 1
    x = mant * 1e5
    scaled = x if math.isinf(x) else float(math.floor(x + 0.5))  # f64::round, x >= 0
    if scaled >= 1e6:
        scaled /= 10.0
        ex += 1
    s = f"{sign}{scaled / 1e5:.5f}E{ex:+03d}"
    return s[:12] if len(s) >= 12 else s.rjust(12)


def write_dat(model, u_full, rf, stress_gp=None, nodal_stress=None, freqs=None):
    """dat.rs:3-94 (static / frequency result of a solid deck)."""
    ids = model.node_ids
    u3 = np.asarray(u_full).reshape(-1, 3)
    r3 = np.asarray(rf).reshape(-1, 3)
    t0 = rust_exp(0.0, 7, 14)
    o = [f"\n displacements (vx,vy,vz) for set NALL and time  {t0}\n\n"]
    for k, nid in enumerate(ids):
        o.append(f"{nid:10d}" + "".join(" " + rust_exp(float(x), 6, 14) for x in u3[k]) + "\n")
    o.append(f"\n forces (fx,fy,fz) for set NALL and time  {t0}\n\n")
    for k, nid in enumerate(ids):
        r = r3[k]
        if abs(r[0]) + abs(r[1]) + abs(r[2]) > 1e-14:
            o.append(f"{nid:10d}" + "".join(" " + rust_exp(float(x), 6, 14) for x in r) + "\n")
    o.append("\n stresses (elem, integ.pnt.,sxx,syy,szz,sxy,syz,szx) for set EALL and time  "
             f"{t0}\n\n")
    if stress_gp is not None:
        for k, el in enumerate(model.elements):
            o.append(f"{el[0]:10d} {1:4d}" +
                     "".join(" " + rust_exp(float(x), 6, 14) for x in stress_gp[k]) + "\n")
    o.append("\n")
    if nodal_stress is not None:
        o.append("\n stresses (node,sxx,syy,szz,sxy,syz,szx) for set NALL\n\n")
        for k, nid in enumerate(ids):
            o.append(f"{nid:10d}" +
                     "".join(" " + rust_exp(float(x), 6, 14) for x in nodal_stress[k]) + "\n")
        o.append("\n")
    if freqs is not None and len(freqs):
        o.append("\n EIGENVALUE OUTPUT\n\n MODE  FREQUENCY (CYCLES/TIME)\n\n")
        for i, f in enumerate(freqs):
            o.append(f"{i + 1:10d} {rust_exp(float(f), 6, 14)}\n")
        o.append("\n")
    return "".join(o)


def _put_int(line, start, width, v):
    t = f"{v:>{width}}"
    t = t[len(t) - width:] if len(t) > width else t
    return line[:start + width - len(t)] + t + line[start + width:]


def frd_1pstep(counter, iinc, istep):
    """frd.rs:57-64."""
    t = "    1PSTEP" + " " * 60
    t = _put_int(t, 24, 12, counter)
    t = _put_int(t, 36, 12, iinc)
    return _put_int(t, 48, 12, istep)


def frd_100cl(iinc, time, nout):
    """frd.rs:81-92."""
    t = "  100CL" + " " * 68
    t = _put_int(t, 7, 5, 100 + iinc)
    t = t[:12] + frd_e12(time) + t[24:]
    t = _put_int(t, 24, 12, nout)
    t = t[:57] + "0" + t[58:]
    t = _put_int(t, 58, 5, iinc)
    return t[:74] + "1"


_FRD20 = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 16, 17, 18, 19, 12, 13, 14, 15]


def write_frd(model, u_full, stress=None, rf=None, strain=None):
    """frd.rs:173-452 for one static frame of a solid deck."""
    h = model.heading if model.heading else "Axia"
    o = [f"    1C{h[:66]}\n",
         "    1UDATE              24.September.2026\n", "    1UTIME              21:00:00\n",
         "    1UHOST              axia\n", "    1UPGM               Axia FEM\n",
         "    1UVERSION           1.31.10\n",
         "    1UCODE              CalculiX-compatible Axia FEM\n"]
    used = set()
    for el in model.elements:
        used.update(el[2])
    out = [(nid, k) for k, nid in enumerate(model.node_ids) if nid in used]
    if not out:
        out = list(enumerate(model.node_ids))
        out = [(nid, k) for k, nid in out]
    nn = model.elements[0][1]

    def header(kind, n):
        return f"    {kind}C{'':18}{n:>12}{1:>38}\n"

    o.append(header("2", len(out)))
    for nid, k in out:
        c = model.coords[k]
        o.append(f" -1{nid:10d}{frd_e12(c[0])}{frd_e12(c[1])}{frd_e12(c[2])}\n")
    o.append(" -3\n")
    o.append(header("3", len(model.elements)))
    ty = 1 if nn == 8 else 4
    for el in model.elements:
        o.append(f" -1{el[0]:10d}{ty:5d}{0:5d}{1:5d}\n")
        nodes = list(el[2])
        if nn == 20:
            nodes = [nodes[a] for a in _FRD20]
        for a0 in range(0, len(nodes), 10):
            o.append(" -2" + "".join(f"{x:10d}" for x in nodes[a0:a0 + 10]) + "\n")
    o.append(" -3\n")
    nnode = len(model.node_ids)
    u3 = np.asarray(u_full).reshape(-1, 3)
    s6 = np.zeros((nnode, 6)) if stress is None else np.asarray(stress).reshape(-1, 6)
    r3 = np.zeros((nnode, 3)) if rf is None else np.asarray(rf).reshape(-1, 3)
    counter = [0]

    def block(name, ncomp, comps, last_all, vals):
        counter[0] += 1
        o.append(frd_1pstep(counter[0], 1, 1) + "\n")
        o.append(frd_100cl(1, 1.0, len(out)) + "\n")
        o.append(f" -4  {name:<8}{ncomp:5d}    1\n")
        for i, (cn, typ, a, b) in enumerate(comps):
            tail = "    1ALL" if last_all and i == len(comps) - 1 else ""
            o.append(f" -5  {cn:<8}    1{typ:5d}{a:5d}{b:5d}{tail}\n")
        nv = ncomp - 1 if last_all else ncomp
        for nid, k in out:
            o.append(f" -1{nid:10d}" + "".join(frd_e12(float(vals[k][q])) for q in range(nv)) + "\n")
        o.append(" -3\n")

    if model.output["u"]:
        block("DISP", 4, [("D1", 2, 1, 0), ("D2", 2, 2, 0), ("D3", 2, 3, 0), ("ALL", 2, 0, 0)],
              True, u3)
    if model.output["s"]:
        block("STRESS", 6, [("SXX", 4, 1, 1), ("SYY", 4, 2, 2), ("SZZ", 4, 3, 3),
                            ("SXY", 4, 1, 2), ("SYZ", 4, 2, 3), ("SZX", 4, 3, 1)], False, s6)
    if model.output["rf"]:
        block("FORC", 4, [("F1", 2, 1, 0), ("F2", 2, 2, 0), ("F3", 2, 3, 0), ("ALL", 2, 0, 0)],
              True, r3)
    if model.output["e"]:
        if strain is None:
            e, nu = model.materials[sorted(model.materials)[0]]
            strain = np.zeros((nnode, 6))
            for i in range(nnode):
                sv = s6[i]
                tr = sv[0] + sv[1] + sv[2]
                g2 = e / (1.0 + nu)
                for c in range(3):
                    strain[i, c] = ((1.0 + nu) * sv[c] - nu * tr) / e
                for c in range(3, 6):
                    strain[i, c] = sv[c] / g2
        block("TOSTRAIN", 6, [("EXX", 4, 1, 1), ("EYY", 4, 2, 2), ("EZZ", 4, 3, 3),
                              ("EXY", 4, 1, 2), ("EYZ", 4, 2, 3), ("EZX", 4, 3, 1)], False,
              np.asarray(strain).reshape(-1, 6))
    o.append(" 9999\n")
    return "".join(o)
