"""Synthetic structured C3D8 / C3D20 meshes and the benchmark load cases.

Node and element numbering follow what the reference does after
``Model::compact`` (model.rs:1234-1280): node index = rank of the node id,
element index = rank of the element id, DOF = 3 * node + d
(analysis.rs:152-154). Free DOFs are numbered in ascending DOF order
(DofMap::build, constraint.rs:99-108). Face P2 of a brick element is its
top face (local nodes 4, 7, 6, 5; elem.rs:894-901).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Problem:
    name: str
    nn: int
    xyz: np.ndarray          # (nnode, 3)
    conn: np.ndarray         # (nelem, nn) int32
    e: np.ndarray            # (nelem,)
    nu: np.ndarray           # (nelem,)
    ind_of: np.ndarray       # (3 nnode,) int64
    f_full: list = field(default_factory=list)  # load cases, each (3 nnode,)
    meta: dict = field(default_factory=dict)

    @property
    def nnode(self):
        return self.xyz.shape[0]

    @property
    def nelem(self):
        return self.conn.shape[0]

    @property
    def n_free(self):
        return int((self.ind_of >= 0).sum())

    def rhs(self, k=0):
        """DofMap::reduce of the load vector (constraint.rs:143-153): f at free DOFs."""
        f = self.f_full[k]
        b = np.zeros(self.n_free)
        free = self.ind_of >= 0
        b[self.ind_of[free]] = f[free]
        return b

    def rhs_all(self):
        return np.stack([self.rhs(k) for k in range(len(self.f_full))])

    def expand(self, x):
        """DofMap::reconstruct with zero prescribed values: u_full from x."""
        u = np.zeros(3 * self.nnode)
        free = self.ind_of >= 0
        u[free] = x[self.ind_of[free]]
        return u


def brick_hex8(nx, ny, nz, h=1.0):
    """Structured brick of unit hexes: node (i,j,k) has index i + (nx+1)(j + (ny+1)k)."""
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    xyz = np.stack([i.ravel(order="F"), j.ravel(order="F"), k.ravel(order="F")], 1) * h
    xyz = xyz.astype(np.float64)

    def nid(a, b, c):
        return a + (nx + 1) * (b + (ny + 1) * c)

    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ex, ey, ez = ex.ravel(order="F"), ey.ravel(order="F"), ez.ravel(order="F")
    conn = np.stack([
        nid(ex, ey, ez), nid(ex + 1, ey, ez), nid(ex + 1, ey + 1, ez), nid(ex, ey + 1, ez),
        nid(ex, ey, ez + 1), nid(ex + 1, ey, ez + 1), nid(ex + 1, ey + 1, ez + 1),
        nid(ex, ey + 1, ez + 1)], 1).astype(np.int32)
    return xyz, conn


# HEX20_XI (quadratic.rs:13-34) as half-grid offsets 0/1/2
_H20 = np.array([
    [0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0], [0, 0, 2], [2, 0, 2], [2, 2, 2], [0, 2, 2],
    [1, 0, 0], [2, 1, 0], [1, 2, 0], [0, 1, 0], [1, 0, 2], [2, 1, 2], [1, 2, 2], [0, 1, 2],
    [0, 0, 1], [2, 0, 1], [2, 2, 1], [0, 2, 1]])


def brick_hex20(nx, ny, nz, h=1.0):
    """Serendipity brick: half-grid points with at most one odd index, lexicographic (k, j, i)."""
    I, J, K = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1
    gi, gj, gk = np.meshgrid(np.arange(I), np.arange(J), np.arange(K), indexing="ij")
    gi, gj, gk = gi.ravel(order="F"), gj.ravel(order="F"), gk.ravel(order="F")
    keep = (gi % 2 + gj % 2 + gk % 2) <= 1
    index = np.full(I * J * K, -1, np.int64)
    index[keep] = np.arange(int(keep.sum()))
    xyz = np.stack([gi[keep], gj[keep], gk[keep]], 1).astype(np.float64) * (0.5 * h)
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ex, ey, ez = ex.ravel(order="F"), ey.ravel(order="F"), ez.ravel(order="F")
    cols = []
    for o in _H20:
        flat = (2 * ex + o[0]) + I * ((2 * ey + o[1]) + J * (2 * ez + o[2]))
        cols.append(index[flat])
    conn = np.stack(cols, 1).astype(np.int32)
    assert (conn >= 0).all()
    return xyz, conn


def free_map(nnode, fixed_nodes, fixed_dofs=(0, 1, 2)):
    """ind_of for nodes whose listed DOFs are prescribed (zero)."""
    pres = np.zeros(3 * nnode, bool)
    for d in fixed_dofs:
        pres[3 * np.asarray(fixed_nodes, np.int64) + d] = True
    ind = np.full(3 * nnode, -1, np.int64)
    ind[~pres] = np.arange(int((~pres).sum()))
    return ind


def _top_pressure(nn, xyz, conn, elems, p):
    from . import pressure_loads
    f = np.zeros(3 * xyz.shape[0])
    faces = np.full(len(elems), 2, np.int32)
    mags = np.full(len(elems), float(p))
    pressure_loads(nn, xyz, conn[elems], faces, mags, f)
    return f


def cantilever(nx=20, ny=4, nz=4, e=210000.0, nu=0.3, total=-1000.0):
    """Cantilever C3D8: x = 0 clamped, total load in z spread equally on the x = nx face."""
    xyz, conn = brick_hex8(nx, ny, nz)
    ind = free_map(len(xyz), np.nonzero(xyz[:, 0] == 0.0)[0])
    tip = np.nonzero(xyz[:, 0] == float(nx))[0]
    f = np.zeros(3 * len(xyz))
    f[3 * tip + 2] = total / len(tip)
    ne = len(conn)
    return Problem(f"cantilever_c3d8_{nx}x{ny}x{nz}", 8, xyz, conn, np.full(ne, e),
                   np.full(ne, nu), ind, [f], {"shape": (nx, ny, nz)})


def brick_lognormal(n=80, seed=0, median=210000.0, sigma=0.5, nu=0.3, pressure=1.0, nloads=1):
    """C3D8 n^3 brick, lognormal E, bottom clamped, top pressure P2 = 1 MPa.

    Extra load cases (nloads > 1) are top-face tractions: shear in x, shear in
    y, and a seeded random normal pressure field, each applied as P2-style
    face loads or nodal forces of the same total magnitude.
    """
    xyz, conn = brick_hex8(n, n, n)
    rng = np.random.default_rng(seed)
    ne = len(conn)
    e = median * np.exp(sigma * rng.standard_normal(ne))
    ind = free_map(len(xyz), np.nonzero(xyz[:, 2] == 0.0)[0])
    top_el = np.nonzero(xyz[conn[:, 4], 2] == float(n))[0]
    loads = [_top_pressure(8, xyz, conn, top_el, pressure)]
    top = np.nonzero(xyz[:, 2] == float(n))[0]
    area = float(n * n) * pressure
    for k in range(1, nloads):
        f = np.zeros(3 * len(xyz))
        if k == 1:
            f[3 * top + 0] = area / len(top)
        elif k == 2:
            f[3 * top + 1] = area / len(top)
        else:
            g = np.abs(rng.standard_normal(len(top)))
            f[3 * top + 2] = -area * g / g.sum()
        loads.append(f)
    return Problem(f"brick_c3d8_{n}^3_lognormal", 8, xyz, conn, e, np.full(ne, nu), ind, loads,
                   {"seed": seed, "sigma": sigma})


def brick_c3d20_jitter(nx=96, ny=24, nz=24, seed=0, e=70000.0, nu=0.3, jitter=0.15):
    """C3D20 brick, interior corner nodes jittered by U(-j, j) h per coordinate and every
    midside node at the midpoint of its (straight) edge, so boundary faces stay planar
    and no element folds; x = 0 clamped; tip shear 1 N/node in z."""
    xyz, conn = brick_hex20(nx, ny, nz)
    rng = np.random.default_rng(seed)
    d = rng.uniform(-jitter, jitter, size=(nx + 1, ny + 1, nz + 1, 3))
    d[0], d[nx] = 0.0, 0.0
    d[:, 0], d[:, ny] = 0.0, 0.0
    d[:, :, 0], d[:, :, nz] = 0.0, 0.0
    lo = np.floor(xyz).astype(np.int64)
    hi = np.ceil(xyz).astype(np.int64)
    # a corner node has lo == hi; a midside node differs in exactly one axis
    disp = 0.5 * (d[lo[:, 0], lo[:, 1], lo[:, 2]] + d[hi[:, 0], hi[:, 1], hi[:, 2]])
    xyz = xyz + disp
    ind = free_map(len(xyz), np.nonzero(xyz[:, 0] == 0.0)[0])
    tip = np.nonzero(xyz[:, 0] == float(nx))[0]
    f = np.zeros(3 * len(xyz))
    f[3 * tip + 2] = 1.0
    ne = len(conn)
    return Problem(f"brick_c3d20_{nx}x{ny}x{nz}_jitter", 20, xyz, conn, np.full(ne, e),
                   np.full(ne, nu), ind, [f], {"seed": seed})


def batch_meshes(count=48, lo=8, hi=48, nloads=4, seed=0, nu=0.3):
    """Cubes m^3 with m ~ U{lo..hi}, E ~ U(5e4, 2.5e5), bottom clamped, nloads top-face
    Gaussian nodal z-load cases scaled to 1 kN total each."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        m = int(rng.integers(lo, hi + 1))
        xyz, conn = brick_hex8(m, m, m)
        ne = len(conn)
        e = float(rng.uniform(50000.0, 250000.0))
        ind = free_map(len(xyz), np.nonzero(xyz[:, 2] == 0.0)[0])
        top = np.nonzero(xyz[:, 2] == float(m))[0]
        loads = []
        for _ in range(nloads):
            g = np.abs(rng.standard_normal(len(top)))
            f = np.zeros(3 * len(xyz))
            f[3 * top + 2] = -1000.0 * g / g.sum()
            loads.append(f)
        out.append(Problem(f"cube_{m}^3_{b}", 8, xyz, conn, np.full(ne, e), np.full(ne, nu), ind,
                           loads, {"m": m}))
    return out


def simp_brick(nx=120, ny=60, nz=60, seed=0, nu=0.3, rho_lo=0.3, rho_hi=1.0, emin=1e-3, e0=1.0):
    """SIMP brick: E_e = emin + rho_e^3 e0, left face clamped, unit point load -z at (nx, ny/2, 0)."""
    xyz, conn = brick_hex8(nx, ny, nz)
    rng = np.random.default_rng(seed)
    ne = len(conn)
    rho = rng.uniform(rho_lo, rho_hi, ne)
    e = emin * e0 + rho ** 3 * e0
    ind = free_map(len(xyz), np.nonzero(xyz[:, 0] == 0.0)[0])
    node = int(np.nonzero((xyz[:, 0] == nx) & (xyz[:, 1] == ny // 2) & (xyz[:, 2] == 0))[0][0])
    f = np.zeros(3 * len(xyz))
    f[3 * node + 2] = -1.0
    return Problem(f"simp_c3d8_{nx}x{ny}x{nz}", 8, xyz, conn, e, np.full(ne, nu), ind, [f],
                   {"rho": rho, "emin": emin, "e0": e0, "load_node": node})


# ------------------------------------------------------------- INP writing
def inp_solid(nn, xyz, conn, e, nu, rho=None):
    """INP lines of a one-material solid: *NODE, *ELEMENT (ELSET=SOLID), *MATERIAL with
    *ELASTIC (and *DENSITY when rho is given), *SOLID SECTION. Ids are index + 1."""
    typ = "C3D8" if nn == 8 else "C3D20"
    out = ["*NODE\n"]
    out += [f"{i + 1}, {float(x)!r}, {float(y)!r}, {float(z)!r}\n" for i, (x, y, z) in enumerate(xyz)]
    out.append(f"*ELEMENT, TYPE={typ}, ELSET=SOLID\n")
    for k, row in enumerate(conn):
        ids = [str(int(a) + 1) for a in row]
        # 16 values per data line at most, continuation lines after a trailing comma
        cur = [str(k + 1)] + ids
        while len(cur) > 16:
            out.append(", ".join(cur[:16]) + ",\n")
            cur = cur[16:]
        out.append(", ".join(cur) + "\n")
    out.append(f"*MATERIAL, NAME=MAT\n*ELASTIC\n{float(e)!r}, {float(nu)!r}\n")
    if rho is not None:
        out.append(f"*DENSITY\n{float(rho)!r}\n")
    out.append("*SOLID SECTION, ELSET=SOLID, MATERIAL=MAT\n")
    return out


def inp_nset(name, nodes):
    """*NSET lines (ids = index + 1, 16 per line)."""
    out = [f"*NSET, NSET={name}\n"]
    ids = [str(int(a) + 1) for a in nodes]
    for i in range(0, len(ids), 16):
        out.append(", ".join(ids[i:i + 16]) + "\n")
    return out


def nlgeom_cantilever_deck(nx=6, ny=2, nz=2, dims=(6.0, 1.0, 1.0), p_total=-1.0, e=1000.0,
                           nu=0.3, ninc=2, nn=8, pressure=None):
    """A clamped cantilever under *STEP, NLGEOM: p_total spread over the tip nodes in z
    (dead load), or a pressure on the top face (P2 of the top layer, a follower load)."""
    xyz, conn = (brick_hex20 if nn == 20 else brick_hex8)(nx, ny, nz)
    xyz = xyz * np.array([dims[0] / nx, dims[1] / ny, dims[2] / nz])
    out = inp_solid(nn, xyz, conn, e, nu, None)
    fixed = np.nonzero(xyz[:, 0] == 0.0)[0]
    tip = np.nonzero(xyz[:, 0] == dims[0])[0]
    out += inp_nset("FIX", fixed) + inp_nset("TIP", tip)
    top = [k for k in range(len(conn)) if k // (nx * ny) == nz - 1]
    if pressure is not None:
        out.append("*ELSET, ELSET=TOP\n")
        ids = [str(k + 1) for k in top]
        for i in range(0, len(ids), 16):
            out.append(", ".join(ids[i:i + 16]) + "\n")
    out.append(f"*STEP, NLGEOM\n*STATIC\n{1.0 / ninc!r}, 1.0\n*BOUNDARY\nFIX, 1, 3\n")
    if pressure is not None:
        out.append(f"*DLOAD\nTOP, P2, {float(pressure)!r}\n")
    else:
        out.append(f"*CLOAD\nTIP, 3, {p_total / len(tip)!r}\n")
    out.append("*END STEP\n")
    return "".join(out)


def lognormal_brick_deck(n=80, seed=0, nu=0.3, pressure=1.0):
    """brick_lognormal(n, seed) as an INP deck the way a per-element modulus has to be written
    for the reference (its material is per element set, model.rs:1391-1407): one *ELSET,
    *MATERIAL and *SOLID SECTION per element; bottom face clamped (*BOUNDARY on NSET BOT),
    P2 = pressure on the top layer (*DLOAD on ELSET TOP). Returns (text, problem)."""
    p = brick_lognormal(n, seed=seed, nu=nu, pressure=pressure)
    out = [f"*HEADING\nC3D8 brick {n}^3, one material per element\n*NODE\n"]
    out += [f"{i + 1}, {x!r}, {y!r}, {z!r}\n" for i, (x, y, z) in enumerate(p.xyz.tolist())]
    out.append("*ELEMENT, TYPE=C3D8\n")
    out += [f"{k + 1}, {r[0]}, {r[1]}, {r[2]}, {r[3]}, {r[4]}, {r[5]}, {r[6]}, {r[7]}\n"
            for k, r in enumerate((p.conn + 1).tolist())]
    out += [f"*ELSET, ELSET=E{k + 1}\n{k + 1}\n*MATERIAL, NAME=M{k + 1}\n*ELASTIC\n{e!r}, {nu!r}\n"
            f"*SOLID SECTION, ELSET=E{k + 1}, MATERIAL=M{k + 1}\n" for k, e in enumerate(p.e.tolist())]
    out += inp_nset("BOT", np.nonzero(p.xyz[:, 2] == 0.0)[0])
    top = np.nonzero(p.xyz[p.conn[:, 4], 2] == float(n))[0] + 1  # the top layer's elements
    out.append("*ELSET, ELSET=TOP\n")
    out += [", ".join(str(int(v)) for v in top[i:i + 16]) + "\n" for i in range(0, len(top), 16)]
    out.append(f"*STEP\n*STATIC\n*BOUNDARY\nBOT, 1, 3\n*DLOAD\nTOP, P2, {float(pressure)!r}\n"
               "*END STEP\n")
    return "".join(out), p


def buckling_column_deck(nx=320, ny=16, nz=16, p_total=-1.0, e=210000.0, nu=0.3, nbuckle=4):
    """A C3D8 column of unit elements along x (length nx, section ny x nz), the x = 0 face
    clamped, p_total spread equally over the x = nx face nodes in x (the *CLOAD magnitude
    applies to each node of the set), and a *BUCKLE step for nbuckle factors."""
    xyz, conn = brick_hex8(nx, ny, nz)
    out = inp_solid(8, xyz, conn, e, nu, None)
    fixed = np.nonzero(xyz[:, 0] == 0.0)[0]
    tip = np.nonzero(xyz[:, 0] == float(nx))[0]
    out += inp_nset("FIX", fixed) + inp_nset("TIP", tip)
    out.append(f"*BOUNDARY\nFIX, 1, 3\n*STEP\n*BUCKLE\n{int(nbuckle)}\n"
               f"*CLOAD\nTIP, 1, {p_total / len(tip)!r}\n*END STEP\n")
    return "".join(out)
