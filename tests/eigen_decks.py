"""INP decks with a *FREQUENCY step for the eigen tests (generated, not taken from files)."""

BAR = """*NODE
{nodes}*ELEMENT, TYPE=C3D8, ELSET=BAR
{elems}*MATERIAL, NAME=STEEL
*ELASTIC
210000., 0.0
*DENSITY
7.85e-9
*SOLID SECTION, ELSET=BAR, MATERIAL=STEEL
*NSET, NSET=ALLN
{alln}*BOUNDARY
FIX, 1, 3
ALLN, 2, 3
*NSET, NSET=FIX
1, 2, 3, 4
*STEP
*FREQUENCY
{nmodes}
*END STEP
"""


def bar_deck(nel, length=100.0, nmodes=3):
    """An x-bar of nel unit-section C3D8 (lateral motion suppressed): axial modes only."""
    h = length / nel
    nodes, elems = [], []
    for i in range(nel + 1):
        x = i * h
        for k, (y, z) in enumerate(((0, 0), (1, 0), (1, 1), (0, 1))):
            nodes.append(f"{4 * i + k + 1}, {x!r}, {y}., {z}.\n")
    for e in range(nel):
        a = 4 * e + 1
        b = a + 4
        elems.append(f"{e + 1}, {a}, {b}, {b + 1}, {a + 1}, {a + 3}, {b + 3}, {b + 2}, {a + 2}\n")
    ids = ", ".join(str(i) for i in range(1, 4 * (nel + 1) + 1))
    return BAR.format(nodes="".join(nodes), elems="".join(elems), alln=ids + "\n", nmodes=nmodes)


def _solid_lines(nn, xyz, conn, e, nu, rho):
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


def _nset(name, nodes):
    out = [f"*NSET, NSET={name}\n"]
    ids = [str(int(a) + 1) for a in nodes]
    for i in range(0, len(ids), 16):
        out.append(", ".join(ids[i:i + 16]) + "\n")
    return out


def mesh_deck(nn, xyz, conn, fixed_nodes, e=210000.0, nu=0.3, rho=7.85e-9, nmodes=4):
    """A solid deck from arrays: one material with density, fixed_nodes clamped, *FREQUENCY."""
    out = _solid_lines(nn, xyz, conn, e, nu, rho) + _nset("FIX", fixed_nodes)
    out.append(f"*BOUNDARY\nFIX, 1, 3\n*STEP\n*FREQUENCY\n{nmodes}\n*END STEP\n")
    return "".join(out)


def column_deck(nn, xyz, conn, fixed_nodes, tip_nodes, p_total=-1.0, e=210000.0, nu=0.3,
                nbuckle=2):
    """A solid column: fixed_nodes clamped, p_total spread equally over tip_nodes in x (the
    *CLOAD magnitude applies to each node of the set), *BUCKLE."""
    out = _solid_lines(nn, xyz, conn, e, nu, None) + _nset("FIX", fixed_nodes)
    out += _nset("TIP", tip_nodes)
    out.append(f"*BOUNDARY\nFIX, 1, 3\n*STEP\n*BUCKLE\n{nbuckle}\n"
               f"*CLOAD\nTIP, 1, {p_total / len(tip_nodes)!r}\n*END STEP\n")
    return "".join(out)


def nl_bar_deck(nel=4, length=4.0, delta=0.5, e=1000.0, nu=0.0, ninc=4, nn=8):
    """A unit-section x-bar of nel hexes stretched by a prescribed end displacement delta
    under *STEP, NLGEOM in ninc increments: x = 0 face ux = 0, the origin node uy = uz = 0,
    node (0, 1, 0) uz = 0 (no rigid motion, lateral contraction free)."""
    from axia_gpu import meshgen as MG
    import numpy as np
    xyz, conn = (MG.brick_hex20 if nn == 20 else MG.brick_hex8)(nel, 1, 1)
    xyz = xyz * np.array([length / nel, 1.0, 1.0])
    out = _solid_lines(nn, xyz, conn, e, nu, None)
    x0 = np.nonzero(xyz[:, 0] == 0.0)[0]
    xl = np.nonzero(xyz[:, 0] == length)[0]
    org = int(np.nonzero((xyz == 0.0).all(1))[0][0])
    y1 = int(np.nonzero((xyz[:, 0] == 0.0) & (xyz[:, 1] == 1.0) & (xyz[:, 2] == 0.0))[0][0])
    out += _nset("X0", x0) + _nset("XL", xl)
    out.append(f"*STEP, NLGEOM\n*STATIC\n{1.0 / ninc!r}, 1.0\n*BOUNDARY\nX0, 1, 1\n"
               f"{org + 1}, 2, 3\n{y1 + 1}, 3, 3\nXL, 1, 1, {float(delta)!r}\n*END STEP\n")
    return "".join(out)


def nl_cantilever_deck(nx=6, ny=2, nz=2, dims=(6.0, 1.0, 1.0), p_total=-1.0, e=1000.0,
                       nu=0.3, ninc=2, nn=8, pressure=None):
    """A clamped cantilever under *STEP, NLGEOM: p_total spread over the tip nodes in z
    (dead load), or a pressure on the top face (P2 of the top layer, a follower load)."""
    from axia_gpu import meshgen as MG
    import numpy as np
    xyz, conn = (MG.brick_hex20 if nn == 20 else MG.brick_hex8)(nx, ny, nz)
    xyz = xyz * np.array([dims[0] / nx, dims[1] / ny, dims[2] / nz])
    out = _solid_lines(nn, xyz, conn, e, nu, None)
    fixed = np.nonzero(xyz[:, 0] == 0.0)[0]
    tip = np.nonzero(xyz[:, 0] == dims[0])[0]
    out += _nset("FIX", fixed) + _nset("TIP", tip)
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
