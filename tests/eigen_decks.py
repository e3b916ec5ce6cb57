"""INP decks for the eigen, buckling and NLGEOM tests (generated, not taken from files)."""

from axia_gpu.meshgen import inp_nset as _nset
from axia_gpu.meshgen import inp_solid as _solid_lines
from axia_gpu.meshgen import nlgeom_cantilever_deck as nl_cantilever_deck

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
