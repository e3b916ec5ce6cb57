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
