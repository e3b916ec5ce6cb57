"""DAT / FRD writers (dat.rs, frd.rs): the oracle's restatement pinned by the reference's
own known answers (frd.rs unit tests: ES12.5 values, CalculiX-identical 100CL records,
record lengths, C3D20 node order), and the library's writers byte-for-byte against it
(CPU only: the writers are host code)."""
import os

import numpy as np
import pytest

from axia_gpu import deck as D
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")

# the deck of frd.rs:528-553 (write_frd_mecway_node_and_headers)
FLACHEISEN = """
*HEADING
Flacheisen
*NODE
1, 0.06, 0.01, 0.21
2, 1, 0, 0
3, 1, 1, 0
4, 0, 1, 0
5, 0, 0, 1
6, 1, 0, 1
7, 1, 1, 1
8, 0, 1, 1
*ELEMENT, TYPE=C3D8
1, 1, 2, 3, 4, 5, 6, 7, 8
*MATERIAL, NAME=S
*ELASTIC
210000, 0.3
*SOLID SECTION, ELSET=EALL, MATERIAL=S
*BOUNDARY
1, 1, 3
*STEP
*STATIC
*CLOAD
2, 1, 1
*END STEP
"""


def hex20_deck():
    """frd.rs:612-638 (hex20_frd_midsides_match_ccx): nodes 1..20 at the origin."""
    nodes = "".join(f"{i}, 0, 0, 0\n" for i in range(1, 21))
    ids = ", ".join(str(i) for i in range(1, 21))
    return (f"*NODE\n{nodes}*ELEMENT, TYPE=C3D20\n1, {ids}\n*MATERIAL, NAME=S\n*ELASTIC\n"
            "210000, 0.3\n*SOLID SECTION, ELSET=EALL, MATERIAL=S\n*BOUNDARY\n1, 1, 3\n*STEP\n"
            "*STATIC\n*CLOAD\n2, 1, 1\n*END STEP\n")


def test_e12_and_records_known_answers():
    """frd.rs:463-520."""
    assert O.frd_e12(0.06) == " 6.00000E-02"
    assert O.frd_e12(0.01) == " 1.00000E-02"
    assert O.frd_e12(0.21) == " 2.10000E-01"
    assert O.frd_e12(0.0) == " 0.00000E+00"
    assert O.frd_e12(1.0) == " 1.00000E+00"
    assert len(O.frd_e12(-1234.5)) == 12 and len(O.frd_e12(1.122e7)) == 12
    assert O.frd_e12(0.06)[8] == "E" and O.frd_e12(0.06).endswith("E-02")
    p = O.frd_1pstep(1, 1, 1)
    assert len(p) == 70 and p.startswith("    1PSTEP")
    assert p[24:36] == p[36:48] == p[48:60] == "           1"
    h = O.frd_100cl(1, 1.0, 8)
    assert len(h) == 75 and h[74] == "1" and h[7:12] == "  101" and h[58:63] == "    1"
    assert h[57] == "0"
    # byte-identical to CalculiX 2.22 (BoltedJoint.frd)
    assert O.frd_100cl(1, 0.1, 730) == (
        "  100CL  101 1.00000E-01         730                     0    1           1")
    assert O.frd_100cl(10, 1.0, 730) == (
        "  100CL  110 1.00000E+00         730                     0   10           1")


def _zeros_frd(writer_model, text):
    n = len(writer_model.node_ids)
    return np.zeros(3 * n), np.zeros((n, 6)), np.zeros(3 * n), np.zeros((n, 6))


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_frd_headers_and_node_record(which):
    """frd.rs:525-585 on the oracle's and the library's output."""
    m = O.parse_inp(FLACHEISEN)
    u, s, r, e = _zeros_frd(m, FLACHEISEN)
    frd = (O.write_frd(m, u, s, r, e) if which == "oracle"
           else D.Deck(FLACHEISEN).write_frd(u, s, r, e))
    lines = frd.split("\n")
    c2 = next(l for l in lines if l.startswith("    2C"))
    c3 = next(l for l in lines if l.startswith("    3C"))
    assert len(c2) == 74 and len(c3) == 74 and c2[73] == "1" and c3[73] == "1"
    assert c2[24:36] == "           8"
    node = next(l for l in lines if l.startswith(" -1") and l[3:13].strip() == "1")
    assert " 6.00000E-02" in node and " 1.00000E-02" in node and " 2.10000E-01" in node
    assert " -3" in lines
    cl = next(l for l in lines if l.startswith("  100CL"))
    assert len(cl) == 75 and cl[74] == "1" and cl[7:12] == "  101" and "DISP" not in cl
    assert lines[0] == "    1CFlacheisen"


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_frd_hex20_node_order(which):
    """frd.rs:600-640: C3D20 midsides in CalculiX order on the -2 records."""
    text = hex20_deck()
    m = O.parse_inp(text)
    u, s, r, e = _zeros_frd(m, text)
    frd = O.write_frd(m, u, s, r, e) if which == "oracle" else D.Deck(text).write_frd(u, s, r, e)
    minus2, inside = [], False
    for line in frd.split("\n"):
        if line.startswith("    3C"):
            inside = True
            continue
        if inside and line == " -3":
            break
        if inside and line.startswith(" -2"):
            minus2.append(line)
    assert minus2 == [
        " -2         1         2         3         4         5         6         7         8"
        "         9        10",
        " -2        11        12        17        18        19        20        13        14"
        "        15        16"]


def _edge_values(rng, n):
    v = rng.standard_normal(n) * 10.0 ** rng.integers(-12, 12, n)
    special = [0.0, -0.0, 1.0, 10.0, 100.0, 1e-5, 9.999995, 9.9999949, 0.00099999995, 123456.5,
               -1e100, 1e-100, 5e-324, 1.7976931348623157e308, 999999.5e-5, 0.1, 0.3, 1 / 3]
    k = min(len(special), n)
    v[:k] = special[:k]
    return v


@pytest.mark.parametrize("name", ["cube_tension", "le10_c3d20"])
def test_dat_and_frd_byte_identical_to_the_oracle(name):
    """The library's writers against the oracle's on the oracle's own linear results."""
    text = open(os.path.join(GOLD, name + ".inp")).read()
    m, ref = O.solve_deck(text)
    sn = ref.elem_nodal_stress
    gp = O.element_mean_stress(sn)
    dk = D.Deck(text)
    dat = dk.write_dat(ref.u_full, ref.rf_full, gp, ref.stress, [12.5, 1e3])
    assert dat == O.write_dat(m, ref.u_full, ref.rf_full, gp, ref.stress, [12.5, 1e3])
    frd = dk.write_frd(ref.u_full, ref.stress, ref.rf_full)
    assert frd == O.write_frd(m, ref.u_full, ref.stress, ref.rf_full)


def test_number_formats_on_edge_values():
    """Random magnitudes across 24 decades plus rounding, power-of-ten, zero, subnormal and
    extreme values through both writers."""
    rng = np.random.default_rng(7)
    text = hex20_deck()
    m = O.parse_inp(text)
    n = len(m.node_ids)
    u = _edge_values(rng, 3 * n)
    s = _edge_values(rng, 6 * n).reshape(n, 6)
    r = _edge_values(rng, 3 * n)
    dk = D.Deck(text)
    assert dk.write_frd(u, s, r) == O.write_frd(m, u, s, r)
    gp = _edge_values(rng, 6).reshape(1, 6)
    assert dk.write_dat(u, r, gp, s) == O.write_dat(m, u, r, gp, s)
    for k in range(20):
        u = _edge_values(np.random.default_rng(100 + k), 3 * n)
        assert dk.write_frd(u) == O.write_frd(m, u)


def test_output_selection_and_heading():
    """*NODE FILE / *EL FILE select the blocks (inp.rs:1989-2021); the heading is cut at 66
    characters; TOSTRAIN from the stress with the material's E, nu."""
    base = FLACHEISEN.replace("Flacheisen", "Ü" * 70)
    for extra in ("*NODE FILE\nRF\n*EL FILE\nE\n", "*NODEFILE\nU, RF\n*ELFILE\nS, E\n",
                  "*NODE OUTPUT\nNT\n"):
        text = base.replace("*STEP\n", "*STEP\n" + extra)
        m = O.parse_inp(text)
        rng = np.random.default_rng(3)
        n = len(m.node_ids)
        u, s, r = rng.standard_normal(3 * n), rng.standard_normal((n, 6)), rng.standard_normal(3 * n)
        out = D.Deck(text).write_frd(u, s, r)
        assert out == O.write_frd(m, u, s, r)
        assert out.split("\n")[0] == "    1C" + "Ü" * 66
    frd = D.Deck(base.replace("*STEP\n", "*STEP\n*NODE FILE\nRF\n*EL FILE\nE\n")).write_frd(
        np.zeros(24), np.ones((8, 6)), np.ones(24))
    assert "DISP" not in frd and "STRESS" not in frd and "FORC" in frd and "TOSTRAIN" in frd


def test_size_query_text_is_not_served_stale():
    """The size query keeps its text for the fill call only while the inputs are unchanged:
    values changed in place between the two calls are formatted anew."""
    import ctypes as C
    from axia_gpu import _lib
    dk = D.Deck(FLACHEISEN)
    n = len(dk.arrays()["node_ids"])
    u = np.zeros(3 * n)
    L = _lib.lib()
    ln = C.c_int64()
    assert L.axg_deck_write_frd(dk._h, u.ctypes.data_as(C.POINTER(C.c_double)), None, None, None,
                                None, 0, C.byref(ln)) == 0
    u[0] = 0.125  # same address, new value
    buf = C.create_string_buffer(ln.value + 64)
    assert L.axg_deck_write_frd(dk._h, u.ctypes.data_as(C.POINTER(C.c_double)), None, None, None,
                                buf, ln.value + 64, C.byref(ln)) == 0
    assert buf.raw[:ln.value].decode() == dk.write_frd(u)
    assert " 1.25000E-01" in buf.raw[:ln.value].decode()


def test_threaded_rows_byte_identical_on_a_large_deck():
    """More than 8,192 node and element rows, so the rows are formatted in chunks on several
    host threads and joined in order: both files equal the oracle's byte for byte."""
    from axia_gpu import meshgen as MG
    xyz, conn = MG.brick_hex8(24, 24, 24)
    text = "".join(MG.inp_solid(8, xyz, conn, 210000.0, 0.3)) + (
        "*STEP\n*STATIC\n*CLOAD\n1, 3, 1.0\n*END STEP\n")
    m = O.parse_inp(text)
    n, ne = len(m.node_ids), len(m.elements)
    assert n > 2 * 4096 and ne > 2 * 4096
    rng = np.random.default_rng(11)
    u = _edge_values(rng, 3 * n)
    s = _edge_values(rng, 6 * n).reshape(n, 6)
    r = _edge_values(rng, 3 * n)
    gp = _edge_values(rng, 6 * ne).reshape(ne, 6)
    dk = D.Deck(text)
    assert dk.write_frd(u, s, r) == O.write_frd(m, u, s, r)
    assert dk.write_dat(u, r, gp, s) == O.write_dat(m, u, r, gp, s)
