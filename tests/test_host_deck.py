"""The native INP reader and linear-static setup (axg_deck_parse) against the
oracle's restatement of inp.rs / model.rs / analysis.rs:548-946 (CPU only).

Every array of the setup (node ids, coordinates, connectivity, moduli, DOF
numbering, prescribed values, load vector) must match bit for bit."""
import os

import numpy as np
import pytest

from axia_gpu import AxgError
from axia_gpu import deck as D
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DECKS = ["cube_tension", "c3d20_patch", "patch_c3d8", "le10_c3d20", "le10_c3d20_fine"]


def read(name):
    return open(os.path.join(GOLD, name + ".inp")).read()


def assert_same_setup(text):
    d = D.Deck(text)
    a = d.arrays()
    m = O.parse_inp(text)
    s = O.deck_setup(m)
    assert d.nn == s.nn and d.nnode == len(m.node_ids) and d.n_free == s.n_free
    assert np.array_equal(a["node_ids"], np.array(m.node_ids))
    assert np.array_equal(a["elem_ids"], np.array([e[0] for e in m.elements]))
    for key, ref in (("xyz", s.xyz), ("conn", s.conn), ("e", s.e), ("nu", s.nu),
                     ("ind_of", s.ind_of), ("u0", s.u0), ("f_full", s.f_full)):
        assert np.array_equal(a[key], ref), key
    return d, a, s


@pytest.mark.parametrize("name", DECKS)
def test_reference_decks_setup_bitexact(name):
    assert_same_setup(read(name))


def test_deferred_elset_and_second_material():
    """lib.rs:410-427: ELSET defined after the elements, a dummy material besides STEEL."""
    text = read("cube_tension").replace("*ELEMENT, TYPE=C3D8, ELSET=SOLID", "*ELEMENT, TYPE=C3D8")
    text = text.replace("*SOLID SECTION, ELSET=SOLID, MATERIAL=STEEL",
                        "*ELSET, ELSET=SOLID\n1\n*MATERIAL, NAME=DUMMY\n*ELASTIC\n1.0, 0.0\n"
                        "*SOLID SECTION, ELSET=SOLID, MATERIAL=STEEL")
    _, a, _ = assert_same_setup(text)
    assert a["e"][0] == 210000.0


BRICK = """*HEADING
two C3D8 bricks
*NODE, NSET=ALLN
1, 0., 0., 0.
2, 1.d0, 0., 0.
3, 1., 1., 0.
4, 0., 1.0D0, 0.
5, 0., 0., 1.
6, 1., 0., 1.
7, 1., 1., 1.
8, 0., 1., 1.
9, 2., 0., 0.
10, 2., 1., 0.
11, 2., 1., 1.
12, 2., 0., 1.
** a comment line
*ELEMENT, TYPE=C3D8, ELSET=BLK
1, 1, 2, 3, 4, 5, 6, 7, 8
2, 2, 9, 10, 3, 6, 12, 11, 7
*NSET, NSET=LEFT, GENERATE
1, 8, 1
*NSET, NSET=LEFTFACE
1, 4, 5, 8
*NSET, NSET=RIGHT
9, 10, 11, 12
*NSET, NSET=BOTH
LEFTFACE, RIGHT
*MATERIAL, NAME=ALU
*ELASTIC
7.0+4, 0.33
*SOLID SECTION, ELSET=BLK, MATERIAL=ALU
*STEP
*STATIC
*BOUNDARY
LEFTFACE, 1, 3
2, 2, 2, 0.0
*CLOAD
RIGHT, 3, -2.5
LATE, 1, 1.0
9, 3, -1.
*DLOAD
BLK, P2, 3.0
2, P4, -1.5
*NSET, NSET=LATE
11
*END STEP
"""


def test_synthetic_deck_keywords():
    """Fortran reals (1.d0, 7.0+4), GENERATE, nested sets, BOUNDARY ranges, CLOAD on a set
    defined later (deferred), DLOAD on a set and on an element id."""
    d, a, s = assert_same_setup(BRICK)
    assert d.nnode == 12 and d.nelem == 2
    assert a["e"][0] == 70000.0 and a["nu"][0] == 0.33
    assert a["xyz"][1, 0] == 1.0 and a["xyz"][3, 1] == 1.0
    # node 11 gets the deferred CLOAD in x
    assert a["f_full"][3 * 10 + 0] != 0.0


def test_boundary_values_and_overrides():
    """Prescribed values, last BC wins, node ids out of order and a repeated node id
    (first definition wins)."""
    text = BRICK.replace("2, 2, 2, 0.0", "2, 2, 2, 0.25\n2, 2, 2, -0.5\n12, 1, 3, 0.125")
    text = text.replace("12, 2., 0., 1.\n", "12, 2., 0., 1.\n3, 9., 9., 9.\n")
    _, a, s = assert_same_setup(text)
    assert a["u0"][3 * 1 + 1] == -0.5
    assert a["xyz"][2, 0] == 1.0


def test_unsupported_and_malformed_decks():
    with pytest.raises(AxgError) as e:
        D.Deck(BRICK.replace("TYPE=C3D8", "TYPE=C3D10"))
    assert "Nicht unterstützter Elementtyp" in str(e.value)
    with pytest.raises(AxgError) as e:
        D.Deck(BRICK.replace("LATE, 1, 1.0", "NOSUCH, 1, 1.0"))
    assert "Unbekanntes Knotenset" in str(e.value)
    with pytest.raises(AxgError):
        D.Deck(BRICK.replace("2, 2, 9, 10, 3", "2, 2, 99, 10, 3"))
    with pytest.raises(AxgError) as e:
        D.Deck(BRICK.replace("BLK, P2, 3.0", "BLK, BX, 3.0"))
    assert "DLOAD" in str(e.value)


def test_procedure_keywords():
    """*STEP, NLGEOM / INC= and *STATIC increments, *FREQUENCY and *BUCKLE counts
    (inp.rs:953-1081) as the reader records them, against the oracle's reader."""
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from eigen_decks import bar_deck, nl_cantilever_deck
    cases = [
        nl_cantilever_deck(ninc=4),
        nl_cantilever_deck(ninc=3).replace("*STEP, NLGEOM", "*STEP, NLGEOM, INC=2"),
        nl_cantilever_deck(ninc=4).replace("*STATIC\n0.25, 1.0\n", "*STATIC\n0.3, 1.0\n"),
        nl_cantilever_deck(ninc=4).replace("*STEP, NLGEOM", "*STEP"),
        bar_deck(3, nmodes=5),
        bar_deck(3).replace("*FREQUENCY\n3\n", "*FREQUENCY\n"),
    ]
    for text in cases:
        m = O.parse_inp(text)
        p = D.Deck(text).procedure()
        assert p["nlgeom"] == m.nlgeom
        assert p["nmodes"] == m.nmodes and p["nbuckle"] == m.nbuckle
        if m.nlgeom:
            assert p["ninc"] == m.ninc
    assert D.Deck(cases[0]).procedure()["ninc"] == 4
    assert D.Deck(cases[1]).procedure()["ninc"] == 2   # INC= caps the increments
    assert D.Deck(cases[2]).procedure()["ninc"] == 3   # round(1 / 0.3)
    assert D.Deck(cases[5]).procedure()["nmodes"] == 10


def test_number_spellings_match_the_oracle():
    """Node ids and coordinates in every spelling the reader treats differently (plain, signed,
    exponent forms, Fortran D and sign-only exponents, trailing dots, leading zeros, huge and
    tiny magnitudes, negative zero): the native reader's fast and full paths against the
    oracle's restatement of parse_f64_inner / parse_usize (inp.rs:100-141). The corners of
    the element get spellings of 0 and 1; nodes no element uses carry the extremes."""
    zero = ["0", "-0", "-0.0", "0.", "+0.0", "0e5", "0D0", "0.0+0", "1.5e-320", "-1.5e-320",
            "000", ".0"]
    one = ["1.", "+1.0", "1.0e+0", "1E0", "10e-1", "1.d0", "1.0D+0", ".1e1", "1.0+0",
           "0.99999999999999999", "9.999999999999999e-1", "1.0000000000000002", "10D-1", "1..",
           "+.1E+1"]
    ids = ["1", "+2", "003", "4.0", "5", "6", "7", "8"]
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    lines = ["*NODE\n"]
    for k, c in enumerate(corners):
        xs = [(one if c[q] else zero)[(3 * k + q) % len(one if c[q] else zero)] for q in range(3)]
        lines.append(f"{ids[k]}, {xs[0]}, {xs[1]}, {xs[2]}\n")
    lines += ["9, 2.5e300, -1e-400, 123456789.123456789\n", "10, 5d-1, 1.0-1, -.25\n",
              "11, 1e400, 0.30000000000000004, 7.0D+2\n"]
    text = "".join(lines) + (
        "*ELEMENT, TYPE=C3D8, ELSET=B\n1, 1, 2, 3, 4, 5, 6, 7, 8\n*MATERIAL, NAME=M\n"
        "*ELASTIC\n2.1d5, .3\n*SOLID SECTION, ELSET=B, MATERIAL=M\n*BOUNDARY\n1, 1, 3\n")
    d, a, s = assert_same_setup(text)
    assert a["e"][0] == 210000.0 and a["nu"][0] == 0.3
    assert a["xyz"][8, 0] == 2.5e300 and a["xyz"][9, 1] == 0.1 and np.isinf(a["xyz"][10, 0])


def test_lognormal_brick_deck_is_the_brick():
    """The per-element-material deck of the deck benchmark describes brick_lognormal exactly:
    coordinates, connectivity, moduli (one material per element), DOF numbering and the top
    pressure load, through the native reader; the oracle's reader agrees on the same text."""
    from axia_gpu import meshgen as MG
    text, p = MG.lognormal_brick_deck(5, seed=2)
    d, a, s = assert_same_setup(text)
    assert np.array_equal(a["xyz"], p.xyz) and np.array_equal(a["conn"], p.conn)
    assert np.array_equal(a["e"], p.e) and np.array_equal(a["ind_of"], p.ind_of)
    assert np.allclose(a["f_full"], p.f_full[0], rtol=0, atol=1e-15)
    assert abs(a["f_full"][2::3].sum() + 25.0) < 1e-12


def test_large_deck_with_sparse_unsorted_ids():
    """A C3D20 deck with more than 32k *NODE and *ELEMENT data lines (the reader splits them
    over threads; every element continues on a second line), node ids sparse (1000 i + 7,
    so the id lookup is the hash map), listed out of order with a repeated id (first
    definition wins), element ids descending: the native setup equals the oracle's bit for
    bit."""
    from axia_gpu import meshgen as MG
    xyz, conn = MG.brick_hex20(40, 20, 21)
    nnode, ne = len(xyz), len(conn)
    assert nnode > 32768 and 2 * ne > 32768
    nid = 1000 * np.arange(nnode) + 7
    order = np.random.default_rng(3).permutation(nnode)
    out = ["*NODE\n"]
    out += [f"{nid[i]}, {float(xyz[i, 0])!r}, {float(xyz[i, 1])!r}, {float(xyz[i, 2])!r}\n"
            for i in order]
    out.append(f"{nid[order[0]]}, 99.0, 99.0, 99.0\n")  # a repeated id: ignored
    out.append("*ELEMENT, TYPE=C3D20, ELSET=SOLID\n")
    for k in range(ne):
        ids = [str(2 * (ne - k) + 1)] + [str(nid[a]) for a in conn[k]]
        out.append(", ".join(ids[:16]) + ",\n" + ", ".join(ids[16:]) + "\n")
    out.append("*MATERIAL, NAME=MAT\n*ELASTIC\n70000., 0.3\n*SOLID SECTION, ELSET=SOLID, MATERIAL=MAT\n")
    fix = nid[xyz[:, 0] == 0.0]
    out.append("*NSET, NSET=FIX\n")
    out += [", ".join(str(v) for v in fix[j:j + 16]) + "\n" for j in range(0, len(fix), 16)]
    out.append("*STEP\n*STATIC\n*BOUNDARY\nFIX, 1, 3\n")
    out.append(f"*CLOAD\n{nid[np.argmax(xyz[:, 0])]}, 3, -2.5\n")
    out.append(f"*DLOAD\n{2 * ne + 1}, P2, 1.5\n3, P1, 0.5\n*END STEP\n")
    text = "".join(out)
    d, a, s = assert_same_setup(text)
    assert np.array_equal(a["node_ids"], np.sort(nid))
    assert np.array_equal(a["elem_ids"], 2 * np.arange(ne) + 3)
    # a bad id deep inside the threaded blocks raises the reader's error for that token
    bad = text.replace(f"\n{nid[order[40000]]}, ", "\nX12, ", 1)
    with pytest.raises(AxgError, match="Keine ganze Zahl: X12"):
        D.Deck(bad)
    bad = text.replace(f", {nid[conn[ne - 100, 18]]}, ", ", 12Y, ", 1)
    assert bad != text
    with pytest.raises(AxgError, match="Keine ganze Zahl: 12Y"):
        D.Deck(bad)
