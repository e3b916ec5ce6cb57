"""NLGEOM on the device against the oracle's restatement of nlgeom.rs and
solve_continuum_newton: element tangents, internal forces, element means and the
assembled K_T bit for bit; Newton results within 1e-8 (FP64 solves) and the
closed form of a stretched bar through the product path alone."""
import math
import sys

import numpy as np
import pytest

import axia_gpu as G
from axia_gpu import deck as D
from axia_gpu import meshgen as MG
from eigen_decks import nl_bar_deck, nl_cantilever_deck
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("nn,shape,seed", [(8, (4, 3, 2), 3), (20, (2, 2, 1), 4)])
def test_nl_assemble_bitexact(ctx, nn, shape, seed):
    say("case: nl_assemble", nn, shape)
    xyz, conn = (MG.brick_hex20 if nn == 20 else MG.brick_hex8)(*shape)
    rng = np.random.default_rng(seed)
    xyz = xyz + rng.uniform(-0.1, 0.1, xyz.shape) * (xyz > 0)
    ind = MG.free_map(len(xyz), np.nonzero(xyz[:, 0] == 0.0)[0])
    e = rng.uniform(800.0, 1200.0, len(conn))
    nu = np.full(len(conn), 0.3)
    m = G.Mesh(nn, xyz, conn, e, nu, ind, ctx=ctx)
    m.assemble("fp64")
    u = rng.uniform(-0.05, 0.05, 3 * len(xyz))
    u[ind < 0] = 0.0
    fint, cau, gl = m.nl_assemble(u)
    ref_f, trips, ref_c, ref_g = O.nl_assemble(nn, xyz, conn, e, nu, u)
    assert np.array_equal(fint, ref_f)
    assert np.array_equal(cau, ref_c) and np.array_equal(gl, ref_g)
    ke = m.ke()
    u3 = u.reshape(-1, 3)
    for el in range(len(conn)):
        k_ref = O.hex_nl(xyz[conn[el]], u3[conn[el]].ravel(), e[el], nu[el])[0]
        assert np.array_equal(ke[el], k_ref), f"element {el}"
    ti, tj, tv = trips
    oi, oj, ov, _ = O.reduce(ind, np.zeros(3 * len(xyz)), m.n_free, ti, tj, tv,
                             np.zeros(3 * len(xyz)))
    ref = O.csr_from_triplets(m.n_free, oi, oj, ov)
    ip, ix, dv = m.csr()
    assert np.array_equal(ip, ref.indptr) and np.array_equal(ix, ref.indices)
    assert np.array_equal(dv, ref.data)


@pytest.mark.parametrize("nn", [8, 20])
def test_deck_nlgeom_bar_closed_form(ctx, nn):
    """Prescribed end stretch in 4 increments: the reaction is A (1 + d) E E11 and the
    displacement field is the homogeneous one, through the product path alone."""
    d, E, L = 0.5, 1000.0, 4.0
    text = nl_bar_deck(nel=4, length=L, delta=d, e=E, nn=nn)
    dk, res = D.nlgeom_deck(text, ctx=ctx, tol=1e-12)
    lam = 1 + d / L
    e11 = (lam ** 2 - 1) / 2
    xyz = dk.arrays()["xyz"]
    end = np.nonzero(xyz[:, 0] == L)[0]
    say("  reaction", res.rf[3 * end].sum(), "closed form", lam * E * e11, res.solver)
    assert math.isclose(res.rf[3 * end].sum(), lam * E * e11, rel_tol=1e-9)
    assert np.allclose(res.u_full[0::3], (lam - 1) * xyz[:, 0], atol=1e-10)
    assert np.allclose(res.stress[:, 0], lam * E * e11, rtol=1e-9)
    assert np.allclose(res.strain[:, 0], e11, rtol=1e-9)
    assert res.ninc == 4


@pytest.mark.parametrize("kind", ["dead", "follower", "c3d20"])
def test_deck_nlgeom_vs_oracle(ctx, kind):
    if kind == "dead":
        text = nl_cantilever_deck(p_total=-2.0, ninc=4)
    elif kind == "follower":
        text = nl_cantilever_deck(pressure=0.05, ninc=3)
    else:
        text = nl_cantilever_deck(nx=4, ny=1, nz=1, p_total=-1.0, ninc=2, nn=20)
    say("case: nlgeom", kind)
    dk, res = D.nlgeom_deck(text, ctx=ctx, tol=1e-12)
    ref = O.nlgeom_deck(text)
    say("  rel u", rel(res.u_full, ref.u_full), "inc", res.ninc, ref.ninc, res.solver)
    assert rel(res.u_full, ref.u_full) < 1e-8
    assert res.ninc == ref.ninc
    rmax = np.abs(ref.rf_full).max()
    assert np.abs(res.rf - ref.rf_full).max() <= 1e-7 * rmax
    smax = np.abs(ref.stress).max()
    assert np.abs(res.stress - ref.stress).max() <= 1e-7 * smax
    assert np.abs(res.strain - ref.strain).max() <= 1e-7 * np.abs(ref.strain).max()


def test_deck_nlgeom_requires_the_step(ctx):
    text = nl_cantilever_deck().replace("*STEP, NLGEOM", "*STEP")
    dk = D.Deck(text)
    assert dk.procedure()["nlgeom"] is False
    with pytest.raises(ValueError):
        dk.nlgeom(ctx=ctx)
    assert D.Deck(nl_cantilever_deck(ninc=4)).procedure() == {
        "nmodes": 0, "nbuckle": 0, "nlgeom": True, "ninc": 4}
