"""Deck -> results on the device (axg_deck_solve) against the oracle's
solve_linear restatement: displacements and nodal stresses within 1e-8
(FP64 mode), reactions bit-exact for the same displacements, and the
reference's own deck checks (patch tests, NAFEMS LE10) through the product
path alone."""
import os
import sys

import numpy as np
import pytest

import axia_gpu as G
from axia_gpu import deck as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
DECKS = ["cube_tension", "c3d20_patch", "patch_c3d8", "le10_c3d20", "le10_c3d20_fine"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def read(name):
    return open(os.path.join(GOLD, name + ".inp")).read()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("name", DECKS)
def test_deck_solve_vs_oracle(ctx, name):
    say("case: deck", name)
    text = read(name)
    d, res = D.solve_deck(text, ctx=ctx, precond="block3", tol=1e-13)
    m, ref = O.solve_deck(text)
    assert np.array_equal(res.node_ids, np.array(m.node_ids))
    assert rel(res.u_full, ref.u_full) < 1e-8
    smax = np.abs(ref.stress).max()
    assert np.abs(res.stress - ref.stress).max() <= 1e-8 * smax
    rmax = max(np.abs(ref.rf_full).max(), 1e-300)
    assert np.abs(res.rf - ref.rf_full).max() <= 1e-8 * rmax
    assert res.nfree == ref.setup.n_free
    assert "GPU PCG" in res.solver


@pytest.mark.parametrize("name", ["cube_tension", "le10_c3d20"])
def test_reactions_and_reduced_rhs_bitexact(ctx, name):
    """Given the oracle's displacements, K_full u - f is the oracle's bit for bit; the
    reduced right-hand side with prescribed values matches DofMap::reduce."""
    text = read(name)
    dk = D.Deck(text)
    a = dk.arrays()
    m, ref = O.solve_deck(text)
    mesh = G.Mesh(dk.nn, a["xyz"], a["conn"], a["e"], a["nu"], a["ind_of"], ctx=ctx)
    mesh.assemble("fp64")
    assert np.array_equal(mesh.reactions(ref.u_full, a["f_full"]), ref.rf_full)
    # prescribed values: move the loaded face instead of pressing it
    u0 = np.zeros(3 * dk.nnode)
    pres = np.nonzero(a["ind_of"] < 0)[0]
    u0[pres[::2]] = 1e-3 * (1 + np.arange(len(pres[::2])))
    ti, tj, tv = O.assemble_triplets(dk.nn, a["xyz"], a["conn"], a["e"], a["nu"])
    _, _, _, rhs = O.reduce(a["ind_of"], u0, dk.n_free, ti, tj, tv, a["f_full"])
    assert np.array_equal(mesh.reduce_rhs(a["f_full"], u0), rhs)
    # the full diagonal: diagonal triplets summed in element order
    diag = np.zeros(3 * dk.nnode)
    for i, j, v in zip(ti, tj, tv):
        if i == j:
            diag[i] += v
    assert np.array_equal(mesh.full_diagonal(), diag)


def test_patch_decks_through_product_path(ctx):
    """lib.rs:429-462 and 751-774 evaluated on the product's own results."""
    _, r8 = D.solve_deck(read("cube_tension"), ctx=ctx, tol=1e-12)
    ux = np.mean([r8.u_full[3 * r8.node_index(i)] for i in (2, 3, 6, 7)])
    assert abs(ux - 0.01) < 1e-6
    assert abs(r8.stress[:, 0].mean() - 210.0) < 1e-3
    _, r20 = D.solve_deck(read("c3d20_patch"), ctx=ctx, tol=1e-12)
    ux = np.mean([r20.u_full[3 * r20.node_index(i)] for i in (2, 3, 6, 7)])
    assert abs(ux + 0.01) < 1e-5
    assert abs(r20.stress[:, 0].mean() + 210.0) < 1.0


@pytest.mark.parametrize("name,axia", [("le10_c3d20", -5.51), ("le10_c3d20_fine", -5.50)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nafems_le10_through_product_path(ctx, name, axia, precision):
    """nafems.rs:50-90: sigma_yy at D within 5 % / 3 % of -5.38, and the reference's values."""
    text = read(name)
    m = O.parse_inp(text)  # only for the node id of point D
    _, r = D.solve_deck(text, ctx=ctx, precision=precision, precond="block3", tol=1e-11)
    syy = r.stress[r.node_index(m.nsets["D"][0]), 1]
    assert round(syy, 2) == axia
