"""Pin the oracle's NLGEOM restatement (nlgeom.rs St. Venant-Kirchhoff hex elements and the
continuum Newton driver) against closed forms and finite differences (CPU only)."""
import math

import numpy as np
import pytest

from axia_gpu import meshgen as MG
from eigen_decks import nl_bar_deck, nl_cantilever_deck
from oracle import oracle as O


def _cube(nn, lx=1.0, ly=1.0, lz=1.0):
    xyz, conn = (MG.brick_hex20 if nn == 20 else MG.brick_hex8)(1, 1, 1)
    return (xyz * np.array([lx, ly, lz]))[conn[0]]


@pytest.mark.parametrize("nn", [8, 20])
def test_zero_displacement_is_the_linear_element(nn):
    """At u = 0: F = I, S = 0, so the tangent is the linear Ke and the internal force 0."""
    rng = np.random.default_rng(nn)
    x = _cube(nn, 2.0, 1.0, 0.7) + rng.uniform(-0.05, 0.05, (nn, 3))
    ke, fe, cau, gl, vol = O.hex_nl(x, np.zeros(3 * nn), 2100.0, 0.3)
    k_lin, v_lin = O.hex_stiffness(x, 2100.0, 0.3)
    assert np.array_equal(ke, k_lin)
    assert np.all(fe == 0.0) and np.all(cau == 0.0) and np.all(gl == 0.0)
    assert vol == v_lin


@pytest.mark.parametrize("nn", [8, 20])
def test_uniaxial_stretch_closed_form(nn):
    """u_x = d x on an L x b x h box, nu = 0: E11 = ((1 + d)^2 - 1) / 2, S11 = E E11, the
    nominal stress (1 + d) S11 on the x = L face and Cauchy (1 + d) S11 (J = 1 + d)."""
    L, b, h, E, d = 2.0, 0.8, 0.5, 1000.0, 0.3
    x = _cube(nn, L, b, h)
    ue = np.zeros((nn, 3))
    ue[:, 0] = d * x[:, 0]
    ke, fe, cau, gl, vol = O.hex_nl(x, ue.ravel(), E, 0.0)
    e11 = ((1 + d) ** 2 - 1) / 2
    s11 = E * e11
    end = x[:, 0] == L
    assert math.isclose(fe[0::3][end].sum(), (1 + d) * s11 * b * h, rel_tol=1e-12)
    assert math.isclose(fe[0::3][~end].sum(), -(1 + d) * s11 * b * h, rel_tol=1e-12)
    assert np.allclose(fe[1::3], 0, atol=1e-12 * s11) and np.allclose(fe[2::3], 0, atol=1e-12 * s11)
    assert math.isclose(cau[0], (1 + d) * s11, rel_tol=1e-12)
    assert math.isclose(gl[0], e11, rel_tol=1e-12)
    if nn == 20:
        cs, gs = O.hex20_nodal_nl(x, ue.ravel(), E, 0.0)
        assert np.allclose(cs[:, 0], (1 + d) * s11, rtol=1e-12)
        assert np.allclose(gs[:, 0], e11, rtol=1e-12)


@pytest.mark.parametrize("nn", [8, 20])
def test_rigid_rotation_is_stress_free(nn):
    """A finite rotation leaves E = 0: no internal force, no stress (objectivity)."""
    x = _cube(nn, 1.5, 1.0, 0.8)
    t = 0.6
    rot = np.array([[math.cos(t), -math.sin(t), 0], [math.sin(t), math.cos(t), 0], [0, 0, 1.0]])
    ue = (x @ rot.T - x).ravel()
    ke, fe, cau, gl, vol = O.hex_nl(x, ue, 1000.0, 0.3)
    assert np.abs(fe).max() < 1e-10 and np.abs(cau).max() < 1e-10 and np.abs(gl).max() < 1e-13


@pytest.mark.parametrize("nn", [8, 20])
def test_tangent_is_the_derivative_of_the_internal_force(nn):
    """K_T du = d f_int / d u (central differences), geometric term included."""
    rng = np.random.default_rng(10 + nn)
    x = _cube(nn, 1.0, 1.2, 0.9) + rng.uniform(-0.05, 0.05, (nn, 3))
    ue = rng.uniform(-0.08, 0.08, 3 * nn)
    ke, fe, *_ = O.hex_nl(x, ue, 1000.0, 0.25)
    du = rng.standard_normal(3 * nn)
    h = 1e-6
    fp = O.hex_nl(x, ue + h * du, 1000.0, 0.25)[1]
    fm = O.hex_nl(x, ue - h * du, 1000.0, 0.25)[1]
    fd = (fp - fm) / (2 * h)
    assert np.linalg.norm(ke @ du - fd) <= 1e-7 * np.linalg.norm(fd)
    assert np.allclose(ke, ke.T, rtol=0, atol=1e-12 * np.abs(ke).max())


def test_inversion_error():
    x = _cube(8)
    ue = np.zeros((8, 3))
    ue[:, 0] = -2.0 * x[:, 0]  # F11 = -1
    with pytest.raises(O.OracleError) as ex:
        O.hex_nl(x, ue.ravel(), 1000.0, 0.3)
    assert ex.value.code == 10


@pytest.mark.parametrize("nn", [8, 20])
def test_newton_bar_stretch_closed_form(nn):
    """A bar stretched by a prescribed end displacement in 4 increments: the homogeneous
    state is exact for both elements, so the reaction equals A (1 + d) E E11."""
    d, E, L = 0.5, 1000.0, 4.0
    res = O.nlgeom_deck(nl_bar_deck(nel=4, length=L, delta=d, e=E, nn=nn))
    lam = 1 + d / L
    e11 = (lam ** 2 - 1) / 2
    model = O.parse_inp(nl_bar_deck(nel=4, length=L, delta=d, e=E, nn=nn))
    nn_, xyz, conn, *_ = model.mesh_arrays()
    end = np.nonzero(xyz[:, 0] == L)[0]
    assert math.isclose(res.rf_full[3 * end].sum(), lam * E * e11, rel_tol=1e-9)
    assert np.allclose(res.u_full[0::3], (lam - 1) * xyz[:, 0], atol=1e-12)
    assert np.allclose(res.stress[:, 0], lam * E * e11, rtol=1e-9)
    assert res.ninc == 4


def test_newton_small_load_is_linear():
    """A tiny tip load: the NLGEOM solution approaches the linear one (relative O(load))."""
    text = nl_cantilever_deck(p_total=-1e-6, ninc=1)
    res = O.nlgeom_deck(text)
    lin_text = text.replace("*STEP, NLGEOM", "*STEP")
    _, lin = O.solve_deck(lin_text)
    rel = np.linalg.norm(res.u_full - lin.u_full) / np.linalg.norm(lin.u_full)
    assert rel < 1e-6


def test_newton_large_deflection_stiffens_and_converges():
    """A large tip load: geometric stiffening shortens the deflection against the linear
    solution, the Newton residual is below tolerance, and the follower pressure differs
    from the dead load."""
    text = nl_cantilever_deck(p_total=-2.0, ninc=4)
    res = O.nlgeom_deck(text)
    _, lin = O.solve_deck(text.replace("*STEP, NLGEOM", "*STEP"))
    model = O.parse_inp(text)
    _, xyz, *_ = model.mesh_arrays()
    tip = np.nonzero(xyz[:, 0] == 6.0)[0]
    w_nl = res.u_full[3 * tip + 2].mean()
    w_lin = lin.u_full[3 * tip + 2].mean()
    assert w_nl < 0 and abs(w_nl) < abs(w_lin)
    # equilibrium: reactions balance the applied load
    assert math.isclose(res.rf_full[2::3][np.nonzero(xyz[:, 0] == 0.0)[0]].sum(), 2.0, rel_tol=1e-6)
    pr = O.nlgeom_deck(nl_cantilever_deck(pressure=0.05, ninc=3))
    assert np.isfinite(pr.u_full).all() and pr.residual < 1e-6
