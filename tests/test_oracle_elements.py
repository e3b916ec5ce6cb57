"""Pin the oracle's element routines (CPU only).

Checks are independent of the oracle's formulas: sympy-derived shape
functions, exact polynomial integration on affine elements, rigid-body null
spaces, ranks, symmetry, constant-strain patch exactness and the total force
of a face pressure.
"""
import itertools

import numpy as np
import pytest
import sympy as sp

from oracle import oracle as O

XI, ETA, ZETA = sp.symbols("xi eta zeta")

HEX8_NODES = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
              (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
HEX20_NODES = HEX8_NODES + [(0, -1, -1), (1, 0, -1), (0, 1, -1), (-1, 0, -1),
                            (0, -1, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1),
                            (-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)]


def sym_shape(nodes):
    """Textbook Lagrange / serendipity shape functions, derived symbolically."""
    out = []
    for (a, b, c) in nodes:
        if len(nodes) == 8:
            out.append(sp.Rational(1, 8) * (1 + a * XI) * (1 + b * ETA) * (1 + c * ZETA))
        elif 0 not in (a, b, c):
            out.append(sp.Rational(1, 8) * (1 + a * XI) * (1 + b * ETA) * (1 + c * ZETA)
                       * (a * XI + b * ETA + c * ZETA - 2))
        else:
            f = sp.Rational(1, 4)
            for v, s in ((a, XI), (b, ETA), (c, ZETA)):
                f *= (1 - s ** 2) if v == 0 else (1 + v * s)
            out.append(f)
    return out


SYM = {8: sym_shape(HEX8_NODES), 20: sym_shape(HEX20_NODES)}
DSYM = {nn: [[sp.lambdify((XI, ETA, ZETA), sp.diff(N, v)) for v in (XI, ETA, ZETA)]
             for N in SYM[nn]] for nn in (8, 20)}
NSYM = {nn: [sp.lambdify((XI, ETA, ZETA), N) for N in SYM[nn]] for nn in (8, 20)}


@pytest.mark.parametrize("nn", [8, 20])
def test_shape_functions_match_textbook(nn):
    rng = np.random.default_rng(0)
    for _ in range(20):
        p = rng.uniform(-1, 1, 3)
        n, dn = O.hex_shape(nn, *p)
        n_ref = np.array([f(*p) for f in NSYM[nn]])
        dn_ref = np.array([[d(*p) for d in row] for row in DSYM[nn]])
        assert np.allclose(n, n_ref, atol=1e-15)
        assert np.allclose(dn, dn_ref, atol=1e-15)
        assert abs(n.sum() - 1.0) < 1e-14          # partition of unity (quadratic.rs:890-894)
        assert np.allclose(dn.sum(0), 0.0, atol=1e-14)


@pytest.mark.parametrize("nn", [8, 20])
def test_kronecker_delta_at_nodes(nn):
    nodes = HEX8_NODES if nn == 8 else HEX20_NODES
    for k, p in enumerate(nodes):
        n, _ = O.hex_shape(nn, *map(float, p))
        e = np.zeros(nn)
        e[k] = 1.0
        assert np.allclose(n, e, atol=1e-15)


def d_iso(e, nu):
    lam = e * nu / ((1 + nu) * (1 - 2 * nu))
    mu = e / (2 * (1 + nu))
    d = np.zeros((6, 6))
    d[:3, :3] = lam
    d[np.arange(3), np.arange(3)] = lam + 2 * mu
    d[np.arange(3, 6), np.arange(3, 6)] = mu
    return d


def test_d_iso_closed_form_and_errors():
    assert np.allclose(O.d_iso_3d(210000.0, 0.3), d_iso(210000.0, 0.3), rtol=1e-15)
    for e, nu in ((1.0, 0.5), (1.0, -1.0), (0.0, 0.3), (-1.0, 0.2)):
        with pytest.raises(O.OracleError):
            O.d_iso_3d(e, nu)


def ke_reference(nn, xyz, e, nu, npts):
    """Independent Ke: textbook B from sympy derivatives, Gauss-Legendre npts^3."""
    g, w = np.polynomial.legendre.leggauss(npts)
    d = d_iso(e, nu)
    m = 3 * nn
    ke = np.zeros((m, m))
    for (i, a), (j, b), (k, c) in itertools.product(enumerate(g), repeat=3):
        dn = np.array([[f(a, b, c) for f in row] for row in DSYM[nn]])  # nn x 3
        J = xyz.T @ dn
        dndx = dn @ np.linalg.inv(J)
        B = np.zeros((6, m))
        for n_ in range(nn):
            dx, dy, dz = dndx[n_]
            B[0, 3 * n_], B[1, 3 * n_ + 1], B[2, 3 * n_ + 2] = dx, dy, dz
            B[3, 3 * n_], B[3, 3 * n_ + 1] = dy, dx
            B[4, 3 * n_ + 1], B[4, 3 * n_ + 2] = dz, dy
            B[5, 3 * n_], B[5, 3 * n_ + 2] = dz, dx
        ke += w[i] * w[j] * w[k] * np.linalg.det(J) * B.T @ d @ B
    return ke


def affine_element(nn, seed):
    rng = np.random.default_rng(seed)
    A = np.eye(3) + rng.uniform(-0.3, 0.3, (3, 3))
    t = rng.uniform(-5, 5, 3)
    nodes = np.array(HEX8_NODES if nn == 8 else HEX20_NODES, float)
    return (0.5 * (nodes + 1.0)) @ A.T * 2.0 + t, abs(np.linalg.det(A)) * 8.0


@pytest.mark.parametrize("nn", [8, 20])
def test_stiffness_exact_on_affine_elements(nn):
    """On affine elements the reference rules are exact; a higher rule must agree."""
    for seed in range(3):
        xyz, vol = affine_element(nn, seed)
        ke, v = O.hex_stiffness(xyz, 70000.0, 0.27)
        ref = ke_reference(nn, xyz, 70000.0, 0.27, 4)
        assert np.allclose(ke, ref, rtol=1e-11, atol=1e-9 * np.abs(ref).max())
        assert abs(v - vol) < 1e-11 * vol


def test_unit_cube_hex8_closed_form_diagonal():
    """Unit cube, nu = 0: K_00 = E * (1/9 + 1/18 + 1/18) = E/4 from the exact integral."""
    xyz = 0.5 * (np.array(HEX8_NODES, float) + 1.0)
    ke, vol = O.hex_stiffness(xyz, 1.0, 0.0)
    # with nu = 0: int (dN0/dx)^2 + 1/2 (dN0/dy)^2 + 1/2 (dN0/dz)^2 over the unit cube
    assert abs(ke[0, 0] - (1 / 9 + 0.5 / 9 + 0.5 / 9)) < 1e-15
    assert abs(vol - 1.0) < 1e-15


def rigid_modes(xyz):
    n = len(xyz)
    R = np.zeros((3 * n, 6))
    for a, (x, y, z) in enumerate(xyz):
        R[3 * a:3 * a + 3, 0:3] = np.eye(3)
        R[3 * a:3 * a + 3, 3] = [0, -z, y]
        R[3 * a:3 * a + 3, 4] = [z, 0, -x]
        R[3 * a:3 * a + 3, 5] = [-y, x, 0]
    return R


@pytest.mark.parametrize("nn", [8, 20])
def test_stiffness_null_space_rank_symmetry_distorted(nn):
    rng = np.random.default_rng(nn)
    nodes = np.array(HEX8_NODES if nn == 8 else HEX20_NODES, float)
    xyz = nodes * [2.0, 1.5, 1.0] + rng.uniform(-0.15, 0.15, nodes.shape)
    ke, vol = O.hex_stiffness(xyz, 210000.0, 0.3)
    scale = np.abs(ke).max()
    assert np.allclose(ke, ke.T, rtol=0, atol=1e-12 * scale)
    assert np.abs(ke @ rigid_modes(xyz)).max() < 1e-9 * scale
    ev = np.linalg.eigvalsh(0.5 * (ke + ke.T))
    assert (ev > 1e-8 * ev.max()).sum() == 3 * nn - 6
    assert ev.min() > -1e-9 * ev.max()
    # volume vs a fine independent integration of det J
    ref = ke_reference(nn, xyz, 1.0, 0.0, 1)  # only to exercise the helper
    assert ref.shape == ke.shape
    g, w = np.polynomial.legendre.leggauss(6)
    v = 0.0
    for (i, a), (j, b), (k, c) in itertools.product(enumerate(g), repeat=3):
        dn = np.array([[f(a, b, c) for f in row] for row in DSYM[nn]])
        v += w[i] * w[j] * w[k] * np.linalg.det(xyz.T @ dn)
    assert abs(vol - v) < 1e-9 * abs(v) if nn == 8 else abs(vol - v) < 1e-6 * abs(v)


def test_negative_jacobian_and_singular():
    xyz = 0.5 * (np.array(HEX8_NODES, float) + 1.0)
    with pytest.raises(O.OracleError) as e:
        O.hex_stiffness(xyz[[4, 5, 6, 7, 0, 1, 2, 3]], 1.0, 0.3)
    assert e.value.code == 3
    with pytest.raises(O.OracleError) as e:
        O.hex_stiffness(np.zeros((8, 3)), 1.0, 0.3)
    assert e.value.code == 2


@pytest.mark.parametrize("nn", [8, 20])
def test_constant_strain_stress_exact(nn):
    """u = eps x reproduces sigma = D eps at every node (patch test of stress recovery)."""
    rng = np.random.default_rng(7)
    nodes = np.array(HEX8_NODES if nn == 8 else HEX20_NODES, float)
    xyz = nodes + rng.uniform(-0.2, 0.2, nodes.shape)
    g = rng.standard_normal((3, 3)) * 1e-3
    u = xyz @ g.T
    eps = np.array([g[0, 0], g[1, 1], g[2, 2], g[0, 1] + g[1, 0], g[1, 2] + g[2, 1],
                    g[2, 0] + g[0, 2]])
    sig = d_iso(1000.0, 0.25) @ eps
    out = O.hex_nodal_stress(xyz, u.ravel(), 1000.0, 0.25)
    assert np.allclose(out, np.tile(sig, (nn, 1)), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("nn,face,normal", [(8, 2, (0, 0, -1)), (8, 1, (0, 0, 1)),
                                            (8, 4, (-1, 0, 0)), (20, 2, (0, 0, -1)),
                                            (20, 3, (0, 1, 0)), (20, 6, (1, 0, 0))])
def test_face_pressure_total_force(nn, face, normal):
    """Positive pressure pushes into the element: total = p * area * inward normal."""
    nodes = np.array(HEX8_NODES if nn == 8 else HEX20_NODES, float)
    xyz = (nodes + 1.0) * [1.5, 2.0, 0.5]  # box 3 x 4 x 1
    area = {1: 12.0, 2: 12.0, 3: 3.0, 4: 4.0, 5: 3.0, 6: 4.0}[face]
    fe = O.hex_face_pressure(xyz, face, 2.5).reshape(nn, 3)
    assert np.allclose(fe.sum(0), 2.5 * area * np.array(normal, float), atol=1e-12)


def test_von_mises():
    assert abs(O.von_mises([1.0, 0, 0, 0, 0, 0]) - 1.0) < 1e-15
    assert abs(O.von_mises([0, 0, 0, 1.0, 0, 0]) - np.sqrt(3.0)) < 1e-15
    assert O.von_mises([2.0, 2.0, 2.0, 0, 0, 0]) == 0.0
