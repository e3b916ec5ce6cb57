"""Pin the oracle's *FREQUENCY restatement (lumped mass, Jacobi eigen, subspace iteration)
against LAPACK and closed forms (CPU only)."""
import math

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import oracle as O

from eigen_decks import bar_deck

def test_jacobi_eigen_vs_lapack():
    rng = np.random.default_rng(0)
    for n in (2, 5, 9):
        a = rng.standard_normal((n, n))
        a = a + a.T
        w, v = O.jacobi_eigen(a)
        ref = np.linalg.eigvalsh(a)
        assert np.allclose(np.sort(w), ref, rtol=1e-12, atol=1e-12)
        assert np.allclose(v.T @ v, np.eye(n), atol=1e-12)
        assert np.allclose(a @ v, v * w[None, :], atol=1e-10)


def test_lumped_mass_closed_form():
    """Each node of an element carries 1/nn of density x volume; the total is 3 rho V."""
    m_full = O.lumped_mass(8, np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0],
                                        [0, 0, 3], [2, 0, 3], [2, 1, 3], [0, 1, 3]], float),
                           np.arange(8)[None, :], [0.5])
    assert np.allclose(m_full, 0.5 * 6.0 / 8.0, rtol=1e-15)


def test_m_orthonormalize():
    rng = np.random.default_rng(1)
    m = rng.uniform(0.5, 2.0, 30)
    x = O.m_orthonormalize(rng.standard_normal((5, 30)), m)
    assert np.allclose((x * m) @ x.T, np.eye(5), atol=1e-13)


def test_subspace_vs_lapack_generalized():
    """The lowest modes of K phi = lambda M phi against scipy's dense generalized eigh."""
    text = bar_deck(6, nmodes=3).replace("ALLN, 2, 3\n", "")  # full 3-D motion
    model, s, csr, m_ind, vals, vecs, freqs, sweeps = O.frequency_deck(text)
    ref = sla.eigh(csr.to_dense(), np.diag(m_ind), eigvals_only=True)[:3]
    assert np.allclose(vals, ref, rtol=1e-8)
    # mode shapes satisfy the eigen-equation
    k = csr.to_dense()
    for lam, v in zip(vals, vecs):
        r = k @ v - lam * m_ind * v
        assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(k @ v)


def test_axial_bar_closed_form():
    """Fixed-free bar, axial modes only: f_k = (2k - 1) c / (4 L), c = sqrt(E / rho); lumped
    mass and 40 linear elements put the first mode within 0.1 %."""
    _, _, _, _, vals, _, freqs, _ = O.frequency_deck(bar_deck(40, nmodes=2))
    c = math.sqrt(210000.0 / 7.85e-9)
    f1 = c / (4 * 100.0)
    assert abs(freqs[0] - f1) / f1 < 1e-3
    assert abs(freqs[1] - 3 * f1) / (3 * f1) < 1e-2


def test_frequency_deck_errors():
    with pytest.raises(ValueError):
        O.frequency_deck(bar_deck(3).replace("*DENSITY\n7.85e-9\n", ""))


# ------------------------------------------------------------------ buckling
def _box(nn, lx, ly, lz):
    from axia_gpu import meshgen as MG
    xyz, conn = (MG.brick_hex20 if nn == 20 else MG.brick_hex8)(1, 1, 1)
    return xyz * np.array([lx, ly, lz]), conn[0]


@pytest.mark.parametrize("nn", [8, 20])
def test_hex_kg_symmetry_and_translations(nn):
    rng = np.random.default_rng(nn)
    xyz, conn = _box(nn, 2.0, 1.0, 0.5)
    xyz = xyz + rng.uniform(-0.05, 0.05, xyz.shape)
    s = rng.standard_normal(6)
    kg = O.hex_kg(xyz[conn], s)
    # symmetric to rounding (gtg(a, b) and gtg(b, a) sum in different orders)
    assert np.allclose(kg, kg.T, rtol=0, atol=1e-14 * np.abs(kg).max())
    for d in range(3):
        t = np.zeros(3 * nn)
        t[d::3] = 1.0
        assert np.abs(kg @ t).max() <= 1e-12 * np.abs(kg).max()
    # entries couple only equal directions
    m = 3 * nn
    i, j = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    assert np.all(kg[(i % 3) != (j % 3)] == 0.0)


@pytest.mark.parametrize("nn", [8, 20])
def test_hex_kg_string_closed_form(nn):
    """Uniaxial sigma_xx on an L x A box: a transverse field u_y = x / L stores
    sigma * A / L in v^T Kg v (the string's geometric stiffness)."""
    L, b, h, sig = 3.0, 0.7, 0.4, 2.5
    xyz, conn = _box(nn, L, b, h)
    kg = O.hex_kg(xyz[conn], [sig, 0, 0, 0, 0, 0])
    v = np.zeros(3 * nn)
    v[1::3] = xyz[conn][:, 0] / L
    assert math.isclose(float(v @ kg @ v), sig * b * h / L, rel_tol=1e-13)


def _column(nn, nx, ny, nz, L=20.0, b=1.0):
    from axia_gpu import meshgen as MG
    from eigen_decks import column_deck
    xyz, conn = (MG.brick_hex20 if nn == 20 else MG.brick_hex8)(nx, ny, nz)
    xyz = xyz * np.array([L / nx, b / ny, b / nz])
    fixed = np.nonzero(xyz[:, 0] == 0.0)[0]
    tip = np.nonzero(xyz[:, 0] == L)[0]
    return column_deck(nn, xyz, conn, fixed, tip, p_total=-1.0, nbuckle=2)


def test_subspace_ab_vs_lapack_and_euler():
    """Cantilever column under tip compression: the lowest buckling factors against
    scipy's dense generalized eigensolver on (A, K), and Euler's pi^2 E I / (4 L^2 P)."""
    text = _column(20, 10, 1, 1)
    model, res, kg, vals, vecs = O.buckle_deck(text)
    a = -kg.to_dense()
    mu = sla.eigh(a, res.csr.to_dense(), eigvals_only=True)
    ref = np.sort(1.0 / mu[np.abs(mu) > 0])
    ref = ref[np.argsort(np.abs(ref), kind="stable")][:2]
    assert np.allclose(vals, ref, rtol=1e-9)
    euler = math.pi ** 2 * 210000.0 * (1.0 / 12.0) / (4 * 20.0 ** 2 * 1.0)
    assert abs(vals[0] / euler - 1.0) < 2e-2
    assert abs(vals[1] / vals[0] - 1.0) < 1e-6  # square section: a degenerate pair
    # Kg_ff: symmetric, zero-free
    assert np.all(kg.data != 0.0)
    d = kg.to_dense()
    assert np.allclose(d, d.T, rtol=0, atol=1e-14 * np.abs(d).max())


def test_buckle_parse_defaults():
    text = _column(8, 4, 1, 1)
    assert O.parse_inp(text).nbuckle == 2
    assert O.parse_inp(text.replace("*BUCKLE\n2\n", "*BUCKLE\n")).nbuckle == 1
    with pytest.raises(ValueError):
        O.subspace_ab(O.Csr(0, np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0)),
                      None, 1)
