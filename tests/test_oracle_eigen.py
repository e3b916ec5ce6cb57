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
