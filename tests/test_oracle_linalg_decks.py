"""Pin the oracle's CSR, PCG, Cholesky and deck pipeline (CPU only)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def brute_csr(n, ti, tj, tv):
    """linalg.rs:10-37 by an ordered dictionary: first-seen order per (i, j), left-to-right sum."""
    acc = {}
    for i, j, v in zip(ti, tj, tv):
        if 0 <= i < n and 0 <= j < n and abs(v) > 0.0:
            acc.setdefault((int(i), int(j)), []).append(float(v))
    keys = sorted(acc)
    indptr = np.zeros(n + 1, np.int64)
    data = []
    for (i, j) in keys:
        s = 0.0
        for v in acc[(i, j)]:
            s += v
        data.append(s)
        indptr[i + 1] += 1
    return np.cumsum(indptr), np.array([k[1] for k in keys]), np.array(data)


def test_csr_from_triplets_matches_brute_force():
    rng = np.random.default_rng(0)
    n, t = 50, 3000
    ti = rng.integers(-1, n + 1, t)
    tj = rng.integers(-1, n + 1, t)
    tv = rng.standard_normal(t)
    tv[rng.random(t) < 0.2] = 0.0
    csr = O.csr_from_triplets(n, ti, tj, tv)
    indptr, indices, data = brute_csr(n, ti, tj, tv)
    assert np.array_equal(csr.indptr, indptr)
    assert np.array_equal(csr.indices, indices)
    assert np.array_equal(csr.data, data)


def test_csr_keeps_cancelled_entries():
    # (0,0) sums to 0.0 from nonzero parts and is kept; an exact-zero triplet is dropped
    csr = O.csr_from_triplets(2, [0, 0, 1], [0, 0, 1], [1.5, -1.5, 0.0])
    assert list(csr.indptr) == [0, 1, 1] and csr.data[0] == 0.0


def spd(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.2)
    a = a @ a.T + n * np.eye(n)
    idx = np.nonzero(a)
    return a, O.csr_from_triplets(n, idx[0], idx[1], a[idx])


def test_pcg_and_cholesky_vs_lapack():
    a, csr = spd(60, 1)
    b = np.random.default_rng(2).standard_normal(60)
    ref = np.linalg.solve(a, b)
    x, it, res = O.pcg(csr, b, tol=1e-13)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-11
    assert 0 < it <= 240 and res <= 1e-13 * np.linalg.norm(b)
    assert np.allclose(O.chol_solve(a, b), ref, rtol=1e-12, atol=1e-14)


def test_pcg_error_paths():
    a, csr = spd(30, 3)
    b = np.ones(30)
    with pytest.raises(O.OracleError) as e:
        O.pcg(csr, b, tol=1e-14, max_iter=2)
    assert e.value.code == 6 and "konvergierte nicht" in str(e.value)
    flip = O.csr_from_triplets(2, [0, 1], [1, 0], [1.0, 1.0])
    with pytest.raises(O.OracleError) as e:
        O.pcg(flip, np.array([1.0, 0.0]))
    assert e.value.code == 5
    x, it, res = O.pcg(csr, np.zeros(30))
    assert it == 0 and not x.any()


def test_cholesky_rejects_indefinite():
    """linalg.rs:284-303: diag(1, -2) is indefinite for the dense Cholesky."""
    with pytest.raises(O.OracleError) as e:
        O.chol_solve(np.diag([1.0, -2.0]), np.array([3.0, -8.0]))
    assert e.value.code == 7


def deck(name):
    return open(os.path.join(GOLD, name + ".inp")).read()


def test_patch_c3d8_reference_values():
    """lib.rs:429-462: mean u_x of the loaded face 0.01 +- 1e-6, mean sigma_xx 210 +- 1e-3."""
    m, res = O.solve_deck(deck("cube_tension"))
    ux = np.mean([res.u_full[3 * m.node_index(i)] for i in (2, 3, 6, 7)])
    assert abs(ux - 0.01) < 1e-6
    assert abs(res.stress[:, 0].mean() - 210.0) < 1e-3
    assert O.von_mises(res.stress[0]) > 0.0
    assert res.solver == "Cholesky"


def test_patch_c3d8_elset_deferred():
    """lib.rs:410-427: ELSET defined after the elements, a second dummy material."""
    text = deck("cube_tension").replace("*ELEMENT, TYPE=C3D8, ELSET=SOLID", "*ELEMENT, TYPE=C3D8")
    text = text.replace("*SOLID SECTION, ELSET=SOLID, MATERIAL=STEEL",
                        "*ELSET, ELSET=SOLID\n1\n*MATERIAL, NAME=DUMMY\n*ELASTIC\n1.0, 0.0\n"
                        "*SOLID SECTION, ELSET=SOLID, MATERIAL=STEEL")
    m, res = O.solve_deck(text)
    ux = np.mean([res.u_full[3 * m.node_index(i)] for i in (2, 3, 6, 7)])
    assert abs(ux - 0.01) < 1e-6


def test_patch_c3d20_reference_values():
    """lib.rs:751-774: P4 = 210 gives u_x = -0.01 +- 1e-5 and sigma_xx = -210 +- 1."""
    m, res = O.solve_deck(deck("c3d20_patch"))
    ux = np.mean([res.u_full[3 * m.node_index(i)] for i in (2, 3, 6, 7)])
    assert abs(ux + 0.01) < 1e-5
    assert abs(res.stress[:, 0].mean() + 210.0) < 1.0


@pytest.mark.parametrize("name,axia,tol", [("le10_c3d20", -5.51, 0.05),
                                           ("le10_c3d20_fine", -5.50, 0.03)])
def test_nafems_le10(name, axia, tol):
    """nafems.rs:50-90 (target -5.38 within 5 % / 3 %) and the README's axia values."""
    m, res = O.solve_deck(deck(name))
    syy = res.stress[m.node_index(m.nsets["D"][0]), 1]
    assert abs((syy + 5.38) / 5.38) < tol
    assert round(syy, 2) == axia


def test_backends_agree_on_patch():
    """lib.rs:464-497: direct and iterative backends give u_x = 0.01 +- 1e-6."""
    for solver in ("cholesky", "lapack", "pcg", "superlu"):
        m, res = O.solve_deck(deck("cube_tension"), solver=solver)
        ux = np.mean([res.u_full[3 * m.node_index(i)] for i in (2, 3, 6, 7)])
        assert abs(ux - 0.01) < 1e-6, solver


def test_goldens_reproduce():
    """The committed fixtures match the oracle (guards against drift in either)."""
    import hashlib
    for name in ("cube_tension", "c3d20_patch", "le10_c3d20"):
        g = np.load(os.path.join(GOLD, name + ".npz"))
        m, res = O.solve_deck(deck(name))
        h = hashlib.sha256(np.ascontiguousarray(res.csr.data.astype(np.float64)).tobytes())
        assert h.hexdigest() == str(g["sha_data"])
        assert np.array_equal(res.stress, g["stress"])
        assert np.allclose(res.x, g["x"], rtol=1e-12, atol=0)


def test_structured_nnz_closed_form():
    """nnz of a clamped-bottom n^3 C3D8 brick: 9 * sum over free nodes of their neighbour
    counts (27 interior); checked against csr_from_triplets on small bricks."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(__file__)), "axia-fem_amd"))
    from axia_gpu import meshgen as MG
    for n in (2, 3, 5):
        p = MG.brick_lognormal(n, seed=0)
        ti, tj, tv = O.assemble_triplets(8, p.xyz, p.conn, p.e, p.nu)
        sel = (p.ind_of[ti] >= 0) & (p.ind_of[tj] >= 0)
        csr = O.csr_from_triplets(p.n_free, p.ind_of[ti[sel]], p.ind_of[tj[sel]], tv[sel])
        # closed form: for node (i,j,k) with k >= 1, neighbours (i',j',k') with k' >= 1
        cnt = 0
        for k in range(1, n + 1):
            nk = (2 if k == n else 3) - (1 if k == 1 else 0)
            for j in range(n + 1):
                nj = 3 - (j == 0) - (j == n)
                for i in range(n + 1):
                    ni = 3 - (i == 0) - (i == n)
                    cnt += 9 * ni * nj * nk
        assert csr.nnz == cnt
