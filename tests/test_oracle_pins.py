"""Pins of the oracle where constant fields and totals leave a mistake undetected
(CPU only).

* Nodal stress extrapolation (elem.rs:743-779 for C3D8, quadratic.rs:148-244 for
  C3D20): a displacement field the element reproduces exactly whose strain varies
  over the element. On an axis-aligned box the C3D8 space holds u = xyz, so the
  strain at the Gauss points is exact and bilinear, and the extrapolation with N
  at +-sqrt(3) must return D eps(x) at every node; C3D20 holds x^2 y z and its 27
  point tensor Lagrange extrapolation must return the (triquadratic) strain. A
  wrong Gauss-point order or extrapolation factor fails these.
* Face pressure (elem.rs:892-950, quadratic.rs:266-315): the first moment
  sum_a x_a (x) F_a of the nodal forces of a planar face that is not a
  parallelogram (quad8: edges curved in the plane) equals p A c (x) n_in, with
  the area A and centroid c from Green's theorem on the edge curves. The Gauss
  rules integrate these forces exactly, so a load distributed to the wrong nodes
  fails even when the total is right.
* DAT numbers (dat.rs:32-68, Rust's "{:14.6E}" and "{:14.7E}"): strings written
  by hand from Rust's documented float formatting (mantissa rounded to the
  precision, exponent as a plain signed integer without '+' or padding, right
  aligned to the width), through the oracle and through the library's writer.
"""

import numpy as np
import pytest

from axia_gpu import deck as D
from oracle import oracle as O

HEX8_NODES = np.array([(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                       (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)], float)
HEX20_NODES = np.concatenate([HEX8_NODES, np.array(
    [(0, -1, -1), (1, 0, -1), (0, 1, -1), (-1, 0, -1), (0, -1, 1), (1, 0, 1), (0, 1, 1),
     (-1, 0, 1), (-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)], float)])
HEX8_FACE = [[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 5, 1], [1, 5, 6, 2], [2, 6, 7, 3], [3, 7, 4, 0]]
HEX20_FACE = [[0, 1, 2, 3, 8, 9, 10, 11], [4, 7, 6, 5, 15, 14, 13, 12],
              [0, 4, 5, 1, 16, 12, 17, 8], [1, 5, 6, 2, 17, 13, 18, 9],
              [2, 6, 7, 3, 18, 14, 19, 10], [3, 7, 4, 0, 19, 15, 16, 11]]


def hooke(e, nu):
    """Isotropic D in Voigt order xx, yy, zz, xy, yz, zx with engineering shear."""
    lam = e * nu / ((1 + nu) * (1 - 2 * nu))
    mu = e / (2 * (1 + nu))
    d = np.zeros((6, 6))
    d[:3, :3] = lam
    d[[0, 1, 2], [0, 1, 2]] += 2 * mu
    d[[3, 4, 5], [3, 4, 5]] = mu
    return d


def strain(grad):
    """Voigt strain from the displacement gradient grad[i][j] = du_i / dx_j."""
    g = grad
    return np.array([g[0][0], g[1][1], g[2][2], g[0][1] + g[1][0], g[1][2] + g[2][1],
                     g[2][0] + g[0][2]])


# ------------------------------------------------ stress extrapolation
def _box(nodes, lo, size):
    return lo + (nodes + 1.0) * 0.5 * size


def test_hex8_extrapolation_trilinear_field():
    """u_x = xyz, u_y = x y - 2 y z, u_z = z x + 3 xyz on a 2 x 1.5 x 3 box off the origin:
    eps_xx = yz, gamma_xy = xz + y, ... all in span{1, x, y, z, xy, yz, zx, xyz}, so the
    nodal stress is D eps at the node itself."""
    lo, size = np.array([0.5, -1.0, 2.0]), np.array([2.0, 1.5, 3.0])
    xyz = _box(HEX8_NODES, lo, size)

    def u(p):
        x, y, z = p
        return np.array([x * y * z, x * y - 2 * y * z, z * x + 3 * x * y * z])

    def grad(p):
        x, y, z = p
        return [[y * z, x * z, x * y], [y, x - 2 * z, -2 * y], [z + 3 * y * z, 3 * x * z, x + 3 * x * y]]

    e, nu = 1000.0, 0.27
    d = hooke(e, nu)
    ue = np.concatenate([u(p) for p in xyz])
    out = O.hex_nodal_stress(xyz, ue, e, nu)
    exact = np.array([d @ strain(grad(p)) for p in xyz])
    assert np.allclose(out, exact, rtol=0, atol=1e-10 * np.abs(exact).max())
    # the field varies: a constant-strain check alone would not see the extrapolation
    assert np.ptp(exact[:, 0]) > 0.5 * np.abs(exact[:, 0]).max()
    # a wrong extrapolation (nodal value = the nearest Gauss point's, i.e. N at
    # +-1 instead of +-sqrt(3)) gives a different answer on this field
    gp = 1.0 / np.sqrt(3.0)
    near = np.array([d @ strain(grad(lo + (n * gp + 1.0) * 0.5 * size)) for n in HEX8_NODES])
    assert np.abs(near - exact).max() > 1e-3 * np.abs(exact).max()


def test_hex20_extrapolation_serendipity_field():
    """u_x = x^2 y z, u_y = x y^2 - z^2 x, u_z = y z^2 + x y on a box: each strain component is
    at most quadratic in each coordinate, which the 27-point tensor Lagrange extrapolation
    reproduces, so every one of the 20 nodes gets D eps(node)."""
    lo, size = np.array([1.0, 0.25, -0.5]), np.array([1.5, 2.0, 1.0])
    xyz = _box(HEX20_NODES, lo, size)

    def u(p):
        x, y, z = p
        return np.array([x * x * y * z, x * y * y - z * z * x, y * z * z + x * y])

    def grad(p):
        x, y, z = p
        return [[2 * x * y * z, x * x * z, x * x * y],
                [y * y - z * z, 2 * x * y, -2 * z * x],
                [y, z * z + x, 2 * y * z]]

    e, nu = 70000.0, 0.3
    d = hooke(e, nu)
    ue = np.concatenate([u(p) for p in xyz])
    out = O.hex_nodal_stress(xyz, ue, e, nu)
    exact = np.array([d @ strain(grad(p)) for p in xyz])
    assert np.allclose(out, exact, rtol=0, atol=1e-10 * np.abs(exact).max())
    assert np.ptp(exact[:, 3]) > 0.2 * np.abs(exact[:, 3]).max()


# ------------------------------------------------------- face pressure
def _lagrange3(t):
    """Quadratic Lagrange basis on t in [-1, 1] with nodes -1, 0, 1."""
    return np.array([0.5 * t * (t - 1), 1 - t * t, 0.5 * t * (t + 1)])


def _green(edges, axes):
    """Area and first moments of a planar region bounded by quadratic curves.

    edges: list of (start, mid, end) points, counter-clockwise in the (u, v) plane given
    by axes = (iu, iv). Green's theorem: A = 1/2 oint (u dv - v du),
    int u dA = 1/2 oint u^2 dv, int v dA = -1/2 oint v^2 du. The integrands are
    polynomials of degree <= 5 in t, so 4-point Gauss-Legendre is exact."""
    iu, iv = axes
    t, w = np.polynomial.legendre.leggauss(4)
    area = mu = mv = 0.0
    for a, m, b in edges:
        pts = np.array([a, m, b])
        for tk, wk in zip(t, w):
            L = _lagrange3(tk)
            dL = np.array([tk - 0.5, -2 * tk, tk + 0.5])
            p = L @ pts
            dp = dL @ pts
            u, v, du, dv = p[iu], p[iv], dp[iu], dp[iv]
            area += wk * 0.5 * (u * dv - v * du)
            mu += wk * 0.5 * u * u * dv
            mv += wk * (-0.5) * v * v * du
    return area, mu, mv


def _face_case(nn, face, rng):
    """An element whose face `face` lies in a coordinate plane with a general (not a
    parallelogram) outline; quad8 faces get midside nodes moved within the plane."""
    base = (HEX8_NODES if nn == 8 else HEX20_NODES).copy()
    fn = (HEX8_FACE if nn == 8 else HEX20_FACE)[face - 1]
    # the face's constant natural coordinate and its axis
    axis = int(np.nonzero(np.all(base[fn] == base[fn][0], axis=0))[0][0])
    xyz = base * 1.0 + rng.uniform(-0.12, 0.12, base.shape)
    xyz[fn, axis] = base[fn[0], axis] * 1.0  # the face stays in the plane
    # corners: a general quadrilateral
    xyz[fn[:4]] += np.where(np.arange(3) == axis, 0.0, 1.0) * rng.uniform(-0.3, 0.3, (4, 3))
    if nn == 20:
        # midsides: the edge midpoint moved within the plane (curved edges)
        for k in range(4):
            a, b = xyz[fn[k]], xyz[fn[(k + 1) % 4]]
            off = np.where(np.arange(3) == axis, 0.0, 1.0) * rng.uniform(-0.15, 0.15, 3)
            xyz[fn[4 + k]] = 0.5 * (a + b) + off
    return xyz, fn, axis


@pytest.mark.parametrize("nn,face", [(8, 2), (8, 1), (8, 4), (8, 5), (20, 2), (20, 3), (20, 6),
                                     (20, 1)])
def test_face_pressure_first_moment(nn, face):
    rng = np.random.default_rng(100 * nn + face)
    xyz, fn, axis = _face_case(nn, face, rng)
    p = 2.75
    fe = O.hex_face_pressure(xyz, face, p).reshape(nn, 3)
    others = [k for k in range(nn) if k not in fn]
    assert np.all(fe[others] == 0.0)
    iu, iv = [a for a in range(3) if a != axis]
    if nn == 8:
        edges = [(xyz[fn[k]], 0.5 * (xyz[fn[k]] + xyz[fn[(k + 1) % 4]]), xyz[fn[(k + 1) % 4]])
                 for k in range(4)]
    else:
        edges = [(xyz[fn[k]], xyz[fn[4 + k]], xyz[fn[(k + 1) % 4]]) for k in range(4)]
    area, mu, mv = _green(edges, (iu, iv))
    # positive pressure pushes into the element: towards the side of the plane that
    # holds the element's other nodes
    n_in = np.zeros(3)
    n_in[axis] = np.sign(xyz.mean(0)[axis] - xyz[fn[0], axis])
    a = abs(area)
    c = np.zeros(3)
    c[iu], c[iv], c[axis] = mu / area, mv / area, xyz[fn[0], axis]
    total = fe.sum(0)
    assert np.allclose(total, p * a * n_in, rtol=0, atol=1e-12 * p * a)
    moment = xyz.T @ fe  # moment[i][j] = sum_a x_a,i F_a,j
    assert np.allclose(moment, p * a * np.outer(c, n_in), rtol=0, atol=1e-11 * p * a)
    # a distribution that only gets the total right (equal shares on the corners) fails
    wrong = np.zeros_like(fe)
    wrong[fn[:4]] = total / 4.0
    assert np.abs(xyz.T @ wrong - moment).max() > 1e-4 * p * a


# ------------------------------------------------------------ DAT numbers
RUST_E6 = [  # value, Rust's format!("{:14.6E}", value), written out by hand
    (0.0015, "   1.500000E-3"),
    (1.0, "    1.000000E0"),
    (123456.0, "    1.234560E5"),
    (-2.5e-10, " -2.500000E-10"),
    (0.0, "    0.000000E0"),
    (6.02214076e23, "   6.022141E23"),
    (-0.1, "  -1.000000E-1"),
    (9.9999996, "    1.000000E1"),
    (9.9999994, "    9.999999E0"),
    (1e-300, " 1.000000E-300"),
    (-123.4567891, "   -1.234568E2"),
]


def test_rust_exp_known_strings():
    for v, s in RUST_E6:
        assert len(s) == 14
        assert O.rust_exp(v, 6, 14) == s, v
    assert O.rust_exp(0.0, 7, 14) == "   0.0000000E0"


def test_dat_lines_known_strings():
    """The displacement block of the DAT text for hand-written values, through the
    oracle and the library's writer."""
    text = ("*NODE\n1, 0, 0, 0\n2, 1, 0, 0\n3, 1, 1, 0\n4, 0, 1, 0\n5, 0, 0, 1\n6, 1, 0, 1\n"
            "7, 1, 1, 1\n8, 0, 1, 1\n*ELEMENT, TYPE=C3D8\n1, 1, 2, 3, 4, 5, 6, 7, 8\n"
            "*MATERIAL, NAME=S\n*ELASTIC\n210000, 0.3\n*SOLID SECTION, ELSET=EALL, MATERIAL=S\n"
            "*BOUNDARY\n1, 1, 3\n*STEP\n*STATIC\n*CLOAD\n2, 1, 1\n*END STEP\n")
    vals = [v for v, _ in RUST_E6] + [0.0] * (24 - len(RUST_E6))
    strs = [s for _, s in RUST_E6] + ["    0.000000E0"] * (24 - len(RUST_E6))
    u = np.array(vals)
    rf = np.zeros(24)
    m = O.parse_inp(text)
    want = ["", " displacements (vx,vy,vz) for set NALL and time     0.0000000E0", ""]
    for k in range(8):
        want.append(f"{k + 1:10d} " + " ".join(strs[3 * k:3 * k + 3]))
    want += ["", " forces (fx,fy,fz) for set NALL and time     0.0000000E0", ""]
    for dat in (O.write_dat(m, u, rf), D.Deck(text).write_dat(u, rf)):
        lines = dat.split("\n")
        assert lines[:len(want)] == want
