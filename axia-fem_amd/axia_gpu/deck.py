"""INP decks through the native reader and the device linear-static driver.

Deck(text) parses the keyword subset of the reference's reader (inp.rs) and
builds solve_linear's setup (analysis.rs:548-946) in libaxia_gpu; Deck.solve
runs the rest of solve_linear on the device (assembly, reduced rhs, PCG,
reconstruction, reactions, nodal stresses) and returns a LinearResult that
mirrors SolveOutput (analysis.rs:24-54). Deck.frequency runs the *FREQUENCY
branch (analysis.rs:955-1005): lumped mass, subspace iteration on the device,
frequencies and mass-normalised mode shapes. Deck.buckle runs the *BUCKLE
branch (analysis.rs:1176-1260): the static solution of the step's loads, the
geometric stiffness of its element mean stresses and the buckling factors.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr, lptr


@dataclass
class LinearResult:
    node_ids: np.ndarray    # (nnode,) ascending node ids
    u_full: np.ndarray      # (3 nnode,) displacements
    rf: np.ndarray          # (3 nnode,) reactions K u - f
    stress: np.ndarray      # (nnode, 6) nodal mean stress, Voigt xx yy zz xy yz zx
    von_mises: np.ndarray   # (nnode,)
    solver: str
    iters: int
    residual: float
    nfree: int
    ndof: int
    stress_gp: np.ndarray = None  # (nelem, 6) element mean stress
    solve_ms: float = 0.0         # device time of the PCG

    def node_index(self, nid):
        i = int(np.searchsorted(self.node_ids, nid))
        if i >= len(self.node_ids) or self.node_ids[i] != nid:
            raise KeyError(nid)
        return i


@dataclass
class FrequencyResult:
    node_ids: np.ndarray     # (nnode,) ascending node ids
    frequencies: np.ndarray  # (k,) cycles per time, sqrt(lambda) / (2 pi)
    eigenvalues: np.ndarray  # (k,) lambda = omega^2
    modes: np.ndarray        # (k, 3 nnode) mass-normalised mode shapes
    sweeps: int
    solver: str
    iters: int
    nfree: int
    ndof: int


@dataclass
class BuckleResult:
    node_ids: np.ndarray     # (nnode,) ascending node ids
    factors: np.ndarray      # (k,) buckling load factors, ascending |lambda|
    modes: np.ndarray        # (k, 3 nnode) buckling shapes (unit length over the free DOFs)
    u_static: np.ndarray     # (3 nnode,) the preload's displacements
    solver: str
    iters: int
    nfree: int
    ndof: int
    solve_ms: float = 0.0    # device time of the solves (the preload's and the sweeps')


@dataclass
class NlResult:
    node_ids: np.ndarray     # (nnode,) ascending node ids
    u_full: np.ndarray       # (3 nnode,) displacements
    rf: np.ndarray           # (3 nnode,) f_int - f_ext
    stress: np.ndarray       # (nnode, 6) nodal Cauchy stress
    von_mises: np.ndarray    # (nnode,)
    strain: np.ndarray       # (nnode, 6) nodal Green-Lagrange strain
    ninc: int                # increments taken
    iters: int               # Newton iterations of the last increment
    residual: float          # last ||r_free||
    solver: str
    pcg_iters: int           # summed PCG iterations of all corrections
    solve_ms: float          # their summed device time
    stress_gp: np.ndarray = None  # (nelem, 6) element mean Cauchy stress


def _text_out(fill, n):
    """The text of a writer's fill call (size n from its size query): written into an
    uninitialised buffer and decoded from it in place (no zero fill, no intermediate copy)."""
    buf = np.empty(n + 1, np.uint8)
    m = C.c_int64()
    check(fill(buf.ctypes.data_as(C.c_char_p), n + 1, C.byref(m)))
    return str(memoryview(buf)[:m.value], "utf-8")


class Deck:
    """A parsed deck and its linear-static setup (host side of libaxia_gpu)."""

    def __init__(self, text):
        raw = text.encode("utf-8") if isinstance(text, str) else bytes(text)
        h = C.c_void_p()
        check(_lib.lib().axg_deck_parse(raw, len(raw), C.byref(h)))
        self._h = h
        nn, nnode, nelem, nfree = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int64()
        check(_lib.lib().axg_deck_info(self._h, C.byref(nn), C.byref(nnode), C.byref(nelem),
                                       C.byref(nfree)))
        self.nn, self.nnode, self.nelem, self.n_free = (nn.value, nnode.value, nelem.value,
                                                        nfree.value)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().axg_deck_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arrays(self):
        """dict of node_ids, elem_ids, xyz, conn, e, nu, ind_of, u0, f_full."""
        ndof = 3 * self.nnode
        out = {
            "node_ids": np.zeros(self.nnode, np.int32), "elem_ids": np.zeros(self.nelem, np.int32),
            "xyz": np.zeros((self.nnode, 3)), "conn": np.zeros((self.nelem, self.nn), np.int32),
            "e": np.zeros(self.nelem), "nu": np.zeros(self.nelem),
            "ind_of": np.zeros(ndof, np.int64), "u0": np.zeros(ndof), "f_full": np.zeros(ndof),
        }
        check(_lib.lib().axg_deck_arrays(
            self._h, iptr(out["node_ids"]), iptr(out["elem_ids"]), dptr(out["xyz"]),
            iptr(out["conn"]), dptr(out["e"]), dptr(out["nu"]), lptr(out["ind_of"]),
            dptr(out["u0"]), dptr(out["f_full"])))
        return out

    def solve(self, ctx=None, precision="fp64", precond="jacobi", tol=1e-8, max_iter=0,
              inner_tol=1e-3):
        from . import _info_dict, _opts, default_context
        ctx = ctx or default_context()
        ndof = 3 * self.nnode
        u = np.zeros(ndof)
        rf = np.zeros(ndof)
        stress = np.zeros((self.nnode, 6))
        vm = np.zeros(self.nnode)
        gp = np.zeros((self.nelem, 6))
        info = _lib.SolveInfo()
        o = _opts(precision, precond, tol, max_iter, inner_tol)
        check(_lib.lib().axg_deck_solve(ctx.handle, self._h, C.byref(o), dptr(u), dptr(rf),
                                        dptr(stress), dptr(vm), dptr(gp), C.byref(info)))
        inf = _info_dict(info)
        ids = self.arrays()["node_ids"]
        self._refresh()
        return LinearResult(ids, u, rf, stress, vm, inf["name"], inf["iters"], inf["residual"],
                            self.n_free, ndof, gp, inf["solve_ms"])


    def procedure(self):
        """dict nmodes (*FREQUENCY), nbuckle (*BUCKLE), nlgeom (*STEP, NLGEOM), ninc."""
        v = [C.c_int32() for _ in range(4)]
        check(_lib.lib().axg_deck_procedure(self._h, *[C.byref(x) for x in v]))
        return {"nmodes": v[0].value, "nbuckle": v[1].value, "nlgeom": bool(v[2].value),
                "ninc": v[3].value}

    @property
    def nmodes(self):
        """Modes the deck's *FREQUENCY step asks for (0 without one)."""
        return self.procedure()["nmodes"]

    @property
    def nbuckle(self):
        """Buckling factors the deck's *BUCKLE step asks for (0 without one)."""
        return self.procedure()["nbuckle"]

    def nlgeom(self, ctx=None, precision="fp64", precond="jacobi", tol=1e-10, max_iter=0,
               inner_tol=1e-3):
        """solve_continuum_newton on the device: NlResult (u, rf, nodal Cauchy stress, von
        Mises, Green-Lagrange strain, increments, Newton iterations of the last increment)."""
        from . import _info_dict, _opts, default_context
        ctx = ctx or default_context()
        if not self.procedure()["nlgeom"]:
            raise ValueError("the deck has no *STEP, NLGEOM")
        ndof = 3 * self.nnode
        u, rf = np.zeros(ndof), np.zeros(ndof)
        stress, strain = np.zeros((self.nnode, 6)), np.zeros((self.nnode, 6))
        vm = np.zeros(self.nnode)
        gp = np.zeros((self.nelem, 6))
        ninc, its = C.c_int32(), C.c_int32()
        info = _lib.SolveInfo()
        o = _opts(precision, precond, tol, max_iter, inner_tol)
        check(_lib.lib().axg_deck_nlgeom(ctx.handle, self._h, C.byref(o), dptr(u), dptr(rf),
                                         dptr(stress), dptr(vm), dptr(strain), dptr(gp),
                                         C.byref(ninc), C.byref(its), C.byref(info)))
        inf = _info_dict(info)
        return NlResult(self.arrays()["node_ids"], u, rf, stress, vm, strain, ninc.value,
                        its.value, inf["residual"], inf["name"], inf["iters"], inf["solve_ms"], gp)

    def buckle(self, ctx=None, precision="fp64", precond="jacobi", tol=1e-10, max_iter=0,
               inner_tol=1e-3, modes=True):
        from . import _info_dict, _opts, default_context
        ctx = ctx or default_context()
        k = self.nbuckle
        if k <= 0:
            raise ValueError("the deck has no *BUCKLE step")
        ndof = 3 * self.nnode
        fac = np.zeros(k)
        shapes = np.zeros((k, ndof)) if modes else None
        u = np.zeros(ndof)
        nf = C.c_int32()
        info = _lib.SolveInfo()
        o = _opts(precision, precond, tol, max_iter, inner_tol)
        check(_lib.lib().axg_deck_buckle(ctx.handle, self._h, C.byref(o), dptr(fac),
                                         dptr(shapes) if modes else None, dptr(u), C.byref(nf),
                                         C.byref(info)))
        inf = _info_dict(info)
        self._refresh()
        n = nf.value
        return BuckleResult(self.arrays()["node_ids"], fac[:n], shapes[:n] if modes else None, u,
                            inf["name"], inf["iters"], self.n_free, ndof, inf["solve_ms"])

    def frequency(self, ctx=None, precision="fp64", precond="jacobi", tol=1e-10, max_iter=0,
                  inner_tol=1e-3, modes=True):
        from . import _info_dict, _opts, default_context
        ctx = ctx or default_context()
        k = self.nmodes
        if k <= 0:
            raise ValueError("the deck has no *FREQUENCY step")
        ndof = 3 * self.nnode
        freqs, vals = np.zeros(k), np.zeros(k)
        shapes = np.zeros((k, ndof)) if modes else None
        nf, sw = C.c_int32(), C.c_int32()
        info = _lib.SolveInfo()
        o = _opts(precision, precond, tol, max_iter, inner_tol)
        check(_lib.lib().axg_deck_frequency(ctx.handle, self._h, C.byref(o), dptr(freqs),
                                            dptr(vals), dptr(shapes) if modes else None,
                                            C.byref(nf), C.byref(sw), C.byref(info)))
        inf = _info_dict(info)
        self._refresh()
        n = nf.value
        return FrequencyResult(self.arrays()["node_ids"], freqs[:n], vals[:n],
                               shapes[:n] if modes else None, sw.value, inf["name"], inf["iters"],
                               self.n_free, ndof)

    def write_dat(self, u_full, rf, stress_gp=None, nodal_stress=None, freqs=None):
        """The reference's DAT text for results on this deck (dat.rs), as a str."""
        keep = [np.ascontiguousarray(a, np.float64) if a is not None else None
                for a in (u_full, rf, stress_gp, nodal_stress, freqs)]
        args = [dptr(a) if a is not None else None for a in keep]
        nf = 0 if freqs is None else len(keep[4])
        n = C.c_int64()
        L = _lib.lib()
        check(L.axg_deck_write_dat(self._h, *args, nf, None, 0, C.byref(n)))
        return _text_out(lambda buf, cap, m: L.axg_deck_write_dat(self._h, *args, nf, buf, cap, m),
                         n.value)

    def write_frd(self, u_full, stress=None, rf=None, strain=None):
        """The reference's FRD text for results on this deck (frd.rs, one frame), as a str."""
        keep = [np.ascontiguousarray(a, np.float64) if a is not None else None
                for a in (u_full, stress, rf, strain)]
        args = [dptr(a) if a is not None else None for a in keep]
        n = C.c_int64()
        L = _lib.lib()
        check(L.axg_deck_write_frd(self._h, *args, None, 0, C.byref(n)))
        return _text_out(lambda buf, cap, m: L.axg_deck_write_frd(self._h, *args, buf, cap, m),
                         n.value)

    def _refresh(self):
        nn, nnode, nelem, nfree = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int64()
        check(_lib.lib().axg_deck_info(self._h, C.byref(nn), C.byref(nnode), C.byref(nelem),
                                       C.byref(nfree)))
        self.n_free = nfree.value


def solve_deck(text, ctx=None, **kw):
    """Parse and solve a deck: (Deck, LinearResult)."""
    d = Deck(text)
    return d, d.solve(ctx=ctx, **kw)


def frequency_deck(text, ctx=None, **kw):
    """Parse a *FREQUENCY deck and run it: (Deck, FrequencyResult)."""
    d = Deck(text)
    return d, d.frequency(ctx=ctx, **kw)


def buckle_deck(text, ctx=None, **kw):
    """Parse a *BUCKLE deck and run it: (Deck, BuckleResult)."""
    d = Deck(text)
    return d, d.buckle(ctx=ctx, **kw)


def nlgeom_deck(text, ctx=None, **kw):
    """Parse a *STEP, NLGEOM deck and run it: (Deck, NlResult)."""
    d = Deck(text)
    return d, d.nlgeom(ctx=ctx, **kw)
