"""ctypes binding of libaxia_gpu.so (include/axia_gpu.h).

The library is loaded from this package's ``lib/`` directory (built in-tree by
``axia-fem_amd/build.py``). There is no fallback: if the library is missing,
loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
# AXIA_GPU_LIB selects another build of the same library (the sanitizer builds of
# sanitize.py); by default the in-tree build
LIB_PATH = os.environ.get("AXIA_GPU_LIB") or os.path.join(os.path.dirname(_PKG), "lib",
                                                           "libaxia_gpu.so")

AXG_OK = 0
AXG_E_INVALID = 1
AXG_E_JAC_SINGULAR = 2
AXG_E_NEG_JACOBIAN = 3
AXG_E_MATERIAL = 4
AXG_E_NOT_SPD = 5
AXG_E_NOT_CONVERGED = 6
AXG_E_SINGULAR = 7
AXG_E_DEVICE = 8
AXG_E_OOM = 9

PREC_FP64 = 0
PREC_MIXED = 1
PRECOND_JACOBI = 0
PRECOND_BLOCK3 = 1
REFINE_FCG = 0
REFINE_PLAIN = 1
REFINE_RELIABLE = 2


class SolveOpts(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("precond", C.c_int32),
        ("tol", C.c_double),
        ("max_iter", C.c_int64),
        ("inner_tol", C.c_double),
        ("max_outer", C.c_int32),
        ("timing", C.c_int32),
        ("refinement", C.c_int32),
        ("batch_groups", C.c_int32),
    ]


class SolveInfo(C.Structure):
    _fields_ = [
        ("iters", C.c_int64),
        ("outer_iters", C.c_int32),
        ("nrhs", C.c_int32),
        ("residual", C.c_double),
        ("rel_residual", C.c_double),
        ("solve_ms", C.c_double),
        ("spmv_ms", C.c_double),
        ("spmv_launches", C.c_int64),
        ("name", C.c_char * 64),
    ]


class AxgError(RuntimeError):
    """A FemError-style failure reported by the C ABI (code + message)."""

    def __init__(self, code, msg):
        self.code = code
        super().__init__(msg)


_lib = None

# Every entry point the header declares, with its signature.
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_int64)
_ip = C.POINTER(C.c_int32)
_i32 = C.c_int32
_i64 = C.c_int64
SIGNATURES = {
    "axg_last_error": (C.c_char_p, []),
    "axg_version": (C.c_char_p, []),
    "axg_default_opts": (None, [C.POINTER(SolveOpts)]),
    "axg_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "axg_destroy": (C.c_int, [_vp]),
    "axg_stream": (_vp, [_vp]),
    "axg_solve_kff": (C.c_int, [_vp, _i64, _i64, _lp, _lp, _dp, _dp, C.POINTER(SolveOpts), _dp,
                                C.POINTER(SolveInfo)]),
    "axg_csr_from_triplets": (C.c_int, [_vp, _i64, _i64, _lp, _lp, _dp, _lp, _lp, _dp]),
    "axg_solve_csr": (C.c_int, [_vp, _i64, _lp, _lp, _dp, _dp, C.POINTER(SolveOpts), _dp,
                                C.POINTER(SolveInfo)]),
    "axg_solve_csr_device_noalloc": (C.c_int, [_vp, _i64, _vp, _vp, _vp, _vp,
                                               C.POINTER(SolveOpts), _vp, C.POINTER(SolveInfo)]),
    "axg_mesh_create": (C.c_int, [_vp, _i32, _i64, _dp, _i64, _ip, _dp, _dp, _lp, C.POINTER(_vp)]),
    "axg_mesh_destroy": (C.c_int, [_vp]),
    "axg_mesh_info": (C.c_int, [_vp, _lp, _lp, _lp, _dp]),
    "axg_mesh_set_modulus": (C.c_int, [_vp, _dp]),
    "axg_assemble": (C.c_int, [_vp, _i32, _dp]),
    "axg_assembly_times": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "axg_csr_export": (C.c_int, [_vp, _lp, _lp, _dp]),
    "axg_scatter_map_export": (C.c_int, [_vp, _lp, _lp]),
    "axg_ke_export": (C.c_int, [_vp, _i64, _i64, _dp]),
    "axg_op32_export": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32), C.POINTER(C.c_float)]),
    "axg_mesh_solve": (C.c_int, [_vp, _dp, _i32, C.POINTER(SolveOpts), _dp, C.POINTER(SolveInfo)]),
    "axg_mesh_solve_device": (C.c_int, [_vp, _vp, _i32, C.POINTER(SolveOpts), _vp,
                                        C.POINTER(SolveInfo)]),
    "axg_nodal_stress": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "axg_compliance_adjoint": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "axg_solve_batch": (C.c_int, [_vp, _i32, C.POINTER(_vp), C.POINTER(_dp), _ip,
                                  C.POINTER(SolveOpts), C.POINTER(_dp), _lp, _dp,
                                  C.POINTER(SolveInfo)]),
    "axg_bench_stream": (C.c_int, [_vp, _i64, _i32, _dp]),
    "axg_mesh_spmm_shape": (C.c_int, [_vp, _i32, C.POINTER(_i32)]),
    "axg_mesh_asm32_mode": (C.c_int, [_vp, C.POINTER(_i32)]),
    "axg_bench_operator": (C.c_int, [_vp, _i32, _i32, _i32, _dp]),
    "axg_pattern_host": (C.c_int, [_i32, _i64, _i64, _ip, _lp, _lp, _lp, _lp, _lp, _lp, _lp, _lp]),
    "axg_pressure_loads": (C.c_int, [_i32, _i64, _dp, _i64, _ip, _ip, _dp, _dp]),
    "axg_reactions": (C.c_int, [_vp, _dp, _dp, _dp]),
    "axg_reduce_rhs": (C.c_int, [_vp, _dp, _dp, _dp]),
    "axg_full_diagonal": (C.c_int, [_vp, _dp]),
    "axg_element_volumes": (C.c_int, [_vp, _dp]),
    "axg_lumped_mass": (C.c_int, [_vp, _dp, _dp]),
    "axg_eigen": (C.c_int, [_vp, _dp, _i32, C.POINTER(SolveOpts), _dp, _dp, _ip, _ip,
                            C.POINTER(SolveInfo)]),
    "axg_deck_parse": (C.c_int, [C.c_char_p, _i64, C.POINTER(_vp)]),
    "axg_deck_destroy": (C.c_int, [_vp]),
    "axg_deck_info": (C.c_int, [_vp, C.POINTER(C.c_int32), _lp, _lp, _lp]),
    "axg_deck_arrays": (C.c_int, [_vp, _ip, _ip, _dp, _ip, _dp, _dp, _lp, _dp, _dp]),
    "axg_deck_procedure": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "axg_deck_nlgeom": (C.c_int, [_vp, _vp, C.POINTER(SolveOpts), _dp, _dp, _dp, _dp, _dp, _dp,
                                  _ip, _ip, C.POINTER(SolveInfo)]),
    "axg_deck_write_dat": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _i32, C.c_char_p, _i64, _lp]),
    "axg_deck_write_frd": (C.c_int, [_vp, _dp, _dp, _dp, _dp, C.c_char_p, _i64, _lp]),
    "axg_nl_assemble": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "axg_deck_buckle": (C.c_int, [_vp, _vp, C.POINTER(SolveOpts), _dp, _dp, _dp, _ip,
                                  C.POINTER(SolveInfo)]),
    "axg_geometric_stiffness": (C.c_int, [_vp, _dp, _dp, _dp]),
    "axg_buckle": (C.c_int, [_vp, _dp, _i32, C.POINTER(SolveOpts), _dp, _dp, _ip,
                             C.POINTER(SolveInfo)]),
    "axg_deck_frequency": (C.c_int, [_vp, _vp, C.POINTER(SolveOpts), _dp, _dp, _dp, _ip, _ip,
                                     C.POINTER(SolveInfo)]),
    "axg_deck_solve": (C.c_int, [_vp, _vp, C.POINTER(SolveOpts), _dp, _dp, _dp, _dp, _dp,
                                 C.POINTER(SolveInfo)]),
}


def lib():
    """Load the library (once). Raises OSError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with python axia-fem_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != AXG_OK:
        msg = lib().axg_last_error().decode("utf-8", "replace")
        raise AxgError(status, msg)


def dptr(a):
    return a.ctypes.data_as(_dp)


def lptr(a):
    return a.ctypes.data_as(_lp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
