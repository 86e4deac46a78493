"""ctypes binding of libcfdv_amd.so (the C ABI in include/cfdv.h).

The library is loaded from this package directory only; a missing or
unloadable library raises immediately -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcfdv_amd.so")

CFDV_OK = 0
CFDV_ERR_ARG = -1
CFDV_ERR_HIP = -2
CFDV_ERR_BC = -3
CFDV_ERR_GEOM = -4
CFDV_ERR_NAN = -5
CFDV_ERR_BREAKDOWN = -6
CFDV_ERR_CFL = -7
CFDV_ERR_NOCONV = -8
CFDV_ERR_REFINE = -9

CFDV_REFINE_PARITY = 0
CFDV_REFINE_3D = 1

CFDV_BLOCKS_FP64 = 1
CFDV_BLOCKS_FP32 = 2
CFDV_BLOCKS_ROWS = 4

CFDV_SOLVE_REFERENCE = 0
CFDV_SOLVE_FP64 = 1
CFDV_SOLVE_MIXED = 2

CFDV_PRECOND_NONE = 0
CFDV_PRECOND_NODE_BLOCK = 1

# every symbol declared in include/cfdv.h
EXPORTS = ["cfdv_create", "cfdv_destroy", "cfdv_last_error", "cfdv_mesh_info", "cfdv_set_state",
           "cfdv_get_state", "cfdv_cfl", "cfdv_build", "cfdv_get_blocks", "cfdv_get_blocks32",
           "cfdv_get_rhs_vector", "cfdv_solve", "cfdv_get_du", "cfdv_update", "cfdv_rms",
           "cfdv_refine_flags", "cfdv_color", "cfdv_matvec", "cfdv_profile_matvec", "cfdv_step",
           "cfdv_set_thermo", "cfdv_adjoint_solve", "cfdv_adjoint_solve_multi", "cfdv_build_vjp",
           "cfdv_matvec_multi",
           "cfdv_profile_matvec_multi", "cfdv_refine_create", "cfdv_refine_destroy",
           "cfdv_refine_last_error", "cfdv_refine_list", "cfdv_refine_apply", "cfdv_refine_size",
           "cfdv_refine_get", "cfdv_reference_trace", "cfdv_refine_flags_noalloc", "cfdv_rows_info"]


class Mesh(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nnod", C.c_int), ("neqn", C.c_int),
                ("n_elem", C.c_int64), ("n_node", C.c_int64),
                ("conn", C.POINTER(C.c_int32)), ("x", C.POINTER(C.c_double)),
                ("bc", C.POINTER(C.c_int32)), ("dirichlet", C.POINTER(C.c_uint8)),
                ("adap", C.POINTER(C.c_uint8)), ("hnm", C.POINTER(C.c_double))]


class ThermoS(C.Structure):
    _fields_ = [("gamma", C.c_double), ("mach", C.c_double), ("reynolds", C.c_double),
                ("prandtl", C.c_double), ("tinf", C.c_double), ("twall", C.c_double),
                ("adiabatic", C.c_int)]


class SolveOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("restarts", C.c_int), ("m", C.c_int), ("rtol", C.c_double),
                ("inner_m", C.c_int), ("inner_rtol", C.c_double), ("inner_restarts", C.c_int),
                ("transpose", C.c_int), ("precond", C.c_int), ("warm_start", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("bnorm", C.c_double), ("resid", C.c_double), ("rel_resid", C.c_double),
                ("est_resid", C.c_double), ("matvecs64", C.c_int), ("matvecs32", C.c_int),
                ("iters", C.c_int), ("restarts", C.c_int), ("breakdown", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [("dt", C.c_double), ("urms", C.c_double * 5), ("ms_cfl", C.c_double),
                ("ms_build", C.c_double), ("ms_solve", C.c_double), ("ms_update", C.c_double),
                ("ms_matvec64", C.c_double), ("ms_matvec32", C.c_double),
                ("n_matvec64", C.c_int), ("n_matvec32", C.c_int),
                ("solve", SolveStats), ("ms_assemble", C.c_double)]


_LIB = None


def load() -> C.CDLL:
    """Load the in-tree library; raises if it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "or python -m cfdv_amd.build_lib")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    L.cfdv_create.argtypes = [C.POINTER(Mesh), C.POINTER(ThermoS), C.c_int, C.POINTER(vp)]
    L.cfdv_destroy.argtypes = [vp]
    L.cfdv_destroy.restype = None
    L.cfdv_last_error.argtypes = [vp]
    L.cfdv_last_error.restype = C.c_char_p
    L.cfdv_mesh_info.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int64)]
    L.cfdv_rows_info.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.cfdv_set_state.argtypes = [vp, dp]
    L.cfdv_get_state.argtypes = [vp, dp, dp]
    L.cfdv_cfl.argtypes = [vp, C.c_double, dp]
    L.cfdv_build.argtypes = [vp, C.c_double]
    L.cfdv_get_blocks.argtypes = [vp, dp, dp, dp]
    L.cfdv_get_blocks32.argtypes = [vp, C.POINTER(C.c_float)]
    L.cfdv_get_rhs_vector.argtypes = [vp, dp]
    L.cfdv_solve.argtypes = [vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]
    L.cfdv_get_du.argtypes = [vp, dp]
    L.cfdv_reference_trace.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), dp, dp,
                                       C.POINTER(C.c_int32)]
    L.cfdv_update.argtypes = [vp]
    L.cfdv_rms.argtypes = [vp, dp]
    L.cfdv_refine_flags.argtypes = [vp, C.POINTER(C.c_uint8), dp, C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32)]
    L.cfdv_refine_flags_noalloc.argtypes = [vp, C.POINTER(C.c_uint8), dp, C.POINTER(C.c_int32),
                                            C.POINTER(C.c_int32)]
    L.cfdv_color.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.cfdv_matvec.argtypes = [vp, dp, dp, C.c_int, C.c_int]
    L.cfdv_profile_matvec.argtypes = [vp, C.c_int, C.c_int, dp]
    L.cfdv_step.argtypes = [vp, C.c_double, C.POINTER(SolveOpts), C.POINTER(StepStats)]
    L.cfdv_set_thermo.argtypes = [vp, C.POINTER(ThermoS)]
    L.cfdv_adjoint_solve.argtypes = [vp, C.POINTER(SolveOpts), dp, dp, C.POINTER(SolveStats)]
    L.cfdv_adjoint_solve_multi.argtypes = [vp, C.POINTER(SolveOpts), dp, dp, C.c_int, dp, C.POINTER(SolveStats)]
    L.cfdv_build_vjp.argtypes = [vp, dp, dp, dp, dp]
    fp = C.POINTER(C.c_float)
    L.cfdv_matvec_multi.argtypes = [vp, fp, fp, C.c_int, C.c_int]
    L.cfdv_profile_matvec_multi.argtypes = [vp, C.c_int, dp]
    ip = C.POINTER(C.c_int32)
    up = C.POINTER(C.c_uint8)
    L.cfdv_refine_create.argtypes = [C.POINTER(Mesh), dp, dp, ip, C.POINTER(vp)]
    L.cfdv_refine_destroy.argtypes = [vp]
    L.cfdv_refine_destroy.restype = None
    L.cfdv_refine_last_error.argtypes = [vp]
    L.cfdv_refine_last_error.restype = C.c_char_p
    L.cfdv_refine_list.argtypes = [vp, up, ip, ip, ip, ip]
    L.cfdv_refine_apply.argtypes = [vp, ip, C.c_int32, ip, C.c_int32, C.c_int, C.c_double]
    L.cfdv_refine_size.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), ip]
    L.cfdv_refine_get.argtypes = [vp, ip, ip, dp, dp, dp, up, up, dp, ip]
    _LIB = L
    return L
