"""ctypes binding of libpsp_hip.so (the C ABI declared in include/psp_hip.h).

The library is loaded from this package directory only; there is no CPU
fallback.  A missing or stale library raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpsp_hip.so")

OK, ERR_INVALID, ERR_SHALLOW, ERR_NAN, ERR_BREAKDOWN, ERR_NOCONV, ERR_STATE, ERR_HIP, ERR_MEMORY = range(9)
PML, DIRICHLET = 0, 1
FRONTS_FP64, FRONTS_FP32 = 0, 1
LDL_1D, LDL_SELINV_1D = 0, 1
DEVICE_POINTER = 1
BATCH_THREADS = 2
OPT_SWEEP_GRAPH = 1
OPT_SOLVE_MIN_TILES = 2
OPT_SOLVE_MAX_SPLIT = 3
OPT_SOLVE_WIDE_WG = 4
OPT_SOLVE_PERSIST = 5
OPT_SOLVE_PERSIST_GRID = 6
OPT_SOLVE_FUSED = 7
OPT_SOLVE_FUSED_QUEUES = 8
OPT_SOLVE_FUSED_ROWS = 9
OPT_SOLVE_FUSED_CUT = 10
OPT_SOLVE_FUSED_WG = 11
OPT_SOLVE_XCD_REDUCE = 12
OPT_FACTOR_OVERLAP = 13

# Every entry point include/psp_hip.h declares.
EXPORTS = (
    "psp_hip_create", "psp_hip_finalize", "psp_hip_destroy", "psp_hip_set_iterate_callback",
    "psp_hip_last_error", "psp_hip_set_stream",
    "psp_hip_set_option",
    "psp_hip_factor", "psp_hip_solve", "psp_hip_multiply", "psp_hip_precondition",
    "psp_hip_precondition_noalloc",
    "psp_hip_solve_panel", "psp_hip_solve_fixed", "psp_hip_adjoint_gradient",
    "psp_hip_profile_sweep", "psp_hip_solve_batch", "psp_hip_autotune",
    "psp_hip_num_panels", "psp_hip_panel_info", "psp_hip_class_num_nodes",
    "psp_hip_class_nodes", "psp_hip_reordered_index", "psp_hip_device_count",
    "psp_hip_grid_last_error", "psp_hip_grid_index", "psp_hip_grid_load",
    "psp_hip_grid_reorder", "psp_hip_grid_interpolate", "psp_hip_grid_write_plane",
    "psp_hip_grid_write_volume",
)
XYZ, XZY, YXZ, YZX, ZXY, ZYX = range(6)
PLANE_XY, PLANE_XZ, PLANE_YZ = range(3)
REAL64, COMPLEX128 = 0, 1


class Desc(C.Structure):
    _fields_ = [
        ("omega", C.c_double),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("wx", C.c_double), ("wy", C.c_double), ("wz", C.c_double),
        ("frontBC", C.c_int), ("rightBC", C.c_int), ("backBC", C.c_int),
        ("leftBC", C.c_int), ("topBC", C.c_int),
        ("bx", C.c_int), ("by", C.c_int), ("bz", C.c_int),
        ("sigmax", C.c_double), ("sigmay", C.c_double), ("sigmaz", C.c_double),
        ("damping", C.c_double),
        ("planesPerPanel", C.c_int),
        ("cutoff", C.c_int),
        ("precision", C.c_int),
    ]


class FactorReport(C.Structure):
    _fields_ = [
        ("seconds_total", C.c_double), ("seconds_fill", C.c_double),
        ("seconds_ldl", C.c_double), ("seconds_inverse", C.c_double),
        ("flops", C.c_double), ("schur_flops", C.c_double),
        ("bytes_factor", C.c_double),
        ("num_panels", C.c_int), ("num_fronts", C.c_int),
        ("seconds_schur", C.c_double),
    ]


class KernelStats(C.Structure):
    _fields_ = [("launches", C.c_int), ("seconds", C.c_double), ("bytes", C.c_double),
                ("flops", C.c_double)]


class SolveReport(C.Structure):
    _fields_ = [
        ("status", C.c_int), ("iterations", C.c_int), ("restarts", C.c_int),
        ("final_max_rel", C.c_double), ("seconds_total", C.c_double),
        ("seconds_precond", C.c_double), ("seconds_multiply", C.c_double),
        ("seconds_arnoldi", C.c_double),
        ("num_precond", C.c_int), ("num_multiply", C.c_int),
    ]


_LIB = None

# void (*)(void* user, int iteration, const void* x, int k)
ITERATE_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_void_p, C.c_int)


def lib():
    """Load libpsp_hip.so and declare every C ABI signature."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C psp_amd/csrc` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    IP = C.POINTER(C.c_int)
    sig = {
        "psp_hip_create": (I, [C.POINTER(Desc), C.POINTER(P)]),
        "psp_hip_destroy": (I, [P]),
        "psp_hip_finalize": (I, [P]),
        "psp_hip_set_iterate_callback": (I, [P, ITERATE_FN, P]),
        "psp_hip_last_error": (C.c_char_p, [P]),
        "psp_hip_set_stream": (I, [P, P]),
        "psp_hip_set_option": (I, [P, I, I]),
        "psp_hip_factor": (I, [P, P, I, C.POINTER(FactorReport)]),
        "psp_hip_solve": (I, [P, P, I, I, D, I, I, C.POINTER(SolveReport)]),
        "psp_hip_solve_batch": (I, [P, P, P, I, I, D, I, I, P]),
        "psp_hip_autotune": (I, [P, I, C.POINTER(D), C.POINTER(D)]),
        "psp_hip_multiply": (I, [P, P, P, I]),
        "psp_hip_precondition": (I, [P, P, I]),
        "psp_hip_precondition_noalloc": (I, [P, P, I]),
        "psp_hip_solve_panel": (I, [P, P, I, I]),
        "psp_hip_solve_fixed": (I, [P, P, I, I, I, I, P]),
        "psp_hip_adjoint_gradient": (I, [P, P, P, I, P, I]),
        "psp_hip_profile_sweep": (I, [P, P, I, C.POINTER(KernelStats)]),
        "psp_hip_num_panels": (I, [P]),
        "psp_hip_panel_info": (I, [P, I, IP, IP, IP, IP]),
        "psp_hip_class_num_nodes": (I, [P, I]),
        "psp_hip_class_nodes": (I, [P, I, P, P, P, P, P]),
        "psp_hip_reordered_index": (I, [P, I, I, I, I]),
        "psp_hip_device_count": (I, []),
        "psp_hip_grid_last_error": (C.c_char_p, []),
        "psp_hip_grid_index": (C.c_longlong, [I, I, I, I, I, I, I]),
        "psp_hip_grid_load": (I, [C.c_char_p, I, I, I, I, I, I, P]),
        "psp_hip_grid_reorder": (I, [P, I, P, I, I, I, I, I, I]),
        "psp_hip_grid_interpolate": (I, [P, I, I, I, P, I, I, I, I, I, I]),
        "psp_hip_grid_write_plane": (I, [P, I, I, I, I, I, I, I, I, C.c_char_p]),
        "psp_hip_grid_write_volume": (I, [P, I, I, I, I, I, I, C.c_char_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L
