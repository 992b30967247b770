"""psp_amd: MI355X-native sweeping-preconditioned Helmholtz solver.

The hot path (multifrontal panel LDL^T, the sweep, the PML stencil and the
GMRES driver) lives in libpsp_hip.so behind the C ABI of include/psp_hip.h;
this package is a thin binding over it.
"""
from .helmholtz import (CLIQUE_LDL_1D, CLIQUE_LDL_SELINV_1D, DIRICHLET, PML, XY, XYZ, XZ,
                        XZY, YXZ, YZ, YZX, ZXY, ZYX, BreakdownError, Discretization,
                        DistHelmholtz, DistUniformGrid, GridData, LogicError, RuntimeError_,
                        SolveBatch, unit_cube_discretization)

__all__ = [
    "CLIQUE_LDL_1D", "CLIQUE_LDL_SELINV_1D", "DIRICHLET", "PML", "BreakdownError",
    "Discretization", "DistHelmholtz", "DistUniformGrid", "GridData", "LogicError",
    "RuntimeError_", "SolveBatch", "unit_cube_discretization",
    "XYZ", "XZY", "YXZ", "YZX", "ZXY", "ZYX", "XY", "XZ", "YZ",
]
