"""ctypes wrapper of the CPU oracle (psp_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by the cpu_baseline leg of bench.py, always as the checker.  The product
package (psp_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PML, DIRICHLET = 0, 1


class Disc(C.Structure):
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
    ]


class Report(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("restarts", C.c_int),
        ("final_max_rel", C.c_double),
        ("t_total", C.c_double), ("t_precond", C.c_double),
        ("t_multiply", C.c_double), ("t_arnoldi", C.c_double),
        ("n_precond", C.c_int), ("n_multiply", C.c_int),
    ]


def lib():
    """The oracle library built by oracle/Makefile (loaded once)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libpsp_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _LIB = load_library(path)
    return _LIB


def load_library(path):
    """Loads an oracle build from `path` with its function signatures (the
    pinning tests load builds with a planted mistake this way)."""
    L = C.CDLL(path)
    P, I, D, V = C.c_void_p, C.c_int, C.c_double, None
    sig = {
        "ora_create": (I, [C.POINTER(Disc), C.POINTER(P)]),
        "ora_destroy": (V, [P]),
        "ora_num_panels": (I, [P]),
        "ora_panel_info": (I, [P, I, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(I)]),
        "ora_global_size": (C.c_long, [P]),
        "ora_local_to_natural": (V, [P, P]),
        "ora_reordered_index": (I, [P, I, I, I, I]),
        "ora_class_vsize": (I, [P, I]),
        "ora_class_num_nodes": (I, [P, I]),
        "ora_class_nodes": (V, [P, I, P, P, P, P]),
        "ora_class_structs": (V, [P, I, P]),
        "ora_global_nnz": (C.c_long, [P]),
        "ora_global_csr": (I, [P, P, P, P]),
        "ora_panel_lower_coo": (C.c_long, [P, I, P, P, P, P]),
        "ora_factor": (I, [P, P]),
        "ora_factor_seconds": (D, [P]),
        "ora_factor_sample": (I, [P, P, I, I, C.POINTER(D)]),
        "ora_multiply": (I, [P, P, I]),
        "ora_precondition": (I, [P, P, I]),
        "ora_solve_panel": (I, [P, P, I, I]),
        "ora_solve": (I, [P, P, I, I, D, I, C.POINTER(Report)]),
        "ora_solve_fixed": (I, [P, P, I, I, I, P]),
        "ora_partial_ldl": (I, [P, I, I]),
        "ora_fill_global": (I, [P, P]),
        "ora_num_threads": (I, []),
        "ora_time_arnoldi": (D, [P, I, I]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleError(RuntimeError):
    def __init__(self, code, where):
        super().__init__(f"oracle {where} failed with status {code}")
        self.code = code


class Oracle:
    """One DistHelmholtz-equivalent CPU solver instance."""

    def __init__(self, omega, nx, ny, nz, wx=None, wy=None, wz=None,
                 bcs=(PML,) * 5, pml=5, sigma=1.5, damping=7.0,
                 planes_per_panel=4, cutoff=12, sigmas=None, pmls=None, library=None):
        h = None
        if wx is None:
            h = 1.0 / (max(nx, ny, nz) + 1)
            wx, wy, wz = (nx + 1) * h, (ny + 1) * h, (nz + 1) * h
        bx, by, bz = pmls if pmls is not None else (pml, pml, pml)
        sx, sy, sz = sigmas if sigmas is not None else (sigma, sigma, sigma)
        self.disc = Disc(omega, nx, ny, nz, wx, wy, wz, *bcs, bx, by, bz,
                         sx, sy, sz, damping, planes_per_panel, cutoff)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.N = nx * ny * nz
        self._L = library if library is not None else lib()
        h_ = C.c_void_p()
        st = self._L.ora_create(C.byref(self.disc), C.byref(h_))
        if st != 0:
            raise OracleError(st, "create")
        self._h = h_

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.ora_destroy(self._h)
            self._h = None

    # geometry -----------------------------------------------------------
    @property
    def num_panels(self):
        return self._L.ora_num_panels(self._h)

    def panel_info(self, p):
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._L.ora_panel_info(self._h, p, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return dict(v=a.value, depth=b.value, padding=c.value, cls=d.value)

    def local_to_natural(self):
        m = np.empty(self.N, dtype=np.int32)
        self._L.ora_local_to_natural(self._h, _ptr(m))
        return m

    def reordered_index(self, x, y, vl, vs):
        return self._L.ora_reordered_index(self._h, x, y, vl, vs)

    def class_vsize(self, cls):
        return self._L.ora_class_vsize(self._h, cls)

    def class_nodes(self, cls):
        n = self._L.ora_class_num_nodes(self._h, cls)
        off, size, parent, usize = (np.empty(n, np.int32) for _ in range(4))
        self._L.ora_class_nodes(self._h, cls, _ptr(off), _ptr(size), _ptr(parent), _ptr(usize))
        st = np.empty(int(usize.sum()), np.int32)
        self._L.ora_class_structs(self._h, cls, _ptr(st))
        starts = np.concatenate([[0], np.cumsum(usize)])
        structs = [st[starts[i]:starts[i + 1]] for i in range(n)]
        return dict(off=off, size=size, parent=parent, usize=usize, structs=structs)

    # numerics -----------------------------------------------------------
    def factor(self, velocity):
        v = np.ascontiguousarray(velocity, dtype=np.float64).reshape(-1)
        assert v.size == self.N
        self._vel = v
        st = self._L.ora_factor(self._h, _ptr(v))
        if st != 0:
            raise OracleError(st, "factor")

    @property
    def factor_seconds(self):
        return self._L.ora_factor_seconds(self._h)

    def factor_sample(self, velocity, first, count):
        v = np.ascontiguousarray(velocity, dtype=np.float64).reshape(-1)
        self._vel = v
        t = C.c_double()
        st = self._L.ora_factor_sample(self._h, _ptr(v), first, count, C.byref(t))
        if st != 0:
            raise OracleError(st, "factor_sample")
        return t.value

    def fill_global(self, velocity):
        v = np.ascontiguousarray(velocity, dtype=np.float64).reshape(-1)
        self._vel = v
        self._L.ora_fill_global(self._h, _ptr(v))

    def time_arnoldi(self, k, m):
        return self._L.ora_time_arnoldi(self._h, k, m)

    def global_csr(self):
        nnz = self._L.ora_global_nnz(self._h)
        rp = np.empty(self.N + 1, np.int32)
        cols = np.empty(nnz, np.int32)
        vals = np.empty(nnz, np.complex128)
        self._L.ora_global_csr(self._h, _ptr(rp), _ptr(cols), _ptr(vals))
        return rp, cols, vals

    def panel_lower_coo(self, panel, velocity):
        v = np.ascontiguousarray(velocity, dtype=np.float64).reshape(-1)
        n = self._L.ora_panel_lower_coo(self._h, panel, _ptr(v), None, None, None)
        rows = np.empty(n, np.int32)
        cols = np.empty(n, np.int32)
        vals = np.empty(n, np.complex128)
        self._L.ora_panel_lower_coo(self._h, panel, _ptr(v), _ptr(rows), _ptr(cols), _ptr(vals))
        return rows, cols, vals

    def _nd_block(self, B):
        B = np.asfortranarray(np.asarray(B, dtype=np.complex128))
        if B.ndim == 1:
            B = B.reshape(-1, 1, order="F")
        assert B.shape[0] == self.N
        return B

    def multiply(self, B):
        B = self._nd_block(B).copy(order="F")
        st = self._L.ora_multiply(self._h, _ptr(B), B.shape[1])
        if st != 0:
            raise OracleError(st, "multiply")
        return B

    def precondition(self, B):
        B = self._nd_block(B).copy(order="F")
        st = self._L.ora_precondition(self._h, _ptr(B), B.shape[1])
        if st != 0:
            raise OracleError(st, "precondition")
        return B

    def solve_panel(self, B, panel):
        B = self._nd_block(B).copy(order="F")
        st = self._L.ora_solve_panel(self._h, _ptr(B), B.shape[1], panel)
        if st != 0:
            raise OracleError(st, "solve_panel")
        return B

    def solve(self, Bgrid, m=20, rel_tol=1e-5, max_its=0):
        """Bgrid: (N, k) in grid order (natural point-major, k interleaved)."""
        B = np.ascontiguousarray(np.asarray(Bgrid, dtype=np.complex128)).copy()
        if B.ndim == 1:
            B = B.reshape(-1, 1)
        k = B.shape[1]
        rep = Report()
        st = self._L.ora_solve(self._h, _ptr(B), k, m, rel_tol, max_its, C.byref(rep))
        return B, st, rep

    def solve_fixed(self, Bgrid, m, steps):
        B = np.ascontiguousarray(np.asarray(Bgrid, dtype=np.complex128)).copy()
        if B.ndim == 1:
            B = B.reshape(-1, 1)
        hist = np.zeros(steps)
        st = self._L.ora_solve_fixed(self._h, _ptr(B), B.shape[1], m, steps, _ptr(hist))
        if st != 0:
            raise OracleError(st, "solve_fixed")
        return B, hist


def num_threads():
    return lib().ora_num_threads()


def partial_ldl(F, s):
    """Dense partial LDL^T of a column-major lower-stored front (tests)."""
    F = np.asfortranarray(np.array(F, dtype=np.complex128))
    st = lib().ora_partial_ldl(_ptr(F), F.shape[0], s)
    if st != 0:
        raise OracleError(st, "partial_ldl")
    return F
