"""Python mirror of PSP's public solver API over the C ABI.

Names, argument meanings, defaults and error behaviour follow
include/psp/discretization.hpp, include/psp/dist_uniform_grid.hpp and
include/psp/dist_helmholtz.hpp of the reference:

    disc = Discretization(omega, nx, ny, nz, wx, wy, wz, PML, ..., pmlSize, sigma)
    helm = DistHelmholtz(disc, damping=7.0, numPlanesPerPanel=4, cutoff=12)
    helm.Initialize(velocity)            # DistUniformGrid of reals
    helm.Solve(B, m=20, relTol=1e-4)     # DistUniformGrid of complex, in place
    helm.Finalize()

Right-hand sides may be host numpy arrays or torch complex128 CUDA tensors
(HBM-resident); both use the DistUniformGrid layout at one process,
(x + y*nx + z*nx*ny) * k + r.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (DIRICHLET, FRONTS_FP32, FRONTS_FP64, LDL_1D, LDL_SELINV_1D,
                   PML, XYZ, XZY, YXZ, YZX, ZXY, ZYX, FactorReport, SolveReport)
from ._lib import PLANE_XY as XY, PLANE_XZ as XZ, PLANE_YZ as YZ

CLIQUE_LDL_1D = LDL_1D
CLIQUE_LDL_SELINV_1D = LDL_SELINV_1D


class LogicError(ValueError):
    """psp::LogicError (std::logic_error)."""


class RuntimeError_(RuntimeError):
    """psp::RuntimeError (std::runtime_error)."""


class BreakdownError(RuntimeError_):
    """Exact Arnoldi breakdown; the reference returns silently here."""


_LOGIC = {_lib.ERR_INVALID, _lib.ERR_SHALLOW, _lib.ERR_STATE}


def _raise(code, ctx, what):
    msg = ""
    if ctx:
        raw = _lib.lib().psp_hip_last_error(ctx)
        msg = raw.decode() if raw else ""
    text = f"{what}: {msg} (status {code})"
    if code in _LOGIC:
        raise LogicError(text)
    if code == _lib.ERR_BREAKDOWN:
        raise BreakdownError(text)
    raise RuntimeError_(text)


@dataclass
class Discretization:
    """Discretization<double> (include/psp/discretization.hpp:36-110)."""
    omega: float
    nx: int
    ny: int
    nz: int
    wx: float
    wy: float
    wz: float
    frontBC: int = PML
    rightBC: int = PML
    backBC: int = PML
    leftBC: int = PML
    topBC: int = PML
    bx: int = 5
    by: int = 5
    bz: int = 5
    sigmax: float | None = None
    sigmay: float | None = None
    sigmaz: float | None = None

    def __post_init__(self):
        # default sigma = 1.2 * max width (discretization.hpp:60-63)
        s = 1.2 * max(self.wx, self.wy, self.wz)
        self.sigmax = s if self.sigmax is None else self.sigmax
        self.sigmay = s if self.sigmay is None else self.sigmay
        self.sigmaz = s if self.sigmaz is None else self.sigmaz

    @classmethod
    def with_pml(cls, omega, nx, ny, nz, wx, wy, wz, front, right, back, left, top,
                 pmlSize, sigma):
        """The (..., pmlSize, sigma) constructor (discretization.hpp:96-109)."""
        return cls(omega, nx, ny, nz, wx, wy, wz, front, right, back, left, top,
                   pmlSize, pmlSize, pmlSize, sigma, sigma, sigma)


class DistUniformGrid:
    """DistUniformGrid<F> at one process (dist_uniform_grid.hpp:48-145).

    numScalars values per point; points stored in the grid's GridDataOrder
    (XYZ by default: buffer[(x + y*nx + z*nx*ny) * numScalars + r]).  data
    has shape (nx*ny*nz, numScalars), row = LocalIndex(x, y, z) / numScalars.
    F is float64 (velocity) or complex128 (right-hand sides, solutions).
    """

    def __init__(self, numScalars, nx, ny, nz, dtype=np.complex128, data=None, order=XYZ):
        if order not in (XYZ, XZY, YXZ, YZX, ZXY, ZYX):
            raise LogicError("unknown GridDataOrder")
        self.nx, self.ny, self.nz, self.numScalars = nx, ny, nz, numScalars
        self.order = order
        if data is None:
            data = np.zeros((nx * ny * nz, numScalars), dtype=dtype)
        self.data = data

    def XSize(self):
        return self.nx

    def YSize(self):
        return self.ny

    def ZSize(self):
        return self.nz

    def NumScalars(self):
        return self.numScalars

    def Order(self):
        return self.order

    def Buffer(self):
        return self.data

    def NaturalIndex(self, x, y, z):
        return x + y * self.nx + z * self.nx * self.ny

    def LocalIndex(self, x, y, z):
        i = _lib.lib().psp_hip_grid_index(self.nx, self.ny, self.nz, self.order, x, y, z)
        if i < 0:
            raise LogicError("point outside the grid")
        return i * self.numScalars

    # ---- on-disk formats (dist_uniform_grid.hpp:349-2198) ----------------
    def _type(self):
        if self.data.dtype == np.complex128:
            return _lib.COMPLEX128
        if self.data.dtype == np.float64:
            return _lib.REAL64
        raise LogicError("grid I/O supports float64 and complex128 grids")

    def _check(self, st, what):
        if st != _lib.OK:
            raise LogicError(f"{what}: {_lib.lib().psp_hip_grid_last_error().decode()}")

    def _buf(self):
        if not (isinstance(self.data, np.ndarray) and self.data.flags.c_contiguous):
            raise LogicError("grid I/O needs a C-contiguous host buffer")
        return self.data.ctypes.data_as(C.c_void_p)

    def SequentialLoad(self, filename):
        """Fill the grid from a raw binary file holding its points in the
        grid's order, numScalars native-endian values each."""
        self._check(_lib.lib().psp_hip_grid_load(str(filename).encode(), self.nx, self.ny,
                                                 self.nz, self.numScalars, self._type(),
                                                 self.order, self._buf()),
                    "SequentialLoad")

    def InterpolateTo(self, nx, ny, nz):
        """Trilinear resampling onto nx x ny x nz points spanning the same box."""
        if (nx, ny, nz) == (self.nx, self.ny, self.nz):
            return
        out = np.zeros((nx * ny * nz, self.numScalars), dtype=self.data.dtype)
        self._check(_lib.lib().psp_hip_grid_interpolate(
            self._buf(), self.nx, self.ny, self.nz, out.ctypes.data_as(C.c_void_p), nx, ny, nz,
            self.numScalars, self._type(), self.order), "InterpolateTo")
        self.nx, self.ny, self.nz, self.data = nx, ny, nz, out

    def WritePlane(self, planeType, whichPlane, basename):
        """ASCII VTK ImageData of one XY, XZ or YZ plane per scalar."""
        self._check(_lib.lib().psp_hip_grid_write_plane(
            self._buf(), self.nx, self.ny, self.nz, self.numScalars, self._type(), self.order,
            planeType, whichPlane, str(basename).encode()), "WritePlane")

    def WriteVolume(self, basename):
        """ASCII VTK parallel ImageData (one piece) of the whole grid per scalar."""
        self._check(_lib.lib().psp_hip_grid_write_volume(
            self._buf(), self.nx, self.ny, self.nz, self.numScalars, self._type(), self.order,
            str(basename).encode()), "WriteVolume")

    def natural(self):
        """The data in natural XYZ order (a copy unless the grid is XYZ)."""
        if self.order == XYZ:
            return self.data
        out = np.empty_like(self.data)
        self._check(_lib.lib().psp_hip_grid_reorder(
            self._buf(), self.order, out.ctypes.data_as(C.c_void_p), XYZ, self.nx, self.ny,
            self.nz, self.numScalars, self._type()), "reorder")
        return out

    def set_natural(self, nat):
        """Store natural XYZ-order data in the grid's order."""
        nat = np.ascontiguousarray(nat, dtype=self.data.dtype).reshape(self.data.shape)
        if self.order == XYZ:
            self.data[...] = nat
            return
        self._check(_lib.lib().psp_hip_grid_reorder(
            nat.ctypes.data_as(C.c_void_p), XYZ, self._buf(), self.order, self.nx, self.ny,
            self.nz, self.numScalars, self._type()), "reorder")


GridData = DistUniformGrid


def _is_torch_cuda(a):
    try:
        import torch
    except ImportError:
        return False
    return isinstance(a, torch.Tensor) and a.is_cuda


class DistHelmholtz:
    """DistHelmholtz<double> on one MI355X (dist_helmholtz.hpp:21-283)."""

    def __init__(self, disc: Discretization, damping=7.0, numPlanesPerPanel=4,
                 cutoff=12, precision="fp64"):
        prec = {"fp64": FRONTS_FP64, "fp32": FRONTS_FP32}[precision]
        self.disc = disc
        self.precision = precision
        d = _lib.Desc(disc.omega, disc.nx, disc.ny, disc.nz, disc.wx, disc.wy, disc.wz,
                      disc.frontBC, disc.rightBC, disc.backBC, disc.leftBC, disc.topBC,
                      disc.bx, disc.by, disc.bz, disc.sigmax, disc.sigmay, disc.sigmaz,
                      damping, numPlanesPerPanel, cutoff, prec)
        self._desc = d
        h = C.c_void_p()
        st = _lib.lib().psp_hip_create(C.byref(d), C.byref(h))
        if st != _lib.OK:
            if st == _lib.ERR_SHALLOW:
                raise LogicError("The domain is very shallow. Please run a sparse-direct "
                                 "factorization instead.")
            _raise(st, None, "DistHelmholtz")
        self._h = h
        self.N = disc.nx * disc.ny * disc.nz
        self.factor_report = None

    # ---- lifecycle -------------------------------------------------------
    def __del__(self):
        try:
            self.destroy()
        except Exception:  # interpreter shutdown: the library may be gone already
            pass

    def Finalize(self):
        """Free the factors and the solve workspace
        (initialize_finalize.hpp:379-396); the object stays usable and a later
        Initialize rebuilds them."""
        st = _lib.lib().psp_hip_finalize(self._ctx())
        if st != _lib.OK:
            _raise(st, self._h, "Finalize")
        self.factor_report = None

    def destroy(self):
        """Release everything, the context included (the C++ destructor)."""
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().psp_hip_destroy(h)
            self._h = None

    def _ctx(self):
        if not self._h:
            raise LogicError("DistHelmholtz was destroyed")
        return self._h

    def _bind_stream(self):
        """Order the library's work after torch's pending work on the device:
        run on torch's current stream, or, when that is the legacy default
        stream, synchronize it and use the context's own stream."""
        import torch
        s = torch.cuda.current_stream().cuda_stream
        if s == 0:
            torch.cuda.current_stream().synchronize()
        _lib.lib().psp_hip_set_stream(self._ctx(), C.c_void_p(s) if s else None)

    def set_sweep_graph(self, enabled: bool):
        """Replay the sweep from a captured hipGraph (default) or launch it
        kernel by kernel."""
        st = _lib.lib().psp_hip_set_option(self._ctx(), _lib.OPT_SWEEP_GRAPH, int(enabled))
        if st != _lib.OK:
            _raise(st, self._h, "set_option")

    def Autotune(self, k):
        """Picks the front-solve tile shapes by timing (see psp_hip_autotune);
        returns the estimated front-solve seconds of one sweep before and
        after."""
        b, a = C.c_double(), C.c_double()
        st = _lib.lib().psp_hip_autotune(self._ctx(), int(k), C.byref(b), C.byref(a))
        if st != _lib.OK:
            _raise(st, self._h, "Autotune")
        return b.value, a.value

    def set_solve_tiling(self, min_tiles=None, max_split=None, wide_wg=None, persist=None,
                         persist_grid=None, fused=None, fused_queues=None, fused_rows=None,
                         fused_cut=None, fused_wg=None, xcd_reduce=None):
        """Default front-solve tile shapes: workgroups wanted per launch, the
        most inner slices per row block, whether levels with few row blocks
        get 16-wave workgroups instead of slices, and whether the solves run
        on persistent pipelined workgroups (8 <= k <= 16); rebuilds the solve
        tiles.  persist_grid caps the workgroups of a persistent launch
        (0: as many as fit on the device).  fused: the tree levels from the
        cut depth down as one persistent launch per direction (8 <= k <= 16),
        with fused_queues subtree queues (one per XCD; more leave queues the
        last workgroup runs alone), fused_rows output rows per tile and the
        cut depth fused_cut (-1: the first depth with 8 nodes); fused_wg: the
        workgroups a queue expects, which sets how far row blocks near the
        cut are cut into inner slices (1: none).  xcd_reduce (default 1): the
        inner slices of a row block in a one-workgroup-per-tile launch run on
        one XCD and the last to finish sums them (0: a reduce launch)."""
        for opt, v in ((_lib.OPT_SOLVE_MIN_TILES, min_tiles), (_lib.OPT_SOLVE_MAX_SPLIT, max_split),
                       (_lib.OPT_SOLVE_WIDE_WG, wide_wg), (_lib.OPT_SOLVE_PERSIST, persist),
                       (_lib.OPT_SOLVE_PERSIST_GRID, persist_grid), (_lib.OPT_SOLVE_FUSED, fused),
                       (_lib.OPT_SOLVE_FUSED_QUEUES, fused_queues),
                       (_lib.OPT_SOLVE_FUSED_ROWS, fused_rows),
                       (_lib.OPT_SOLVE_FUSED_CUT, fused_cut), (_lib.OPT_SOLVE_FUSED_WG, fused_wg),
                       (_lib.OPT_SOLVE_XCD_REDUCE, xcd_reduce)):
            if v is None:
                continue
            st = _lib.lib().psp_hip_set_option(self._ctx(), opt, int(v))
            if st != _lib.OK:
                _raise(st, self._h, "set_option")

    def set_factor_overlap(self, on):
        """on (default): Initialize factors the leftover and top panels on
        their own streams beside the batch of full panels; off: one panel
        class after another, so the report's Schur-complement GEMM time
        covers those launches alone."""
        st = _lib.lib().psp_hip_set_option(self._ctx(), _lib.OPT_FACTOR_OVERLAP, int(bool(on)))
        if st != _lib.OK:
            _raise(st, self._h, "set_option")

    # ---- Initialize ------------------------------------------------------
    def Initialize(self, velocity, panelScheme=CLIQUE_LDL_SELINV_1D):
        """Factor every panel (initialize_finalize.hpp:12-377)."""
        v = velocity
        if isinstance(velocity, DistUniformGrid):
            if (velocity.nx, velocity.ny, velocity.nz) != (self.disc.nx, self.disc.ny, self.disc.nz):
                raise LogicError("Velocity grid is incorrect")
            if velocity.numScalars != 1:
                raise LogicError("Velocity grid should have one entry per point")
            v = velocity.natural()
        v = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))
        if v.size != self.N:
            raise LogicError("Velocity grid is incorrect")
        rep = FactorReport()
        st = _lib.lib().psp_hip_factor(self._ctx(), v.ctypes.data_as(C.c_void_p), panelScheme,
                                       C.byref(rep))
        if st != _lib.OK:
            _raise(st, self._h, "Initialize")
        self.factor_report = rep
        return rep

    # ---- Solve -----------------------------------------------------------
    def Solve(self, B, m=20, relTol=1e-4, viewIterates=False, *, maxIts=0,
              raise_on_error=True):
        """Solve in place (solve.hpp:12-56); returns the SolveReport.

        B: a DistUniformGrid of complex128 (any GridDataOrder), a host
        (N, k) complex128 array, or a torch complex128 CUDA tensor, the last
        two in natural grid order.  viewIterates: every unconverged GMRES
        iterate is written as the VTK volume "iterates-<it>"
        (solve.hpp:579-585).  maxIts > 0 caps the GMRES steps (the reference
        has no cap)."""
        grid = B if isinstance(B, DistUniformGrid) else None
        if grid is not None:
            if (grid.nx, grid.ny, grid.nz) != (self.disc.nx, self.disc.ny, self.disc.nz):
                raise LogicError("B grid is incorrect")
            arr = np.ascontiguousarray(grid.natural())
        else:
            arr = B
        cb = None
        if viewIterates:
            k_ = arr.shape[1] if arr.ndim == 2 else 1
            view = grid if grid is not None else DistUniformGrid(
                k_, self.disc.nx, self.disc.ny, self.disc.nz)

            def _dump(user, it, x, k):
                nat = np.ctypeslib.as_array(C.cast(x, C.POINTER(C.c_double)),
                                            shape=(self.N * k * 2,)).view(np.complex128)
                snap = DistUniformGrid(k, view.nx, view.ny, view.nz, order=view.order)
                snap.set_natural(nat.reshape(self.N, k))
                snap.WriteVolume(f"iterates-{it}")

            cb = _lib.ITERATE_FN(_dump)
        _lib.lib().psp_hip_set_iterate_callback(self._ctx(), cb if cb else _lib.ITERATE_FN(),
                                                None)
        try:
            rep = self._solve(arr, m, relTol, maxIts, raise_on_error)
        finally:
            _lib.lib().psp_hip_set_iterate_callback(self._ctx(), _lib.ITERATE_FN(), None)
        if grid is not None:
            grid.set_natural(arr)
        return rep

    def _solve(self, arr, m, relTol, maxIts, raise_on_error):
        rep = SolveReport()
        if _is_torch_cuda(arr):
            import torch
            assert arr.dtype == torch.complex128 and arr.is_contiguous()
            k = arr.shape[1] if arr.dim() == 2 else 1
            self._bind_stream()
            st = _lib.lib().psp_hip_solve(self._ctx(), C.c_void_p(arr.data_ptr()), k, m, relTol,
                                          maxIts, _lib.DEVICE_POINTER, C.byref(rep))
        else:
            if not (isinstance(arr, np.ndarray) and arr.dtype == np.complex128
                    and arr.flags.c_contiguous):
                raise LogicError("B must be a C-contiguous complex128 array")
            k = arr.shape[1] if arr.ndim == 2 else 1
            st = _lib.lib().psp_hip_solve(self._ctx(), arr.ctypes.data_as(C.c_void_p), k, m,
                                          relTol, maxIts, 0, C.byref(rep))
        if st != _lib.OK and raise_on_error:
            _raise(st, self._h, "Solve")
        return rep

    SolveWithGMRES = Solve

    def SolveFixed(self, B, m, steps):
        """GMRES with exactly `steps` inner steps; returns the residual history."""
        hist = np.zeros(steps)
        arr = B
        if _is_torch_cuda(arr):
            k = arr.shape[1]
            self._bind_stream()
            st = _lib.lib().psp_hip_solve_fixed(self._ctx(), C.c_void_p(arr.data_ptr()), k, m,
                                                steps, _lib.DEVICE_POINTER,
                                                hist.ctypes.data_as(C.c_void_p))
        else:
            k = arr.shape[1]
            st = _lib.lib().psp_hip_solve_fixed(self._ctx(), arr.ctypes.data_as(C.c_void_p), k,
                                                m, steps, 0, hist.ctypes.data_as(C.c_void_p))
        if st != _lib.OK:
            _raise(st, self._h, "SolveFixed")
        return hist

    # ---- device operators (torch complex128 CUDA tensors, grid layout) ---
    def Multiply(self, x, y):
        self._bind_stream()
        st = _lib.lib().psp_hip_multiply(self._ctx(), C.c_void_p(x.data_ptr()),
                                         C.c_void_p(y.data_ptr()), x.shape[1])
        if st != _lib.OK:
            _raise(st, self._h, "Multiply")

    def Precondition(self, B):
        self._bind_stream()
        st = _lib.lib().psp_hip_precondition(self._ctx(), C.c_void_p(B.data_ptr()), B.shape[1])
        if st != _lib.OK:
            _raise(st, self._h, "Precondition")

    def PreconditionNoAlloc(self, B):
        """Precondition with the sweep's buffers in the library's fixed
        device scratch (psp_hip_precondition_noalloc): allocates nothing;
        grids whose sweep needs more than the scratch are a LogicError."""
        self._bind_stream()
        st = _lib.lib().psp_hip_precondition_noalloc(self._ctx(), C.c_void_p(B.data_ptr()),
                                                     B.shape[1])
        if st != _lib.OK:
            _raise(st, self._h, "PreconditionNoAlloc")

    def SolvePanel(self, B, panel):
        self._bind_stream()
        st = _lib.lib().psp_hip_solve_panel(self._ctx(), C.c_void_p(B.data_ptr()), B.shape[1],
                                            panel)
        if st != _lib.OK:
            _raise(st, self._h, "SolvePanel")

    def ProfileSweep(self, B):
        """One sweep launched kernel by kernel with HIP events around every
        front-solve product launch; returns (launches, seconds, bytes, flops)."""
        self._bind_stream()
        st_ = _lib.KernelStats()
        st = _lib.lib().psp_hip_profile_sweep(self._ctx(), C.c_void_p(B.data_ptr()), B.shape[1],
                                              C.byref(st_))
        if st != _lib.OK:
            _raise(st, self._h, "ProfileSweep")
        return st_.launches, st_.seconds, st_.bytes, st_.flops

    def AdjointGradient(self, u, lam, grad, accumulate=False):
        self._bind_stream()
        st = _lib.lib().psp_hip_adjoint_gradient(self._ctx(), C.c_void_p(u.data_ptr()),
                                                 C.c_void_p(lam.data_ptr()), u.shape[1],
                                                 C.c_void_p(grad.data_ptr()), int(accumulate))
        if st != _lib.OK:
            _raise(st, self._h, "AdjointGradient")

    # ---- symbolic queries (host only) ------------------------------------
    @property
    def num_panels(self):
        return _lib.lib().psp_hip_num_panels(self._ctx())

    def panel_info(self, p):
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        st = _lib.lib().psp_hip_panel_info(self._ctx(), p, C.byref(a), C.byref(b), C.byref(c),
                                           C.byref(d))
        if st != _lib.OK:
            _raise(st, self._h, "panel_info")
        return dict(v=a.value, depth=b.value, padding=c.value, cls=d.value)

    def class_nodes(self, cls):
        L = _lib.lib()
        n = L.psp_hip_class_num_nodes(self._ctx(), cls)
        off, size, parent, usize = (np.empty(n, np.int32) for _ in range(4))
        L.psp_hip_class_nodes(self._ctx(), cls, off.ctypes.data_as(C.c_void_p),
                              size.ctypes.data_as(C.c_void_p), parent.ctypes.data_as(C.c_void_p),
                              usize.ctypes.data_as(C.c_void_p), None)
        st = np.empty(int(usize.sum()), np.int32)
        L.psp_hip_class_nodes(self._ctx(), cls, None, None, None, None,
                              st.ctypes.data_as(C.c_void_p))
        starts = np.concatenate([[0], np.cumsum(usize)])
        structs = [st[starts[i]:starts[i + 1]] for i in range(n)]
        return dict(off=off, size=size, parent=parent, usize=usize, structs=structs)

    def reordered_index(self, x, y, vl, vs):
        return _lib.lib().psp_hip_reordered_index(self._ctx(), x, y, vl, vs)


def SolveBatch(helms, Bs, m=20, relTol=1e-4, maxIts=0, raise_on_error=True, threads=False):
    """Solves many independent problems at once: problem i gets the result of
    helms[i].Solve(Bs[i], m, relTol, maxIts).  By default the GMRES steps of
    all problems run in lockstep and their sweeps are merged: one launch per
    sweep step (a panel copy, a tree level's products or its reduce) covers
    every unfinished problem, replayed from one captured graph.  Problems
    with FP32 and FP64 fronts form two such batches.  threads=True: each
    problem runs its own Solve on its own host thread and stream instead.
    The blocks are all numpy arrays (host) or all torch complex128 CUDA
    tensors (HBM).  Returns the SolveReports in order."""
    n = len(helms)
    if len(Bs) != n:
        raise LogicError("SolveBatch: one right-hand-side block per problem")
    if len({id(h) for h in helms}) != n:
        raise LogicError("SolveBatch: problems must be distinct DistHelmholtz objects")
    dev = [_is_torch_cuda(b) for b in Bs]
    if n and any(d != dev[0] for d in dev):
        raise LogicError("SolveBatch: blocks must be all host or all device")
    ctxs = (C.c_void_p * n)()
    ptrs = (C.c_void_p * n)()
    ks = (C.c_int * n)()
    for i, (h, b) in enumerate(zip(helms, Bs)):
        ctxs[i] = h._ctx()
        _lib.lib().psp_hip_set_stream(ctxs[i], None)  # each problem on its own stream
        if dev[i]:
            import torch
            assert b.dtype == torch.complex128 and b.is_contiguous()
            ptrs[i] = b.data_ptr()
            ks[i] = b.shape[1] if b.dim() == 2 else 1
        else:
            if not (isinstance(b, np.ndarray) and b.dtype == np.complex128
                    and b.flags.c_contiguous):
                raise LogicError("B must be a C-contiguous complex128 array")
            ptrs[i] = b.ctypes.data
            ks[i] = b.shape[1] if b.ndim == 2 else 1
    if n and dev[0]:
        import torch
        torch.cuda.current_stream().synchronize()  # the blocks are ready
    reps = (SolveReport * n)()
    flags = (_lib.DEVICE_POINTER if (n and dev[0]) else 0) | (_lib.BATCH_THREADS if threads else 0)
    st = _lib.lib().psp_hip_solve_batch(ctxs, ptrs, ks, n, m, relTol, maxIts, flags, reps)
    if st != _lib.OK and raise_on_error:
        bad = next((i for i in range(n) if reps[i].status != _lib.OK), 0)
        _raise(st, helms[bad]._h, "SolveBatch")
    return list(reps)


def unit_cube_discretization(n, nz=None, ppw=10.0, cmin=1.0, pml=5, sigma=1.5, ny=None):
    """Helper: grid spacing h = 1/(n+1) and omega from PSP's points-per-
    wavelength definition (initialize_finalize.hpp:60-69):
    omega = 2 pi cmin n / (ppw w) with w = 1."""
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    h = 1.0 / (max(n, ny, nz) + 1)
    wx, wy, wz = (n + 1) * h, (ny + 1) * h, (nz + 1) * h
    wmax = max(wx, wy, wz)
    nmax = max(n, ny, nz)
    omega = 2 * np.pi * cmin * nmax / (ppw * wmax)
    return Discretization.with_pml(omega, n, ny, nz, wx, wy, wz, PML, PML, PML, PML, PML,
                                   pml, sigma)
