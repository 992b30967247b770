"""Independent numpy/scipy restatement of PSP's operators, used to pin the
CPU oracle (test infrastructure only).

Written vectorized over the natural grid, in the opposite style to the
oracle's per-row C loops, so that a transposed index, a dropped term or a
wrong sign in either shows up as a mismatch:

- PML profiles s1Inv/s2Inv/s3Inv/s3InvArtificial
  (initialize_finalize.hpp:415-525);
- the global operator of FormGlobalRow (:527-587);
- the padded panel operator of FormLowerColumnOfNode (:589-806): the global
  stencil with omega -> omega + i*damping, s3 -> s3InvArtificial and the
  v-couplings cut at the panel ends;
- the panel split (main.hpp:35-57, 728-789);
- the sweep of Precondition (solve.hpp:681-734) with scipy sparse LU panel
  solves;
- textbook restarted GMRES(m) (Saad & Schultz 1986) with the reference's
  stopping rule.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

PML = 0


class Disc:
    def __init__(self, omega, nx, ny, nz, wx, wy, wz, bcs=(PML,) * 5, b=(5, 5, 5),
                 sigma=(1.5, 1.5, 1.5), damping=7.0, P=4):
        self.omega, self.nx, self.ny, self.nz = omega, nx, ny, nz
        self.wx, self.wy, self.wz = wx, wy, wz
        self.front, self.right, self.back, self.left, self.top = bcs
        self.bx, self.by, self.bz = b
        self.sx, self.sy, self.sz = sigma
        self.damping, self.P = damping, P
        self.hx, self.hy, self.hz = wx / (nx + 1), wy / (ny + 1), wz / (nz + 1)


def _profile(i, lo_on, hi_on, b, n, sigma, h, omega, lo_lim=None):
    i = np.asarray(i, dtype=np.float64)
    eta = b * h
    lo = (b - 1) if lo_lim is None else lo_lim
    out = np.ones(i.shape, dtype=np.complex128)
    mlo = (i < lo) & lo_on
    mhi = (i > n - b) & hi_on & ~mlo
    d_lo = lo - i
    d_hi = i - (n - b)
    out[mlo] = 1 + 1j * (sigma / eta) * (d_lo[mlo] / b) ** 2 * (2 * np.pi / omega)
    out[mhi] = 1 + 1j * (sigma / eta) * (d_hi[mhi] / b) ** 2 * (2 * np.pi / omega)
    return out


def s1inv(d, x):
    return _profile(x, d.front == PML, d.back == PML, d.bx, d.nx, d.sx, d.hx, d.omega)


def s2inv(d, y):
    return _profile(y, d.left == PML, d.right == PML, d.by, d.ny, d.sy, d.hy, d.omega)


def s3inv(d, v, vOffset=None):
    lim = None if vOffset is None else vOffset + d.bz - 1
    return _profile(v, True, d.top == PML, d.bz, d.nz, d.sz, d.hz, d.omega, lo_lim=lim)


def _terms(d, X, Y, V, k2, vOffset=None):
    """Stencil terms on coordinate arrays; returns centre and the six
    coupling values (x-1, x+1, y-1, y+1, v-1, v+1)."""
    s1L, s1M, s1R = s1inv(d, X - 1), s1inv(d, X), s1inv(d, X + 1)
    s2L, s2M, s2R = s2inv(d, Y - 1), s2inv(d, Y), s2inv(d, Y + 1)
    s3L, s3M, s3R = s3inv(d, V - 1, vOffset), s3inv(d, V, vOffset), s3inv(d, V + 1, vOffset)
    xL = (s2M * s3M / s1L + s2M * s3M / s1M) / (2 * d.hx ** 2)
    xR = (s2M * s3M / s1R + s2M * s3M / s1M) / (2 * d.hx ** 2)
    yL = (s1M * s3M / s2L + s1M * s3M / s2M) / (2 * d.hy ** 2)
    yR = (s1M * s3M / s2R + s1M * s3M / s2M) / (2 * d.hy ** 2)
    vL = (s1M * s2M / s3L + s1M * s2M / s3M) / (2 * d.hz ** 2)
    vR = (s1M * s2M / s3R + s1M * s2M / s3M) / (2 * d.hz ** 2)
    centre = xL + xR + yL + yR + vL + vR - k2 * s1M * s2M * s3M
    return centre, [-xL, -xR, -yL, -yR, -vL, -vR]


def global_operator(d, vel):
    """Global operator in natural order (x + y*nx + z*nx*ny), v = nz-1-z."""
    nx, ny, nz = d.nx, d.ny, d.nz
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X, Y, Z = x.ravel(), y.ravel(), z.ravel()
    V = nz - 1 - Z
    k2 = (d.omega / vel) ** 2
    c, off = _terms(d, X, Y, V, k2)
    idx = X + Y * nx + Z * nx * ny
    rows, cols, vals = [idx], [idx], [c]
    nbr = [(X > 0, idx - 1), (X + 1 < nx, idx + 1), (Y > 0, idx - nx), (Y + 1 < ny, idx + nx),
           (V > 0, idx + nx * ny), (V + 1 < nz, idx - nx * ny)]
    for (m, j), val in zip(nbr, off):
        rows.append(idx[m])
        cols.append(j[m])
        vals.append(val[m])
    N = nx * ny * nz
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(N, N))


class Panels:
    """Panel split (main.hpp:35-57) and helpers (main.hpp:728-789)."""

    def __init__(self, d):
        P, bz, nz = d.P, d.bz, d.nz
        self.bottom = bz + P
        self.topOrig = bz + P if d.top == PML else P
        if nz <= self.bottom + self.topOrig:
            raise ValueError("shallow")
        self.inner = nz - self.bottom - self.topOrig
        self.left = self.inner % P
        self.nfull = self.inner // P
        self.n = 1 + self.nfull + (self.left != 0) + 1
        self.P, self.bz = P, bz

    def depth(self, p):
        if p == 0:
            return self.bottom
        if p < 1 + self.nfull:
            return self.P
        if self.left and p == 1 + self.nfull:
            return self.left
        return self.topOrig

    def padding(self, p):
        return 0 if p == 0 else self.bz

    def v0(self, p):
        if p == 0:
            return 0
        return self.bottom + sum(self.depth(q) for q in range(1, p))


def panel_operator(d, vel, pan, p):
    """Padded panel operator in panel-natural order x + y*nx + vl*nx*ny."""
    nx, ny, nz = d.nx, d.ny, d.nz
    vOff = pan.v0(p) - pan.padding(p)
    vs = pan.depth(p) + pan.padding(p)
    vl, y, x = np.meshgrid(np.arange(vs), np.arange(ny), np.arange(nx), indexing="ij")
    X, Y, VL = x.ravel(), y.ravel(), vl.ravel()
    V = vOff + VL
    Z = nz - 1 - V
    alpha = vel[X + Y * nx + Z * nx * ny]
    so = d.omega + 1j * d.damping
    c, off = _terms(d, X, Y, V, (so / alpha) ** 2, vOffset=vOff)
    idx = X + Y * nx + VL * nx * ny
    rows, cols, vals = [idx], [idx], [c]
    nbr = [(X > 0, idx - 1), (X + 1 < nx, idx + 1), (Y > 0, idx - nx), (Y + 1 < ny, idx + nx),
           (VL > 0, idx - nx * ny), (VL + 1 < vs, idx + nx * ny)]
    for (m, j), val in zip(nbr, off):
        rows.append(idx[m])
        cols.append(j[m])
        vals.append(val[m])
    n = nx * ny * vs
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n)), vOff, vs


def panel_natural_rows(d, pan, p):
    """Natural-order global rows of the non-padding planes of panel p, in
    panel-natural order, plus their positions in the padded panel."""
    nx, ny, nz = d.nx, d.ny, d.nz
    pad, dep = pan.padding(p), pan.depth(p)
    vl, y, x = np.meshgrid(np.arange(pad, pad + dep), np.arange(ny), np.arange(nx), indexing="ij")
    v = pan.v0(p) + (vl - pad)
    z = nz - 1 - v
    glob = (x + y * nx + z * nx * ny).ravel()
    loc = (x + y * nx + vl * nx * ny).ravel()
    return glob, loc


class Sweep:
    """Precondition (solve.hpp:681-734) in natural order with exact panel
    solves T_i (scipy SuperLU on the padded panel operators)."""

    def __init__(self, d, vel):
        self.d, self.pan = d, Panels(d)
        self.A = global_operator(d, vel).tocsr()
        self.lu, self.maps = [], []
        for p in range(self.pan.n):
            Ap, _, vs = panel_operator(d, vel, self.pan, p)
            self.lu.append(spl.splu(Ap.tocsc()))
            self.maps.append(panel_natural_rows(d, self.pan, p) + (Ap.shape[0],))

    def solve_panel(self, B, p):
        glob, loc, n = self.maps[p]
        xp = np.zeros((n, B.shape[1]), dtype=np.complex128)
        xp[loc] = B[glob]
        xp = self.lu[p].solve(xp)
        B[glob] = xp[loc]

    def _plane(self, v):
        d = self.d
        z = d.nz - 1 - v
        return np.arange(d.nx * d.ny) + z * d.nx * d.ny

    def precondition(self, B):
        B = np.array(B, dtype=np.complex128)
        pan, A = self.pan, self.A
        for i in range(pan.n - 1):
            self.solve_panel(B, i)
            r = self._plane(pan.v0(i + 1))            # first plane of panel i+1
            c = self._plane(pan.v0(i + 1) - 1)        # last plane of panel i
            B[r] -= A[r][:, c].diagonal()[:, None] * B[c]
        self.solve_panel(B, pan.n - 1)
        for i in range(pan.n - 2, -1, -1):
            glob, _, _ = self.maps[i]
            Z = B[glob].copy()
            B[glob] = 0
            r = self._plane(pan.v0(i) + pan.depth(i) - 1)  # last plane of panel i
            c = self._plane(pan.v0(i) + pan.depth(i))      # first plane of panel i+1
            B[r] = -A[r][:, c].diagonal()[:, None] * B[c]
            self.solve_panel(B, i)
            B[glob] += Z
        return B


def gmres(A, M, b, m, steps):
    """Left-preconditioned restarted GMRES(m) for one right-hand side,
    x0 = 0, with the true residual ||b - A x||/||b|| after every step."""
    N = b.shape[0]
    x = np.zeros(N, dtype=np.complex128)
    nb = np.linalg.norm(b)
    hist = []
    while len(hist) < steps:
        r = M(b - A @ x)
        beta = np.linalg.norm(r)
        V = np.zeros((N, m + 1), dtype=np.complex128)
        H = np.zeros((m + 1, m), dtype=np.complex128)
        V[:, 0] = r / beta
        for j in range(m):
            w = M(A @ V[:, j])
            for i in range(j + 1):
                H[i, j] = np.vdot(V[:, i], w)
                w = w - H[i, j] * V[:, i]
            H[j + 1, j] = np.linalg.norm(w)
            V[:, j + 1] = w / H[j + 1, j]
            e1 = np.zeros(j + 2, dtype=np.complex128)
            e1[0] = beta
            y = np.linalg.lstsq(H[:j + 2, :j + 1], e1, rcond=None)[0]
            xj = x + V[:, :j + 1] @ y
            hist.append(np.linalg.norm(b - A @ xj) / nb)
            if len(hist) >= steps:
                return xj, np.array(hist)
        x = xj
    return x, np.array(hist)
