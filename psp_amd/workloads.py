"""Synthetic workloads: velocity models, sources and discretizations.

Conventions follow the reference's UnitCube driver (tests/UnitCube.cpp):
the unit cube has widths 1 and spacing h = 1/(n+1); coordinates are
X = x/(n+1) (UnitCube.cpp:105).  Non-cubic grids keep h uniform with
w_i = (n_i + 1) h.  omega comes from PSP's points-per-wavelength definition
(initialize_finalize.hpp:60-69): PPW = n * lambda / w with lambda =
2 pi c_min / omega, i.e. omega = 2 pi c_min n / (PPW w) on the longest axis.

All randomness is seeded (numpy PCG64): seed = 20261202 + config index, with
a separate stream per purpose.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .helmholtz import PML, Discretization

SEED0 = 20261202


def _rng(cfg, purpose):
    return np.random.default_rng([SEED0 + cfg, purpose])


def grid_coords(nx, ny, nz):
    """Physical coordinates of grid points, arrays of shape (nz, ny, nx)."""
    h = 1.0 / (max(nx, ny, nz) + 1)
    X = np.arange(nx) * h
    Y = np.arange(ny) * h
    Z = np.arange(nz) * h
    zz, yy, xx = np.meshgrid(Z, Y, X, indexing="ij")
    return xx, yy, zz, h


def discretization(nx, ny, nz, c_min, ppw, pml=5, sigma=1.5, top=PML):
    h = 1.0 / (max(nx, ny, nz) + 1)
    wx, wy, wz = (nx + 1) * h, (ny + 1) * h, (nz + 1) * h
    nmax = max(nx, ny, nz)
    wmax = max(wx, wy, wz)
    omega = 2 * np.pi * c_min * nmax / (ppw * wmax)
    return Discretization.with_pml(omega, nx, ny, nz, wx, wy, wz, PML, PML, PML, PML, top,
                                   pml, sigma)


def gaussian_sources(nx, ny, nz, centers):
    """Gaussian point sources exp(-|X - X0|^2 / (2 (2h)^2)), grid layout
    (N, k) with N in natural order x + y*nx + z*nx*ny."""
    xx, yy, zz, h = grid_coords(nx, ny, nz)
    s = 2.0 * h
    cols = []
    for (cx, cy, cz) in centers:
        g = np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2) / (2 * s * s))
        cols.append(g.reshape(-1))
    return np.ascontiguousarray(np.stack(cols, axis=1).astype(np.complex128))


def bump_velocity(nx, ny, nz, cfg, nbumps=32, background=3.0, lo=1.5, hi=4.5):
    """Smooth random velocity: background + sum of seeded Gaussian bumps
    (amplitudes U[-1.5, 1.5], widths U[5, 20] cells), clipped to [lo, hi]."""
    rng = _rng(cfg, 1)
    xx, yy, zz, h = grid_coords(nx, ny, nz)
    c = np.full(xx.shape, background)
    for _ in range(nbumps):
        a = rng.uniform(-1.5, 1.5)
        w = rng.uniform(5, 20) * h
        cx, cy, cz = rng.uniform(0.1, 0.9, 3) * np.array([nx, ny, nz]) * h
        c += a * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2) / (2 * w * w))
    return np.clip(c, lo, hi).reshape(-1)


def interior_points(nx, ny, nz, count, cfg, pml=5):
    rng = _rng(cfg, 2)
    h = 1.0 / (max(nx, ny, nz) + 1)
    def pick(n):
        lo = min(pml + 2, (n - 1) // 2)
        return rng.integers(lo, max(lo + 1, n - lo))
    pts = []
    for _ in range(count):
        ix, iy, iz = pick(nx), pick(ny), pick(nz)
        pts.append((ix * h, iy * h, iz * h))
    return pts


def thrust_belt_model(nx, ny, nz, cfg, nlayers=12, nfaults=3):
    """Seeded parameters of the layered thrust-belt model, lengths in grid
    cells: faults (strike, offset, throw) and interfaces (depth z0 at the
    model centre, dip angle, dip direction), both drawn from PCG64 stream
    (cfg, 1).  Dips are U[0, 30] degrees, throws U[5, 15] cells."""
    rng = _rng(cfg, 1)
    faults = []
    for _ in range(nfaults):
        strike = rng.uniform(0, np.pi)
        off = rng.uniform(0.3, 0.7)
        throw = rng.uniform(5, 15)
        faults.append((strike, off, throw))
    z0s = np.sort(rng.uniform(0.05, 0.95, nlayers)) * nz
    ifaces = []
    for z0 in z0s:
        dip = np.deg2rad(rng.uniform(0, 30))
        azim = rng.uniform(0, 2 * np.pi)
        ifaces.append((float(z0), float(dip), float(azim)))
    return dict(faults=faults, interfaces=ifaces, rng=rng)


def thrust_belt_velocity(nx, ny, nz, cfg, nlayers=12, nfaults=3, noise_amp=0.05):
    """Layered thrust-belt model: velocity rising 2.0 -> 6.0 with depth (z)
    through seeded dipping interfaces, offset by planar thrust faults, with
    +-noise_amp smooth noise.  Interfaces are planes
    z = z0 + tan(dip) * (horizontal distance from the model centre along the
    dip direction), all lengths in grid cells, so a drawn dip is the physical
    dip of the interface."""
    mdl = thrust_belt_model(nx, ny, nz, cfg, nlayers, nfaults)
    rng = mdl["rng"]
    xx, yy, zz, h = grid_coords(nx, ny, nz)
    gz, gy, gx = zz / h, yy / h, xx / h
    X = gx / nx
    Y = gy / ny
    shift = np.zeros_like(gz)
    for strike, off, throw in mdl["faults"]:
        side = (X * np.cos(strike) + Y * np.sin(strike) - off + 0.3 * gz / nz) > 0
        shift = shift + np.where(side, throw, 0.0)
    d = gz - shift  # depth in cells in the unfaulted frame
    layer = np.zeros_like(d)
    for z0, dip, azim in mdl["interfaces"]:
        plane = z0 + np.tan(dip) * ((gx - (nx - 1) / 2) * np.cos(azim)
                                    + (gy - (ny - 1) / 2) * np.sin(azim))
        layer = layer + (d > plane)
    c = 2.0 + 4.0 * layer / max(nlayers, 1)
    noise = np.zeros_like(c)
    for _ in range(16):
        cx, cy, cz = rng.uniform(0, 1, 3) * np.array([nx, ny, nz]) * h
        w = rng.uniform(8, 30) * h
        noise += rng.uniform(-1, 1) * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2) / (2 * w * w))
    noise /= max(1e-12, np.abs(noise).max())
    return (c * (1 + noise_amp * noise)).reshape(-1)


@dataclass
class Workload:
    name: str
    disc: Discretization
    velocity: np.ndarray
    B: np.ndarray


def config1():
    """40^3, c = 1, 10 PPW, one Gaussian source at the centre."""
    n = 40
    disc = discretization(n, n, n, 1.0, 10.0)
    vel = np.ones(n ** 3)
    B = gaussian_sources(n, n, n, [(0.5, 0.5, 0.5)])
    return Workload("unit-cube 40^3 c=1 10ppw 1rhs", disc, vel, B)


def config2(n=100, k=8):
    """100^3 Gaussian-bump velocity in [1.5, 4.5], 10 PPW, 8 sources."""
    vel = bump_velocity(n, n, n, cfg=2)
    disc = discretization(n, n, n, float(vel.min()), 10.0)
    B = gaussian_sources(n, n, n, interior_points(n, n, n, k, cfg=2))
    return Workload(f"bumps {n}^3 10ppw {k}rhs", disc, vel, B)


def velocity_scaled_sigma(velocity, sigma_at_unit_speed=1.5):
    """PML magnitude for a medium up to c_max: the one-way attenuation of the
    reference's profile through its layer is exp(-2 pi sigma / (3 c))
    (initialize_finalize.hpp:415-525), so sigma = 1.5 c_max keeps the
    attenuation of the reference's unit-speed default (sigma = 1.5 at c = 1,
    tests/UnitCube.cpp:28), exp(-pi), at the fastest velocity."""
    return sigma_at_unit_speed * float(np.max(velocity))


def config3(nx=200, ny=200, nz=100, k=64):
    """Thrust belt 200x200x100 (depth = z = sweep axis), 8 PPW at c_min=2,
    64 sources on an 8x8 lattice of the plane z = bz + 2.  Velocities reach
    6, so the PML magnitude scales with c_max (velocity_scaled_sigma): at
    sigma = 1.5 the layers pass exp(-2 pi 1.5 / 18) = 0.59 of a wave at
    c = 6 one way, and GMRES stagnates at relative residual 1.0."""
    vel = thrust_belt_velocity(nx, ny, nz, cfg=3)
    disc = discretization(nx, ny, nz, 2.0, 8.0, sigma=velocity_scaled_sigma(vel))
    h = 1.0 / (max(nx, ny, nz) + 1)
    side = int(round(np.sqrt(k)))
    xs = np.linspace(0.15, 0.85, side) * nx * h
    ys = np.linspace(0.15, 0.85, side) * ny * h
    z0 = (5 + 2) * h  # the surface is z = 0 (PSP sweeps up from z = nz-1)
    centers = [(x, y, z0) for y in ys for x in xs][:k]
    B = gaussian_sources(nx, ny, nz, centers)
    return Workload(f"thrust-belt {nx}x{ny}x{nz} 8ppw {k}rhs", disc, vel, B)


@dataclass
class InversionWorkload:
    """Misfit workload for the adjoint-state gradient: squared slowness m of
    the current model, the true model's m_true, sources (N, k) and receiver
    indices (natural order) on a surface plane."""
    name: str
    disc: Discretization
    m: np.ndarray
    m_true: np.ndarray
    B: np.ndarray
    receivers: np.ndarray


def surface_survey(nx, ny, nz, k, pml=5, rec_stride=2):
    """k sources on a near-square lattice of the plane z = (pml + 2) h and
    receivers on every rec_stride-th point of the plane z = pml + 3, both
    kept off the lateral PML where the grid is wide enough."""
    h = 1.0 / (max(nx, ny, nz) + 1)
    sx = int(np.ceil(np.sqrt(k)))
    sy = int(np.ceil(k / sx))
    xs = np.linspace(0.2, 0.8, sx) * nx * h
    ys = np.linspace(0.2, 0.8, sy) * ny * h
    centers = [(x, y, (pml + 2) * h) for y in ys for x in xs][:k]
    B = gaussian_sources(nx, ny, nz, centers)
    zr = pml + 3
    lx, ly = min(pml + 1, nx // 3), min(pml + 1, ny // 3)
    rx = np.arange(lx, nx - lx, rec_stride)
    ry = np.arange(ly, ny - ly, rec_stride)
    yy, xx = np.meshgrid(ry, rx, indexing="ij")
    receivers = (xx + yy * nx + zr * nx * ny).reshape(-1).astype(np.int64)
    return B, receivers


def config5(n=80, k=16, pml=5):
    """80^3 adjoint-gradient case.  Current model: Gaussian bumps in
    [1.5, 4.5] (the 100^3 recipe, its own seed); true model: the same with a
    +10% velocity anomaly, a Gaussian of radius 10 cells at the centre.
    10 PPW at the current model's c_min.  16 sources on a 4x4 lattice of the
    first non-PML plane below the top (z = pml; the top face is z = 0, PSP's
    v = nz - 1), receivers at every (x, y) of the last non-PML plane above
    the bottom (z = nz - 1 - pml)."""
    vel = bump_velocity(n, n, n, cfg=5)
    xx, yy, zz, h = grid_coords(n, n, n)
    c0 = (n - 1) / 2 * h
    r = 10 * h
    anomaly = np.exp(-((xx - c0) ** 2 + (yy - c0) ** 2 + (zz - c0) ** 2) / (2 * r * r)).reshape(-1)
    vel_true = vel * (1 + 0.1 * anomaly)
    disc = discretization(n, n, n, float(vel.min()), 10.0)
    side = int(np.ceil(np.sqrt(k)))
    xs = np.linspace(0.2, 0.8, side) * n * h
    centers = [(x, y, pml * h) for y in xs for x in xs][:k]
    B = gaussian_sources(n, n, n, centers)
    zr = n - 1 - pml
    receivers = np.arange(zr * n * n, (zr + 1) * n * n, dtype=np.int64)
    return InversionWorkload(f"adjoint bumps {n}^3 10ppw {k}src", disc, 1.0 / vel ** 2,
                             1.0 / vel_true ** 2, B, receivers)


def config4(count=32):
    """32 independent problems: n_i uniform in {40, 48, ..., 96} (cubes),
    bump velocities in [1.5, 4.5] (the 100^3 recipe, seeded per problem),
    10 PPW at each problem's c_min, k_i uniform in {4, ..., 16} Gaussian
    sources at seeded interior points."""
    rng = _rng(4, 3)
    sizes = rng.choice(np.arange(40, 97, 8), size=count)
    ks = rng.integers(4, 17, size=count)
    out = []
    for i, (n, k) in enumerate(zip(sizes, ks)):
        n, k = int(n), int(k)
        vel = bump_velocity(n, n, n, cfg=4000 + i)
        disc = discretization(n, n, n, float(vel.min()), 10.0)
        B = gaussian_sources(n, n, n, interior_points(n, n, n, k, cfg=4000 + i))
        out.append(Workload(f"batch[{i}] bumps {n}^3 10ppw {k}rhs", disc, vel, B))
    return out
