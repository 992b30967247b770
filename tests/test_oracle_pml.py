"""Pins the oracle's PML region (oracle/psp_oracle.c, the restatement of
initialize_finalize.hpp:415-525, 589-640) by what the mathematics fixes,
independently of any restatement of the profile formula:

  * the reflection coefficient of a layer backed by a zero wall, measured
    from a discrete 1-D solution, equals the continuum value
    exp(-2 kappa Im int s) = exp(-4 pi sigma kappa / (3 omega)) of a
    quadratic profile (within the discretization error, shrinking with the
    layer's cells);
  * a point source in a homogeneous medium converges at O(h^2) to the
    outgoing Green's function exp(i k r) / (4 pi r) inside the domain;
  * the operator of a mirror-symmetric medium commutes with the mirrors;
  * with no damping, panel 0's operator is the global operator's principal
    block, and panel p's is the principal block of the global operator of
    the sub-problem that starts at the panel's padding (the artificial PML
    is the bottom PML moved up);
  * damping enters only through omega -> omega + i damping in the mass term.

Each check is also run against an oracle build with a planted mistake
(psp_oracle.c compiled with -DORA_PLANT=n, into a temporary directory) and
must fail there: a one-sided off-by-one of the x profile, a flipped
imaginary sign, a linear instead of quadratic profile, an artificial PML
one plane too deep, and a flipped damping sign.
"""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from oracle.oracle import DIRICHLET, PML, Oracle, load_library

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "oracle", "psp_oracle.c")


@pytest.fixture(scope="session")
def mutant(tmp_path_factory):
    """mutant(n): the oracle built with planted mistake n (cached)."""
    built = {}
    d = tmp_path_factory.mktemp("oracle_mutants")

    def get(n):
        if n not in built:
            so = str(d / f"libpsp_oracle_plant{n}.so")
            subprocess.run(["gcc", "-O2", "-fcx-fortran-rules", "-fopenmp", "-fPIC",
                            "-std=gnu11", "-shared", f"-DORA_PLANT={n}", "-o", so, ORACLE_SRC,
                            "-lm"], check=True)
            built[n] = load_library(so)
        return built[n]

    return get


def natural_operator(o):
    """The oracle's global operator with rows and columns in natural order."""
    rp, cols, vals = o.global_csr()
    l2n = o.local_to_natural()
    A = sp.csr_matrix((vals, l2n[cols], rp), shape=(o.N, o.N))
    P = sp.csr_matrix((np.ones(o.N), (np.arange(o.N), l2n)), shape=(o.N, o.N))
    return (P.T @ A).tocsr()


# ------------------------------------------------------------ reflection --
def reflection(axis, b, n, sigma=0.5, c=1.0, ppw=40.0, library=None):
    """Reflection coefficients |R| of the low and high layers of `axis`
    (0: x, 1: y, 2: v) and the continuum value.

    The other two axes get layers of one cell (no PML: b = 1 puts no point
    inside a layer) and widths of 1e3, so their lowest sine mode has a
    negligible eigenvalue lam and the operator acting on g(axis) * mode is
    the 1-D PML operator with kappa^2 = (omega/c)^2 - lam.  A plane source
    at the centre sends a wave into each layer; fitting the discrete waves
    exp(+-i kh j) (2 (1 - cos kh) = kappa^2 h^2) between the source and a
    layer gives incident and reflected amplitudes."""
    P = 1
    dims = [1, 1, 2 * (1 + P) + 1]
    pmls = [1, 1, 1]
    widths = [1e3, 1e3, 1e3]
    dims[axis], pmls[axis], widths[axis] = n, b, 1.0
    h = 1.0 / (n + 1)
    omega = 2 * np.pi * c / (ppw * h)
    nx, ny, nz = dims
    o = Oracle(omega, nx, ny, nz, *widths, pmls=pmls, sigmas=(sigma,) * 3,
               planes_per_panel=P, library=library)
    o.fill_global(np.full(o.N, c))
    A = natural_operator(o)
    hs = [widths[i] / (dims[i] + 1) for i in range(3)]
    modes, lam = [], 0.0
    for ax in range(3):
        j = np.arange(dims[ax])
        if ax == axis:
            modes.append(np.ones(dims[ax]))
        else:
            modes.append(np.sin(np.pi * (j + 1) / (dims[ax] + 1)))
            lam += (2 - 2 * np.cos(np.pi / (dims[ax] + 1))) / hs[ax] ** 2
    # natural order x + y nx + z nx ny with z = nz - 1 - v
    phi = modes[2][::-1][:, None, None] * modes[1][None, :, None] * modes[0][None, None, :]
    src = np.zeros(dims[axis])
    src[n // 2] = 1.0
    shp = [1, 1, 1]
    shp[2 - axis] = n
    f = phi * (src[::-1] if axis == 2 else src).reshape(shp)
    u = spl.spsolve(A.tocsc(), f.ravel().astype(complex)).reshape(nz, ny, nx) / phi
    line = np.moveaxis(u, 2 - axis, 0).reshape(n, -1)[:, 0]
    if axis == 2:
        line = line[::-1]  # index by v
    kap2 = (omega / c) ** 2 - lam
    kh = np.arccos(1 - kap2 * h * h / 2)
    j = np.arange(n)
    i0 = n // 2
    out = []
    for rng, lo in ((np.arange(b + 2, i0 - 2), True), (np.arange(i0 + 3, n - b - 2), False)):
        M = np.stack([np.exp(1j * kh * j[rng]), np.exp(-1j * kh * j[rng])], axis=1)
        fwd, bwd = np.linalg.lstsq(M, line[rng], rcond=None)[0]
        out.append(abs(fwd / bwd) if lo else abs(bwd / fwd))
    theory = np.exp(-4 * np.pi * sigma * np.sqrt(kap2) / (3 * omega))
    return out, theory


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_pml_reflection_matches_continuum_closed_form(axis):
    for b, tol in ((12, 0.02), (24, 0.006)):
        (lo, hi), theory = reflection(axis, b, 10 * b)
        for r in (lo, hi):
            assert abs(np.log(r) / np.log(theory) - 1) < tol, (axis, b, r, theory)


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_reflection_check_catches_planted_profile_mistakes(mutant, plant):
    worst = 0.0
    for axis in range(3):
        (lo, hi), theory = reflection(axis, 12, 120, library=mutant(plant))
        for r in (lo, hi):
            worst = max(worst, abs(np.log(r) / np.log(theory) - 1))
    assert worst > 0.2, (plant, worst)


# ---------------------------------------------------------- Green's function --
def green_error(n, b, omega, library=None, max_its=0):
    """Relative L2 error of a unit point source's solution against
    exp(i k r) / (4 pi r) over 0.15 <= r <= 0.3 (c = 1, unit cube)."""
    h = 1.0 / (n + 1)
    o = Oracle(omega, n, n, n, 1.0, 1.0, 1.0, pml=b, sigma=2.0, library=library)
    o.factor(np.ones(o.N))
    c = n // 2
    f = np.zeros((n, n, n), complex)
    f[c, c, c] = 1.0 / h ** 3
    u, _, _ = o.solve(f.reshape(-1, 1), 20, 1e-8, max_its)
    u = u.reshape(n, n, n)
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    r = h * np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    G = np.exp(1j * omega * r) / (4 * np.pi * np.where(r == 0, 1, r))
    m = (r >= 0.15) & (r <= 0.3)
    return np.linalg.norm((u - G)[m]) / np.linalg.norm(G[m])


def test_point_source_converges_to_outgoing_greens_function(mutant):
    omega = 2 * np.pi * 2.5  # 2.5 wavelengths across the box
    coarse = green_error(19, 4, omega)
    fine = green_error(39, 8, omega)  # same physical PML thickness
    assert fine < 0.03 and coarse / fine > 3.5, (coarse, fine)  # O(h^2): ratio 4
    flipped = green_error(19, 4, omega, library=mutant(2), max_its=40)
    assert flipped > 0.5, flipped


# ------------------------------------------------------------------ mirrors --
def mirror_velocity(nx, ny, nz, seed=0):
    rng = np.random.default_rng(seed)
    c = 1.0 + 0.4 * rng.random((nz, ny, nx))
    c = c + c[:, :, ::-1]
    c = c + c[:, ::-1, :]
    c = c + c[::-1, :, :]
    return (c / 8).ravel()


def mirror_defect(library=None):
    nx, ny, nz = 11, 9, 27
    h = 1.0 / 28
    o = Oracle(2 * np.pi * 27 / 10, nx, ny, nz, (nx + 1) * h, (ny + 1) * h, (nz + 1) * h,
               library=library)
    o.fill_global(mirror_velocity(nx, ny, nz))
    A = natural_operator(o)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = []
    for fx, fy, fz in ((True, False, False), (False, True, False), (False, False, True)):
        xr = nx - 1 - x if fx else x
        yr = ny - 1 - y if fy else y
        zr = nz - 1 - z if fz else z
        perm = (xr + yr * nx + zr * nx * ny).ravel()
        R = sp.csr_matrix((np.ones(o.N), (np.arange(o.N), perm)), shape=(o.N, o.N))
        out.append(abs(R @ A @ R.T - A).max() / abs(A).max())
    return out


def test_operator_commutes_with_mirrors(mutant):
    """A mirror-symmetric medium with PML on every face (the v layers are
    the bottom and the top PML) gives an operator that commutes with the
    reflections x -> nx-1-x, y -> ny-1-y and z -> nz-1-z (a one-sided index
    slip in any profile breaks this)."""
    assert max(mirror_defect()) < 1e-14
    assert mirror_defect(mutant(1))[0] > 1e-3


# ----------------------------------------------------------- panel operators --
def panel_lower_natural(o, panel, vel, vsize):
    """Panel `panel`'s lower operator as a csr matrix in panel-natural order
    x + y nx + vLocal nx ny."""
    nx, ny = o.nx, o.ny
    rows, cols, vals = o.panel_lower_coo(panel, vel)
    nd2pn = np.empty(nx * ny * vsize, dtype=np.int64)
    for v in range(vsize):
        for y in range(ny):
            for x in range(nx):
                nd2pn[o.reordered_index(x, y, v, vsize)] = x + y * nx + v * nx * ny
    r, c = nd2pn[rows], nd2pn[cols]
    lo = r >= c
    r, c = np.where(lo, r, c), np.where(lo, c, r)
    return sp.csr_matrix((vals, (r, c)), shape=(nx * ny * vsize,) * 2)


def global_block_lower(o, planes):
    """Lower triangle of the global operator's principal block over the
    planes v < planes, in the order x + y nx + v nx ny (v = nz - 1 - z)."""
    nx, ny, nz = o.nx, o.ny, o.nz
    A = natural_operator(o)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    v = nz - 1 - z
    new = (x + y * nx + v * nx * ny).ravel()
    keep = (v < planes).ravel()
    idx = np.nonzero(keep)[0]
    B = A[idx][:, idx].tocoo()
    r, c = new[idx][B.row], new[idx][B.col]
    lo = r >= c
    return sp.csr_matrix((B.data[lo], (r[lo], c[lo])), shape=(nx * ny * planes,) * 2)


def velocity_v(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    return (1.0 + 0.5 * rng.random(nx * ny * nz))


GEOM = dict(nx=7, ny=6, nz=31, bz=5, P=4)


def make_panel_oracle(nz=GEOM["nz"], damping=0.0, top=PML, library=None):
    nx, ny = GEOM["nx"], GEOM["ny"]
    h = 1.0 / 32
    return Oracle(2 * np.pi * 31 / 10, nx, ny, nz, (nx + 1) * h, (ny + 1) * h, (nz + 1) * h,
                  bcs=(PML, PML, PML, PML, top), damping=damping, library=library)


def test_undamped_panel0_is_global_principal_block():
    nx, ny, nz = GEOM["nx"], GEOM["ny"], GEOM["nz"]
    o = make_panel_oracle()
    vel = velocity_v(nx, ny, nz, 1)
    o.fill_global(vel)
    vs = o.class_vsize(0)
    assert o.panel_info(0)["padding"] == 0
    Pn = panel_lower_natural(o, 0, vel, vs)
    G = global_block_lower(o, vs)
    assert (Pn != G).nnz == 0  # bit for bit


def panel_vs_subproblem(panel, library=None):
    """Max difference between the undamped operator of panel `panel` and the
    principal block of the global operator of the sub-problem made of the
    panel's padded planes and the planes above them (top boundary
    Dirichlet, so only the sub-problem's bottom layer is a PML)."""
    nx, ny, nz = GEOM["nx"], GEOM["ny"], GEOM["nz"]
    o = make_panel_oracle(library=library)
    vel = velocity_v(nx, ny, nz, 2)
    info = o.panel_info(panel)
    vOff = info["v"] - info["padding"]
    vs = o.class_vsize(info["cls"])
    Pn = panel_lower_natural(o, panel, vel, vs)
    nzs = nz - vOff  # planes v >= vOff
    h = 1.0 / 32
    sub = Oracle(2 * np.pi * 31 / 10, nx, ny, nzs, (nx + 1) * h, (ny + 1) * h, (nzs + 1) * h,
                 bcs=(PML, PML, PML, PML, DIRICHLET), damping=0.0, library=library)
    # natural z of the sub-problem: z_sub = nzs - 1 - (v - vOff) = z
    sub.fill_global(vel.reshape(nz, ny, nx)[:nzs].ravel())
    G = global_block_lower(sub, vs)
    return abs(Pn - G).max() / abs(G).max()


@pytest.mark.parametrize("panel", [1, 2, 3])
def test_artificial_pml_is_the_bottom_pml_moved_up(panel, mutant):
    assert panel_vs_subproblem(panel) == 0.0
    assert panel_vs_subproblem(panel, mutant(4)) > 1e-3


def damping_defect(library=None):
    nx, ny, nz = GEOM["nx"], GEOM["ny"], GEOM["nz"]
    vel = velocity_v(nx, ny, nz, 3)
    d = 7.0
    o0 = make_panel_oracle(library=library)
    od = make_panel_oracle(damping=d, library=library)
    vs = o0.class_vsize(0)
    worst = 0.0
    for panel in (0, 2):
        P0 = panel_lower_natural(o0, panel, vel, vs).tocsr()
        Pd = panel_lower_natural(od, panel, vel, vs).tocsr()
        # the mass term's s1 s2 s3 read off the operator: it is affine in
        # m = 1/c^2 with slope -omega^2 s1 s2 s3 on the diagonal
        P2 = panel_lower_natural(o0, panel, vel / np.sqrt(2.0), vs).tocsr()
        info = o0.panel_info(panel)
        vOff = info["v"] - info["padding"]
        nat = vel.reshape(nz, ny, nx)[::-1][vOff:vOff + vs].ravel()  # by (x, y, vLocal)
        m = 1.0 / nat ** 2
        w = o0.disc.omega
        S = (P2.diagonal() - P0.diagonal()) / (-(w ** 2) * m)
        want = -(((w + 1j * d) ** 2) - w ** 2) * m * S
        off0 = P0 - sp.diags(P0.diagonal())
        offd = Pd - sp.diags(Pd.diagonal())
        assert (off0 != offd).nnz == 0
        worst = max(worst, abs(Pd.diagonal() - P0.diagonal() - want).max() / abs(want).max())
    return worst


def test_damping_enters_only_as_shifted_frequency(mutant):
    """Panel operator(d) - panel operator(0) = -((omega + i d)^2 - omega^2)
    m s1 s2 s3 on the diagonal and nothing elsewhere
    (initialize_finalize.hpp:633-636)."""
    assert damping_defect() < 1e-13
    assert damping_defect(mutant(5)) > 0.1
