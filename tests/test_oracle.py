"""Pins the CPU oracle (oracle/psp_oracle.c) before anything is checked
against it.  The reference ships no golden vectors or assertions
(its tests/UnitCube.cpp only prints), so each oracle function is pinned by
closed forms, invariants, brute force on tiny inputs, scipy direct solves and
an independent vectorized restatement (tests/npref.py), plus the committed
fixtures of tests/golden/ (made by tests/golden/make_golden.py from npref
and scipy, never from the oracle).
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import npref
from oracle.oracle import Oracle, partial_ldl

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def make(nx, ny, nz, omega=None, sigma=1.5, cutoff=12, damping=7.0, P=4, bcs=(0,) * 5, pml=5):
    h = 1.0 / (max(nx, ny, nz) + 1)
    wx, wy, wz = (nx + 1) * h, (ny + 1) * h, (nz + 1) * h
    omega = 2 * np.pi * max(nx, ny, nz) / 10.0 if omega is None else omega
    o = Oracle(omega, nx, ny, nz, wx, wy, wz, bcs=bcs, pml=pml, sigma=sigma, damping=damping,
               planes_per_panel=P, cutoff=cutoff)
    d = npref.Disc(omega, nx, ny, nz, wx, wy, wz, bcs=bcs, b=(pml,) * 3, sigma=(sigma,) * 3,
                   damping=damping, P=P)
    return o, d


def smooth_velocity(nx, ny, nz, seed=0):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = 1.0 + 0.3 * np.sin(0.3 * x + rng.uniform()) * np.cos(0.2 * y + 0.1 * z)
    return c.ravel()


# ---------------------------------------------------------------- geometry --
@pytest.mark.parametrize("nz,panels", [(40, 8), (100, 23), (80, 18), (19, 3)])
def test_panel_count_closed_form(nz, panels):
    # nz = 40: bottom 9 + 5 full inner + leftover 2 + top 9 -> 8 panels
    o, _ = make(6, 6, nz)
    assert o.num_panels == panels


def test_shallow_domain_rejected():
    with pytest.raises(Exception):
        make(6, 6, 18)  # nz <= (bz+P) + (bz+P)


def test_panel_geometry_matches_restatement():
    o, d = make(7, 5, 37, P=3)
    pan = npref.Panels(d)
    assert pan.n == o.num_panels
    for p in range(pan.n):
        info = o.panel_info(p)
        assert info["v"] == pan.v0(p)
        assert info["depth"] == pan.depth(p)
        assert info["padding"] == pan.padding(p)


@pytest.mark.parametrize("nx,ny,vs,cut", [(9, 7, 3, 12), (16, 16, 9, 12), (5, 11, 2, 3), (8, 8, 2, 1)])
def test_nested_dissection_is_a_bijection(nx, ny, vs, cut):
    o, _ = make(nx, ny, 30, cutoff=cut)
    idx = [o.reordered_index(x, y, v, vs) for v in range(vs) for y in range(ny) for x in range(nx)]
    assert sorted(idx) == list(range(nx * ny * vs))


def _panel_graph(o, nx, ny, vs):
    edges = []
    for v in range(vs):
        for y in range(ny):
            for x in range(nx):
                i = o.reordered_index(x, y, v, vs)
                for (a, b, c) in ((x + 1, y, v), (x, y + 1, v), (x, y, v + 1)):
                    if a < nx and b < ny and c < vs:
                        edges.append((i, o.reordered_index(a, b, c, vs)))
    return edges


@pytest.mark.parametrize("nx,ny,cut", [(9, 8, 12), (12, 5, 3), (6, 6, 2)])
def test_elimination_tree_properties(nx, ny, cut):
    o, _ = make(nx, ny, 30, cutoff=cut)
    vs = o.class_vsize(0)
    T = o.class_nodes(0)
    off, size, parent = T["off"], T["size"], T["parent"]
    n = nx * ny * vs
    owner = np.full(n, -1)
    for t in range(len(off)):
        owner[off[t]:off[t] + size[t]] = t
    assert (owner >= 0).all()
    # post-order: children before parents, offsets increasing, root last
    nonempty = [t for t in range(len(off)) if size[t] > 0]
    assert all(off[a] + size[a] <= off[b] for a, b in zip(nonempty, nonempty[1:]))
    assert parent[-1] == -1 and all(parent[t] > t for t in range(len(off) - 1))

    def ancestors(t):
        out = set()
        while t >= 0:
            out.add(t)
            t = parent[t]
        return out

    for (i, j) in _panel_graph(o, nx, ny, vs):
        a, b = owner[min(i, j)], owner[max(i, j)]
        assert b in ancestors(a)  # every edge goes to the node itself or an ancestor


@pytest.mark.parametrize("nx,ny,cut", [(7, 6, 12), (10, 4, 3)])
def test_front_structure_equals_brute_force_fill(nx, ny, cut):
    """struct(t) = rows of the Cholesky factor below node t's columns, from a
    dense elimination of a random SPD matrix with the panel's pattern."""
    o, _ = make(nx, ny, 30, cutoff=cut)
    vs = o.class_vsize(0)
    n = nx * ny * vs
    rng = np.random.default_rng(3)
    A = np.zeros((n, n))
    for (i, j) in _panel_graph(o, nx, ny, vs):
        A[i, j] = A[j, i] = rng.uniform(0.1, 1.0)
    A += np.diag(np.abs(A).sum(1) + 1.0)
    Lc = np.linalg.cholesky(A)
    T = o.class_nodes(0)
    for t in range(len(T["off"])):
        o0, s = T["off"][t], T["size"][t]
        if s == 0:
            continue
        cols = Lc[:, o0:o0 + s]
        below = np.nonzero(np.abs(cols[o0 + s:]).max(axis=1) > 1e-14)[0] + o0 + s
        assert np.array_equal(below, np.sort(T["structs"][t])), t


def test_front_structure_is_box_ring_for_cutoff_12():
    """For the analytic ND with cutoff 12 the structure of a leaf is its ring
    (4 sides, no corners) times vSize, in-domain only."""
    nx = ny = 16
    o, _ = make(nx, ny, 30)
    vs = o.class_vsize(0)
    T = o.class_nodes(0)
    parents = set(T["parent"].tolist())
    leaves = [t for t in range(len(T["off"])) if t not in parents and T["size"][t] > 0]
    assert len(leaves) > 4
    for t in leaves:
        own = set(range(T["off"][t], T["off"][t] + T["size"][t]))
        pts = [(x, y) for y in range(ny) for x in range(nx)
               if o.reordered_index(x, y, 0, vs) in own]
        if len(pts) * vs != T["size"][t]:
            continue
        xs, ys = [p[0] for p in pts], [p[1] for p in pts]
        if len(set(pts)) != (max(xs) - min(xs) + 1) * (max(ys) - min(ys) + 1):
            continue  # a separator, not a leaf box
        x0, x1, y0, y1 = min(xs), max(xs), min(ys), max(ys)
        ring = set()
        for v in range(vs):
            for y in range(y0, y1 + 1):
                for x in (x0 - 1, x1 + 1):
                    if 0 <= x < nx:
                        ring.add(o.reordered_index(x, y, v, vs))
            for x in range(x0, x1 + 1):
                for y in (y0 - 1, y1 + 1):
                    if 0 <= y < ny:
                        ring.add(o.reordered_index(x, y, v, vs))
        assert set(T["structs"][t].tolist()) == ring


# ---------------------------------------------------------------- operator --
def test_global_operator_known_answer_interior():
    """sigma = 0 removes the PML: centre 2/hx^2+2/hy^2+2/hz^2-(omega/c)^2,
    couplings -1/h^2 (initialize_finalize.hpp:527-587)."""
    nx, ny, nz = 8, 7, 25
    o, d = make(nx, ny, nz, sigma=0.0)
    c = 1.7
    o.factor(np.full(nx * ny * nz, c))
    rp, cols, vals = o.global_csr()
    h2 = [d.hx ** 2, d.hy ** 2, d.hz ** 2]
    want_c = 2 / h2[0] + 2 / h2[1] + 2 / h2[2] - (d.omega / c) ** 2
    assert np.allclose(vals[rp[:-1]], want_c, rtol=1e-14)
    l2n = o.local_to_natural()
    for g in range(0, o.N, 7):
        for e in range(rp[g] + 1, rp[g + 1]):
            dn = abs(int(l2n[cols[e]]) - int(l2n[g]))
            h = d.hx if dn == 1 else (d.hy if dn == nx else d.hz)
            assert abs(vals[e] - (-1 / h ** 2)) < 1e-10 * abs(1 / h ** 2)


def test_global_operator_symmetric_bitwise():
    nx, ny, nz = 9, 8, 26
    o, _ = make(nx, ny, nz)
    o.factor(smooth_velocity(nx, ny, nz))
    rp, cols, vals = o.global_csr()
    A = sp.csr_matrix((vals, cols, rp), shape=(o.N, o.N))
    D = (A - A.T).tocoo()
    assert np.all(D.data == 0)


def test_global_operator_matches_restatement():
    nx, ny, nz = 9, 8, 26
    o, d = make(nx, ny, nz)
    vel = smooth_velocity(nx, ny, nz)
    o.factor(vel)
    rp, cols, vals = o.global_csr()
    l2n = o.local_to_natural()
    A = sp.csr_matrix((vals, l2n[cols], rp), shape=(o.N, o.N))  # rows ND, cols natural
    P = sp.csr_matrix((np.ones(o.N), (np.arange(o.N), l2n)), shape=(o.N, o.N))
    An = (P.T @ A).tocsr()  # rows natural
    R = npref.global_operator(d, vel)
    assert abs(An - R).max() <= 1e-12 * abs(R).max()


@pytest.mark.parametrize("panel", [0, 1, 3, 4])
def test_panel_operator_matches_restatement(panel):
    nx, ny, nz = 7, 6, 31
    o, d = make(nx, ny, nz)
    vel = smooth_velocity(nx, ny, nz, seed=2)
    rows, cols, vals = o.panel_lower_coo(panel, vel)
    pan = npref.Panels(d)
    R, vOff, vs = npref.panel_operator(d, vel, pan, panel)
    # padded ND index -> panel-natural x + y*nx + vl*nx*ny
    nd2pn = np.empty(nx * ny * vs, dtype=np.int64)
    for v in range(vs):
        for y in range(ny):
            for x in range(nx):
                nd2pn[o.reordered_index(x, y, v, vs)] = x + y * nx + v * nx * ny
    Lo = sp.csr_matrix((vals, (nd2pn[rows], nd2pn[cols])), shape=R.shape)
    Full = Lo + (Lo.T - sp.diags(Lo.diagonal()))  # lower storage -> symmetric
    assert abs(Full - R).max() <= 1e-12 * abs(R).max()


# -------------------------------------------------------------------- LDL^T --
@pytest.mark.parametrize("fs,s", [(70, 50), (130, 97), (40, 40), (33, 0)])
def test_partial_ldl_reconstructs_front(fs, s):
    rng = np.random.default_rng(fs)
    B = rng.standard_normal((fs, fs)) + 1j * rng.standard_normal((fs, fs))
    F = B + B.T + np.diag(np.full(fs, 4.0 * fs))  # complex symmetric
    out = partial_ldl(np.tril(F), s)
    L = np.tril(out[:, :s], -1) + np.eye(fs, s)
    D = np.diag(np.diag(out)[:s])
    S = np.tril(out[s:, s:])
    recon = L @ D @ L.T
    recon[s:, s:] += S + np.tril(S, -1).T
    assert np.abs(np.tril(recon - F)).max() < 1e-10 * np.abs(F).max()


# ------------------------------------------------------------------- solves --
@pytest.mark.parametrize("panel", [0, 1, 2, 4])
def test_panel_solve_matches_scipy(panel):
    nx, ny, nz = 7, 6, 31
    o, d = make(nx, ny, nz)
    vel = smooth_velocity(nx, ny, nz, seed=4)
    o.factor(vel)
    l2n = o.local_to_natural()
    rng = np.random.default_rng(5)
    Bn = rng.standard_normal((o.N, 2)) + 1j * rng.standard_normal((o.N, 2))  # natural
    got = o.solve_panel(Bn[l2n], panel)
    want = Bn.copy()
    npref.Sweep(d, vel).solve_panel(want, panel)
    assert np.linalg.norm(got - want[l2n]) <= 1e-11 * np.linalg.norm(want)


def test_precondition_matches_sweep_restatement():
    nx, ny, nz = 6, 7, 33
    o, d = make(nx, ny, nz, P=3)
    vel = smooth_velocity(nx, ny, nz, seed=6)
    o.factor(vel)
    l2n = o.local_to_natural()
    rng = np.random.default_rng(7)
    Bn = rng.standard_normal((o.N, 3)) + 1j * rng.standard_normal((o.N, 3))
    got = o.precondition(Bn[l2n])
    want = npref.Sweep(d, vel).precondition(Bn)
    assert np.linalg.norm(got - want[l2n]) <= 1e-11 * np.linalg.norm(want)


def test_multiply_matches_restatement():
    nx, ny, nz = 8, 6, 27
    o, d = make(nx, ny, nz)
    vel = smooth_velocity(nx, ny, nz, seed=8)
    o.factor(vel)
    l2n = o.local_to_natural()
    rng = np.random.default_rng(9)
    Bn = rng.standard_normal((o.N, 2)) + 1j * rng.standard_normal((o.N, 2))
    got = o.multiply(Bn[l2n])
    want = npref.global_operator(d, vel) @ Bn
    assert np.linalg.norm(got - want[l2n]) <= 1e-13 * np.linalg.norm(want)


def test_gmres_history_matches_textbook():
    nx, ny, nz = 7, 6, 28
    o, d = make(nx, ny, nz)
    vel = smooth_velocity(nx, ny, nz, seed=10)
    o.factor(vel)
    sw = npref.Sweep(d, vel)
    rng = np.random.default_rng(11)
    b = rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)
    m, steps = 3, 7
    x, hist = o.solve_fixed(b.reshape(-1, 1), m, steps)
    M = lambda v: sw.precondition(v.reshape(-1, 1)).ravel()
    xr, hr = npref.gmres(sw.A, M, b, m, steps)
    assert np.allclose(hist, hr, rtol=1e-8, atol=1e-14)
    assert np.linalg.norm(x.ravel() - xr) <= 1e-9 * np.linalg.norm(xr)


def test_gmres_converges_to_direct_solution():
    nx, ny, nz = 8, 7, 26
    o, d = make(nx, ny, nz)
    vel = smooth_velocity(nx, ny, nz, seed=12)
    o.factor(vel)
    rng = np.random.default_rng(13)
    B = rng.standard_normal((o.N, 2)) + 1j * rng.standard_normal((o.N, 2))
    x, st, rep = o.solve(B, 20, 1e-11)
    assert st == 0 and rep.final_max_rel < 1e-11
    A = npref.global_operator(d, vel).tocsc()
    want = spl.spsolve(A, B)
    assert np.linalg.norm(x - want) <= 1e-8 * np.linalg.norm(want)


def test_nan_rhs_is_runtime_error():
    o, _ = make(6, 6, 25)
    o.factor(np.ones(o.N))
    B = np.ones((o.N, 1), dtype=np.complex128)
    B[3] = np.nan
    _, st, _ = o.solve(B, 5, 1e-5)
    assert st == 3  # "Original residual norms had a NaN" (solve.hpp:280-282)


# ----------------------------------------------------------------- fixtures --
def test_oracle_against_golden_fixtures():
    path = os.path.join(GOLDEN, "small_problem.npz")
    g = np.load(path)  # allow_pickle=False
    nx, ny, nz = (int(v) for v in g["grid"])
    o, d = make(nx, ny, nz, omega=float(g["omega"]))
    vel = g["velocity"]
    o.factor(vel)
    l2n = o.local_to_natural()
    rp, cols, vals = o.global_csr()
    A = sp.csr_matrix((vals, l2n[cols], rp), shape=(o.N, o.N))
    P = sp.csr_matrix((np.ones(o.N), (np.arange(o.N), l2n)), shape=(o.N, o.N))
    An = (P.T @ A).tocoo()
    R = sp.coo_matrix((g["A_vals"], (g["A_rows"], g["A_cols"])), shape=An.shape)
    assert abs(An - R).max() <= 1e-12 * abs(R).max()
    got = o.precondition(g["rhs"][l2n])
    assert np.linalg.norm(got - g["precond_rhs"][l2n]) <= 1e-11 * np.linalg.norm(g["precond_rhs"])
    x, st, rep = o.solve(g["rhs"], 20, 1e-11)
    assert st == 0
    assert np.linalg.norm(x - g["solution"]) <= 1e-8 * np.linalg.norm(g["solution"])
