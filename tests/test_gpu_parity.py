"""GPU parity: the HIP path against the CPU oracle on the same inputs.

Tolerances (FP64 fronts): operator application 1e-13 relative; panel solve
and sweep 1e-10 relative; GMRES iterates 1e-10 relative at a fixed
iteration count.  FP32 fronts: the sweep agrees with the FP64 sweep to
1e-4 relative and the outer FP64 GMRES still reaches the requested
tolerance.
"""
import numpy as np
import pytest

from conftest import rand_block

pytestmark = pytest.mark.gpu


def _problem(nx, ny, nz, cfg_seed=7, cutoff=12, bumps=True):
    from psp_amd import workloads
    if bumps:
        vel = workloads.bump_velocity(nx, ny, nz, cfg=cfg_seed)
    else:
        vel = np.ones(nx * ny * nz)
    disc = workloads.discretization(nx, ny, nz, float(vel.min()), 10.0)
    return disc, vel


def _oracle(disc, vel, cutoff=12):
    from oracle.oracle import Oracle
    o = Oracle(disc.omega, disc.nx, disc.ny, disc.nz, disc.wx, disc.wy, disc.wz,
               sigma=disc.sigmax, cutoff=cutoff)
    o.factor(vel)
    return o


def _helm(disc, vel, precision="fp64", cutoff=12):
    import psp_amd
    h = psp_amd.DistHelmholtz(disc, cutoff=cutoff, precision=precision)
    h.Initialize(vel)
    return h


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


GRIDS = [(12, 10, 24, 12), (16, 16, 20, 5), (9, 13, 27, 3)]


@pytest.mark.parametrize("nx,ny,nz,cut", GRIDS)
def test_multiply_matches_oracle(torch_cuda, nx, ny, nz, cut):
    torch = torch_cuda
    disc, vel = _problem(nx, ny, nz)
    o = _oracle(disc, vel, cut)
    h = _helm(disc, vel, cutoff=cut)
    rng = np.random.default_rng(1)
    k = 3
    X = rand_block(rng, o.N, k)
    xd = torch.from_numpy(X).cuda()
    yd = torch.empty_like(xd)
    h.Multiply(xd, yd)
    torch.cuda.synchronize()
    Y = yd.cpu().numpy()
    l2n = o.local_to_natural()
    ref = o.multiply(X[l2n])  # ND order
    got = Y[l2n]
    assert _rel(got, ref) < 1e-13


@pytest.mark.parametrize("nx,ny,nz,cut", GRIDS)
def test_panel_solves_match_oracle(torch_cuda, nx, ny, nz, cut):
    torch = torch_cuda
    disc, vel = _problem(nx, ny, nz)
    o = _oracle(disc, vel, cut)
    h = _helm(disc, vel, cutoff=cut)
    rng = np.random.default_rng(2)
    k = 2
    l2n = o.local_to_natural()
    for p in range(o.num_panels):
        X = rand_block(rng, o.N, k)
        xd = torch.from_numpy(X).cuda()
        h.SolvePanel(xd, p)
        torch.cuda.synchronize()
        got = xd.cpu().numpy()[l2n]
        ref = o.solve_panel(X[l2n], p)
        assert _rel(got, ref) < 1e-10, p


@pytest.mark.parametrize("nx,ny,nz,cut", GRIDS)
def test_precondition_matches_oracle(torch_cuda, nx, ny, nz, cut):
    torch = torch_cuda
    disc, vel = _problem(nx, ny, nz)
    o = _oracle(disc, vel, cut)
    h = _helm(disc, vel, cutoff=cut)
    rng = np.random.default_rng(3)
    X = rand_block(rng, o.N, 4)
    xd = torch.from_numpy(X).cuda()
    h.Precondition(xd)
    torch.cuda.synchronize()
    l2n = o.local_to_natural()
    got = xd.cpu().numpy()[l2n]
    ref = o.precondition(X[l2n])
    assert _rel(got, ref) < 1e-10


def test_gmres_fixed_iterations_match_oracle(torch_cuda):
    disc, vel = _problem(14, 12, 26)
    o = _oracle(disc, vel)
    h = _helm(disc, vel)
    rng = np.random.default_rng(4)
    B = rand_block(rng, o.N, 3)
    steps, m = 7, 4  # crosses a restart
    xo, ho = o.solve_fixed(B, m, steps)
    Bg = B.copy()
    hg = h.SolveFixed(Bg, m, steps)
    assert np.allclose(hg, ho, rtol=1e-9, atol=0)
    assert _rel(Bg, xo) < 1e-10


def test_solve_converges_and_matches_oracle(torch_cuda):
    disc, vel = _problem(20, 18, 30)
    o = _oracle(disc, vel)
    h = _helm(disc, vel)
    from psp_amd import workloads
    B = workloads.gaussian_sources(disc.nx, disc.ny, disc.nz,
                                   workloads.interior_points(disc.nx, disc.ny, disc.nz, 2, cfg=9))
    xo, st, rep_o = o.solve(B, 20, 1e-8)
    assert st == 0
    Bg = B.copy()
    rep = h.Solve(Bg, 20, 1e-8)
    assert rep.status == 0 and rep.final_max_rel < 1e-8
    assert rep.iterations == rep_o.iterations
    assert _rel(Bg, xo) < 1e-10


def test_fp32_fronts_reach_tolerance(torch_cuda):
    disc, vel = _problem(20, 18, 30)
    h64 = _helm(disc, vel, "fp64")
    h32 = _helm(disc, vel, "fp32")
    from psp_amd import workloads
    B = workloads.gaussian_sources(disc.nx, disc.ny, disc.nz,
                                   workloads.interior_points(disc.nx, disc.ny, disc.nz, 3, cfg=9))
    x64 = B.copy()
    r64 = h64.Solve(x64, 20, 1e-8)
    x32 = B.copy()
    r32 = h32.Solve(x32, 20, 1e-8)
    assert r64.status == 0 and r32.status == 0
    assert r32.final_max_rel < 1e-8
    # same operator, both converged to 1e-8: solutions agree to ~cond * 1e-8
    assert _rel(x32, x64) < 1e-5


def test_sweep_graph_replays_match_eager_bitwise(torch_cuda):
    """The captured sweep is replayed several times on one buffer, with the
    program's own eager work and synchronizations between replays; every
    replay must equal the eager sweep bit for bit (all reductions in the
    sweep have a fixed order)."""
    torch = torch_cuda
    disc, vel = _problem(16, 14, 26)
    h = _helm(disc, vel)
    rng = np.random.default_rng(21)
    X = torch.from_numpy(rand_block(rng, disc.nx * disc.ny * disc.nz, 3)).cuda()
    h.set_sweep_graph(False)
    eager = X.clone()
    h.Precondition(eager)
    torch.cuda.synchronize()
    h.set_sweep_graph(True)
    y = torch.empty_like(X)
    for rep in range(4):
        y.copy_(X * 1.0)
        junk = torch.randn(1000, device="cuda").sum()  # eager work between replays
        h.Precondition(y)
        torch.cuda.synchronize()
        assert torch.equal(y, eager), rep
        del junk


WIDE = [(8, (12, 10, 24, 12)), (13, (16, 16, 20, 5)), (40, (9, 13, 27, 3)), (64, (12, 10, 24, 12))]


@pytest.mark.parametrize("k,grid", WIDE)
def test_wide_block_panel_solves_and_sweep_match_oracle(torch_cuda, k, grid):
    """k >= 8 right-hand sides run the front solves on the matrix cores
    (FP64 MFMA for FP64 fronts): panel solves and the sweep against the
    oracle at 1e-10, including padded passes (k = 13, 40)."""
    torch = torch_cuda
    nx, ny, nz, cut = grid
    disc, vel = _problem(nx, ny, nz)
    o = _oracle(disc, vel, cut)
    h = _helm(disc, vel, cutoff=cut)
    rng = np.random.default_rng(30 + k)
    l2n = o.local_to_natural()
    for p in (0, o.num_panels // 2, o.num_panels - 1):
        X = rand_block(rng, o.N, k)
        xd = torch.from_numpy(X).cuda()
        h.SolvePanel(xd, p)
        torch.cuda.synchronize()
        assert _rel(xd.cpu().numpy()[l2n], o.solve_panel(X[l2n], p)) < 1e-10, p
    X = rand_block(rng, o.N, k)
    xd = torch.from_numpy(X).cuda()
    h.Precondition(xd)
    torch.cuda.synchronize()
    assert _rel(xd.cpu().numpy()[l2n], o.precondition(X[l2n])) < 1e-10


@pytest.mark.parametrize("k", [8, 64])
def test_wide_block_fp32_sweep_close_to_fp64(torch_cuda, k):
    """FP32 fronts on FP32 MFMA (FP32 accumulation): the sweep agrees with the
    FP64 sweep to FP32 accuracy, column by column."""
    torch = torch_cuda
    disc, vel = _problem(16, 14, 26)
    h64 = _helm(disc, vel, "fp64")
    h32 = _helm(disc, vel, "fp32")
    rng = np.random.default_rng(40 + k)
    X = torch.from_numpy(rand_block(rng, disc.nx * disc.ny * disc.nz, k)).cuda()
    y64, y32 = X.clone(), X.clone()
    h64.Precondition(y64)
    h32.Precondition(y32)
    torch.cuda.synchronize()
    err = (torch.linalg.norm(y32 - y64, dim=0) / torch.linalg.norm(y64, dim=0)).max().item()
    assert err < 1e-4, err
    # narrow blocks (vector path) and wide blocks (matrix cores) agree
    y8 = X[:, :4].contiguous()
    h32.Precondition(y8)
    torch.cuda.synchronize()
    err2 = (torch.linalg.norm(y8 - y32[:, :4]) / torch.linalg.norm(y32[:, :4])).item()
    assert err2 < 1e-5, err2


def test_threaded_solve_batch_matches_individual_solves(torch_cuda):
    """Uneven problems solved together, one host thread and stream each
    (SolveBatch(threads=True)), give exactly the results and iteration counts
    of one-by-one solves."""
    torch = torch_cuda
    import psp_amd
    from psp_amd import workloads
    probs = [(14, 12, 24, 3, "fp64"), (16, 16, 20, 9, "fp32"), (12, 10, 26, 5, "fp64"),
             (18, 14, 22, 33, "fp32")]
    helms, blocks, ref, its = [], [], [], []
    for j, (nx, ny, nz, k, prec) in enumerate(probs):
        disc, vel = _problem(nx, ny, nz, cfg_seed=50 + j)
        h = _helm(disc, vel, prec)
        B = workloads.gaussian_sources(nx, ny, nz, workloads.interior_points(nx, ny, nz, k, cfg=60 + j))
        x = torch.from_numpy(B).cuda()
        r = h.Solve(x, 20, 1e-9)
        torch.cuda.synchronize()
        helms.append(h)
        blocks.append(torch.from_numpy(B).cuda())
        ref.append(x.cpu())
        its.append(r.iterations)
    reps = psp_amd.SolveBatch(helms, blocks, 20, 1e-9, threads=True)
    torch.cuda.synchronize()
    for j in range(len(probs)):
        assert reps[j].status == 0 and reps[j].final_max_rel < 1e-9
        assert reps[j].iterations == its[j], j
        assert torch.equal(blocks[j].cpu(), ref[j]), j


def test_autotuned_shapes_keep_parity(torch_cuda):
    """Tile shapes picked by the autotuner change only the summation order:
    the sweep still matches the oracle."""
    torch = torch_cuda
    disc, vel = _problem(16, 14, 26)
    o = _oracle(disc, vel)
    h = _helm(disc, vel)
    before, after = h.Autotune(8)
    assert 0 < after <= before
    rng = np.random.default_rng(71)
    l2n = o.local_to_natural()
    for k in (3, 8, 40):
        X = rand_block(rng, o.N, k)
        xd = torch.from_numpy(X).cuda()
        h.Precondition(xd)
        torch.cuda.synchronize()
        assert _rel(xd.cpu().numpy()[l2n], o.precondition(X[l2n])) < 1e-10, k


@pytest.mark.parametrize("k", [8, 13])
def test_wide_workgroup_shapes_match_oracle(torch_cuda, k):
    """16-wave workgroups (inner range spread over waves, summed by a fixed
    LDS tree) instead of inner slices: panel solves and the sweep still
    match the oracle."""
    torch = torch_cuda
    disc, vel = _problem(16, 16, 20)
    o = _oracle(disc, vel, 5)
    h = _helm(disc, vel, cutoff=5)
    # every level counts as "few tiles"; one workgroup per tile
    h.set_solve_tiling(min_tiles=100000, wide_wg=True, persist=False)
    rng = np.random.default_rng(90 + k)
    l2n = o.local_to_natural()
    X = rand_block(rng, o.N, k)
    xd = torch.from_numpy(X).cuda()
    h.Precondition(xd)
    torch.cuda.synchronize()
    assert _rel(xd.cpu().numpy()[l2n], o.precondition(X[l2n])) < 1e-10


@pytest.mark.parametrize("k", [8, 13, 16])
@pytest.mark.parametrize("grid,min_tiles", [(0, None), (2, None), (3, 100000)])
def test_persistent_solves_match_oracle(torch_cuda, k, grid, min_tiles):
    """Persistent pipelined workgroups: with a device-filling grid, with two
    or three workgroups (each works through many tiles, so the step pipeline
    crosses tiles and nodes), and with inner slices (min_tiles forces every
    level to split): panel solves and the sweep match the oracle."""
    torch = torch_cuda
    disc, vel = _problem(16, 16, 20)
    o = _oracle(disc, vel, 5)
    h = _helm(disc, vel, cutoff=5)
    h.set_solve_tiling(min_tiles=min_tiles, persist=True, persist_grid=grid)
    rng = np.random.default_rng(120 + k + grid)
    l2n = o.local_to_natural()
    X = rand_block(rng, o.N, k)
    xd = torch.from_numpy(X).cuda()
    h.Precondition(xd)
    torch.cuda.synchronize()
    assert _rel(xd.cpu().numpy()[l2n], o.precondition(X[l2n])) < 1e-10


@pytest.mark.parametrize("k", [8, 16])
def test_persistent_fp32_sweep_close_to_fp64(torch_cuda, k):
    """FP32 fronts on the persistent path agree with the FP64 sweep to FP32
    accuracy, column by column."""
    torch = torch_cuda
    disc, vel = _problem(16, 14, 26)
    h64 = _helm(disc, vel, "fp64")
    h32 = _helm(disc, vel, "fp32")
    h32.set_solve_tiling(persist=True, persist_grid=4)
    rng = np.random.default_rng(60 + k)
    X = torch.from_numpy(rand_block(rng, disc.nx * disc.ny * disc.nz, k)).cuda()
    y64, y32 = X.clone(), X.clone()
    h64.Precondition(y64)
    h32.Precondition(y32)
    torch.cuda.synchronize()
    err = (torch.linalg.norm(y32 - y64, dim=0) / torch.linalg.norm(y64, dim=0)).max().item()
    assert err < 1e-4, err


@pytest.mark.parametrize("k", [8, 13, 16, 24])
def test_tile_per_workgroup_solves_match_oracle(torch_cuda, k):
    """One workgroup per tile (the non-persistent matrix-core kernels), with
    default shapes: the sweep matches the oracle (k = 24: two passes of 16
    right-hand sides, a block the persistent path does not take)."""
    torch = torch_cuda
    disc, vel = _problem(16, 16, 20)
    o = _oracle(disc, vel, 5)
    h = _helm(disc, vel, cutoff=5)
    h.set_solve_tiling(persist=False)
    rng = np.random.default_rng(150 + k)
    l2n = o.local_to_natural()
    X = rand_block(rng, o.N, k)
    xd = torch.from_numpy(X).cuda()
    h.Precondition(xd)
    torch.cuda.synchronize()
    assert _rel(xd.cpu().numpy()[l2n], o.precondition(X[l2n])) < 1e-10


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_factor_class_overlap_gives_the_same_factors(torch_cuda, prec):
    """The leftover and top panels factored on side streams beside the full
    panels (default) or one class after another: the same factors, so the
    same sweep bit for bit; both against the oracle (FP64 fronts)."""
    torch = torch_cuda
    disc, vel = _problem(16, 16, 33)  # panel classes 0 (bottom, inner), 1 (leftover), 2 (top)
    X = rand_block(np.random.default_rng(240), disc.nx * disc.ny * disc.nz, 8)
    out = []
    for overlap in (True, False):
        import psp_amd
        h = psp_amd.DistHelmholtz(disc, precision=prec)
        h.set_factor_overlap(overlap)
        h.Initialize(vel)
        xd = torch.from_numpy(X).cuda()
        h.Precondition(xd)
        torch.cuda.synchronize()
        out.append(xd.cpu().numpy())
    n = h.num_panels() if callable(h.num_panels) else h.num_panels
    assert {h.panel_info(p)["cls"] for p in range(n)} == {0, 1, 2}
    assert np.array_equal(out[0], out[1])
    if prec == "fp64":
        o = _oracle(disc, vel)
        l2n = o.local_to_natural()
        assert _rel(out[0][l2n], o.precondition(X[l2n])) < 1e-10
