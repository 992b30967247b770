"""Fused subtree solves (k_mvq, PSP_HIP_OPT_SOLVE_FUSED): the tree levels
from the cut depth down run as one persistent launch per direction, a
subtree queue per XCD, with per-node counters instead of level launches.  Checked through the C ABI
against the oracle (FP64 fronts, 1e-10 relative) for 8 to 16 right-hand
sides, every tile height, queues that no XCD serves (run by the last
workgroup out), FP32 fronts against FP64, repeat sweeps bitwise, and a
converged solve in the oracle's iteration count."""
import numpy as np
import pytest

from conftest import rand_block
from test_gpu_parity import _oracle, _problem, _rel

pytestmark = pytest.mark.gpu


def _helm(disc, vel, precision="fp64", **tiling):
    import psp_amd
    h = psp_amd.DistHelmholtz(disc, precision=precision)
    h.set_solve_tiling(**{"fused": 1, **tiling})
    h.Initialize(vel)
    return h


def _sweep(torch, h, X):
    xd = torch.from_numpy(X).cuda()
    h.Precondition(xd)
    torch.cuda.synchronize()
    return xd.cpu().numpy()


def _sweep_vs_oracle(torch, grid, k, seed, **tiling):
    disc, vel = _problem(*grid)
    o = _oracle(disc, vel)
    h = _helm(disc, vel, **tiling)
    X = rand_block(np.random.default_rng(seed), o.N, k)
    got = _sweep(torch, h, X)
    l2n = o.local_to_natural()
    return _rel(got[l2n], o.precondition(X[l2n]))


@pytest.mark.parametrize("k", [8, 9, 12, 13, 16])
@pytest.mark.parametrize("grid", [(24, 22, 30), (40, 40, 24)])
def test_fused_sweep_matches_oracle(torch_cuda, k, grid):
    # k = 9, 12, 13: vector rows that are not whole 128-byte lines
    assert _sweep_vs_oracle(torch_cuda, grid, k, 40 + k) < 1e-10


@pytest.mark.parametrize("k", [8, 13])
def test_fused_fp32_sweeps_repeat_and_match_level_launches(torch_cuda, k):
    """FP32 fronts (8-byte elements: rows of k = 8 are half a cache line):
    six repeated fused sweeps are bitwise equal and within FP32 rounding of
    the level launches' sweep."""
    torch = torch_cuda
    disc, vel = _problem(24, 22, 30)
    X = rand_block(np.random.default_rng(70 + k), disc.nx * disc.ny * disc.nz, k)
    ref = _sweep(torch, _helm(disc, vel, "fp32", fused=0), X)
    h = _helm(disc, vel, "fp32")
    first = _sweep(torch, h, X)
    for _ in range(5):
        assert np.array_equal(_sweep(torch, h, X), first)
    assert _rel(first, ref) < 1e-5


@pytest.mark.parametrize("rows", [16, 32, 64])
def test_fused_tile_heights_match_oracle(torch_cuda, rows):
    assert _sweep_vs_oracle(torch_cuda, (40, 40, 24), 8, 50 + rows, fused_rows=rows) < 1e-10


def test_fused_queues_without_an_xcd_run_by_last_workgroup(torch_cuda):
    # 12 queues under the 16 nodes of depth 4: queues 8-11 have work that no
    # XCD serves
    err = _sweep_vs_oracle(torch_cuda, (40, 40, 24), 8, 61, fused_queues=12, fused_cut=4)
    assert err < 1e-10


def test_fused_deeper_cut_matches_oracle(torch_cuda):
    assert _sweep_vs_oracle(torch_cuda, (40, 40, 24), 8, 62, fused_cut=5) < 1e-10


def test_fused_fp32_sweep_close_to_fp64(torch_cuda):
    torch = torch_cuda
    disc, vel = _problem(40, 40, 24)
    X = rand_block(np.random.default_rng(63), disc.nx * disc.ny * disc.nz, 8)
    y64 = _sweep(torch, _helm(disc, vel, "fp64"), X)
    y32 = _sweep(torch, _helm(disc, vel, "fp32"), X)
    err = np.linalg.norm(y32 - y64, axis=0) / np.linalg.norm(y64, axis=0)
    assert err.max() < 1e-4


def test_fused_sweeps_repeat_bitwise(torch_cuda):
    """Dynamic tile order changes no sum: the graph-replayed sweep, a second
    replay and an eager sweep agree bit for bit."""
    torch = torch_cuda
    disc, vel = _problem(40, 40, 24)
    h = _helm(disc, vel)
    X = rand_block(np.random.default_rng(64), disc.nx * disc.ny * disc.nz, 8)
    a = _sweep(torch, h, X)
    b = _sweep(torch, h, X)
    h.set_sweep_graph(False)
    c = _sweep(torch, h, X)
    assert np.array_equal(a, b) and np.array_equal(a, c)


def test_fused_and_level_launches_agree(torch_cuda):
    torch = torch_cuda
    disc, vel = _problem(40, 40, 24)
    X = rand_block(np.random.default_rng(65), disc.nx * disc.ny * disc.nz, 8)
    a = _sweep(torch, _helm(disc, vel), X)
    b = _sweep(torch, _helm(disc, vel, fused=0), X)
    assert _rel(a, b) < 1e-12


def test_fused_solve_matches_oracle(torch_cuda):
    torch = torch_cuda
    from psp_amd import workloads
    disc, vel = _problem(40, 40, 40)
    B = workloads.gaussian_sources(40, 40, 40, workloads.interior_points(40, 40, 40, 8, cfg=66))
    xo, st, rep_o = _oracle(disc, vel).solve(B, 20, 1e-8)
    assert st == 0
    h = _helm(disc, vel)
    Bd = torch.from_numpy(B).cuda()
    rep = h.Solve(Bd, 20, 1e-8)
    assert rep.status == 0 and rep.iterations == rep_o.iterations
    assert _rel(Bd.cpu().numpy(), xo) < 1e-10
