"""Split row blocks summed inside the product launch (PSP_HIP_OPT_SOLVE_XCD_REDUCE):
one-workgroup-per-tile levels place every slice of a row block on one XCD,
and the slice that finishes last sums the block's slots in slot order.  The
sums are those of the reduce launch, so sweeps with the option on and off
are bitwise equal; both are checked against the oracle (FP64 fronts, 1e-10
relative).  Shapes with up to 16 slices per row block at every level,
8 <= k < 32 (the matrix-core path that sums in the launch), and k < 8 and
k >= 32 (paths that keep the reduce launch and skip the placement's holes)."""
import numpy as np
import pytest

from conftest import rand_block
from test_gpu_parity import _oracle, _problem, _rel

pytestmark = pytest.mark.gpu


def _helm(disc, vel, precision="fp64", **tiling):
    import psp_amd
    h = psp_amd.DistHelmholtz(disc, precision=precision)
    h.set_solve_tiling(**{"persist": 0, "min_tiles": 4096, **tiling})
    h.Initialize(vel)
    return h


def _sweep(torch, h, X):
    xd = torch.from_numpy(X).cuda()
    h.Precondition(xd)
    torch.cuda.synchronize()
    return xd.cpu().numpy()


@pytest.mark.parametrize("k", [8, 12, 16, 24])
@pytest.mark.parametrize("split", [8, 16])
def test_in_launch_sums_equal_reduce_launch_and_oracle(torch_cuda, k, split):
    torch = torch_cuda
    disc, vel = _problem(24, 22, 30)
    o = _oracle(disc, vel)
    X = rand_block(np.random.default_rng(80 + k + split), o.N, k)
    on = _sweep(torch, _helm(disc, vel, max_split=split, xcd_reduce=1), X)
    off = _sweep(torch, _helm(disc, vel, max_split=split, xcd_reduce=0), X)
    assert np.array_equal(on, off)
    l2n = o.local_to_natural()
    assert _rel(on[l2n], o.precondition(X[l2n])) < 1e-10


@pytest.mark.parametrize("k", [3, 40])
def test_placed_levels_on_the_other_paths(torch_cuda, k):
    # k < 8 (vector path) and k >= 32 (tile engine) keep the reduce launch
    # and must skip the placement's holes
    torch = torch_cuda
    disc, vel = _problem(24, 22, 30)
    o = _oracle(disc, vel)
    X = rand_block(np.random.default_rng(90 + k), o.N, k)
    got = _sweep(torch, _helm(disc, vel, max_split=8), X)
    l2n = o.local_to_natural()
    assert _rel(got[l2n], o.precondition(X[l2n])) < 1e-10


def test_fp32_sweeps_repeat_bitwise(torch_cuda):
    """The slot order fixes every sum: graph replays and an eager sweep agree
    bit for bit, with FP32 fronts and 16-wave and 4-wave workgroups."""
    torch = torch_cuda
    disc, vel = _problem(40, 40, 24)
    X = rand_block(np.random.default_rng(95), disc.nx * disc.ny * disc.nz, 8)
    for wide in (0, 1):
        h = _helm(disc, vel, "fp32", max_split=16, wide_wg=wide)
        a = _sweep(torch, h, X)
        b = _sweep(torch, h, X)
        h.set_sweep_graph(False)
        c = _sweep(torch, h, X)
        assert np.array_equal(a, b) and np.array_equal(a, c)
        ref = _sweep(torch, _helm(disc, vel, "fp32", max_split=16, wide_wg=wide, xcd_reduce=0), X)
        assert np.array_equal(a, ref)


def test_autotuned_solve_matches_oracle(torch_cuda):
    torch = torch_cuda
    import psp_amd
    from psp_amd import workloads
    disc, vel = _problem(40, 40, 40)
    B = workloads.gaussian_sources(40, 40, 40, workloads.interior_points(40, 40, 40, 8, cfg=96))
    xo, st, rep_o = _oracle(disc, vel).solve(B, 20, 1e-8)
    assert st == 0
    h = psp_amd.DistHelmholtz(disc)
    h.Initialize(vel)
    h.Autotune(8)
    Bd = torch.from_numpy(B).cuda()
    rep = h.Solve(Bd, 20, 1e-8)
    assert rep.status == 0 and rep.iterations == rep_o.iterations
    assert _rel(Bd.cpu().numpy(), xo) < 1e-10
