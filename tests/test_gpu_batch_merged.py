"""Batched solves with merged sweeps (psp_hip_solve_batch's default): the
GMRES steps of all problems in lockstep, one launch per sweep step for every
unfinished problem.  Each problem must get exactly the iterations and, for
8 <= k < 32 (the matrix-core kernel its own sweep runs), exactly the bits of
its one-by-one Solve; k < 8 and k >= 32 sum in another order and are held to
the oracle (FP64 fronts, 1e-10 relative).  Problems of uneven sizes, tile
shapes (default, 16-wave, autotuned), both front precisions, and problems
that stop early (iteration cap) or late."""
import numpy as np
import pytest

from test_gpu_parity import _helm, _oracle, _problem, _rel

pytestmark = pytest.mark.gpu


def _sources(nx, ny, nz, k, seed):
    from psp_amd import workloads
    return workloads.gaussian_sources(nx, ny, nz,
                                      workloads.interior_points(nx, ny, nz, k, cfg=seed))


def _alone(torch, h, B, tol, maxIts=0):
    x = torch.from_numpy(B).cuda()
    r = h.Solve(x, 20, tol, maxIts=maxIts, raise_on_error=False)
    torch.cuda.synchronize()
    return x.cpu(), r


def _batch(torch, helms, Bs, tol, maxIts=0, threads=False):
    import psp_amd
    xs = [torch.from_numpy(B).cuda() for B in Bs]
    reps = psp_amd.SolveBatch(helms, xs, 20, tol, maxIts=maxIts, raise_on_error=False,
                              threads=threads)
    torch.cuda.synchronize()
    return [x.cpu() for x in xs], reps


PROBS = [(14, 12, 24, 8), (16, 16, 20, 12), (12, 10, 26, 16), (18, 14, 22, 9), (10, 12, 28, 24)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_merged_batch_equals_one_by_one_solves(torch_cuda, prec):
    torch = torch_cuda
    helms, Bs = [], []
    for j, (nx, ny, nz, k) in enumerate(PROBS):
        disc, vel = _problem(nx, ny, nz, cfg_seed=150 + j)
        h = _helm(disc, vel, prec)
        if j == 1:
            h.set_solve_tiling(wide_wg=1)  # 16-wave tiles on the levels with few row blocks
        if j == 3:
            h.Autotune(k)                  # tuned shapes (persistent tiles among them)
        helms.append(h)
        Bs.append(_sources(nx, ny, nz, k, 160 + j))
    ref = [_alone(torch, h, B, 1e-9) for h, B in zip(helms, Bs)]
    xs, reps = _batch(torch, helms, Bs, 1e-9)
    for j, (x, r) in enumerate(zip(xs, reps)):
        xr, rr = ref[j]
        assert rr.status == 0 and r.status == 0, j
        assert r.iterations == rr.iterations and r.restarts == rr.restarts, j
        assert r.final_max_rel == rr.final_max_rel, j
        assert torch.equal(x, xr), j


def test_merged_batch_other_kernels_match_oracle(torch_cuda):
    """k = 3 (vector kernel alone) and k = 40 (wide-block kernel alone) run
    the matrix-core tile kernel in the merged sweep."""
    torch = torch_cuda
    probs = [(12, 10, 26, 3), (14, 12, 24, 40), (16, 14, 22, 8)]
    helms, Bs, oracles = [], [], []
    for j, (nx, ny, nz, k) in enumerate(probs):
        disc, vel = _problem(nx, ny, nz, cfg_seed=170 + j)
        helms.append(_helm(disc, vel))
        Bs.append(_sources(nx, ny, nz, k, 180 + j))
        oracles.append(_oracle(disc, vel))
    xs, reps = _batch(torch, helms, Bs, 1e-8)
    for j, (x, r) in enumerate(zip(xs, reps)):
        xo, st, ro = oracles[j].solve(Bs[j], 20, 1e-8)
        assert st == 0 and r.status == 0, j
        assert r.iterations == ro.iterations, j
        assert _rel(x.numpy(), xo) < 1e-10, j


def test_merged_batch_mixed_precisions_and_iteration_cap(torch_cuda):
    """FP32 and FP64 problems in one call (two merged batches); with an
    iteration cap every problem stops where its own Solve stops, with the
    same status and iterate."""
    torch = torch_cuda
    helms, Bs = [], []
    for j, (nx, ny, nz, k) in enumerate(PROBS[:4]):
        disc, vel = _problem(nx, ny, nz, cfg_seed=190 + j)
        helms.append(_helm(disc, vel, "fp32" if j % 2 else "fp64"))
        Bs.append(_sources(nx, ny, nz, k, 200 + j))
    for tol, cap in ((1e-12, 3), (1e-7, 0)):
        ref = [_alone(torch, h, B, tol, cap) for h, B in zip(helms, Bs)]
        xs, reps = _batch(torch, helms, Bs, tol, cap)
        for j, (x, r) in enumerate(zip(xs, reps)):
            xr, rr = ref[j]
            assert r.status == rr.status and r.iterations == rr.iterations, (tol, j)
            assert torch.equal(x, xr), (tol, j)


def test_merged_and_threaded_batches_agree(torch_cuda):
    torch = torch_cuda
    helms, Bs = [], []
    for j, (nx, ny, nz, k) in enumerate(PROBS):
        disc, vel = _problem(nx, ny, nz, cfg_seed=210 + j)
        helms.append(_helm(disc, vel, "fp32"))
        Bs.append(_sources(nx, ny, nz, k, 220 + j))
    xm, rm = _batch(torch, helms, Bs, 1e-6)
    xt, rt = _batch(torch, helms, Bs, 1e-6, threads=True)
    for j in range(len(PROBS)):
        assert rm[j].status == 0 and rt[j].status == 0
        assert rm[j].iterations == rt[j].iterations, j
        assert torch.equal(xm[j], xt[j]), j


def test_merged_batch_host_blocks_and_single_problem(torch_cuda):
    """Host (numpy) right-hand-side blocks, and a batch of one problem, give
    the bits of one-by-one solves."""
    torch = torch_cuda
    import psp_amd
    helms, Bs = [], []
    for j, (nx, ny, nz, k) in enumerate(PROBS[:3]):
        disc, vel = _problem(nx, ny, nz, cfg_seed=230 + j)
        helms.append(_helm(disc, vel, "fp32"))
        Bs.append(_sources(nx, ny, nz, k, 235 + j))
    ref = [_alone(torch, h, B, 1e-8) for h, B in zip(helms, Bs)]
    host = [B.copy() for B in Bs]
    reps = psp_amd.SolveBatch(helms, host, 20, 1e-8)
    for j, r in enumerate(reps):
        assert r.status == 0 and r.iterations == ref[j][1].iterations, j
        assert np.array_equal(host[j], ref[j][0].numpy()), j
    xs, reps1 = _batch(torch, helms[:1], Bs[:1], 1e-8)
    assert reps1[0].iterations == ref[0][1].iterations
    assert torch.equal(xs[0], ref[0][0])
