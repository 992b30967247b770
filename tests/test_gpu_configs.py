"""GPU parity at the benchmark configurations, through the C ABI, against the
CPU oracle on the same seeded inputs:

  1. unit cube 40^3, c = 1, one source: the FP64-front solve converges in
     the oracle's iteration count to its solution within 1e-10;
  2. bumps 100^3, 8 sources: the FP64 sweep within 1e-10 of the oracle's,
     three GMRES(20) steps within 1e-10 (history 1e-9), and the converged
     FP32-front and FP64-front solutions against the oracle's converged
     solution (a committed sample of it, tests/golden/), FP32 error reported;
  3. thrust belt (reduced to 64 x 64 x 32, 4 sources, the velocity-scaled
     PML of the full model): converged FP64 solve within 1e-10 of the
     oracle's with the same iteration count, FP32 reaching the tolerance;
  4. the 32 uneven problems solved together: every problem's true residual
     (recomputed with the operator) below 1e-5, and two of them against the
     oracle's solutions;
  and the C++ drop-in (psp_amd/examples/unit_cube, the reference's UnitCube
  driver on psp::DistHelmholtz) run as a child process, against the Python
  path and the oracle.
"""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _oracle(disc, vel):
    from oracle.oracle import Oracle
    o = Oracle(disc.omega, disc.nx, disc.ny, disc.nz, disc.wx, disc.wy, disc.wz,
               sigma=disc.sigmax)
    o.factor(vel)
    return o


def _helm(disc, vel, precision="fp64"):
    import psp_amd
    h = psp_amd.DistHelmholtz(disc, precision=precision)
    h.Initialize(vel)
    return h


def _solve_gpu(torch, h, B, tol=1e-5):
    Bd = torch.from_numpy(B).cuda()
    rep = h.Solve(Bd, 20, tol)
    torch.cuda.synchronize()
    return Bd.cpu().numpy(), rep


# ------------------------------------------------------------------ config 1 --
def test_config1_unit_cube_matches_oracle(torch_cuda):
    from psp_amd import workloads
    wl = workloads.config1()
    o = _oracle(wl.disc, wl.velocity)
    xo, st, rep_o = o.solve(wl.B, 20, 1e-5)
    assert st == 0
    x, rep = _solve_gpu(torch_cuda, _helm(wl.disc, wl.velocity), wl.B)
    assert rep.status == 0 and rep.final_max_rel < 1e-5
    assert rep.iterations == rep_o.iterations
    assert _rel(x, xo) < 1e-10
    x32, rep32 = _solve_gpu(torch_cuda, _helm(wl.disc, wl.velocity, "fp32"), wl.B)
    err = _rel(x32, xo)
    print(f"config 1: {rep.iterations} its; FP32 fronts {rep32.iterations} its, "
          f"|u32 - u_oracle| / |u_oracle| = {err:.2e}")
    assert rep32.status == 0 and err < 1e-5


# ------------------------------------------------------------------ config 2 --
@pytest.fixture(scope="module")
def config2(torch_cuda):
    from psp_amd import workloads
    wl = workloads.config2()
    o = _oracle(wl.disc, wl.velocity)
    return wl, o


def test_config2_sweep_fp64_matches_oracle(torch_cuda, config2):
    torch = torch_cuda
    wl, o = config2
    h = _helm(wl.disc, wl.velocity)
    xd = torch.from_numpy(wl.B).cuda()
    h.Precondition(xd)
    torch.cuda.synchronize()
    l2n = o.local_to_natural()
    got = xd.cpu().numpy()[l2n]
    ref = o.precondition(wl.B[l2n])
    assert _rel(got, ref) < 1e-10


def test_config2_three_gmres_steps_match_oracle(torch_cuda, config2):
    wl, o = config2
    h = _helm(wl.disc, wl.velocity)
    xo, ho = o.solve_fixed(wl.B, 20, 3)
    Bg = wl.B.copy()
    hg = h.SolveFixed(Bg, 20, 3)
    assert np.allclose(hg, ho, rtol=1e-9, atol=0)
    assert _rel(Bg, xo) < 1e-10


def test_config2_converged_solutions_against_oracle(torch_cuda):
    """The oracle's converged solution of the whole 8-source block is a
    committed sample (tests/golden/make_config2_sample.py): its iteration
    count, its norm per source and its values at seeded points."""
    from psp_amd import workloads
    g = np.load(os.path.join(GOLDEN, "config2_oracle_solution_sample.npz"))
    wl = workloads.config2()
    idx = g["points"]
    ref = g["values"]
    out = {}
    for prec in ("fp64", "fp32"):
        x, rep = _solve_gpu(torch_cuda, _helm(wl.disc, wl.velocity, prec), wl.B)
        assert rep.status == 0 and rep.final_max_rel < 1e-5
        out[prec] = (rep.iterations, _rel(x[idx], ref),
                     np.abs(np.linalg.norm(x, axis=0) / g["col_norms"] - 1).max())
    its64, err64, nrm64 = out["fp64"]
    its32, err32, nrm32 = out["fp32"]
    print(f"config 2: oracle {int(g['iterations'])} its; FP64 fronts {its64} its, sample error "
          f"{err64:.2e}; FP32 fronts {its32} its, |u32 - u_oracle| / |u_oracle| = {err32:.2e} "
          f"(column norms {nrm32:.1e})")
    assert its64 == int(g["iterations"])
    assert err64 < 1e-10 and nrm64 < 1e-10
    assert err32 < 1e-6 and nrm32 < 1e-6


# ------------------------------------------------------------------ config 3 --
def test_config3_reduced_thrust_belt_matches_oracle(torch_cuda):
    from psp_amd import workloads
    wl = workloads.config3(64, 64, 32, k=4)
    o = _oracle(wl.disc, wl.velocity)
    xo, st, rep_o = o.solve(wl.B, 20, 1e-5)
    assert st == 0
    x, rep = _solve_gpu(torch_cuda, _helm(wl.disc, wl.velocity), wl.B)
    assert rep.status == 0 and rep.iterations == rep_o.iterations
    assert _rel(x, xo) < 1e-10
    x32, rep32 = _solve_gpu(torch_cuda, _helm(wl.disc, wl.velocity, "fp32"), wl.B)
    print(f"config 3 (64x64x32): {rep.iterations} its; FP32 fronts {rep32.iterations} its, "
          f"error {_rel(x32, xo):.2e}")
    assert rep32.status == 0 and _rel(x32, xo) < 1e-5


# ------------------------------------------------------------------ config 4 --
def test_config4_batch_residuals_and_oracle_samples(torch_cuda):
    torch = torch_cuda
    import psp_amd
    from psp_amd import workloads
    wls = workloads.config4()
    helms = [_helm(w.disc, w.velocity, "fp32") for w in wls]
    bufs = [torch.from_numpy(w.B).cuda() for w in wls]
    reps = psp_amd.SolveBatch(helms, bufs, 20, 1e-5)
    torch.cuda.synchronize()
    worst = 0.0
    for h, w, x, r in zip(helms, wls, bufs, reps):
        assert r.status == 0
        ax = torch.empty_like(x)
        h.Multiply(x, ax)
        b = torch.from_numpy(w.B).cuda()
        res = (torch.linalg.norm(b - ax, dim=0) / torch.linalg.norm(b, dim=0)).max().item()
        assert res < 1e-5, (w.name, res)
        worst = max(worst, res)
    print(f"config 4: 32 problems, worst true residual {worst:.2e}")
    # the two smallest problems against the oracle
    order = sorted(range(len(wls)), key=lambda i: (wls[i].disc.nx, wls[i].B.shape[1]))[:2]
    for i in order:
        w = wls[i]
        xo, st, _ = _oracle(w.disc, w.velocity).solve(w.B, 20, 1e-5)
        assert st == 0
        x64, _ = _solve_gpu(torch, _helm(w.disc, w.velocity), w.B)
        assert _rel(x64, xo) < 1e-10, w.name
        assert _rel(bufs[i].cpu().numpy(), xo) < 1e-5, w.name
    del helms, bufs


# --------------------------------------------------------------- C++ drop-in --
def test_cpp_unit_cube_driver_matches_python_path_and_oracle(torch_cuda, tmp_path):
    """psp_amd/examples/unit_cube (the reference's tests/UnitCube.cpp
    structure on psp::DistHelmholtz: Initialize, Finalize, Initialize, Solve
    of 4 right-hand sides) in a child process; its solution equals the
    Python path's on the inputs it wrote, with the same iteration count, and
    the oracle's within 1e-10."""
    from psp_amd.helmholtz import PML, Discretization
    exe = os.path.join(ROOT, "psp_amd", "examples", "unit_cube")
    assert os.path.exists(exe), "build the example first (make -C psp_amd/csrc)"
    n, omega = 24, 20.0
    prefix = str(tmp_path / "uc")
    out = subprocess.run([exe, str(n), str(omega), "2", "0", prefix], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    its = int(re.search(r"(\d+) GMRES iterations", out.stdout).group(1))
    vel = np.fromfile(prefix + "_velocity.bin", dtype=np.float64)
    B = np.fromfile(prefix + "_rhs.bin", dtype=np.complex128).reshape(n ** 3, 4)
    X = np.fromfile(prefix + "_solution.bin", dtype=np.complex128).reshape(n ** 3, 4)
    disc = Discretization.with_pml(omega, n, n, n, 1.0, 1.0, 1.0, PML, PML, PML, PML, PML, 5,
                                   1.5)
    x, rep = _solve_gpu(torch_cuda, _helm(disc, vel), B)
    assert rep.iterations == its
    assert _rel(X, x) < 1e-12
    xo, st, rep_o = _oracle(disc, vel).solve(B, 20, 1e-5)
    assert st == 0 and rep_o.iterations == its
    assert _rel(X, xo) < 1e-10
