"""GPU adjoint-state gradient (psp_amd.adjoint.HelmholtzMisfit).

- the gradient kernel against -Re(dA/dm * sum_s lambda_s u_s), with the
  diagonal derivative read off the oracle's operator: 1e-13 relative;
- the whole gradient (forward solve, adjoint solve reusing the factors,
  kernel) against the FP64 CPU reference of tests/adjref.py (exact sparse
  LU solves): 1e-7 relative with FP64 fronts and GMRES to 1e-11;
- the same survey at 40^3 against the FP64 CPU gradient built on the
  oracle's GMRES solves: 1e-8 relative;
- the 80^3, 16-source survey: directional derivatives along 8 seeded unit
  directions against central finite differences of the GPU misfit, 1e-4
  relative, with FP64 fronts; FP32 fronts (FP64 GMRES) agree with FP64
  fronts to 1e-4.
"""
import numpy as np
import pytest

import adjref

pytestmark = pytest.mark.gpu


def _small(nx=14, ny=12, nz=24, k=3):
    from psp_amd import workloads
    from oracle.oracle import Oracle
    vel_true = workloads.bump_velocity(nx, ny, nz, cfg=11)
    vel = np.full(nx * ny * nz, float(np.median(vel_true)))
    disc = workloads.discretization(nx, ny, nz, float(vel_true.min()), 10.0)
    o = Oracle(disc.omega, nx, ny, nz, disc.wx, disc.wy, disc.wz, sigma=disc.sigmax)
    F, rec = workloads.surface_survey(nx, ny, nz, k, rec_stride=1)
    m, m_true = 1.0 / vel ** 2, 1.0 / vel_true ** 2
    data = adjref.spl.splu(adjref.natural_operator(o, vel_true)).solve(F)[rec]
    return disc, o, m, m_true, F, rec, data


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_gradient_kernel_formula(torch_cuda):
    torch = torch_cuda
    import psp_amd
    disc, o, m, _, _, _, _ = _small()
    h = psp_amd.DistHelmholtz(disc)
    h.Initialize(1.0 / np.sqrt(m))
    rng = np.random.default_rng(5)
    k = 4
    U = rng.standard_normal((o.N, k)) + 1j * rng.standard_normal((o.N, k))
    L = rng.standard_normal((o.N, k)) + 1j * rng.standard_normal((o.N, k))
    g = torch.zeros(o.N, dtype=torch.float64, device="cuda")
    h.AdjointGradient(torch.from_numpy(U).cuda(), torch.from_numpy(L).cuda(), g)
    torch.cuda.synchronize()
    ref = -(adjref.diag_derivative(o, m) * (L * U).sum(axis=1)).real
    assert _rel(g.cpu().numpy(), ref) < 1e-13
    # accumulate adds onto what is there
    h.AdjointGradient(torch.from_numpy(U).cuda(), torch.from_numpy(L).cuda(), g, accumulate=True)
    torch.cuda.synchronize()
    assert _rel(g.cpu().numpy(), 2 * ref) < 1e-13


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-7), ("fp32", 1e-7)])
def test_gradient_matches_cpu_reference(torch_cuda, precision, tol):
    torch = torch_cuda
    from psp_amd.adjoint import misfit_and_gradient
    disc, o, m, m_true, F, rec, data = _small()
    loss_ref, g_ref = adjref.gradient(o, m, F, rec, data)
    opts = {"m": 20, "tol": 1e-11, "maxIts": 400, "precision": precision}
    loss, g = misfit_and_gradient(torch.from_numpy(m).cuda(), torch.from_numpy(F).cuda(),
                                  torch.from_numpy(rec).cuda(), torch.from_numpy(data).cuda(),
                                  disc, opts)
    print(f"[adjoint small {precision}] loss {loss.item():.6e} ref {loss_ref:.6e} "
          f"grad rel err {_rel(g.cpu().numpy(), g_ref):.2e}", flush=True)
    assert abs(loss.item() - loss_ref) <= tol * loss_ref
    assert _rel(g.cpu().numpy(), g_ref) < tol


def _survey_data(torch, wl, tol=1e-11):
    """Receiver data of the true model, solved on the GPU in FP64."""
    import psp_amd
    ht = psp_amd.DistHelmholtz(wl.disc)
    ht.Initialize(1.0 / np.sqrt(wl.m_true))
    Ut = torch.from_numpy(wl.B).cuda()
    ht.Solve(Ut, 20, tol, maxIts=1000)
    rec = torch.from_numpy(wl.receivers).cuda()
    return Ut[rec].clone()


def test_gradient_40cube_matches_oracle(torch_cuda):
    """The 80^3 survey recipe at 40^3, 4 sources: GPU gradient against the
    FP64 CPU gradient of the oracle's own GMRES solves (both to 1e-12)."""
    torch = torch_cuda
    from psp_amd import workloads
    from psp_amd.adjoint import misfit_and_gradient
    from oracle.oracle import Oracle
    wl = workloads.config5(n=40, k=4)
    d = wl.disc
    data = _survey_data(torch, wl, 1e-12)
    o = Oracle(d.omega, d.nx, d.ny, d.nz, d.wx, d.wy, d.wz, sigma=d.sigmax)
    loss_ref, g_ref = adjref.gradient_oracle(o, wl.m, wl.B, wl.receivers, data.cpu().numpy())
    opts = {"m": 20, "tol": 1e-12, "maxIts": 800, "precision": "fp64"}
    loss, g = misfit_and_gradient(torch.from_numpy(wl.m).cuda(), torch.from_numpy(wl.B).cuda(),
                                  torch.from_numpy(wl.receivers).cuda(), data, d, opts)
    err = _rel(g.cpu().numpy(), g_ref)
    print(f"[adjoint 40^3] loss {loss.item():.9e} oracle {loss_ref:.9e} grad rel err {err:.2e}",
          flush=True)
    assert abs(loss.item() - loss_ref) <= 1e-8 * loss_ref
    assert err < 1e-8


def test_gradient_80cube_central_fd(torch_cuda):
    """80^3, 16 sources: central differences with a step of 1e-4 relative
    (||eps d|| = 1e-4 ||m||) along 8 seeded unit directions."""
    torch = torch_cuda
    from psp_amd import workloads
    from psp_amd.adjoint import HelmholtzMisfit, misfit_and_gradient
    wl = workloads.config5()
    disc = wl.disc
    F = torch.from_numpy(wl.B).cuda()
    rec = torch.from_numpy(wl.receivers).cuda()
    opts = {"m": 20, "tol": 1e-12, "maxIts": 1000, "precision": "fp64"}
    data = _survey_data(torch, wl, 1e-12)
    m = torch.from_numpy(wl.m).cuda()
    loss, g = misfit_and_gradient(m, F, rec, data, disc, opts)
    print(f"[adjoint 80^3] loss {loss.item():.6e} |g| {g.norm().item():.3e}", flush=True)

    def L(mv):
        with torch.no_grad():
            return HelmholtzMisfit.apply(mv, F, rec, data, disc, opts).item()

    eps = 1e-4 * torch.linalg.norm(m).item()
    worst = 0.0
    for j in range(8):
        r = np.random.default_rng([workloads.SEED0 + 5, 3, j])
        d = r.standard_normal(m.shape[0])
        d = torch.from_numpy(d / np.linalg.norm(d)).cuda()
        fd = (L(m + eps * d) - L(m - eps * d)) / (2 * eps)
        an = torch.dot(g, d).item()
        rel = abs(fd - an) / abs(fd)
        worst = max(worst, rel)
        print(f"[adjoint 80^3] direction {j}: fd {fd:.9e} adjoint {an:.9e} rel {rel:.2e}",
              flush=True)
    assert worst <= 1e-4

    opts32 = dict(opts, precision="fp32")
    _, g32 = misfit_and_gradient(m, F, rec, data, disc, opts32)
    err = (torch.linalg.norm(g32 - g) / torch.linalg.norm(g)).item()
    print(f"[adjoint 80^3] fp32 fronts vs fp64 fronts gradient rel err {err:.2e}", flush=True)
    assert err < 1e-4
