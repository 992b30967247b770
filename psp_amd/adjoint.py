"""Adjoint-state gradient of a receiver misfit with respect to squared
slowness, as a torch.autograd.Function over the C ABI.

    L(m) = 1/2 sum_s sum_{r in receivers} |u_s(x_r) - d_s(x_r)|^2,
    A(m) u_s = f_s,   m = 1/c^2 (squared slowness).

A is complex symmetric (A^T = A, FormGlobalRow, initialize_finalize.hpp:
527-587), and m enters only through the centre term -omega^2 m s1 s2 s3
(initialize_finalize.hpp:567-568).  With r_s = P u_s - d_s:

    lambda_s = A^{-T} P^T conj(r_s) = A^{-1} P^T conj(r_s),
    dL/dm_i  = Re( omega^2 s1 s2 s3 (i) sum_s lambda_s(i) u_s(i) ).

The backward pass solves for every lambda_s as one right-hand-side block
with the factors the forward pass built; no refactorization.
"""
from __future__ import annotations

import torch

from .helmholtz import DistHelmholtz


class HelmholtzMisfit(torch.autograd.Function):
    """loss = HelmholtzMisfit.apply(m, sources, receivers, data, disc, opts)

    m         (N,) float64 CUDA tensor, squared slowness in natural order
    sources   (N, k) complex128 CUDA tensor, right-hand sides f_s
    receivers (R,) int64 CUDA tensor of natural grid indices
    data      (R, k) complex128 CUDA tensor, observed u_s(x_r)
    disc      psp_amd.Discretization
    opts      dict: m (GMRES restart, 20), tol (1e-10), maxIts (0 = no cap),
              precision ("fp64" or "fp32" fronts; GMRES is FP64 either way)

    The forward pass factors A(m) once; backward reuses those factors for the
    adjoint block.  Differentiable in m only.
    """

    @staticmethod
    def forward(ctx, m, sources, receivers, data, disc, opts=None):
        opts = dict(opts or {})
        restart = opts.get("m", 20)
        tol = opts.get("tol", 1e-10)
        max_its = opts.get("maxIts", 0)
        helm = DistHelmholtz(disc, precision=opts.get("precision", "fp64"))
        vel = (1.0 / torch.sqrt(m)).detach().cpu().numpy()
        helm.Initialize(vel)
        U = sources.detach().clone().contiguous()
        helm.Solve(U, restart, tol, maxIts=max_its)
        r = U[receivers] - data
        loss = 0.5 * (r.real.square().sum() + r.imag.square().sum())
        ctx.helm, ctx.restart, ctx.tol, ctx.max_its = helm, restart, tol, max_its
        ctx.save_for_backward(U, r, receivers)
        ctx.N = m.shape[0]
        return loss

    @staticmethod
    def backward(ctx, grad_loss):
        U, r, receivers = ctx.saved_tensors
        helm = ctx.helm
        Lam = torch.zeros_like(U)
        Lam[receivers] = torch.conj(r)
        helm.Solve(Lam, ctx.restart, ctx.tol, maxIts=ctx.max_its)
        grad = torch.zeros(ctx.N, dtype=torch.float64, device=U.device)
        helm.AdjointGradient(U, Lam, grad)
        ctx.helm = None
        return grad * grad_loss, None, None, None, None, None


def misfit_and_gradient(m, sources, receivers, data, disc, opts=None):
    """Convenience: loss and dL/dm for a float64 CUDA tensor m."""
    m = m.detach().clone().requires_grad_(True)
    loss = HelmholtzMisfit.apply(m, sources, receivers, data, disc, opts)
    loss.backward()
    return loss.detach(), m.grad.detach()

