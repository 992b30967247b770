"""Pins the FP64 CPU reference gradient (tests/adjref.py) that the GPU
adjoint is checked against: central finite differences with exact sparse
solves, the diagonal derivative against the independent restatement, and a
planted mistake that the finite-difference check must catch."""
import numpy as np
import pytest

import adjref
import npref
from test_oracle import make

NX, NY, NZ, K = 12, 11, 20, 3


def problem(seed=1):
    o, d = make(NX, NY, NZ)
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    vel = (1.0 + 0.2 * np.sin(0.4 * x) * np.cos(0.3 * y + 0.2 * z)).ravel()
    m = 1.0 / vel ** 2
    F = np.zeros((o.N, K), np.complex128)
    for s in range(K):
        F[rng.integers(o.N), s] = 1.0
    F += 0.1 * (rng.standard_normal(F.shape) + 1j * rng.standard_normal(F.shape))
    rec = rng.choice(o.N, 40, replace=False)
    m_true = m * (1 + 0.1 * np.exp(-((x - 6.0) ** 2 + (y - 5.0) ** 2 + (z - 10.0) ** 2) / 8.0).ravel())
    A = adjref.natural_operator(o, 1.0 / np.sqrt(m_true))
    data = adjref.spl.splu(A).solve(F)[rec]
    return o, d, m, F, rec, data, rng


def test_diag_derivative_matches_restatement():
    o, d, m, *_ = problem()
    dd = adjref.diag_derivative(o, m)
    A0 = npref.global_operator(d, 1.0 / np.sqrt(m)).diagonal()
    A1 = npref.global_operator(d, 1.0 / np.sqrt(m + 1.0)).diagonal()
    assert np.allclose(dd, A1 - A0, rtol=1e-9, atol=1e-9 * abs(dd).max())
    # -omega^2 s1 s2 s3 is -omega^2 exactly away from every PML
    z, y, x = np.unravel_index(np.arange(o.N), (NZ, NY, NX))
    inner = (x >= 5) & (x < NX - 5) & (y >= 5) & (y < NY - 5) & (z >= 5) & (z < NZ - 5)
    assert inner.any()
    assert np.allclose(dd[inner], -d.omega ** 2, rtol=1e-12)


@pytest.mark.parametrize("dseed", [2, 3])
def test_reference_gradient_matches_central_fd(dseed):
    o, d, m, F, rec, data, _ = problem()
    loss, g = adjref.gradient(o, m, F, rec, data)
    rng = np.random.default_rng(dseed)
    delta = m * rng.standard_normal(m.shape)
    eps = 1e-5
    fd = (adjref.misfit(o, m + eps * delta, F, rec, data)
          - adjref.misfit(o, m - eps * delta, F, rec, data)) / (2 * eps)
    assert abs(fd - g @ delta) <= 1e-7 * abs(fd)


def test_fd_check_catches_missing_conjugate():
    o, d, m, F, rec, data, rng = problem()
    _, g_bad = adjref.gradient(o, m, F, rec, data, conj_residual=False)
    delta = m * np.random.default_rng(2).standard_normal(m.shape)
    eps = 1e-5
    fd = (adjref.misfit(o, m + eps * delta, F, rec, data)
          - adjref.misfit(o, m - eps * delta, F, rec, data)) / (2 * eps)
    assert abs(fd - g_bad @ delta) > 1e-2 * abs(fd)
