#!/usr/bin/env python3
"""Writes tests/golden/small_problem.npz from the independent numpy/scipy
restatement (tests/npref.py), never from the oracle or the product:

- the global operator of an 8x7x26 grid (natural order, COO);
- Precondition applied to two fixed right-hand sides (exact SuperLU panel
  solves inside the sweep of solve.hpp:681-734);
- the exact solution A^{-1} b by scipy's sparse direct solver.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.sparse.linalg as spl

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import npref  # noqa: E402


def main():
    nx, ny, nz = 8, 7, 26
    h = 1.0 / (max(nx, ny, nz) + 1)
    omega = 2 * np.pi * max(nx, ny, nz) / 10.0
    d = npref.Disc(omega, nx, ny, nz, (nx + 1) * h, (ny + 1) * h, (nz + 1) * h)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    vel = (1.2 + 0.25 * np.cos(0.4 * x) * np.sin(0.3 * y + 0.2 * z)).ravel()
    A = npref.global_operator(d, vel).tocoo()
    rng = np.random.default_rng(20261202)
    N = nx * ny * nz
    rhs = rng.standard_normal((N, 2)) + 1j * rng.standard_normal((N, 2))
    sw = npref.Sweep(d, vel)
    pre = sw.precondition(rhs)
    sol = spl.spsolve(A.tocsc(), rhs)
    np.savez_compressed(os.path.join(HERE, "small_problem.npz"), grid=np.array([nx, ny, nz]),
                        omega=omega, velocity=vel, A_rows=A.row, A_cols=A.col, A_vals=A.data,
                        rhs=rhs, precond_rhs=pre, solution=sol)
    print("wrote small_problem.npz", N, "unknowns")


if __name__ == "__main__":
    main()
