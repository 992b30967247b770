#!/usr/bin/env python3
"""Writes tests/golden/config2_oracle_solution_sample.npz: the oracle's
converged FP64 solution of the 100^3 bump workload with its 8 sources
(psp_amd.workloads.config2, GMRES(20) to 1e-5), kept as a sample small
enough to commit:

- iterations: the oracle's GMRES iteration count;
- col_norms: the 2-norm of each of the 8 solution columns;
- points, values: the solution at 4096 seeded grid points (natural order).

The oracle is pinned by tests/test_oracle.py and tests/test_oracle_pml.py.
Takes about 5 minutes on 8 cores.  Run from the repository root:
    python tests/golden/make_config2_sample.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    from oracle.oracle import Oracle
    from psp_amd import workloads
    wl = workloads.config2()
    d = wl.disc
    o = Oracle(d.omega, d.nx, d.ny, d.nz, d.wx, d.wy, d.wz, sigma=d.sigmax)
    t0 = time.time()
    o.factor(wl.velocity)
    x, st, rep = o.solve(wl.B, 20, 1e-5)
    assert st == 0, st
    rng = np.random.default_rng(20261202)
    pts = np.sort(rng.choice(x.shape[0], 4096, replace=False)).astype(np.int64)
    np.savez_compressed(os.path.join(HERE, "config2_oracle_solution_sample.npz"),
                        iterations=rep.iterations, final_max_rel=rep.final_max_rel,
                        col_norms=np.linalg.norm(x, axis=0), points=pts, values=x[pts])
    print(f"wrote config2_oracle_solution_sample.npz: {rep.iterations} its, rel "
          f"{rep.final_max_rel:.3e}, {time.time() - t0:.0f}s")


if __name__ == "__main__":
    main()
