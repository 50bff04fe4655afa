"""Write the oracle's converged discrete solutions of problems too large for the oracle to solve
inside a test run (test infrastructure; run here, on the CPU, with the oracle built).

sinker32_oracle.npz -- the 32^3 Q1/dQ0 sinker: unit cube, free-slip walls, PerCellRandomLayout
    30 particles per cell (seed 0), sphere of radius 0.1 at (0.5, 0.5, 0.6) with viscosity 1e3
    and density 1 in a background of viscosity 1 and density 0, gravity (0, 0, -1), DVC
    integration weights.  Pipeline: oracle.full_pipeline (layout, material, coincident map,
    DVC, particle integration, assembly).  Solve: the oracle's restatement of the reference's
    Schur-complement solve (oracle/cpu_baseline.c through System.bsscr_solve) with outer rtol
    1e-12 and inner rtol 1e-13, capped at 5,000 outer iterations (the cap is reached; the stored
    solution's relative residual in the oracle's assembled system is 1.5e-12, momentum, and
    2.7e-13, continuity); a sparse direct factorisation of this 3D saddle system does not finish
    in useful time on the CPU (about 85 minutes on 8 threads for this run).
    Stored: velocity [N][3], mean-free pressure [E], the relative residuals of the stored
    solution in the oracle's assembled system (|f - K u - G p| / |f|, |G^T u| / |f|), the
    particle count and (sum, sum of squares) of the particle coordinates, which the test
    recomputes from the GPU swarm.

Usage: python tests/golden/make_oracle_solutions.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    from sinker import build_oracle
    t = time.time()
    o = build_oracle((32, 32, 32), 30, 0, (0.5, 0.5, 0.6), 0.1, eta=(1.0, 1000.0), solve=False)
    print(f"sinker32: pipeline {time.time() - t:.1f} s", flush=True)
    sysm = o["system"]
    u, p, st = sysm.bsscr_solve(outer_rtol=1e-12, inner_rtol=1e-13, max_outer=5000)
    p = p - p.mean()
    c = o["coords"]
    r1, r2 = sysm.residuals(u, p)
    fn = np.linalg.norm(sysm.f)
    print(f"sinker32: {time.time() - t:.1f} s, {st}, relative residuals {r1 / fn:.2e} {r2 / fn:.2e}", flush=True)
    np.savez_compressed(os.path.join(HERE, "sinker32_oracle.npz"), u=u, p=p, residuals=np.array([r1 / fn, r2 / fn]),
                        nparticles=np.int64(c.shape[0]), coord_sums=np.array([c.sum(), (c * c).sum()]))


if __name__ == "__main__":
    main()
