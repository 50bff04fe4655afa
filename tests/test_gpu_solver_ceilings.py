"""The reference's solver iteration ceilings (docs/test/solver_1.py) on the GPU Schur path.

solver_1.py: Q1/dQ0 32 x 32 on the unit square, free-slip walls, the Gauss swarm (no voronoi
swarm), SolCx with eta_B = 1e5 and its defaults n_z = 3, x_c = 0.75, eta_A = 1 (viscosity
eta_A for x < x_c else eta_B, body force (0, cos(pi x) sin(n_z pi y)); AnalyticSolCx.hpp),
the default BSSCR solve (outer/scr rtol 1e-5, inner/A11 rtol 1e-6).  The reference raises when
the pressure solve takes more than 5 iterations, the velocity pre-solve or back-solve more than
8, or the inner solves of the pressure solve more than 30 in total.

The same case through the CPU restatement of the reference's solve (oracle/cpu_baseline.c)
takes 4 pressure iterations.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _walls(uw, mesh, vel):
    return uw.conditions.DirichletCondition(vel, (mesh.specialSets["MinI_VertexSet"] + mesh.specialSets["MaxI_VertexSet"],
                                                  mesh.specialSets["MinJ_VertexSet"] + mesh.specialSets["MaxJ_VertexSet"]))


def test_solver_1_iteration_ceilings():
    import underworld2_amd as uw
    print("solver_1.py: Q1/dQ0 32^2 SolCx eta_B 1e5, Gauss swarm, default tolerances", flush=True)
    mesh = uw.mesh.FeMesh_Cartesian("Q1/DQ0", (32, 32), (0.0, 0.0), (1.0, 1.0))
    vel = uw.mesh.MeshVariable(mesh, 2)
    vel.data[:] = (0.0, 0.0)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    p.data[:] = 0.0
    fn = uw.function
    x = fn.input()
    eta = fn.branching.conditional([(x[0] < 0.75, 1.0), (True, 1.0e5)])
    force = (0.0, fn.math.cos(np.pi * x[0]) * fn.math.sin(3.0 * np.pi * x[1]))
    stokes = uw.systems.Stokes(vel, p, eta, force, conditions=[_walls(uw, mesh, vel)])
    solver = uw.systems.Solver(stokes)
    solver.solve()
    stats = solver.get_stats()
    solver.print_stats()
    assert stats.pressure_its <= 5
    assert stats.velocity_presolve_its <= 8
    assert stats.velocity_pressuresolve_its <= 30
    assert stats.velocity_backsolve_its <= 8
