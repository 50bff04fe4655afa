"""PerCellSpaceFillerLayout on the GPU and the reference's own end-to-end Q1/dQ0 check.

* The Sobol layout (MeshParticleLayout.c:194-229, filltype 0) is bit-exact against the oracle
  (2D 64^2 x 20 on [-1,1] x [0,1], 3D 8^3 x 30) and equals the reference docstring
  (PY/swarm/layouts.py:160-172).
* docs/examples/04_StokesSinker.ipynb: 64^2 Q1/dQ0 on [-1,1] x [0,1], free-slip walls, 20
  particles per cell by PerCellSpaceFillerLayout, a sphere of radius 0.1 at (0, 0.7) with
  viscosity 10 and density 10 in a background of viscosity 1 and density 1, body force
  -density (0, 1), the swarm as the voronoi integration swarm, the default solver, and ten
  steps of solve + RK2 advection at dt = get_max_dt().  The notebook raises unless the v_rms of
  the tenth solve is within rtol 1e-3 of 1.0135e-02 (04_StokesSinker.ipynb, testing cell).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _swarm(uw, res, lo, hi, ppc):
    mesh = uw.mesh.FeMesh_Cartesian("Q1/dQ0", res, lo, hi)
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellSpaceFillerLayout(swarm, particlesPerCell=ppc))
    return mesh, swarm


def test_sobol_layout_bitexact():
    import underworld2_amd as uw
    from oracle import oracle as O
    print("PerCellSpaceFillerLayout: docstring, 2D 64^2 x 20, 3D 8^3 x 30", flush=True)
    _, sw = _swarm(uw, (1, 1), (0.0, 0.0), (1.0, 1.0), 4)
    np.testing.assert_array_equal(sw.particleCoordinates.data,
                                  [[0.5, 0.5], [0.25, 0.75], [0.75, 0.25], [0.375, 0.625]])
    for res, lo, hi, ppc in (((64, 64), (-1.0, 0.0), (1.0, 1.0), 20),
                             ((8, 8, 8), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 30)):
        _, sw = _swarm(uw, res, lo, hi, ppc)
        ref = O.layout_sobol(O.Mesh(res, lo, hi), ppc)
        assert np.array_equal(sw.particleCoordinates.data, ref), res
        assert np.array_equal(sw.owningCell.data[:, 0], np.repeat(np.arange(int(np.prod(res))), ppc))


def test_stokes_sinker_notebook_vrms():
    import underworld2_amd as uw
    fn = uw.function
    print("04_StokesSinker.ipynb: 64^2, Sobol swarm 20 ppc, 10 solve + advect steps", flush=True)
    mesh = uw.mesh.FeMesh_Cartesian(elementType="Q1/dQ0", elementRes=(64, 64), minCoord=(-1.0, 0.0),
                                    maxCoord=(1.0, 1.0))
    velocityField = mesh.add_variable(nodeDofCount=2)
    pressureField = mesh.subMesh.add_variable(nodeDofCount=1)
    velocityField.data[:] = [0.0, 0.0]
    pressureField.data[:] = 0.0
    iWalls = mesh.specialSets["MinI_VertexSet"] + mesh.specialSets["MaxI_VertexSet"]
    jWalls = mesh.specialSets["MinJ_VertexSet"] + mesh.specialSets["MaxJ_VertexSet"]
    freeslipBC = uw.conditions.DirichletCondition(variable=velocityField, indexSetsPerDof=(iWalls, jWalls))
    swarm = uw.swarm.Swarm(mesh=mesh)
    advector = uw.systems.SwarmAdvector(swarm=swarm, velocityField=velocityField, order=2)
    materialIndex = swarm.add_variable(dataType="int", count=1)
    swarm.populate_using_layout(layout=uw.swarm.layouts.PerCellSpaceFillerLayout(swarm=swarm, particlesPerCell=20))
    coord = fn.input() - (0.0, 0.7)
    materialIndex.data[:] = fn.branching.conditional([(fn.math.dot(coord, coord) < 0.1 ** 2, 1),
                                                      (True, 0)]).evaluate(swarm)
    viscosityMapFn = fn.branching.map(fn_key=materialIndex, mapping={0: 1.0, 1: 10.0})
    densityFn = fn.branching.map(fn_key=materialIndex, mapping={0: 1.0, 1: 10.0})
    buoyancyFn = -densityFn * (0.0, 1.0)
    stokes = uw.systems.Stokes(velocityField=velocityField, pressureField=pressureField, voronoi_swarm=swarm,
                               conditions=freeslipBC, fn_viscosity=viscosityMapFn, fn_bodyforce=buoyancyFn)
    solver = uw.systems.Solver(stokes)
    vrms = None
    for step in range(10):
        solver.solve()
        vrms = stokes.velocity_rms()
        s = solver.get_stats()
        print(f"step {step}: v_rms {vrms:.6e}, pressure its {s.pressure_its}", flush=True)
        advector.integrate(advector.get_max_dt())
    assert np.isclose(vrms, 1.0135e-02, rtol=1e-3), vrms
