"""Device evaluation of the viscosity / body-force function graph (fn_kernels.hip) against the
host evaluator of the same graph (the reference's function semantics: PY/function/_function.py,
branching.py, math.py), and a solve with a temperature-dependent viscosity
eta = exp(-ln(1e4) T) against the oracle's direct solve."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(res=(8, 8, 8), ppc=10, seed=2):
    import underworld2_amd as uw
    dim = len(res)
    mesh = uw.mesh.FeMesh_Cartesian("Q1/dQ0", res, (0.0,) * dim, (1.0,) * dim)
    vel = uw.mesh.MeshVariable(mesh, dim)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(swarm, particlesPerCell=ppc, seed=seed))
    names = "IJK"
    walls = tuple(mesh.specialSets[f"Min{names[d]}_VertexSet"] + mesh.specialSets[f"Max{names[d]}_VertexSet"]
                  for d in range(dim))
    bc = uw.conditions.DirichletCondition(vel, walls)
    return uw, mesh, vel, p, swarm, bc


def _device_values(uw, st, swarm, dim):
    import underworld2_amd._lib as L
    st._prepare_values()
    P = swarm.particleLocalCount
    eta = np.zeros(P)
    force = np.zeros((P, dim))
    L.check(L.lib().uwg_stokes_get_particle_values(st._h, P, L.dptr(eta), L.dptr(force)))
    return eta, force


def test_device_functions_match_host_evaluation():
    uw, mesh, vel, p, swarm, bc = _setup()
    fn = uw.function
    T = uw.mesh.MeshVariable(mesh, 1)
    X = mesh.data
    T.data[:, 0] = 1.0 - X[:, 2] + 0.05 * np.sin(np.pi * X[:, 0]) * np.cos(np.pi * X[:, 1])
    mat = swarm.add_variable("int", 1)
    c = fn.input() - (0.5, 0.5, 0.6)
    mat.data[:] = fn.branching.conditional([(fn.math.dot(c, c) < 0.2 ** 2, 1), (True, 0)]).evaluate(swarm)
    phi = swarm.add_variable("double", 1)
    phi.data[:, 0] = np.random.default_rng(0).random(swarm.particleLocalCount)
    cases = {
        "sinker": (fn.branching.map(fn_key=mat, mapping={0: 1.0, 1: 1.0e3}),
                   fn.branching.map(fn_key=mat, mapping={0: 0.0, 1: 1.0}) * (0.0, 0.0, -1.0)),
        "arrhenius": (fn.math.exp(-np.log(1.0e4) * T), T * (0.0, 0.0, 1.0e2)),
        "mixed": (fn.math.sqrt(fn.math.abs(fn.math.sin(3.0 * fn.input()[0]))) + 2.0 * phi ** 2 + 0.5,
                  fn.branching.conditional([((fn.input()[2] > 0.5) & (phi < 0.5), (1.0, -2.0, 0.5)),
                                            (True, (0.0, 0.0, -1.0))])),
    }
    for name, (etaf, forcef) in cases.items():
        st = uw.systems.Stokes(vel, p, fn_viscosity=etaf, fn_bodyforce=forcef, voronoi_swarm=swarm, conditions=[bc])
        if name == "sinker":
            # a pure material map takes the table path; force the per-particle program instead
            st._table_variable = lambda: None
        eta, force = _device_values(uw, st, swarm, 3)
        he = np.asarray(etaf.evaluate(swarm), dtype=np.float64).reshape(-1)
        hf = np.asarray(forcef.evaluate(swarm), dtype=np.float64).reshape(len(he), -1)
        hf = np.broadcast_to(hf, (len(he), 3))
        print(f"{name}: max rel eta diff {np.max(np.abs(eta - he) / np.abs(he)):.2e}", flush=True)
        np.testing.assert_allclose(eta, he, rtol=1e-14, atol=0)
        np.testing.assert_allclose(force, hf, rtol=1e-14, atol=1e-300)
        if name == "sinker":
            # maps and arithmetic only (no transcendental functions): bit-identical to the host
            assert np.array_equal(eta, he) and np.array_equal(force, hf)


def test_device_function_errors():
    uw, mesh, vel, p, swarm, bc = _setup()
    fn = uw.function
    import underworld2_amd._lib as L
    st = uw.systems.Stokes(vel, p, fn_viscosity=fn.branching.conditional([(fn.input()[0] < 0.5, 1.0)]),
                           fn_bodyforce=(0.0, 0.0, -1.0), voronoi_swarm=swarm, conditions=[bc])
    with pytest.raises(L.UWGError):
        st._prepare_values()
    st2 = uw.systems.Stokes(vel, p, fn_viscosity=fn.input()[0] - 0.5, fn_bodyforce=(0.0, 0.0, -1.0),
                            voronoi_swarm=swarm, conditions=[bc])
    with pytest.raises(L.UWGError):
        st2._prepare_values()


def test_temperature_dependent_viscosity_solve_matches_oracle():
    """eta = exp(-ln(1e4) T) with T a nodal field: device program -> assembly -> FP64 solve,
    against the oracle's assembly and direct solve with eta evaluated on the host."""
    from oracle import oracle as O
    uw, mesh, vel, p, swarm, bc = _setup((8, 8, 8), 12, 5)
    fn = uw.function
    T = uw.mesh.MeshVariable(mesh, 1)
    X = mesh.data
    T.data[:, 0] = 1.0 - X[:, 2] + 0.1 * np.cos(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 2])
    etaf = fn.math.exp(-np.log(1.0e4) * T)
    forcef = T * (0.0, 0.0, 10.0)
    st = uw.systems.Stokes(vel, p, fn_viscosity=etaf, fn_bodyforce=forcef, voronoi_swarm=swarm, conditions=[bc])
    solver = uw.systems.Solver(st)
    solver.set_inner_rtol(1e-12)
    solver.set_outer_rtol(1e-11)
    solver.solve()
    u = vel.data.copy()
    pr = p.data[:, 0].copy()
    # oracle: same layout, DVC, host-evaluated function values in slot order
    om = O.Mesh((8, 8, 8), (0.0,) * 3, (1.0,) * 3)
    coords = O.layout_random(om, 12, 5)
    assert np.array_equal(coords, swarm.particleCoordinates.data)
    ptr, pid = O.initial_cells(om, 12)
    xi = O.coincident_map(om, ptr, pid, coords)
    xw, w = O.dvc_all(3, ptr, xi)
    he = np.asarray(etaf.evaluate(swarm), dtype=np.float64).reshape(-1)
    hf = np.asarray(forcef.evaluate(swarm), dtype=np.float64).reshape(-1, 3)
    Ke, fe, Ge, S = O.assemble_elements(om, ptr, xw, w, he[pid], hf[pid])
    sysm = O.System(om, om.freeslip_mask(), Ke, fe, Ge, S)
    uo, po = sysm.solve()
    eu = np.linalg.norm(u - uo) / np.linalg.norm(uo)
    ep = np.linalg.norm((pr - pr.mean()) - (po - po.mean())) / np.linalg.norm(po - po.mean())
    print(f"velocity error {eu:.2e}, pressure error {ep:.2e}", flush=True)
    assert eu < 1e-8 and ep < 1e-8
