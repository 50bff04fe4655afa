"""Q2/dPc1 element pair on the GPU against the CPU oracle (oracle/q2_oracle.py).

* operators: K, G, G^T, the body force and the approxS blocks of the GPU path equal the
  oracle's assembled matrices, for the Gauss-swarm path (Q2 default: 5^dim points) and for
  particle quadrature with DVC weights (the 5^dim-weight reduction of the particle sums is
  exact, so the products agree to rounding);
* solves: the reference's SolCx Q2/dPc1 32^2 case (docs/examples/10_Analytic_Solutions.ipynb:
  v_rms 0.005578279626793198, rtol 1e-7) through the Stokes / Solver API, for the coupled
  FP64, the Schur complement and the mixed-precision paths; a 3D particle sinker against the
  oracle's direct solve (velocity and pressure relative L2 within 1e-8);
* interpolation of a quadratic field is exact.
"""
import numpy as np
import pytest

from oracle import q2_oracle as Q

pytestmark = pytest.mark.gpu

SOLCX_VRMS = 0.005578279626793198


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _walls(uw, mesh, vel):
    sets = []
    for d, n in enumerate("IJK"[:mesh.dim]):
        sets.append(mesh.specialSets[f"Min{n}_VertexSet"] + mesh.specialSets[f"Max{n}_VertexSet"])
    return uw.conditions.DirichletCondition(vel, tuple(sets))


def _full_ops(om, Ke, fe, Ge):
    """Dense full-size K (identity Dirichlet rows/columns), G (zero Dirichlet rows), f."""
    d = om.dim
    en = om.element_nodes()
    nd = om.nen * d
    dofs = (en[:, :, None] * d + np.arange(d)[None, None, :]).reshape(om.E, nd)
    bc = om.freeslip_mask().reshape(-1)
    n = om.N * d
    K = np.zeros((n, n))
    for e in range(om.E):
        K[np.ix_(dofs[e], dofs[e])] += Ke[e]
    K[bc, :] = 0.0
    K[:, bc] = 0.0
    K[bc, bc] = 1.0
    npr = om.npe
    G = np.zeros((n, om.E * npr))
    for e in range(om.E):
        G[dofs[e], e * npr:(e + 1) * npr] += Ge
    G[bc, :] = 0.0
    f = np.zeros(n)
    np.add.at(f, dofs.reshape(-1), fe.reshape(-1))
    f[bc] = 0.0
    return K, G, f


def _apply(st, which, level, x, nout):
    import underworld2_amd._lib as L
    y = np.zeros(nout)
    L.check(L.lib().uwg_stokes_apply(st._h, which, level, L.dptr(np.ascontiguousarray(x, np.float64)), L.dptr(y)),
            "apply")
    return y


def _check_operators(stokes, om, Ke, fe, tol=1e-12):
    import underworld2_amd._lib as L
    K, G, f = _full_ops(om, Ke, fe, Q.element_G(om))
    rng = np.random.default_rng(7)
    n, npt = om.N * om.dim, om.E * om.npe
    x = rng.standard_normal(n)
    p = rng.standard_normal(npt)
    assert _rel(_apply(stokes, 0, 0, x, n), K @ x) < tol
    assert _rel(_apply(stokes, 1, 0, p, n), G @ p) < tol
    assert _rel(_apply(stokes, 2, 0, x, npt), G.T @ x) < tol
    fr = np.zeros(n)
    L.check(L.lib().uwg_stokes_get_rhs(stokes._h, L.dptr(fr)), "rhs")
    assert _rel(fr, f) < tol
    S = np.zeros((om.E, om.npe, om.npe))
    L.check(L.lib().uwg_stokes_get_assembly(stokes._h, 0, None, None, L.dptr(S)), "assembly")
    assert _rel(S, Q.element_approxS(Ke, Q.element_G(om))) < tol


@pytest.mark.parametrize("res", [(4, 3), (3, 2, 2)])
def test_q2_gauss_operators_match_oracle(res):
    import underworld2_amd as uw
    dim = len(res)
    print(f"Q2 Gauss-path operators {res}", flush=True)
    hi = (1.0, 0.7, 1.3)[:dim]
    mesh = uw.mesh.FeMesh_Cartesian("Q2/dPc1", res, (0.0,) * dim, hi)
    assert mesh.nodesGlobal == int(np.prod([2 * r + 1 for r in res]))
    vel = uw.mesh.MeshVariable(mesh, dim)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    assert p.data.shape == (mesh.elementsGlobal * (dim + 1), 1)
    fn = uw.function
    x = fn.input()
    eta = 1.0 + x[0] * x[0] + 2.0 * x[1]
    force = (fn.math.sin(3.0 * x[1]), x[0] * x[1]) if dim == 2 else (x[2], fn.math.cos(x[0]), x[0] * x[1])
    stokes = uw.systems.Stokes(vel, p, fn_viscosity=eta, fn_bodyforce=force, conditions=[_walls(uw, mesh, vel)])
    stokes._prepare_values()
    om = Q.Q2Mesh(res, (0.0,) * dim, hi)
    assert np.array_equal(mesh.data, om.coords())

    def feta(X):
        return 1.0 + X[:, 0] * X[:, 0] + 2.0 * X[:, 1]

    def fforce(X):
        if dim == 2:
            return np.stack([np.sin(3.0 * X[:, 1]), X[:, 0] * X[:, 1]], -1)
        return np.stack([X[:, 2], np.cos(X[:, 0]), X[:, 0] * X[:, 1]], -1)
    Ke, fe = Q.gauss_assembly(om, feta, fforce)
    _check_operators(stokes, om, Ke, fe)


@pytest.mark.parametrize("res,ppc", [((4, 4), 20), ((3, 3, 2), 30)])
def test_q2_particle_operators_match_oracle(res, ppc):
    import underworld2_amd as uw
    dim = len(res)
    print(f"Q2 particle-path operators {res} ppc {ppc}", flush=True)
    mesh = uw.mesh.FeMesh_Cartesian("Q2/dPc1", res, (0.0,) * dim, (1.0,) * dim)
    vel = uw.mesh.MeshVariable(mesh, dim)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(swarm, particlesPerCell=ppc, seed=3))
    mat = swarm.add_variable("int", 1)
    fn = uw.function
    c = fn.input() - (0.5,) * dim
    mat.data[:] = fn.branching.conditional([(fn.math.dot(c, c) < 0.3 ** 2, 1), (True, 0)]).evaluate(swarm)
    etaf = fn.branching.map(fn_key=mat, mapping={0: 1.0, 1: 50.0})
    rhof = fn.branching.map(fn_key=mat, mapping={0: 0.0, 1: 1.0})
    g = (0.0, -1.0) if dim == 2 else (0.0, 0.0, -1.0)
    stokes = uw.systems.Stokes(vel, p, fn_viscosity=etaf, fn_bodyforce=rhof * g, voronoi_swarm=swarm,
                               conditions=[_walls(uw, mesh, vel)])
    stokes._prepare_values()
    # oracle: brute-force particle sums with the swarm's own integration points (slot order)
    xi, w = swarm.integration_points(dvc=True)
    ptr, pid = swarm.cell_tables()
    assert np.all(np.diff(ptr) == ppc)
    m = mat.data[:, 0][pid]
    om = Q.Q2Mesh(res, (0.0,) * dim, (1.0,) * dim)
    E = om.E
    xi = xi.reshape(E, ppc, dim)
    wd = w.reshape(E, ppc) * om.detJ
    c_eta = np.where(m == 1, 50.0, 1.0).reshape(E, ppc) * wd
    frc = np.where(m == 1, 1.0, 0.0).reshape(E, ppc)[:, :, None] * np.asarray(g)[None, None, :]
    Ke = Q.element_K(om, xi, c_eta)
    fe = Q.element_f(om, xi, frc * wd[:, :, None])
    _check_operators(stokes, om, Ke, fe, tol=1e-11)


def test_q2_apply_kernels_match_oracle():
    """The three 3D apply kernels (sum-factorised, matrix cores, dense vector) on a ragged
    (5, 4, 3) grid whose parity classes hold 1-12 elements, Gauss path with eta varying inside
    the elements: FP64 K x within 1e-12 of the oracle's assembled K; the FP32 apply (FP32
    weights and vectors) within 2e-6."""
    import underworld2_amd as uw
    import underworld2_amd._lib as L
    res, hi = (5, 4, 3), (1.0, 0.8, 0.7)
    print(f"Q2 apply kernels {res}", flush=True)
    mesh = uw.mesh.FeMesh_Cartesian("Q2/dPc1", res, (0.0,) * 3, hi)
    vel = uw.mesh.MeshVariable(mesh, 3)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    fn = uw.function
    x = fn.input()
    eta = 1.0 + 3.0 * x[0] * x[1] + x[2] * x[2]
    stokes = uw.systems.Stokes(vel, p, fn_viscosity=eta, fn_bodyforce=(0.0, 0.0, -1.0),
                               conditions=[_walls(uw, mesh, vel)])
    stokes._prepare_values()
    om = Q.Q2Mesh(res, (0.0,) * 3, hi)
    Ke, fe = Q.gauss_assembly(om, lambda X: 1.0 + 3.0 * X[:, 0] * X[:, 1] + X[:, 2] * X[:, 2],
                              lambda X: np.stack([0 * X[:, 0], 0 * X[:, 0], -1.0 + 0 * X[:, 0]], -1))
    K, _, _ = _full_ops(om, Ke, fe, Q.element_G(om))
    n = om.N * 3
    xv = np.random.default_rng(11).standard_normal(n)
    ref = K @ xv
    try:
        for kind, name in ((2, "sum-factorised"), (1, "matrix cores"), (0, "vector")):
            L.check(L.lib().uwg_set_q2_apply_kernel(kind), "kernel")
            e64 = _rel(_apply(stokes, 0, 0, xv, n), ref)
            e32 = _rel(_apply(stokes, 4, 0, xv, n), ref)
            print(f"  {name}: FP64 {e64:.1e}, FP32 {e32:.1e}", flush=True)
            assert e64 < 1e-12 and e32 < 2e-6
    finally:
        L.check(L.lib().uwg_set_q2_apply_kernel(2), "kernel")


def _solcx_uw(res=(32, 32)):
    import underworld2_amd as uw
    mesh = uw.mesh.FeMesh_Cartesian("Q2/dPc1", res, (0.0, 0.0), (1.0, 1.0))
    vel = uw.mesh.MeshVariable(mesh, 2)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    vel.data[:] = 0.0
    p.data[:] = 0.0
    fn = uw.function
    x = fn.input()
    eta = fn.branching.conditional([(x[0] < 0.75, 1.0), (True, 1.0e5)])
    force = (0.0, fn.math.cos(np.pi * x[0]) * fn.math.sin(np.pi * x[1]))
    stokes = uw.systems.Stokes(vel, p, fn_viscosity=eta, fn_bodyforce=force, conditions=[_walls(uw, mesh, vel)])
    return uw, mesh, vel, p, stokes


@pytest.fixture(scope="module")
def solcx_oracle():
    eta, f = Q.solcx_functions(eta_a=1.0, eta_b=1.0e5, x_c=0.75, n_z=1)
    return Q.solve_gauss_problem((32, 32), eta, f)


def test_solcx_q2_vrms_coupled_fp64(solcx_oracle):
    print("SolCx Q2/dPc1 32^2, coupled FGMRES FP64", flush=True)
    uw, mesh, vel, p, stokes = _solcx_uw()
    solver = uw.systems.Solver(stokes)
    solver.set_method("block")
    solver.set_outer_rtol(1e-11)
    solver.solve()
    st = solver.get_stats()
    v = stokes.velocity_rms()
    print(f"v_rms {v:.16g} (reference {SOLCX_VRMS}), outer its {st.outer_its}", flush=True)
    assert v == pytest.approx(SOLCX_VRMS, rel=1e-7)
    o = solcx_oracle
    assert _rel(vel.data, o["u"]) < 1e-8
    pg = p.data[:, 0]
    assert _rel(pg - pg.mean(), o["p"] - o["p"].mean()) < 1e-8
    r1, r2 = stokes.eqResiduals
    print(f"true FP64 residuals {r1:.2e} {r2:.2e}", flush=True)


def test_solcx_q2_vrms_schur(solcx_oracle):
    print("SolCx Q2/dPc1 32^2, Schur complement (BSSCR) FP64", flush=True)
    uw, mesh, vel, p, stokes = _solcx_uw()
    solver = uw.systems.Solver(stokes)
    solver.set_inner_rtol(1e-11)
    solver.set_outer_rtol(1e-10)
    solver.solve()
    v = stokes.velocity_rms()
    st = solver.get_stats()
    print(f"v_rms {v:.16g}, pressure its {st.pressure_its}, velocity its {st.velocity_total_its}", flush=True)
    assert v == pytest.approx(SOLCX_VRMS, rel=1e-7)
    assert _rel(vel.data, solcx_oracle["u"]) < 1e-8


def test_solcx_q2_mixed_precision(solcx_oracle):
    print("SolCx Q2/dPc1 32^2, mixed precision", flush=True)
    uw, mesh, vel, p, stokes = _solcx_uw()
    solver = uw.systems.Solver(stokes)
    solver.set_method("block")
    solver.set_precision("mixed")
    solver.set_outer_rtol(1e-8)
    solver.solve()
    v = stokes.velocity_rms()
    e = _rel(vel.data, solcx_oracle["u"])
    print(f"v_rms {v:.16g}, velocity error {e:.2e}, outer its {solver.get_stats().outer_its}", flush=True)
    # FP32 cycle inside FP64 refinement, stopped at FP64 relative residual 1e-8
    assert v == pytest.approx(SOLCX_VRMS, rel=1e-6)
    assert e < 1e-5


def test_q2_sinker_3d_particles_vs_oracle():
    import underworld2_amd as uw
    res, ppc = (6, 6, 6), 20
    print(f"Q2/dPc1 3D sinker {res}, {ppc} ppc, against the oracle's direct solve", flush=True)
    mesh = uw.mesh.FeMesh_Cartesian("Q2/dPc1", res, (0.0,) * 3, (1.0,) * 3)
    vel = uw.mesh.MeshVariable(mesh, 3)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(swarm, particlesPerCell=ppc, seed=1))
    mat = swarm.add_variable("int", 1)
    fn = uw.function
    c = fn.input() - (0.5, 0.5, 0.6)
    mat.data[:] = fn.branching.conditional([(fn.math.dot(c, c) < 0.2 ** 2, 1), (True, 0)]).evaluate(swarm)
    etaf = fn.branching.map(fn_key=mat, mapping={0: 1.0, 1: 1000.0})
    rhof = fn.branching.map(fn_key=mat, mapping={0: 0.0, 1: 1.0})
    stokes = uw.systems.Stokes(vel, p, fn_viscosity=etaf, fn_bodyforce=rhof * (0.0, 0.0, -1.0), voronoi_swarm=swarm,
                               conditions=[_walls(uw, mesh, vel)])
    solver = uw.systems.Solver(stokes)
    solver.set_method("block")
    solver.set_outer_rtol(1e-11)
    solver.solve()
    print(f"outer its {solver.get_stats().outer_its}", flush=True)
    xi, w = swarm.integration_points(dvc=True)
    ptr, pid = swarm.cell_tables()
    m = mat.data[:, 0][pid]
    om = Q.Q2Mesh(res, (0.0,) * 3, (1.0,) * 3)
    wd = w.reshape(om.E, ppc) * om.detJ
    Ke = Q.element_K(om, xi.reshape(om.E, ppc, 3), np.where(m == 1, 1000.0, 1.0).reshape(om.E, ppc) * wd)
    frc = np.where(m == 1, 1.0, 0.0).reshape(om.E, ppc)[:, :, None] * np.array([0.0, 0.0, -1.0])[None, None, :]
    fe = Q.element_f(om, xi.reshape(om.E, ppc, 3), frc * wd[:, :, None])
    sysm = Q.Q2System(om, om.freeslip_mask(), Ke, fe, Q.element_G(om))
    u, po = sysm.solve()
    eu = _rel(vel.data, u)
    pg = p.data[:, 0]
    ep = _rel(pg - pg.mean(), po - po.mean())
    print(f"velocity error {eu:.2e}, pressure error {ep:.2e}", flush=True)
    assert eu < 1e-8 and ep < 1e-8


def test_q2_field_interpolation_exact():
    import underworld2_amd as uw
    mesh = uw.mesh.FeMesh_Cartesian("Q2/dPc1", (3, 4), (0.0, 0.0), (1.5, 1.0))
    f = uw.mesh.MeshVariable(mesh, 2)
    X = mesh.data
    f.data[:] = np.stack([X[:, 0] ** 2 + X[:, 1], X[:, 0] * X[:, 1] - X[:, 1] ** 2], -1)
    pts = np.random.default_rng(2).uniform((0.0, 0.0), (1.5, 1.0), (200, 2))
    v = f.evaluate(pts)
    want = np.stack([pts[:, 0] ** 2 + pts[:, 1], pts[:, 0] * pts[:, 1] - pts[:, 1] ** 2], -1)
    assert np.allclose(v, want, rtol=1e-13, atol=1e-13)
