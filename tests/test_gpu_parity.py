"""GPU parity tests: the HIP path (through the C ABI, via the underworld2_amd API) against the
CPU oracle on identical seeded meshes, swarms and material fields.

Bars: bit-exact for layout coordinates, local coordinates, DVC weights, advected coordinates,
owners and cell tables; 1e-12 relative for element integration; 1e-8 relative (L2) for the
converged FP64 velocity and mean-free pressure; mixed precision within the stated tolerances.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from sinker import K_from_W, build_oracle, build_uw, sphere_material


def _flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def case2d():
    _flush("case2d: 2D 24x24 sinker, 20 ppc, seed 0")
    g = build_uw((24, 24), 20, 0, (0.5, 0.6), 0.1)
    o = build_oracle((24, 24), 20, 0, (0.5, 0.6), 0.1)
    return g, o


@pytest.fixture(scope="module")
def case3d():
    _flush("case3d: 3D 8^3 sinker, 12 ppc, seed 3, eta contrast 1e3")
    g = build_uw((8, 8, 8), 12, 3, (0.5, 0.5, 0.6), 0.2, eta=(1.0, 1000.0))
    o = build_oracle((8, 8, 8), 12, 3, (0.5, 0.5, 0.6), 0.2, eta=(1.0, 1000.0))
    return g, o


@pytest.mark.parametrize("case", ["case2d", "case3d"])
def test_layout_material_bitexact(case, request):
    g, o = request.getfixturevalue(case)
    c = g["swarm"].particleCoordinates.data
    assert c.shape == o["coords"].shape
    assert np.array_equal(c, o["coords"])
    assert np.array_equal(g["mat"].data[:, 0], o["mat"])


@pytest.mark.parametrize("case", ["case2d", "case3d"])
def test_integration_points_bitexact(case, request):
    """Coincident map (Newton) and DVC weights, slot order."""
    g, o = request.getfixturevalue(case)
    ptr, pid = g["swarm"].cell_tables()
    assert np.array_equal(ptr, o["ptr"]) and np.array_equal(pid, o["pid"])
    xi, w = g["swarm"].integration_points(dvc=True)
    assert np.array_equal(xi, o["xi"])
    assert np.array_equal(w, o["w"])


@pytest.mark.parametrize("case", ["case2d", "case3d"])
def test_element_integration(case, request):
    """K_e rebuilt from the device's Gauss weights equals the oracle's particle K_e; f_e and
    approxS match (relative 1e-12)."""
    from underworld2_amd import _lib as L
    g, o = request.getfixturevalue(case)
    st = g["stokes"]
    st._prepare_values()
    mesh = o["mesh"]
    d = mesh.dim
    Q = 3 ** d
    W = np.zeros((Q, mesh.E))
    fe = np.zeros((mesh.E, 2 ** d * d))
    S = np.zeros(mesh.E)
    L.check(L.lib().uwg_stokes_get_assembly(st._h, 0, L.dptr(W), L.dptr(fe), L.dptr(S)))
    h = [(mesh.hi[i] - mesh.lo[i]) / mesh.res[i] for i in range(d)]
    K = K_from_W(W, d, h)
    scale = np.abs(o["Ke"]).max()
    assert np.abs(K - o["Ke"]).max() < 1e-12 * scale
    assert np.abs(fe - o["fe"]).max() < 1e-12 * max(np.abs(o["fe"]).max(), 1e-300)
    np.testing.assert_allclose(S, o["S"], rtol=1e-11)


@pytest.mark.parametrize("case", ["case2d", "case3d"])
def test_operator_applies(case, request):
    from underworld2_amd import _lib as L
    g, o = request.getfixturevalue(case)
    st = g["stokes"]
    st._prepare_values()
    sysm = o["system"]
    mesh = o["mesh"]
    rng = np.random.default_rng(9)
    x = np.zeros(mesh.N * mesh.dim)
    x[sysm.free] = rng.standard_normal(sysm.nfree)
    y = np.zeros_like(x)
    L.check(L.lib().uwg_stokes_apply(st._h, 0, 0, L.dptr(x), L.dptr(y)))
    ref = sysm.K @ x[sysm.free]
    assert np.abs(y[sysm.free] - ref).max() < 1e-12 * np.abs(ref).max()
    pv = rng.standard_normal(mesh.E)
    yg = np.zeros(mesh.N * mesh.dim)
    L.check(L.lib().uwg_stokes_apply(st._h, 1, 0, L.dptr(pv), L.dptr(yg)))
    refg = sysm.G @ pv
    assert np.abs(yg[sysm.free] - refg).max() < 1e-12 * np.abs(refg).max()
    q = np.zeros(mesh.E)
    L.check(L.lib().uwg_stokes_apply(st._h, 2, 0, L.dptr(x), L.dptr(q)))
    refq = sysm.G.T @ x[sysm.free]
    assert np.abs(q - refq).max() < 1e-12 * np.abs(refq).max()


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("case", ["case2d", "case3d"])
def test_solve_fp64_matches_oracle(case, request):
    """Plain FP64 Schur-complement path (the reference's BSSCR algorithm) converged tightly:
    velocity and mean-free pressure within 1e-8 of the oracle's direct solve."""
    g, o = request.getfixturevalue(case)
    solver = g["uw"].systems.Solver(g["stokes"])
    solver.set_inner_rtol(1e-12)
    solver.set_outer_rtol(1e-11)
    solver.solve()
    u = g["vel"].data
    p = g["p"].data[:, 0]
    assert _rel(u, o["u"]) < 1e-8
    assert _rel(p - p.mean(), o["p"] - o["p"].mean()) < 1e-8
    s = solver.get_stats()
    assert s.pressure_its > 0 and s.velocity_presolve_its > 0


@pytest.mark.parametrize("case", ["case2d", "case3d"])
def test_solve_block_fp64_and_mixed(case, request):
    """Coupled FGMRES: FP64 cycle within 1e-8; FP32 multigrid cycle inside FP64 Krylov
    reaches the same 1e-8 (the Krylov residual is FP64, only the preconditioner is FP32)."""
    g, o = request.getfixturevalue(case)
    for prec in ("fp64", "mixed"):
        solver = g["uw"].systems.Solver(g["stokes"])
        solver.set_method("block")
        solver.set_precision(prec)
        solver.set_outer_rtol(1e-12)
        solver.solve()
        u = g["vel"].data
        p = g["p"].data[:, 0]
        assert _rel(u, o["u"]) < 1e-8, prec
        assert _rel(p - p.mean(), o["p"] - o["p"].mean()) < 1e-8, prec


def test_solve_deterministic(case3d):
    g, o = case3d
    res = []
    for _ in range(2):
        solver = g["uw"].systems.Solver(g["stokes"])
        solver.set_method("block")
        solver.set_precision("mixed")
        solver.set_outer_rtol(1e-6)
        solver.solve()
        res.append((g["vel"].data.copy(), g["p"].data.copy()))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("case", ["case2d", "case3d"])
def test_advection_and_relocation_bitexact(case, request):
    """RK2 with the oracle's velocity field loaded on both sides: identical coordinates,
    owners and per-cell tables (swap-with-last / append order)."""
    from oracle import oracle as O
    g, o = request.getfixturevalue(case)
    mesh = o["mesh"]
    vel = o["u"].copy()
    g["vel"].data[:] = vel
    adv = g["uw"].systems.SwarmAdvector(g["vel"], g["swarm"], order=2)
    dt = adv.get_max_dt()
    assert dt == O.max_dt(mesh, vel)
    dt = dt * 4.0  # several cells per step for a few particles: exercises relocation
    coords0 = g["swarm"].particleCoordinates.data.copy()
    assert np.array_equal(coords0, o["coords"])
    adv.integrate(dt)
    c1, bad, fb = O.advect_rk2(mesh, vel, o["coords"], dt)
    assert bad == 0
    got = g["swarm"].particleCoordinates.data
    assert np.array_equal(got, c1)
    T = O.CellTables(mesh, o["ptr"], o["pid"])
    assert T.update(c1) == 0
    ptr, pid = g["swarm"].cell_tables()
    eptr, epid = T.csr()
    assert np.array_equal(ptr, eptr)
    assert np.array_equal(pid, epid)
    assert np.array_equal(g["swarm"].owningCell.data[:, 0], T.owner)
    ppc = int(o["ptr"][1] - o["ptr"][0])
    assert (T.owner != np.arange(len(o["pid"])) // ppc).sum() > 0


def test_rk2_doctest_through_api():
    """PY/systems/_timeintegration.py:131-147."""
    import underworld2_amd as uw
    mesh = uw.mesh.FeMesh_Cartesian(elementType="Q1/dQ0", elementRes=(9, 9), minCoord=(-1., -1.), maxCoord=(1., 1.))
    v = uw.mesh.MeshVariable(mesh=mesh, nodeDofCount=2)
    swarm = uw.swarm.Swarm(mesh=mesh)
    swarm.add_particles_with_coordinates(np.array([[0.2, -0.2]]))
    v.data[:] = [1.0, 1.0]
    adv = uw.systems.SwarmAdvector(velocityField=v, swarm=swarm, order=2)
    adv.integrate(adv.get_max_dt())
    assert np.allclose(swarm.particleCoordinates.data[0], [0.27856742, -0.12143258], rtol=1e-4)


def test_random_layout_docstring_through_api():
    import underworld2_amd as uw
    mesh = uw.mesh.FeMesh_Cartesian('Q1/dQ0', (1, 1), (0., 0.), (1., 1.))
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(swarm, particlesPerCell=4))
    assert swarm.particleLocalCount == 4
    np.testing.assert_allclose(swarm.particleCoordinates.data,
                               [[0.24261743, 0.67115852], [0.16116546, 0.70790335],
                                [0.73160516, 0.08792286], [0.71953113, 0.15966135]], atol=5e-9)


def test_layout_large_bitexact_sample():
    """3D 32^3 x 30 ppc: a sample of particles (jump-ahead generator) equals the oracle."""
    from oracle import oracle as O
    import underworld2_amd as uw
    mesh = uw.mesh.FeMesh_Cartesian('Q1/dQ0', (32, 32, 32), (0., 0., 0.), (1., 1., 1.))
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(swarm, particlesPerCell=30, seed=0))
    c = swarm.particleCoordinates.data
    ref = O.layout_random(O.Mesh((32, 32, 32), (0., 0., 0.), (1., 1., 1.)), 30, 0)
    assert np.array_equal(c, ref)


@pytest.mark.parametrize("dim", [2, 3])
def test_dvc_unbalanced_cells_match_oracle(dim):
    """DVC weights and centroids after an advection step has unbalanced the cells (a converging
    velocity field: 6 to 73 particles per cell in 3D), on a ragged 3D grid (10 x 7 x 6, 30 ppc)
    and a 2D grid (24 x 17, 20 ppc): bit-identical with the oracle for every element."""
    from oracle import oracle as O
    import underworld2_amd as uw
    res = (10, 7, 6) if dim == 3 else (24, 17)
    ppc = 30 if dim == 3 else 20
    print(f"DVC after advection {res} x {ppc}", flush=True)
    mesh = uw.mesh.FeMesh_Cartesian('Q1/dQ0', res, (0.,) * dim, (1.,) * dim)
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(swarm, particlesPerCell=ppc, seed=4))
    v = uw.mesh.MeshVariable(mesh, dim)
    v.data[:] = -np.sin(np.pi * mesh.data) * 0.5
    adv = uw.systems.SwarmAdvector(v, swarm, order=2)
    adv.integrate(3.2 * adv.get_max_dt())
    ptr, pid = swarm.cell_tables()
    counts = np.diff(ptr)
    print(f"  particles per cell {counts.min()}..{counts.max()}", flush=True)
    assert counts.min() > 0
    xi, w = swarm.integration_points(dvc=True)
    xi0 = O.coincident_map(O.Mesh(res, (0.,) * dim, (1.,) * dim), ptr, pid, swarm.particleCoordinates.data)
    xw, wo = O.dvc_all(dim, ptr, xi0)
    assert np.array_equal(xi, xw) and np.array_equal(w, wo)


def test_underresolved_cell_raises():
    import underworld2_amd as uw
    mesh = uw.mesh.FeMesh_Cartesian('Q1/dQ0', (4, 4), (0., 0.), (1., 1.))
    swarm = uw.swarm.Swarm(mesh)
    swarm.add_particles_with_coordinates(np.array([[0.1, 0.1], [0.9, 0.9]]))
    with pytest.raises(Exception):
        swarm.integration_points(dvc=True)


def test_particles_leaving_domain_raise():
    import underworld2_amd as uw
    mesh = uw.mesh.FeMesh_Cartesian('Q1/dQ0', (4, 4), (0., 0.), (1., 1.))
    v = uw.mesh.MeshVariable(mesh, 2)
    swarm = uw.swarm.Swarm(mesh)
    swarm.add_particles_with_coordinates(np.array([[0.95, 0.5]]))
    v.data[:] = [1.0, 0.0]
    adv = uw.systems.SwarmAdvector(v, swarm, order=2)
    with pytest.raises(RuntimeError):
        adv.integrate(1.0)
