"""The benchmark's other configurations as parity cases (GPU, through the C ABI).

* 2D 64^2 sinker, the reference's own CPU-runnable case: 20 particles per cell (seed 0), circle
  r = 0.1 at (0.5, 0.6) with eta 100 and rho 1 in eta 1 / rho 0, gravity (0, -1); Schur solve to
  rtol 1e-5; one RK2 step at dt = 0.5 get_max_dt().
* 2D layered viscosity (8 horizontal layers, log10 eta ~ U(0, 4) seeded, extremes forced to span
  1e4; a buoyant circle as above drives the flow), 40 particles per cell, mixed precision: the
  error of every field against the FP64 oracle at 128^2, and at the full 1024^2 the
  size-independent checks (true FP64 residual, mixed vs FP64 agreement).
"""
import numpy as np
import pytest

from sinker import build_oracle, build_uw, sphere_material

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_config_2d_sinker_64():
    print("config 2D 64x64 sinker, 20 ppc, seed 0", flush=True)
    g = build_uw((64, 64), 20, 0, (0.5, 0.6), 0.1)
    o = build_oracle((64, 64), 20, 0, (0.5, 0.6), 0.1)
    uw = g["uw"]
    # integration swarm bit-exact
    xi, w = g["swarm"].integration_points(dvc=True)
    assert np.array_equal(xi, o["xi"]) and np.array_equal(w, o["w"])
    # the configuration's solve: Schur complement to rtol 1e-5
    solver = uw.systems.Solver(g["stokes"])
    solver.set_outer_rtol(1e-5)
    solver.solve()
    u5 = g["vel"].data.copy()
    p5 = g["p"].data[:, 0].copy()
    eu, ep = _rel(u5, o["u"]), _rel(p5 - p5.mean(), o["p"] - o["p"].mean())
    print(f"rtol 1e-5: velocity error {eu:.2e}, pressure error {ep:.2e}", flush=True)
    # at the configuration's rtol 1e-5 the Q1/dQ0 pressure carries slowly converging
    # checkerboard modes (measured 1.5e-3); velocity is within 1e-5
    assert eu < 1e-4 and ep < 1e-2
    # tight solve: the plain FP64 path matches the oracle to 1e-8
    solver.set_inner_rtol(1e-12)
    solver.set_outer_rtol(1e-11)
    solver.solve()
    u = g["vel"].data.copy()
    p = g["p"].data[:, 0]
    assert _rel(u, o["u"]) < 1e-8
    assert _rel(p - p.mean(), o["p"] - o["p"].mean()) < 1e-8
    # one RK2 step at dt = 0.5 get_max_dt() with the oracle's advection on the same field
    from oracle import oracle as O
    adv = uw.systems.SwarmAdvector(g["vel"], g["swarm"], order=2)
    dt = 0.5 * adv.get_max_dt()
    assert dt == 0.5 * O.max_dt(o["mesh"], u)
    adv.integrate(dt)
    c1, bad, fb = O.advect_rk2(o["mesh"], u, o["coords"], dt)
    assert bad == 0
    assert np.array_equal(g["swarm"].particleCoordinates.data, c1)
    T = O.CellTables(o["mesh"], o["ptr"], o["pid"])
    assert T.update(c1) == 0
    ptr, pid = g["swarm"].cell_tables()
    eptr, epid = T.csr()
    assert np.array_equal(ptr, eptr) and np.array_equal(pid, epid)


def _layers(seed=11):
    rng = np.random.default_rng(seed)
    le = rng.uniform(0.0, 4.0, 8)
    le[np.argmin(le)] = 0.0
    le[np.argmax(le)] = 4.0
    return 10.0 ** le


def _layer_material(coords, n_layers=8):
    lay = np.minimum((coords[:, 1] * n_layers).astype(np.int32), n_layers - 1)
    inc = sphere_material(coords, (0.5, 0.6), 0.1)
    return np.where(inc == 1, n_layers, lay).astype(np.int32)


def _build_layered(res, ppc, seed=0):
    import underworld2_amd as uw
    mesh = uw.mesh.FeMesh_Cartesian("Q1/dQ0", res, (0.0, 0.0), (1.0, 1.0))
    vel = uw.mesh.MeshVariable(mesh, 2)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(swarm, particlesPerCell=ppc, seed=seed))
    mat = swarm.add_variable("int", 1)
    mat.data[:, 0] = _layer_material(swarm.particleCoordinates.data)
    eta = np.append(_layers(), 100.0)
    rho = np.append(np.zeros(8), 1.0)
    fn = uw.function
    etaf = fn.branching.map(fn_key=mat, mapping={i: float(eta[i]) for i in range(9)})
    rhof = fn.branching.map(fn_key=mat, mapping={i: float(rho[i]) for i in range(9)})
    walls = (mesh.specialSets["MinI_VertexSet"] + mesh.specialSets["MaxI_VertexSet"],
             mesh.specialSets["MinJ_VertexSet"] + mesh.specialSets["MaxJ_VertexSet"])
    bc = uw.conditions.DirichletCondition(vel, walls)
    st = uw.systems.Stokes(vel, p, fn_viscosity=etaf, fn_bodyforce=rhof * (0.0, -1.0), voronoi_swarm=swarm,
                           conditions=[bc])
    return dict(uw=uw, mesh=mesh, vel=vel, p=p, swarm=swarm, stokes=st, eta=eta, rho=rho)


def test_config_layered_2d_parity_128():
    print("config 2D layered viscosity 128x128, 40 ppc, mixed precision vs oracle", flush=True)
    from oracle import oracle as O
    g = _build_layered((128, 128), 40)
    o = O.full_pipeline(2, (128, 128), (0.0, 0.0), (1.0, 1.0), 40, 0, _layer_material, g["eta"], g["rho"],
                        (0.0, -1.0))
    for prec in ("fp64", "mixed"):
        s = g["uw"].systems.Solver(g["stokes"])
        s.set_method("block")
        s.set_precision(prec)
        # contrast 1e4 makes the solution error ~1e3 x the residual: converge to 1e-12
        s.set_outer_rtol(1e-12)
        s.solve()
        u = g["vel"].data
        p = g["p"].data[:, 0]
        eu, ep = _rel(u, o["u"]), _rel(p - p.mean(), o["p"] - o["p"].mean())
        print(f"{prec}: velocity error {eu:.2e}, pressure error {ep:.2e}, its {s.get_stats().outer_its}", flush=True)
        # the pressure of this 1e4-contrast Q1/dQ0 system is ill-conditioned (near-null
        # checkerboard modes): both the oracle's direct solve and the Krylov solve carry
        # ~1e-7 relative pressure error (measured 2.2e-7 at rtol 1e-12)
        assert eu < 1e-8 and ep < 1e-6, prec


def test_config_layered_2d_1024_properties():
    print("config 2D layered viscosity 1024x1024, 40 ppc (41.9M particles)", flush=True)
    g = _build_layered((1024, 1024), 40)
    sols = {}
    for prec in ("fp64", "mixed"):
        s = g["uw"].systems.Solver(g["stokes"])
        s.set_method("block")
        s.set_precision(prec)
        s.set_outer_rtol(1e-6)
        s.solve()
        st = s.get_stats()
        rel = float(np.hypot(st.res_momentum, st.res_continuity) / st.norm_f)
        print(f"{prec}: its {st.outer_its}, solve {st.solve_time:.3f}s, FP64 residual {rel:.2e}", flush=True)
        assert st.converged and rel <= 1.05e-6
        sols[prec] = (g["vel"].data.copy(), g["p"].data[:, 0].copy())
    du = _rel(sols["mixed"][0], sols["fp64"][0])
    pm, pf = sols["mixed"][1], sols["fp64"][1]
    dp = _rel(pm - pm.mean(), pf - pf.mean())
    print(f"mixed vs fp64: velocity {du:.2e}, pressure {dp:.2e}", flush=True)
    assert du < 1e-3 and dp < 1e-3


def _oracle_convection(res, ppc, steps, T0, Ra=1.0e6):
    """The 3D convection loop restated on the CPU: the Stokes solve with particle quadrature
    (coincident map, DVC, eta and buoyancy from T interpolated at the particles, assembly, exact
    solve), dt from the new velocity, the SUPG step of T (oracle/supg_oracle.py), then RK2
    advection and relocation."""
    from oracle import oracle as O
    from oracle import supg_oracle as SO
    mesh = O.Mesh(res, (0.0,) * 3, (1.0,) * 3)
    sm = SO.Q1Mesh(res, (0.0,) * 3, (1.0,) * 3)
    X = sm.coords()
    Tb = (np.abs(X[:, 2]) < 1e-12) | (np.abs(X[:, 2] - 1.0) < 1e-12)
    sup = SO.SUPG(sm, Tb)
    kap = np.ones((sm.E, 27))
    src = np.zeros((sm.E, 27))
    coords = O.layout_random(mesh, ppc, 0)
    ptr, pid = O.initial_cells(mesh, ppc)
    tables = O.CellTables(mesh, ptr, pid)
    T, Td = T0.copy(), np.zeros(sm.N)
    vel = np.zeros((sm.N, 3))
    bc = mesh.freeslip_mask()
    out = []
    for _ in range(steps):
        ptr, pid = tables.csr()
        xi = O.coincident_map(mesh, ptr, pid, coords)
        xw, w = O.dvc_all(3, ptr, xi)
        Tf = np.stack([T, np.zeros_like(T), np.zeros_like(T)], 1)
        Tp = O.interpolate(mesh, Tf, coords)[0][:, 0]
        eta = np.exp(-np.log(1.0e4) * Tp)[pid]
        force = (Ra * Tp)[pid][:, None] * np.array([0.0, 0.0, 1.0])[None, :]
        Ke, fe, Ge, S = O.assemble_elements(mesh, ptr, xw, w, eta, force)
        sysm = O.System(mesh, bc, Ke, fe, Ge, S)
        vel, _ = sysm.solve()
        dt = min(sup.max_dt(vel, kap), O.max_dt(mesh, vel))
        T, Td = sup.integrate(T, Td, vel, kap, src, dt)
        coords, bad, _ = O.advect_rk2(mesh, vel, coords, dt)
        assert bad == 0
        tables.update(coords)
        out.append(dict(dt=dt, T=T.copy(), vel=vel.copy(), coords=coords.copy()))
    return out


def test_config_thermal_convection_small():
    """The 3D thermal-convection configuration (Ra 1e6, eta = exp(-ln(1e4) T), Stokes + SUPG +
    RK2 per step) at 8^3 with 8 particles per cell: two FP64 steps against the CPU restatement of
    the whole loop, then FP64 against mixed precision and against the learned stencil with mixed
    precision (the configuration's learned mode) over eight steps on Nu and v_rms."""
    import bench
    res, ppc = (8, 8, 8), 8
    print(f"thermal convection {res}, {ppc} ppc: GPU FP64 loop against the oracle", flush=True)
    C = bench.build_convection(res, ppc)
    T0 = C["T"].data[:, 0].copy()
    ref = _oracle_convection(res, ppc, 2, T0)
    s = bench.make_solver({"uw": C["uw"], "stokes": C["stokes"]}, "fp64", 1e-11)
    for k in range(2):
        s.solve()
        dt = min(C["ad"].get_max_dt(), C["adv"].get_max_dt())
        assert dt == pytest.approx(ref[k]["dt"], rel=1e-7)
        C["ad"].integrate(dt)
        C["adv"].integrate(dt)
        eT = _rel(C["T"].data[:, 0], ref[k]["T"])
        eu = _rel(C["vel"].data, ref[k]["vel"])
        ec = float(np.abs(C["swarm"].particleCoordinates.data - ref[k]["coords"]).max())
        print(f"step {k}: dt {dt:.4e}, T error {eT:.2e}, velocity error {eu:.2e}, max particle offset {ec:.2e}",
              flush=True)
        # the GPU solve stops at rtol 1e-11 (the oracle solves exactly); T and the particles inherit
        # the velocity's error through dt and the advection
        assert eT < 1e-8 and eu < 1e-8 and ec < 1e-8
    print("FP64 against mixed precision and against learned + mixed over 8 steps", flush=True)
    runs = {}
    for m in ("fp64", "mixed", "learned"):
        C = bench.build_convection(res, ppc)
        runs[m] = bench.run_convection(C, 8, m, rtol=1e-8)
        assert runs[m] is not None
    nu0, vr0 = np.array(runs["fp64"]["nu"]), np.array(runs["fp64"]["vrms"])
    for m in ("mixed", "learned"):
        nu, vr = np.array(runs[m]["nu"]), np.array(runs[m]["vrms"])
        print(f"{m}: Nu {nu0[-1]:.6f} / {nu[-1]:.6f}, v_rms {vr0[-1]:.4f} / {vr[-1]:.4f}, Stokes iterations "
              f"{sum(runs['fp64']['its'])} / {sum(runs[m]['its'])}", flush=True)
        assert np.max(np.abs(nu - nu0) / nu0) < 1e-6
        assert np.max(np.abs(vr - vr0) / vr0) < 1e-6
    assert nu0[-1] > 1.0 and vr0[-1] > 0.0
