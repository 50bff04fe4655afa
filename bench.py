"""Benchmark: Stokes solves/sec on the 3D 128^3 Q1/dQ0 sinker (eta contrast 1e3,
30 particles per cell, free-slip box), relative residual 1e-6, on one MI355X.

One step = one end-to-end ``solver.solve()`` on a fixed swarm: particle-quadrature element
integration (A6-A9: moments, approxS, right-hand side), multigrid hierarchy setup, and the
Krylov solve to outer rtol 1e-6 -- the same work the reference's ``StokesSolver.solve`` does
with ``reinitialise=True`` (the voronoi swarm is unchanged, so DVC is not re-run).

The headline ``value`` is the learned+mixed configuration (the learned Schur stencil in the
coupled solver's Schur block, FP32 multigrid cycle, FP64 Krylov) when a trained model is
present, otherwise the mixed-precision coupled solver.  Every mode's own timing, iteration
counts, FP64 residuals and error against an FP64 reference solution (coupled solve in FP64
throughout, learned stencil, rtol 1e-11) are reported in ``modes``, with a per-solve roofline
from the logged counts;
``advection`` holds the RK2 step and DVC pass times.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--res 128] [--ppc 30]
                       [--mode M] [--modes fp64,mixed,learned[,learned-coarse,schur]]
                       [--no-cpu-baseline] [--no-advection] [--no-reference] [--no-floor]
                       [--apply-only] [--advection-only]
       python bench.py --q2 --res 64 --ppc 20 [--modes fp64,mixed]
       python bench.py --learned-eval 64 --res 128
       python bench.py --convection --res 96 --ppc 25 --steps 40 [--modes fp64,mixed,learned]
           [--series-out F.json] [--compare-series F.json]
"""
import argparse
import ctypes
import json
import os
import sys
import time
import warnings

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_problem(res, ppc, seed=0):
    import underworld2_amd as uw
    fn = uw.function
    dim = len(res)
    mesh = uw.mesh.FeMesh_Cartesian("Q1/dQ0", res, (0.0,) * dim, (1.0,) * dim)
    vel = uw.mesh.MeshVariable(mesh, dim)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    swarm = uw.swarm.Swarm(mesh)
    t0 = time.time()
    swarm.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(swarm, particlesPerCell=ppc, seed=seed))
    mat = swarm.add_variable("int", 1)
    centre = (0.5, 0.5, 0.6) if dim == 3 else (0.5, 0.6)
    c = fn.input() - centre
    mat.data[:] = fn.branching.conditional([(fn.math.dot(c, c) < 0.1 ** 2, 1), (True, 0)]).evaluate(swarm)
    log(f"populate + materials: {time.time() - t0:.2f}s ({swarm.particleLocalCount} particles)")
    eta = fn.branching.map(fn_key=mat, mapping={0: 1.0, 1: 1.0e3})
    rho = fn.branching.map(fn_key=mat, mapping={0: 0.0, 1: 1.0})
    g = (0.0, 0.0, -1.0) if dim == 3 else (0.0, -1.0)
    names = "IJK"
    walls = tuple(mesh.specialSets[f"Min{names[d]}_VertexSet"] + mesh.specialSets[f"Max{names[d]}_VertexSet"]
                  for d in range(dim))
    bc = uw.conditions.DirichletCondition(vel, walls)
    stokes = uw.systems.Stokes(vel, p, fn_viscosity=eta, fn_bodyforce=rho * g, voronoi_swarm=swarm, conditions=[bc])
    t0 = time.time()
    swarm.integration_points(dvc=True)
    log(f"coincident map + DVC weights: {time.time() - t0:.2f}s")
    return dict(uw=uw, mesh=mesh, vel=vel, p=p, swarm=swarm, stokes=stokes)


def make_solver(P, mode, rtol):
    s = P["uw"].systems.Solver(P["stokes"])
    if mode == "schur":
        s.set_method("schur")
        s.set_precision("fp64")
        s.set_inner_rtol(rtol * 0.1)
        s.set_outer_rtol(rtol)
    else:
        s.set_method("block")
        s.set_precision({"fp64": "fp64", "mixed-cycle": "mixed-cycle"}.get(mode, "mixed"))
        s.set_outer_rtol(rtol)
        if mode in ("learned", "learned-coarse"):
            from underworld2_amd import learned
            model = learned.load_default_schur() if mode == "learned" else learned.load_default(P["stokes"])
            if model is None:
                return None
            s.set_learned(model)
    return s


def time_mode(P, mode, steps, warmup, rtol):
    import underworld2_amd._lib as L
    s = make_solver(P, mode, rtol)
    if s is None:
        return None
    for _ in range(warmup):
        s.solve()
    L.lib()  # noqa
    t0 = time.perf_counter()
    stats = []
    for _ in range(steps):
        s.solve()
        stats.append(s.get_stats().as_dict())
    t = time.perf_counter() - t0
    st = stats[-1]
    rel = float(np.hypot(st["res_momentum"], st["res_continuity"]) / st["norm_f"])
    out = dict(mode=mode, seconds_per_solve=t / steps, solves_per_sec=steps / t, outer_its=st["outer_its"],
               pressure_its=st["pressure_its"], velocity_its=st["velocity_total_its"], vcycles=st["vcycles"],
               levels=st["levels"], setup_s=st["setup_time"], solve_s=st["solve_time"], rel_residual_fp64=rel,
               converged=bool(st["converged"]), fallback_used=int(st.get("fallback_used", 0)))
    sol = (np.array(P["vel"].data), np.array(P["p"].data[:, 0]))
    return out, sol


def reference_solution(P, rtol=1e-11, restart=200, max_its=6000):
    """FP64 reference: the coupled solve in FP64 throughout (FP64 cycle, FP64 Krylov, restart
    200) with the learned Schur stencil, to rtol 1e-11.  At 128^3 the plain approxS FP64 solve
    stagnates near 1e-8 (a few percent per 250 iterations), the learned one keeps a steady rate
    (about 3x per 250 iterations near 1e-10).  Q1/dQ0 pressure is ill-conditioned: solutions
    at residual 1e-8 still differ by ~1e-1 in pressure, so the reference must be tight.  Without
    a trained model the plain FP64 solve is used with the same settings.  Both FP64 paths are
    tied to the oracle's converged solution by tests/test_gpu_sinker32.py at 32^3."""
    uw = P["uw"]
    from underworld2_amd import learned
    model = learned.load_default_schur()
    s = uw.systems.Solver(P["stokes"])
    s.set_method("block")
    s.set_precision("fp64")
    s.set_outer_rtol(rtol)
    s.set_restart(restart)
    s.options.scr.ksp_max_it = max_its
    s.set_learned(model)
    import underworld2_amd._lib as L
    # no fallback for the reference: the stencil keeps converging in the slow tail, where the
    # restarts' true-residual steps can look like a stall to the monitor
    L.check(L.lib().uwg_stokes_set_nn_monitor(P["stokes"]._h, 1 << 30, 0.01), "monitor")
    L.check(L.lib().uwg_stokes_set_nn_restart_rise(P["stokes"]._h, 0.0), "monitor")
    t0 = time.perf_counter()
    try:
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            s.solve()
    finally:
        L.check(L.lib().uwg_stokes_set_nn_monitor(P["stokes"]._h, 20, 0.01), "monitor")
        L.check(L.lib().uwg_stokes_set_nn_restart_rise(P["stokes"]._h, 1.25), "monitor")
    t = time.perf_counter() - t0
    st = s.get_stats()
    rel = float(np.hypot(st.res_momentum, st.res_continuity) / st.norm_f)
    return (np.array(P["vel"].data), np.array(P["p"].data[:, 0])), {
        "solver": "coupled FP64, " + ("learned stencil" if model is not None else "approxS"), "restart": restart,
        "rtol": rtol, "outer_its": int(st.outer_its), "converged": bool(st.converged), "rel_residual_fp64": rel,
        "fallback_used": int(st.fallback_used), "seconds": t}


def iteration_floor(P, rtol=1e-6, vel_rtol=1e-8):
    """Outer FGMRES count of the coupled FP64 solve when the preconditioner's velocity block is
    solved (inner FGMRES + V-cycle to vel_rtol) instead of one V-cycle: the count that the Schur
    block alone sets, for the plain approxS diagonal and for the learned stencil.  When it sits
    near the V-cycle count, the Schur block, not the velocity cycle, is what sets the iterations."""
    from underworld2_amd import learned
    out = {"velocity_block_rtol": vel_rtol}
    model = learned.load_default_schur()
    for name, m in (("approxS", None), ("learned_stencil", model)):
        if name == "learned_stencil" and model is None:
            continue
        s = P["uw"].systems.Solver(P["stokes"])
        s.set_method("block")
        s.set_precision("fp64")
        s.set_outer_rtol(rtol)
        s.set_learned(m)
        s.set_velocity_block_solve(vel_rtol)
        try:
            t0 = time.perf_counter()
            s.solve()
            t = time.perf_counter() - t0
        finally:
            s.set_velocity_block_solve(0.0)
        st = s.get_stats()
        out[name] = {"outer_its": int(st.outer_its), "vcycles": int(st.vcycles), "seconds": t,
                     "converged": bool(st.converged), "fallback_used": int(st.fallback_used)}
    return out


def solution_error(sol, ref):
    (u, p), (ur, pr) = sol, ref
    pm, prm = p - p.mean(), pr - pr.mean()
    return {"velocity_rel_l2": float(np.linalg.norm(u - ur) / np.linalg.norm(ur)),
            "pressure_rel_l2_mean_free": float(np.linalg.norm(pm - prm) / np.linalg.norm(prm))}


def per_solve_roofline(P, mode, r, restart=60):
    """Algorithmic HBM bytes of one solve from its logged counts, divided by the solve time and
    the 8 TB/s peak (the per-solve roofline of the path).  Per FGMRES iteration: the outer FP64
    operator (K: E 27 8 + 2 3 N 8 bytes; G and G^T: E 8 + 3 N 8 each), one V-cycle (every
    level: 5 K applies -- Chebyshev(2) before and after the coarse correction and one residual
    -- with bf16 weights and FP32 vectors in the mixed modes, FP64 otherwise, plus restriction
    and prolongation of FP32/FP64 vectors), the Schur block (diagonal: 3 E s; learned stencil:
    27 E 4 + 2 E 8), and classical Gram-Schmidt against the j + 1 basis vectors of the cycle's
    step j (read twice: dots and update; FP32 basis in the mixed modes) plus w and the new basis
    vector.  Setup: the particle-to-weights assembly (P (3 8 + 8 + 4) + E (27 + 8 + 1) 8)."""
    mesh = P["mesh"]
    res = np.array(mesh.elementRes)
    mixed = mode not in ("fp64", "schur")
    s_v, s_w = (4, 2) if mixed else (8, 8)
    E, N = int(np.prod(res)), int(np.prod(res + 1))
    n = 3 * N + E
    kb = lambda e, nn, sw, sv: e * 27 * sw + 2 * 3 * nn * sv
    outer = kb(E, N, 8, 8) + 2 * (E * 8 + 3 * N * 8)
    vc = 0
    lev = res.copy()
    for l in range(int(r["levels"])):
        e, nn = int(np.prod(lev)), int(np.prod(lev + 1))
        if l < int(r["levels"]) - 1:
            vc += 5 * kb(e, nn, s_w, s_v)
            nc = int(np.prod(lev // 2 + 1))
            vc += 2 * (3 * nn * s_v + 3 * nc * s_v)
        lev = lev // 2
    schur = 27 * E * 4 + 2 * E * 8 if mode == "learned" else 3 * E * 8
    its = int(r["outer_its"])
    s_b = 4 if mixed else 8
    gs = 0
    for k in range(its):
        j = k % restart
        gs += 2 * (j + 1) * n * s_b + 3 * n * 8
    setup = int(P["swarm"].particleLocalCount) * (3 * 8 + 8 + 4) + E * (27 + 8 + 1) * 8
    total = its * (outer + vc + schur) + gs + setup
    t = r["seconds_per_solve"]
    return {"bytes": total, "bytes_per_iteration": (total - setup) / max(1, its), "achieved_GB_per_s": total / t / 1e9,
            "frac": total / t / HBM_PEAK, "gram_schmidt_share_of_bytes": gs / total}


def roofline(P, precision, reps=50):
    """Average duration of the K apply (the dominant kernel) measured with HIP events on the
    stream it runs on.  precision 0: FP64 apply (k_apply3<double>); 1: the FP32 multigrid
    cycle's apply (k_apply3f with bfloat16 weights on 3D levels); 2: FP32 apply with FP32
    weights.  Algorithmic bytes = E*27*s_w (weights) + 2*3*N*s_v (read x, write y) with
    s_w = 8 / 2 / 4 and s_v = 8 / 4 / 4.  ``traffic`` (HBM bytes per launch) comes from the
    FETCH_SIZE / WRITE_SIZE counter runs summarised in profiles/k_apply3_sinker128_traffic.json
    (128^3 only)."""
    import underworld2_amd._lib as L
    st = P["stokes"]
    ms = ctypes.c_double()
    L.check(L.lib().uwg_stokes_time_apply(st._h, precision, 0, reps, ctypes.byref(ms)), "time_apply")
    mesh = P["mesh"]
    E, N, d = mesh.elementsGlobal, mesh.nodesGlobal, mesh.dim
    bf16 = precision == 1 and d == 3
    s_w = {0: 8, 1: 2 if bf16 else 4, 2: 4}[precision]
    s_v = 8 if precision == 0 else 4
    Q = 3 ** d
    bytes_ = E * Q * s_w + 2 * d * N * s_v
    ach = bytes_ / (ms.value * 1e-3)
    key = {0: "fp64", 1: "bf16" if bf16 else "fp32", 2: "fp32"}[precision]
    kernel = {0: "k_apply3<double>", 1: "k_apply3f<bf16 weights>" if bf16 else "k_apply3f<fp32 weights>",
              2: "k_apply3f<fp32 weights>"}[precision]
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "k_apply3_sinker128_traffic.json")
    if os.path.exists(tfile) and mesh.elementRes == (128, 128, 128):
        with open(tfile) as f:
            t = json.load(f)
        traffic = t.get(key, {}).get("traffic_bytes")
    return {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": ach / HBM_PEAK,
            "traffic": traffic, "kernel": kernel, "kernel_ms": ms.value, "algorithmic_bytes": bytes_}


def cpu_baseline(res_sample=48, ppc=30, rtol=1e-6, max_outer=40, schur_its_128=4635):
    """The CPU restatement of the reference's solve (oracle/cpu_baseline.c: assembled CSR K, G,
    approxS from the particle quadrature; Schur FGMRES with the approxS preconditioner; inner
    FGMRES + Galerkin multigrid with the SROpGenerator prolongation; the same rtols), OpenMP on
    every core this process may use, on a bounded sample of the same sinker: the 48^3 sinker
    (30 ppc), element integration + CSR assembly timed in full, the Schur solve for its first
    ``max_outer`` outer iterations.  Scaled to 128^3 by the element ratio (assembly) and by the
    DOF ratio times the outer iteration count of the same algorithm at 128^3 (``schur_its_128``:
    4,635 pressure iterations, measured with the GPU Schur path, which runs the reference's
    algorithm) -- the per-iteration cost of the inner multigrid solves grows with the DOFs."""
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from sinker import sphere_material
    res = (res_sample,) * 3
    mesh = O.Mesh(res, (0.0,) * 3, (1.0,) * 3)
    coords = O.layout_random(mesh, ppc, 0)
    mat = sphere_material(coords, (0.5, 0.5, 0.6), 0.1)
    ptr, pid = O.initial_cells(mesh, ppc)
    xi = O.coincident_map(mesh, ptr, pid, coords)
    xw, w = O.dvc_all(3, ptr, xi)
    eta = np.array([1.0, 1.0e3])[mat[pid]]
    force = np.array([0.0, 1.0])[mat[pid]][:, None] * np.array([0.0, 0.0, -1.0])[None, :]
    t0 = time.perf_counter()
    Ke, fe, Ge, S = O.assemble_elements(mesh, ptr, xw, w, eta, force)
    sysm = O.System(mesh, mesh.freeslip_mask(), Ke, fe, Ge, S)
    t1 = time.perf_counter()
    u, p, st = sysm.bsscr_solve(rtol, rtol, max_outer=max_outer)
    t2 = time.perf_counter()
    dof = sysm.nfree + mesh.E
    dof128 = 6340221 + 128 ** 3
    per_outer = (t2 - t1) / max(1, st["outer_its"])
    t128 = (t1 - t0) * (128 ** 3 / mesh.E) + per_outer * (dof128 / dof) * schur_its_128
    try:
        cpu = [l for l in open("/proc/cpuinfo") if l.startswith("model name")][0].split(":", 1)[1].strip()
    except Exception:
        cpu = "unknown"
    return {"value": 1.0 / t128, "unit": "solves/s", "cores": st["threads"], "kind": "port",
            "sample": f"CPU restatement (not PETSc) of the reference solve on the {res_sample}^3 sinker, {ppc} ppc: "
                      f"assembly {t1 - t0:.2f}s + the first {st['outer_its']} Schur iterations {t2 - t1:.2f}s "
                      f"({st['inner_its']} inner its, {st['levels']} levels) on {st['threads']} OpenMP threads "
                      f"({cpu}); scaled to 128^3 by the element ratio (assembly) and the DOF ratio x "
                      f"{schur_its_128} Schur iterations (the same algorithm's count at 128^3)",
            "seconds_per_solve_128_estimate": t128, "seconds_per_outer_iteration_sample": per_outer}


def config1(ppc=20, rtol=1e-5):
    """The 2D 64^2 sinker of the plain-method configuration (20 ppc, seed 0, inclusion r 0.1 at
    (0.5, 0.6): viscosity 100, density 1; background 1 / 0; gravity (0, -1)): the Schur solve to
    rtol 1e-5 and one RK2 step at dt = 0.5 get_max_dt(), timed on the GPU (FP64, the reference's
    algorithm) and in the CPU restatement (element integration + assembly + Schur solve + RK2;
    layout and DVC excluded on both)."""
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from sinker import build_oracle, build_uw
    g = build_uw((64, 64), ppc, 0, (0.5, 0.6), 0.1)
    solver = g["uw"].systems.Solver(g["stokes"])
    solver.set_outer_rtol(rtol)
    g["swarm"].integration_points(dvc=True)
    solver.solve()  # warm-up (first-call set-up of the handle)
    adv = g["uw"].systems.SwarmAdvector(g["vel"], g["swarm"], order=2)
    t0 = time.perf_counter()
    solver.solve()
    dt = 0.5 * adv.get_max_dt()
    adv.integrate(dt)
    g["swarm"].particleCoordinates.data  # noqa: B018 (sync)
    tg = time.perf_counter() - t0
    gst = solver.get_stats()
    o = build_oracle((64, 64), ppc, 0, (0.5, 0.6), 0.1, solve=False)
    m = o["mesh"]
    t0 = time.perf_counter()
    Ke, fe, Ge, S = O.assemble_elements(m, o["ptr"], o["xi"], o["w"], o["eta"], o["force"])
    sysm = O.System(m, m.freeslip_mask(), Ke, fe, Ge, S)
    u, p, st = sysm.bsscr_solve(rtol, 1e-6)
    dtc = 0.5 * O.max_dt(m, u)
    O.advect_rk2(m, u, o["coords"], dtc)
    tc = time.perf_counter() - t0
    return {"workload": "2D 64^2 sinker, 20 ppc, Schur solve rtol 1e-5 + one RK2 step",
            "gpu_seconds": tg, "gpu_pressure_its": int(gst.pressure_its), "cpu_seconds": tc,
            "cpu_outer_its": st["outer_its"], "cpu_threads": st["threads"], "cpu_kind": "port"}


def time_advection(P, steps=10):
    """RK2 advection + relocation steps at dt = 0.5 * get_max_dt() and one DVC pass afterwards
    (the integration-point refresh the next solve needs)."""
    import underworld2_amd._lib as L
    uw = P["uw"]
    adv = uw.systems.SwarmAdvector(P["vel"], P["swarm"], order=2)
    dt = 0.5 * adv.get_max_dt()
    t0 = time.perf_counter()
    for _ in range(steps):
        adv.integrate(dt)
    t1 = time.perf_counter()
    sw = P["swarm"]
    sw._push()
    t2 = time.perf_counter()
    L.check(L.lib().uwg_swarm_integration_points(sw._h, 1, 15), "DVC")
    t3 = time.perf_counter()
    return {"rk2_steps": steps, "dt": dt, "ms_per_rk2_step": (t1 - t0) / steps * 1e3, "ms_per_dvc_pass": (t3 - t2) * 1e3,
            "particles": int(sw.particleLocalCount)}


METRIC = "Stokes solves/sec at 128\u00b3 Q1/dQ0, \u03b7-contrast 1e3, rtol 1e-6; learned vs plain"


def nusselt_3d(mesh, T):
    """-int_top dT/dz dA / int_bottom T dA for a nodal Q1 temperature (exact for the trilinear
    field: on a face the normal derivative is bilinear, its integral the mean of the corners)."""
    nx, ny, nz = (r + 1 for r in mesh.elementRes)
    Tg = T.reshape(nz, ny, nx)
    X = mesh.data
    x = X[:nx, 0]
    y = X[::nx, 1][:ny]
    z = X[::nx * ny, 2][:nz]
    A = np.outer(np.diff(y), np.diff(x))
    d = (Tg[-1] - Tg[-2]) / (z[-1] - z[-2])
    top = np.sum(A * 0.25 * (d[1:, 1:] + d[1:, :-1] + d[:-1, 1:] + d[:-1, :-1]))
    b = Tg[0]
    bot = np.sum(A * 0.25 * (b[1:, 1:] + b[1:, :-1] + b[:-1, 1:] + b[:-1, :-1]))
    return float(-top / bot)


def build_convection(res, ppc, seed=0, Ra=1.0e6):
    """3D thermal convection: unit box, free-slip walls, T = 1 at the bottom and 0 at the top,
    eta = exp(-ln(1e4) T) and buoyancy Ra T z evaluated at the particles, initial T = 1 - z plus
    a seeded 1e-2 sinusoidal perturbation, kappa = 1."""
    import underworld2_amd as uw
    fn = uw.function
    mesh = uw.mesh.FeMesh_Cartesian("Q1/dQ0", res, (0.0,) * 3, (1.0,) * 3)
    vel = uw.mesh.MeshVariable(mesh, 3)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    T = uw.mesh.MeshVariable(mesh, 1)
    Tdot = uw.mesh.MeshVariable(mesh, 1)
    vel.data[:] = 0.0
    p.data[:] = 0.0
    Tdot.data[:] = 0.0
    X = mesh.data
    rng = np.random.default_rng(seed)
    kx, ky = rng.integers(1, 3, 2)
    phase = rng.uniform(0.0, 2.0 * np.pi)
    pert = np.cos(kx * np.pi * X[:, 0] + phase) * np.cos(ky * np.pi * X[:, 1]) * np.sin(np.pi * X[:, 2])
    T.data[:, 0] = np.clip(1.0 - X[:, 2] + 1.0e-2 * pert, 0.0, 1.0)
    bot, top = mesh.specialSets["MinK_VertexSet"], mesh.specialSets["MaxK_VertexSet"]
    T.data[bot.data, 0] = 1.0
    T.data[top.data, 0] = 0.0
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(swarm, particlesPerCell=ppc, seed=seed))
    walls = tuple(mesh.specialSets[f"Min{n}_VertexSet"] + mesh.specialSets[f"Max{n}_VertexSet"] for n in "IJK")
    stokes = uw.systems.Stokes(vel, p, fn_viscosity=fn.math.exp(-np.log(1.0e4) * T),
                               fn_bodyforce=Ra * T * (0.0, 0.0, 1.0), voronoi_swarm=swarm,
                               conditions=[uw.conditions.DirichletCondition(vel, walls)])
    ad = uw.systems.AdvectionDiffusion(phiField=T, phiDotField=Tdot, velocityField=vel, fn_diffusivity=1.0,
                                       conditions=[uw.conditions.DirichletCondition(T, (bot + top,))])
    adv = uw.systems.SwarmAdvector(vel, swarm, order=2)
    return dict(uw=uw, mesh=mesh, vel=vel, p=p, T=T, Tdot=Tdot, swarm=swarm, stokes=stokes, ad=ad, adv=adv)


def run_convection(C, steps, mode, rtol=1e-6, log_every=0):
    """Each step (the loop of the reference's convection examples): Stokes solve from the current
    T (mode fp64 | mixed | learned); dt = min of the advection-diffusion and swarm limits for
    the new velocity; SUPG step of T; RK2 swarm advection.  Returns per-step series and timings."""
    s = make_solver({"uw": C["uw"], "stokes": C["stokes"]}, mode, rtol)
    if s is None:
        return None
    out = dict(mode=mode, t=[], dt=[], nu=[], vrms=[], its=[], s_supg=0.0, s_stokes=0.0, s_adv=0.0)
    t = 0.0
    for k in range(steps):
        t0 = time.perf_counter()
        s.solve()
        st = s.get_stats()
        t1 = time.perf_counter()
        dt = min(C["ad"].get_max_dt(), C["adv"].get_max_dt())
        t2 = time.perf_counter()
        C["ad"].integrate(dt)
        C["T"].data  # noqa: B018 (sync point for the timing)
        t3 = time.perf_counter()
        C["adv"].integrate(dt)
        C["swarm"].particleCoordinates.data  # noqa: B018
        t4 = time.perf_counter()
        t += dt
        out["t"].append(t)
        out["dt"].append(dt)
        out["nu"].append(nusselt_3d(C["mesh"], C["T"].data[:, 0]))
        out["vrms"].append(C["stokes"].velocity_rms())
        out["its"].append(int(st.outer_its))
        out["s_stokes"] += t1 - t0
        out["s_supg"] += t3 - t2
        out["s_adv"] += t4 - t3
        if log_every and (k % log_every == 0 or k == steps - 1):
            log(f"  {mode} step {k}: t {t:.4e} Nu {out['nu'][-1]:.5f} v_rms {out['vrms'][-1]:.4f} its {st.outer_its}")
    return out


def stream_copy(n=1 << 27, reps=20):
    """Attainable HBM bandwidth: a device copy of n doubles (read + write 16 n bytes)."""
    import underworld2_amd._lib as L
    ms = ctypes.c_double()
    L.check(L.lib().uwg_stream_copy(n, reps, ctypes.byref(ms)), "stream_copy")
    return {"bytes_per_copy": 16 * n, "ms": ms.value, "GB_per_s": 16 * n / (ms.value * 1e-3) / 1e9}


def apply_only(P):
    """K applies alone (FP32 and FP64) plus the library's plain copy kernel of known size, for
    hardware counter runs (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) that calibrate HBM traffic."""
    out = {"bf16": roofline(P, 1, reps=20), "fp32": roofline(P, 2, reps=20), "fp64": roofline(P, 0, reps=20),
           "copy": stream_copy(1 << 27, 5)}
    print(json.dumps(out), flush=True)


def advection_only(P):
    """One mixed-precision solve for the velocity, then the RK2 + relocation steps and the DVC
    pass of time_advection, plus the library's plain copy kernel of known size, for kernel-trace
    and hardware counter runs (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) of the advection step."""
    s = make_solver(P, "mixed", 1e-6)
    s.solve()
    out = {"advection": time_advection(P), "copy": stream_copy(1 << 27, 5)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--res", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=30)
    ap.add_argument("--rtol", type=float, default=1e-6)
    ap.add_argument("--mode", default="auto", help="mode whose throughput is the headline value")
    ap.add_argument("--modes", default="fp64,mixed,learned",
                    help="comma list of fp64 (coupled FGMRES, FP64), mixed (FP32 cycle in FP64 Krylov), "
                         "learned (mixed + learned Schur stencil), learned-coarse (mixed + the experimental "
                         "learned coarse correction), schur (the reference's Schur complement algorithm in "
                         "FP64; minutes at 128^3)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-advection", action="store_true")
    ap.add_argument("--no-reference", action="store_true", help="skip the FP64 reference solve and error report")
    ap.add_argument("--no-floor", action="store_true",
                    help="skip the iteration-floor solves (velocity block solved instead of one V-cycle)")
    ap.add_argument("--apply-only", action="store_true")
    ap.add_argument("--advection-only", action="store_true",
                    help="one solve, then 10 RK2 steps + a DVC pass and a copy kernel (for counter runs)")
    ap.add_argument("--q2", action="store_true",
                    help="Q2/dPc1 3D sinker: K apply (matrix cores / vector) and solves at --res, --ppc")
    ap.add_argument("--learned-eval", type=int, default=0, metavar="N",
                    help="evaluate the learned Schur stencil on N unseen random --res^3 problems (FP64, mixed, "
                         "learned + mixed per field) instead")
    ap.add_argument("--learned-eval-start", type=int, default=0, metavar="K",
                    help="learned-eval: first field index (seed 700000 + K), to split the fields over runs")
    ap.add_argument("--floor-fields", type=int, default=8,
                    help="learned-eval: fields on which the iteration floor is also measured")
    ap.add_argument("--convection", action="store_true",
                    help="run the 3D thermal-convection time series (--res, --ppc, --steps, --modes) instead")
    ap.add_argument("--series-out", default=None,
                    help="convection: write each mode's per-step series (t, dt, Nu, v_rms, its) to this JSON file")
    ap.add_argument("--compare-series", default=None,
                    help="convection: FP64 series from an earlier --series-out run to compare against (runs "
                         "too long for one process can be split by mode)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the engine runs on one GPU; --gpus must be 1 (multi-GPU sharding of the "
                                   "element grid is not implemented)"}), flush=True)
        sys.exit(2)
    import underworld2_amd as uw
    log(uw.device_info())
    if args.convection:
        convection_main(args)
        return
    if args.q2:
        q2_main(args)
        return
    if args.learned_eval:
        learned_eval_main(args)
        return
    res = (args.res,) * 3
    P = build_problem(res, args.ppc)
    if args.apply_only:
        P["stokes"]._prepare_values()
        apply_only(P)
        return
    if args.advection_only:
        advection_only(P)
        return
    modes = {}
    sols = {}
    for m in args.modes.split(","):
        m = m.strip()
        if not m:
            continue
        log(f"timing mode {m} ...")
        r = time_mode(P, m, args.steps, args.warmup, args.rtol)
        if r is not None:
            modes[m], sols[m] = r
            modes[m]["per_solve_roofline"] = per_solve_roofline(P, m, modes[m])
            log(json.dumps(modes[m]))
    reference = None
    if not args.no_reference:
        log("FP64 reference solution (coupled FP64 with the learned stencil, rtol 1e-11) ...")
        ref, reference = reference_solution(P)
        for m in modes:
            modes[m]["error_vs_fp64_reference"] = solution_error(sols[m], ref)
        log(json.dumps(reference))
    floor = None
    if not args.no_floor:
        log("iteration floor (coupled FP64, velocity block solved to 1e-8) ...")
        floor = iteration_floor(P, args.rtol)
        log(json.dumps(floor))
    head = args.mode
    if head == "auto":
        head = "learned" if "learned" in modes else ("mixed" if "mixed" in modes else next(iter(modes)))
    H = modes[head]
    roof = roofline(P, 0 if head in ("fp64", "schur") else 1)
    roof["stream_copy_GB_per_s"] = stream_copy()["GB_per_s"]
    # the same apply with FP32 weights and the FP64 apply of the outer Krylov iteration
    others = {"fp32_weights": roofline(P, 2), "fp64": roofline(P, 0)} if head not in ("fp64", "schur") else {}
    adv = None if args.no_advection else time_advection(P)
    cpu = None
    cfg1 = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline()
        except Exception as e:  # the baseline is reported, never required
            cpu = {"error": str(e)}
        try:
            cfg1 = config1()
        except Exception as e:
            cfg1 = {"error": str(e)}
    dtype = "f64" if head in ("fp64", "schur") else "f32+f64"
    line = {
        "metric": METRIC, "value": H["solves_per_sec"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": H["seconds_per_solve"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic: seeded PerCellRandomLayout sinker (sphere r=0.1 at (0.5,0.5,0.6), eta 1e3, rho 1)",
        "config": {"workload": f"3D sinker {args.res}^3 Q1/dQ0, {args.ppc} ppc, free-slip, outer rtol {args.rtol}, "
                               f"end-to-end solve() on a fixed swarm (particle-quadrature assembly + solve), "
                               f"mode={head}" + ("" if head in ("fp64", "schur") else
                                                 " (FP32 cycle with bf16 element weights, FP64 Krylov)")},
        "roofline": roof, "cpu_baseline": cpu, "modes": modes, "advection": adv, "fp64_reference": reference,
        "iteration_floor": floor,
        "config1_plain_method": cfg1,
    }
    if others:
        line["roofline_other_applies"] = others
    for plain in ("fp64", "schur", "mixed"):
        if plain in modes and plain != head:
            line[f"speedup_vs_{plain}"] = H["solves_per_sec"] / modes[plain]["solves_per_sec"]
    print(json.dumps(line), flush=True)


def learned_eval_main(args):
    """The learned Schur stencil on unseen random problems: GRF viscosity 10^(c g), c ~ U(0, 6),
    correlation length ~ U(0.1, 0.3), density 1 in 1-5 random spheres (learned.random_problem,
    seeds 700000 + k, disjoint from the training seeds 5000-5511), at --res.  Each field is
    assembled once (Gauss quadrature; the host evaluation of the random fields is not timed) and
    solved to outer rtol --rtol by the plain FP64 coupled solver, the mixed-precision one, and the
    mixed one with the learned stencil (with the plain approxS preconditioner as its fallback);
    the solver time is the library's own (set-up + Krylov solve) of a second solve in the same
    precision (the first allocates that precision's Krylov basis); the learned solve, which runs
    after the mixed ones, is timed on its first call and includes the network's forward pass
    (once per assembly).  On the first --floor-fields
    fields the coupled FP64 solve is repeated with the velocity block solved (inner solve to
    1e-8), with and without the stencil: the iteration floor the pressure block sets."""
    from underworld2_amd import learned
    model = learned.load_default_schur()
    if model is None:
        raise SystemExit("no trained Schur-stencil model")
    rows = []
    for k in range(args.learned_eval_start, args.learned_eval_start + args.learned_eval):
        seed = 700000 + k
        t0 = time.perf_counter()
        st, v, p, desc = learned.random_problem(args.res, seed)
        st._prepare_values()
        row = dict(desc, build_s=time.perf_counter() - t0)
        for name, prec, m in (("fp64", "fp64", None), ("mixed", "mixed", None), ("learned", "mixed", model)):
            s = P_solver(st, prec, args.rtol, m)
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                if name != "learned":
                    # untimed: a solve in a new precision allocates that precision's Krylov basis
                    # (and frees the other's); the learned solve follows the mixed one, and is
                    # timed on its first call, which includes the network's forward pass
                    P_solver(st, prec, args.rtol, None).solve(reinitialise=False)
                s.solve(reinitialise=False)
            stt = s.get_stats()
            row[name] = {"its": int(stt.outer_its), "s": float(stt.setup_time + stt.solve_time),
                         "converged": bool(stt.converged),
                         "rel_residual_fp64": float(np.hypot(stt.res_momentum, stt.res_continuity) / stt.norm_f),
                         "fallback_used": int(stt.fallback_used)}
        if k < args.floor_fields:
            for name, m in (("floor_approxS", None), ("floor_learned", model)):
                s = P_solver(st, "fp64", args.rtol, m)
                s.set_velocity_block_solve(1e-8)
                try:
                    s.solve(reinitialise=False)
                finally:
                    s.set_velocity_block_solve(0.0)
                row[name] = int(s.get_stats().outer_its)
        rows.append(row)
        fl = f", floor {row['floor_approxS']} / {row['floor_learned']}" if "floor_approxS" in row else ""
        log(f"field {k} seed {seed} c={desc['log10_contrast']:.2f} ell={desc['corr_len']:.2f} "
            f"spheres={desc['spheres']}: its fp64 {row['fp64']['its']} mixed {row['mixed']['its']} "
            f"learned {row['learned']['its']} (fallback {row['learned']['fallback_used']}){fl}, solver s "
            f"{row['fp64']['s']:.3f} / {row['mixed']['s']:.3f} / {row['learned']['s']:.3f}")
        del st, v, p
    its_ratio = np.array([r["learned"]["its"] / r["mixed"]["its"] for r in rows])
    c = np.array([r["log10_contrast"] for r in rows])
    summ = {
        "fields": len(rows), "res": args.res, "rtol": args.rtol,
        "all_converged": all(r[m]["converged"] for r in rows for m in ("fp64", "mixed", "learned")),
        "max_rel_residual_fp64": max(r[m]["rel_residual_fp64"] for r in rows for m in ("fp64", "mixed", "learned")),
        "fallbacks": int(sum(r["learned"]["fallback_used"] for r in rows)),
        "mean_its": {m: float(np.mean([r[m]["its"] for r in rows])) for m in ("fp64", "mixed", "learned")},
        "its_ratio_learned_over_mixed": {"median": float(np.median(its_ratio)), "mean": float(its_ratio.mean()),
                                         "min": float(its_ratio.min()), "max": float(its_ratio.max()),
                                         "mean_log10_contrast_ge_4": float(its_ratio[c >= 4].mean())
                                         if np.any(c >= 4) else None},
        "solver_s": {m: float(sum(r[m]["s"] for r in rows)) for m in ("fp64", "mixed", "learned")},
    }
    fl = [r for r in rows if "floor_approxS" in r]
    if fl:
        summ["floor_fields"] = len(fl)
        summ["floor_mean_its"] = {"vcycle_approxS_fp64": float(np.mean([r["fp64"]["its"] for r in fl])),
                                  "solved_velocity_approxS": float(np.mean([r["floor_approxS"] for r in fl])),
                                  "solved_velocity_learned": float(np.mean([r["floor_learned"] for r in fl]))}
    summ["speedup_learned_vs_fp64"] = summ["solver_s"]["fp64"] / summ["solver_s"]["learned"]
    summ["speedup_learned_vs_mixed"] = summ["solver_s"]["mixed"] / summ["solver_s"]["learned"]
    print(json.dumps({"learned_eval": summ, "fields": rows}), flush=True)


def P_solver(st, precision, rtol, model):
    import underworld2_amd as uw
    s = uw.systems.Solver(st)
    s.set_method("block")
    s.set_precision(precision)
    s.set_outer_rtol(rtol)
    s.set_learned(model)
    return s


def q2_main(args):
    """Q2/dPc1 3D sinker (sphere r 0.1, eta 1e3, particle quadrature, free slip): the Q2 K apply by
    the sum-factorised, matrix-core and dense vector kernels (FP64 and FP32, HIP-event timed), and
    one coupled solve per precision mode with the default (sum-factorised) kernel."""
    import underworld2_amd as uw
    import underworld2_amd._lib as L
    fn = uw.function
    res = (args.res,) * 3
    mesh = uw.mesh.FeMesh_Cartesian("Q2/dPc1", res, (0.0,) * 3, (1.0,) * 3)
    vel = uw.mesh.MeshVariable(mesh, 3)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(swarm, particlesPerCell=args.ppc, seed=0))
    mat = swarm.add_variable("int", 1)
    c = fn.input() - (0.5, 0.5, 0.6)
    mat.data[:] = fn.branching.conditional([(fn.math.dot(c, c) < 0.1 ** 2, 1), (True, 0)]).evaluate(swarm)
    walls = tuple(mesh.specialSets[f"Min{n}_VertexSet"] + mesh.specialSets[f"Max{n}_VertexSet"] for n in "IJK")
    stokes = uw.systems.Stokes(vel, p, fn_viscosity=fn.branching.map(fn_key=mat, mapping={0: 1.0, 1: 1.0e3}),
                               fn_bodyforce=fn.branching.map(fn_key=mat, mapping={0: 0.0, 1: 1.0}) * (0.0, 0.0, -1.0),
                               voronoi_swarm=swarm, conditions=[uw.conditions.DirichletCondition(vel, walls)])
    stokes._prepare_values()
    E = mesh.elementsGlobal
    out = {"workload": f"3D sinker {args.res}^3 Q2/dPc1, {args.ppc} ppc, particle quadrature", "apply": {}, "solve": {}}
    # flops per element of each kernel's algorithm: the sum-factorised passes (12,420
    # multiply-adds) plus the stress (18 per point); the dense contractions
    kflop = {2: 27.09, 1: 121.5, 0: 121.5}
    for kind, name in ((2, "sumfact"), (1, "mfma"), (0, "vector")):
        L.check(L.lib().uwg_set_q2_apply_kernel(kind), "kernel")
        for prec in (0, 1):
            ms = ctypes.c_double()
            L.check(L.lib().uwg_stokes_time_apply(stokes._h, prec, 0, 20, ctypes.byref(ms)), "time_apply")
            out["apply"][f"{name}_{'fp32' if prec else 'fp64'}"] = {
                "ms": ms.value, "kflop_per_element": kflop[kind],
                "TFLOP_per_s": E * kflop[kind] * 1e3 / (ms.value * 1e-3) / 1e12}
    L.check(L.lib().uwg_set_q2_apply_kernel(2), "kernel")
    for m in [x.strip() for x in args.modes.split(",") if x.strip() in ("fp64", "mixed")]:
        s = make_solver({"uw": uw, "stokes": stokes}, m, args.rtol)
        s.solve()
        t0 = time.perf_counter()
        s.solve()
        st = s.get_stats()
        out["solve"][m] = {"s": time.perf_counter() - t0, "outer_its": int(st.outer_its),
                           "rel_residual_fp64": float(np.hypot(st.res_momentum, st.res_continuity) / st.norm_f)}
    print(json.dumps(out), flush=True)


def convection_main(args):
    """3D thermal convection (Ra 1e6, eta = exp(-ln(1e4) T)): the same initial state run once per
    mode; Nu and v_rms series of each mode and their largest relative difference from the
    FP64 series."""
    res = (args.res,) * 3
    runs = {}
    for m in [x.strip() for x in args.modes.split(",") if x.strip()]:
        C = build_convection(res, args.ppc)
        log(f"convection {res} {args.ppc} ppc, mode {m}, {args.steps} steps")
        r = run_convection(C, args.steps, m, rtol=args.rtol, log_every=max(1, args.steps // 10))
        if r is not None:
            runs[m] = r
            if args.series_out:
                with open(args.series_out, "w") as f:
                    json.dump({k: {q: v[q] for q in ("t", "dt", "nu", "vrms", "its", "s_stokes", "s_supg", "s_adv")}
                               for k, v in runs.items()}, f)
    line = {"workload": f"3D thermal convection {args.res}^3 Q1/dQ0, Ra 1e6, eta = exp(-ln(1e4) T), {args.ppc} ppc, "
                        f"{args.steps} steps of SUPG + Stokes (rtol {args.rtol}) + RK2", "modes": {}}
    ref = runs.get("fp64")
    if ref is None and args.compare_series:
        with open(args.compare_series) as f:
            ref = json.load(f).get("fp64")
        if ref is not None:
            n = min(len(ref["nu"]), args.steps)
            ref = {"nu": ref["nu"][:n], "vrms": ref["vrms"][:n]}
            line["compared_with"] = f"FP64 series from {os.path.basename(args.compare_series)}"
    for m, r in runs.items():
        e = {"final_nu": r["nu"][-1], "final_vrms": r["vrms"][-1], "final_time": r["t"][-1],
             "mean_stokes_its": float(np.mean(r["its"])), "s_per_step": (r["s_supg"] + r["s_stokes"] + r["s_adv"]) / len(r["t"]),
             "s_stokes_per_step": r["s_stokes"] / len(r["t"]), "s_supg_per_step": r["s_supg"] / len(r["t"]),
             "s_rk2_per_step": r["s_adv"] / len(r["t"])}
        if ref is not None and m != "fp64":
            n = min(len(ref["nu"]), len(r["nu"]))
            a_nu, r_nu = np.array(r["nu"][:n]), np.array(ref["nu"][:n])
            a_v, r_v = np.array(r["vrms"][:n]), np.array(ref["vrms"][:n])
            e["max_rel_diff_nu_vs_fp64"] = float(np.max(np.abs(a_nu - r_nu) / np.abs(r_nu)))
            e["max_rel_diff_vrms_vs_fp64"] = float(np.max(np.abs(a_v - r_v) / np.abs(r_v)))
            e["final_rel_diff_nu_vs_fp64"] = float(abs(a_nu[-1] - r_nu[-1]) / abs(r_nu[-1]))
            e["final_rel_diff_vrms_vs_fp64"] = float(abs(a_v[-1] - r_v[-1]) / abs(r_v[-1]))
        line["modes"][m] = e
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
