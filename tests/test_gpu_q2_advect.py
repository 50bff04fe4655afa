"""Swarm advection through a Q2 velocity field on the GPU against the CPU oracle.

``SwarmAdvector(vel, swarm).integrate(dt)`` with ``vel`` on a Q2/dPc1 mesh: the lookup goes
through the Biquadratic / Triquadratic element types (Newton on the element's 3^d vertices),
so coordinates after RK2 must equal the oracle's bit for bit (both follow the reference's
operation order), and the relocated cell tables must match the oracle's swap-with-last /
append edits.  ``get_max_dt`` uses the Q2 mesh's minimum vertex separation.  A tensor-quadratic
velocity field is advected exactly (closed form).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(res, ppc, seed):
    import underworld2_amd as uw
    d = len(res)
    hi = (1.0, 0.8, 1.2)[:d]
    mesh = uw.mesh.FeMesh_Cartesian("Q2/dPc1", res, (0.0,) * d, hi)
    vel = uw.mesh.MeshVariable(mesh, d)
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(swarm, particlesPerCell=ppc, seed=seed))
    om = O.Mesh(res, (0.0,) * d, hi)
    return uw, mesh, vel, swarm, om


@pytest.mark.parametrize("res", [(6, 5), (4, 3, 3)])
def test_q2_advection_bitexact(res):
    print(f"Q2 advection {res}", flush=True)
    uw, mesh, vel, swarm, om = _setup(res, 6, 3)
    X = mesh.data
    d = len(res)
    assert np.array_equal(X, O.q2_vertex_coords(om))
    rng = np.random.default_rng(7)
    # cellular flow with zero normal velocity on the walls (particles stay inside), plus noise
    # that also vanishes on each component's own walls
    S = X / np.array(mesh.maxCoord)
    sx, sy = np.sin(np.pi * S[:, 0]), np.sin(np.pi * S[:, 1])
    comps = [0.8 * sx * np.cos(np.pi * S[:, 1]), -0.8 * np.cos(np.pi * S[:, 0]) * sy]
    if d == 3:
        comps.append(0.4 * sx * np.sin(np.pi * S[:, 2]))
    V = np.stack(comps, -1)
    for q in range(d):
        V[:, q] += 0.05 * np.sin(np.pi * S[:, q]) * rng.standard_normal(X.shape[0])
    vel.data[:] = V
    adv = uw.systems.SwarmAdvector(vel, swarm, order=2)
    dt = adv.get_max_dt()
    assert dt == O.max_dt_q2(om, V)
    dt *= 6.0  # a few particles cross several cells: exercises relocation
    c0 = swarm.particleCoordinates.data.copy()
    ptr0, pid0 = swarm.cell_tables()
    adv.integrate(dt)
    c1, bad, fb = O.advect_rk2_q2(om, V, c0, dt)
    print(f"{c0.shape[0]} particles, {fb} first-order fallbacks", flush=True)
    assert bad == 0
    got = swarm.particleCoordinates.data
    assert np.array_equal(got, c1)
    T = O.CellTables(om, ptr0, pid0)
    assert T.update(c1) == 0
    ptr, pid = swarm.cell_tables()
    eptr, epid = T.csr()
    assert np.array_equal(ptr, eptr) and np.array_equal(pid, epid)
    assert np.array_equal(swarm.owningCell.data[:, 0], T.owner)


@pytest.mark.parametrize("res", [(5, 4), (3, 3, 2)])
def test_q2_advection_quadratic_field_closed_form(res):
    uw, mesh, vel, swarm, om = _setup(res, 5, 11)
    d = len(res)
    rng = np.random.default_rng(8)
    c = rng.uniform(-0.5, 0.5, d)
    hi = np.array(mesh.maxCoord)

    def v(x):
        # tensor-quadratic (degree 2 in x_q, 1 in the next axis), zero normal velocity on the walls
        return np.array([c[q] * x[q] * (hi[q] - x[q]) * (1.0 + 0.5 * x[(q + 1) % d]) for q in range(d)])

    vel.data[:] = np.array([v(x) for x in mesh.data])
    c0 = swarm.particleCoordinates.data.copy()
    dt = 0.1
    uw.systems.SwarmAdvector(vel, swarm, order=2).integrate(dt)
    ref = np.array([p + dt * v(p + 0.5 * dt * v(p)) for p in c0])
    assert np.allclose(swarm.particleCoordinates.data, ref, rtol=0, atol=1e-13)
