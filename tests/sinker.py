"""Shared problem builders for the parity tests: the sinker of the benchmark configurations
built once through the ``underworld2_amd`` API (GPU) and once through the CPU oracle."""
import numpy as np


def sphere_material(coords, centre, radius):
    """Material rule fn.branching.conditional([(dot(c, c) < r**2, 1), (True, 0)]) evaluated in
    the same IEEE order as the reference's Binary/Relational closures."""
    c = coords - np.asarray(centre)[None, :]
    d2 = c[:, 0] * c[:, 0] + c[:, 1] * c[:, 1]
    if coords.shape[1] == 3:
        d2 = d2 + c[:, 2] * c[:, 2]
    return (d2 < radius ** 2).astype(np.int32)


def build_uw(res, ppc, seed, centre, radius, eta=(1.0, 100.0), rho=(0.0, 1.0), lo=None, hi=None):
    import underworld2_amd as uw
    dim = len(res)
    lo = lo or (0.0,) * dim
    hi = hi or (1.0,) * dim
    mesh = uw.mesh.FeMesh_Cartesian("Q1/dQ0", res, lo, hi)
    vel = uw.mesh.MeshVariable(mesh, dim)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(swarm, particlesPerCell=ppc, seed=seed))
    mat = swarm.add_variable("int", 1)
    fn = uw.function
    c = fn.input() - tuple(centre)
    mat.data[:] = fn.branching.conditional([(fn.math.dot(c, c) < radius ** 2, 1), (True, 0)]).evaluate(swarm)
    etaf = fn.branching.map(fn_key=mat, mapping={0: eta[0], 1: eta[1]})
    rhof = fn.branching.map(fn_key=mat, mapping={0: rho[0], 1: rho[1]})
    g = (0.0, -1.0) if dim == 2 else (0.0, 0.0, -1.0)
    force = rhof * g
    walls = []
    names = "IJK"
    for d in range(dim):
        walls.append(mesh.specialSets[f"Min{names[d]}_VertexSet"] + mesh.specialSets[f"Max{names[d]}_VertexSet"])
    bc = uw.conditions.DirichletCondition(vel, tuple(walls))
    stokes = uw.systems.Stokes(vel, p, fn_viscosity=etaf, fn_bodyforce=force, voronoi_swarm=swarm, conditions=[bc])
    return dict(uw=uw, mesh=mesh, vel=vel, p=p, swarm=swarm, mat=mat, stokes=stokes)


def build_oracle(res, ppc, seed, centre, radius, eta=(1.0, 100.0), rho=(0.0, 1.0), lo=None, hi=None, solve=True):
    from oracle import oracle as O
    dim = len(res)
    lo = lo or (0.0,) * dim
    hi = hi or (1.0,) * dim
    g = (0.0, -1.0) if dim == 2 else (0.0, 0.0, -1.0)
    return O.full_pipeline(dim, res, lo, hi, ppc, seed, lambda x: sphere_material(x, centre, radius), eta, rho, g,
                           solve=solve)


def K_from_W(W, dim, h):
    """Explicit element matrices from per-element Gauss weights W[Q][E]:
    K_e[ai][bj] = sum_g W_g (delta_ij dN_a.dN_b + dN_a/dx_j dN_b/dx_i)."""
    a = np.sqrt(0.6)
    x1 = [-a, 0.0, a]
    nn = 2 ** dim
    Q = 3 ** dim
    E = W.shape[1]
    K = np.zeros((E, nn * dim, nn * dim))
    for q in range(Q):
        X = [x1[q % 3], x1[(q // 3) % 3], x1[q // 9]]
        dN = np.zeros((nn, dim))
        for an in range(nn):
            s = [1.0 if (an >> d) & 1 else -1.0 for d in range(3)]
            for j in range(dim):
                v = (0.5 ** dim) * s[j]
                for o in range(dim):
                    if o != j:
                        v *= 1 + s[o] * X[o]
                dN[an, j] = v * 2.0 / h[j]
        B = np.zeros((nn, dim, nn, dim))
        dot = dN @ dN.T
        for i in range(dim):
            for j in range(dim):
                B[:, i, :, j] = (dot if i == j else 0.0) + np.outer(dN[:, j], dN[:, i])
        K += W[q][:, None, None] * B.reshape(nn * dim, nn * dim)[None]
    return K
