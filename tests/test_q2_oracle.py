"""Pins of the Q2/dPc1 CPU oracle (oracle/q2_oracle.py) -- no GPU.

The oracle is checked against the reference's own number (SolCx Q2/dPc1 32^2 v_rms,
docs/examples/10_Analytic_Solutions.ipynb) and against what the mathematics fixes:
nodality and partition of unity of both bases, rigid-body modes of K_e, the closed form of
G_e^T u for polynomial velocities, and the particle-sum K_e against the Gauss-rule K_e.
"""
import numpy as np
import pytest

from oracle import q2_oracle as Q

SOLCX_VRMS = 0.005578279626793198  # reference value, rtol 1e-7


def test_solcx_q2_vrms_matches_reference():
    eta, f = Q.solcx_functions(eta_a=1.0, eta_b=1.0e5, x_c=0.75, n_z=1)
    r = Q.solve_gauss_problem((32, 32), eta, f)
    v = Q.velocity_rms(r["mesh"], r["u"])
    assert v == pytest.approx(SOLCX_VRMS, rel=1e-7)
    r1, r2 = r["system"].residuals(r["u"], r["p"])
    assert r1 < 1e-12 and r2 < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_bases_nodal_and_partition_of_unity(dim):
    # Q2: N_a(node_b) = delta_ab at the 3^d local vertices
    nodes = np.stack(np.meshgrid(*([np.array([-1.0, 0.0, 1.0])] * dim), indexing="ij"), -1)
    nodes = nodes.transpose(tuple(range(dim))[::-1] + (dim,)).reshape(-1, dim)
    N, dN = Q.q2_shape(dim, nodes)
    assert np.allclose(N, np.eye(3 ** dim), atol=1e-15)
    x = np.random.default_rng(0).uniform(-1, 1, (20, dim))
    N, dN = Q.q2_shape(dim, x)
    assert np.allclose(N.sum(1), 1.0, atol=1e-14)
    assert np.allclose(dN.sum(2), 0.0, atol=1e-13)
    # derivative against a central difference
    eps = 1e-6
    for k in range(dim):
        xp, xm = x.copy(), x.copy()
        xp[:, k] += eps
        xm[:, k] -= eps
        fd = (Q.q2_shape(dim, xp)[0] - Q.q2_shape(dim, xm)[0]) / (2 * eps)
        assert np.allclose(fd, dN[:, k, :], atol=1e-8)
    # dPc1: nodal at the inner vertices, partition of unity, spans {1, x, y(, z)}
    Np = Q.dpc1_shape(dim, Q.DPC1_NODES[dim])
    assert np.allclose(Np, np.eye(dim + 1), atol=1e-15)
    Np = Q.dpc1_shape(dim, x)
    assert np.allclose(Np.sum(1), 1.0, atol=1e-14)
    M = np.concatenate([np.ones((dim + 1, 1)), Q.DPC1_NODES[dim]], 1)
    assert abs(np.linalg.det(M)) > 1e-3


def _rand_points(mesh, P, seed):
    rng = np.random.default_rng(seed)
    xi = rng.uniform(-1, 1, (mesh.E, P, mesh.dim))
    c = rng.uniform(0.1, 2.0, (mesh.E, P))
    return xi, c


@pytest.mark.parametrize("dim", [2, 3])
def test_rigid_body_modes_in_kernel(dim):
    mesh = Q.Q2Mesh((2,) * dim, (0.0,) * dim, (1.0, 2.0, 0.5)[:dim])
    xi, c = _rand_points(mesh, 30, 1)
    Ke = Q.element_K(mesh, xi, c)
    X = mesh.coords()[mesh.element_nodes()[0]]  # element 0 vertices [nen][d]
    modes = [np.eye(dim)[k][None, :].repeat(mesh.nen, 0) for k in range(dim)]
    modes.append(np.stack([-X[:, 1], X[:, 0]] + ([np.zeros(mesh.nen)] if dim == 3 else []), -1))
    if dim == 3:
        modes.append(np.stack([np.zeros(mesh.nen), -X[:, 2], X[:, 1]], -1))
        modes.append(np.stack([X[:, 2], np.zeros(mesh.nen), -X[:, 0]], -1))
    for m in modes:
        assert np.abs(Ke[0] @ m.reshape(-1)).max() < 1e-11 * np.abs(Ke[0]).max()
    # symmetric positive semi-definite with exactly the rigid modes as kernel
    w = np.linalg.eigvalsh(Ke[0])
    assert np.allclose(Ke[0], Ke[0].T, atol=1e-13)
    assert (w > 1e-10 * w.max()).sum() == mesh.nen * dim - len(modes)


@pytest.mark.parametrize("dim", [2, 3])
def test_G_closed_form_for_polynomial_velocity(dim):
    """G_e^T u = -int_e div(u) N^p_r ; for u = (x^2, x y (, z^2)) div u is linear and the
    integral is fixed by a tensor Gauss rule of high order."""
    mesh = Q.Q2Mesh((1,) * dim, (0.0,) * dim, (2.0, 1.0, 3.0)[:dim])
    Ge = Q.element_G(mesh)
    X = mesh.coords()
    if dim == 2:
        u = np.stack([X[:, 0] ** 2, X[:, 0] * X[:, 1]], -1)

        def div(x):
            return 2 * x[:, 0] + x[:, 0]
    else:
        u = np.stack([X[:, 0] ** 2, X[:, 0] * X[:, 1], X[:, 2] ** 2], -1)

        def div(x):
            return 3 * x[:, 0] + 2 * x[:, 2]
    got = Ge.T @ u[mesh.element_nodes()[0]].reshape(-1)
    g, w = np.polynomial.legendre.leggauss(8)
    pts = np.stack(np.meshgrid(*([g] * dim), indexing="ij"), -1).reshape(-1, dim)
    wts = np.prod(np.stack(np.meshgrid(*([w] * dim), indexing="ij"), -1).reshape(-1, dim), 1)
    xp = mesh.local_to_global(pts)[0]
    want = -np.einsum("g,g,gr->r", wts * mesh.detJ, div(xp), Q.dpc1_shape(dim, pts))
    assert np.allclose(got, want, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_particle_K_equals_gauss_K_at_gauss_points(dim):
    mesh = Q.Q2Mesh((2,) * dim, (0.0,) * dim, (1.0,) * dim)
    pts, wts = Q.gauss5(dim)
    eta = np.random.default_rng(3).uniform(1, 10, (mesh.E, 1))
    c = eta * (wts * mesh.detJ)[None, :]
    K1 = Q.element_K(mesh, pts, c)
    # the same integral from the closed form: int eta 2 eps(v):eps(u) for constant eta is
    # exact with the 5-point rule, and with an 8-point rule
    g, w = np.polynomial.legendre.leggauss(8)
    p8 = np.stack(np.meshgrid(*([g] * dim), indexing="ij"), -1).reshape(-1, dim)
    w8 = np.prod(np.stack(np.meshgrid(*([w] * dim), indexing="ij"), -1).reshape(-1, dim), 1)
    K2 = Q.element_K(mesh, p8, eta * (w8 * mesh.detJ)[None, :])
    assert np.allclose(K1, K2, rtol=1e-11, atol=1e-12 * np.abs(K1).max())


def test_vrms_of_quadratic_field():
    mesh = Q.Q2Mesh((3, 2), (0.0, 0.0), (1.0, 1.0))
    X = mesh.coords()
    u = np.stack([X[:, 0] ** 2, np.ones(len(X))], -1)
    # int x^4 + 1 over the unit square = 1/5 + 1
    assert Q.velocity_rms(mesh, u) == pytest.approx(np.sqrt(1.2), rel=1e-14)


def test_approxS_blocks_symmetric_positive():
    mesh = Q.Q2Mesh((2, 2, 2), (0.0,) * 3, (1.0,) * 3)
    xi, c = _rand_points(mesh, 25, 5)
    Ke = Q.element_K(mesh, xi, c)
    S = Q.element_approxS(Ke, Q.element_G(mesh))
    assert np.allclose(S, S.transpose(0, 2, 1), rtol=1e-13)
    assert np.all(np.linalg.eigvalsh(S) > 0)
