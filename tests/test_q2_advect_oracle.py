"""Pins of the oracle's lookups in a Q2 velocity field (oracle/uwg_oracle.c, Q2 section):
the Biquadratic / Triquadratic bases (Lagrange property, partition of unity, quadratic
completeness, derivatives against central differences), Newton on affine cells (closed form),
interpolation (exact for tensor-quadratic fields), RK2 (closed form for quadratic velocity) and
the Q2 minimum vertex separation.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O


def _nodes(dim):
    """element-local vertex positions in incidence order a = ax + 3 ay (+ 9 az)"""
    t = (-1.0, 0.0, 1.0)
    return np.array([[t[a % 3], t[(a // 3) % 3]] + ([t[a // 9]] if dim == 3 else []) for a in range(3 ** dim)])


@pytest.mark.parametrize("dim", [2, 3])
def test_basis_lagrange_and_completeness(dim):
    X = _nodes(dim)
    for b, xb in enumerate(X):
        N, _ = O.q2_basis(dim, xb)
        e = np.zeros(3 ** dim)
        e[b] = 1.0
        assert np.allclose(N, e, rtol=0, atol=1e-15)
    rng = np.random.default_rng(1)
    C = rng.standard_normal((3,) * dim)  # f = sum_k C[k] prod_d xi_d^k_d (degree <= 2 per axis)

    def f(x):
        return sum(C[k] * np.prod([x[d] ** k[::-1][d] for d in range(dim)]) for k in itertools.product(range(3), repeat=dim))

    fX = np.array([f(x) for x in X])
    for _ in range(20):
        xi = rng.uniform(-1, 1, dim)
        N, _ = O.q2_basis(dim, xi)
        assert N.sum() == pytest.approx(1.0, abs=1e-14)
        assert N @ X == pytest.approx(xi, abs=1e-14)
        assert N @ fX == pytest.approx(f(xi), abs=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_basis_derivatives_central_differences(dim):
    rng = np.random.default_rng(2)
    h = 1e-6
    for _ in range(10):
        xi = rng.uniform(-0.9, 0.9, dim)
        _, G = O.q2_basis(dim, xi)
        for d in range(dim):
            e = np.zeros(dim)
            e[d] = h
            Np, _ = O.q2_basis(dim, xi + e)
            Nm, _ = O.q2_basis(dim, xi - e)
            assert np.allclose(G[d], (Np - Nm) / (2 * h), rtol=0, atol=1e-8)


@pytest.mark.parametrize("res", [(5, 4), (3, 4, 2)])
def test_newton_affine_closed_form(res):
    dim = len(res)
    hi = np.array([1.0, 0.7, 1.3][:dim])
    m = O.Mesh(res, np.zeros(dim), hi)
    V = O.q2_vertex_coords(m)
    rng = np.random.default_rng(3)
    h = hi / np.array(res)
    for _ in range(30):
        ijk = np.array([rng.integers(0, r) for r in res])
        nx = 2 * np.array(res) + 1
        verts = []
        for a in range(3 ** dim):
            s = np.array([a % 3, (a // 3) % 3, a // 9][:dim])
            g = 2 * ijk + s
            n = g[0] + nx[0] * (g[1] + (nx[1] * g[2] if dim == 3 else 0))
            verts.append(V[n])
        verts = np.array(verts)
        x = (ijk + rng.uniform(0, 1, dim)) * h
        xi, it = O.q2_newton_local(dim, verts, x)
        assert it <= 3
        assert np.allclose(xi, 2 * (x - (ijk + 0.5) * h) / h, rtol=0, atol=1e-12)


def _quadratic_velocity(dim, rng):
    C = rng.standard_normal((dim,) + (3,) * dim) * 0.3

    def v(x):
        out = np.zeros(dim)
        for k in itertools.product(range(3), repeat=dim):
            out += C[(slice(None),) + k] * np.prod([x[d] ** k[::-1][d] for d in range(dim)])
        return out

    return v


@pytest.mark.parametrize("res", [(6, 5), (3, 3, 2)])
def test_interpolation_and_rk2_exact_for_quadratic_fields(res):
    dim = len(res)
    m = O.Mesh(res, np.zeros(dim), np.ones(dim))
    rng = np.random.default_rng(4)
    v = _quadratic_velocity(dim, rng)
    V = O.q2_vertex_coords(m)
    vel = np.array([v(x) for x in V])
    x = rng.uniform(0.05, 0.95, (200, dim))
    vi, ok = O.interpolate_q2(m, vel, x)
    assert ok.all()
    assert np.allclose(vi, [v(p) for p in x], rtol=0, atol=1e-13)
    dt = 0.05
    x1, bad, fb = O.advect_rk2_q2(m, vel, x, dt)
    assert bad == 0 and fb == 0
    ref = np.array([p + dt * v(p + 0.5 * dt * v(p)) for p in x])
    assert np.allclose(x1, ref, rtol=0, atol=1e-13)


def test_rk2_fallback_and_outside():
    m = O.Mesh((4, 4), (0.0, 0.0), (1.0, 1.0))
    V = O.q2_vertex_coords(m)
    vel = np.zeros_like(V)
    vel[:, 0] = 1.0
    # midpoint leaves the box -> first-order step from the start; start outside -> counted
    x = np.array([[0.99, 0.5], [1.5, 0.5]])
    x1, bad, fb = O.advect_rk2_q2(m, vel, x, 0.1)
    assert bad == 1 and fb == 1
    assert x1[0, 0] == pytest.approx(0.99 + 0.1)


@pytest.mark.parametrize("res,hi", [((6, 5), (1.0, 0.8)), ((3, 4, 2), (1.0, 1.2, 0.5))])
def test_min_separation_and_max_dt(res, hi):
    dim = len(res)
    m = O.Mesh(res, np.zeros(dim), np.array(hi))
    ms = O.lib().orc_q2_min_sep(*m.args())
    assert ms == pytest.approx(min(h / r for h, r in zip(hi, res)) / 2, rel=1e-12)
    V = O.q2_vertex_coords(m)
    vel = np.ones_like(V) * np.arange(1, dim + 1)
    assert O.max_dt_q2(m, vel) == pytest.approx(0.5 * ms / np.sqrt(np.sum(np.arange(1, dim + 1) ** 2)), rel=1e-14)
