"""Pins of the SUPG advection-diffusion oracle (oracle/supg_oracle.py) -- no GPU.

The reference holds no fixture for this path short of its convection notebooks, so the oracle
is pinned by what the mathematics fixes: the upwind functions' closed forms, the lumped mass,
the multicorrector's exact action on a linear field under uniform velocity (a trapezoidal-rule
translation at interior nodes), and relaxation of pure diffusion with a source to the nodally
exact parabola.
"""
import numpy as np
import pytest

from oracle import supg_oracle as S


def test_upwind_functions_closed_forms():
    for a in (0.3, 1.7, 5.0, -2.5, 15.0):
        assert S.upwind_xi(a, 0) == pytest.approx(1.0 / np.tanh(a) - 1.0 / a, rel=1e-12)
    assert S.upwind_xi(1e-9, 0) == pytest.approx(1e-9 / 3.0, rel=1e-12)
    assert S.upwind_xi(40.0, 0) == 1.0 - 1.0 / 40.0
    assert S.upwind_xi(-40.0, 0) == -1.0 + 1.0 / 40.0
    assert S.upwind_xi(-5.0, 1) == -1.0 and S.upwind_xi(5.0, 1) == 1.0
    assert S.upwind_xi(1.5, 1) == pytest.approx(0.5, rel=1e-12)
    assert S.upwind_xi(0.5, 2) == 0.0 and S.upwind_xi(4.0, 2) == 0.75 and S.upwind_xi(-4.0, 2) == -0.75
    # the exact function is odd, increasing, bounded by 1 and matches its series a/3 - a^3/45
    xs = np.linspace(-30, 30, 601)
    ys = np.array([S.upwind_xi(x, 0) for x in xs])
    assert np.all(np.diff(ys) > 0) and np.all(np.abs(ys) < 1.0)
    assert S.upwind_xi(1e-3, 0) == pytest.approx(1e-3 / 3 - 1e-9 / 45, rel=1e-8)


@pytest.mark.parametrize("res", [(5, 3), (3, 2, 4)])
def test_lumped_mass_sums_to_volume(res):
    d = len(res)
    m = S.Q1Mesh(res, (0.0,) * d, (1.0, 2.0, 0.5)[:d])
    M = S.lumped_mass(m)
    assert M.sum() == pytest.approx(float(np.prod(m.hi - m.lo)), rel=1e-13)
    assert np.all(M > 0)
    # corner / edge / interior weights of a uniform mesh: volume share of the node
    h = np.prod((m.hi - m.lo) / np.array(res))
    assert M.max() == pytest.approx(h, rel=1e-13)


def test_linear_field_translates_by_trapezoid_rule():
    res = (14, 12)
    m = S.Q1Mesh(res, (0.0, 0.0), (1.0, 1.0))
    X = m.coords()
    a = np.array([1.3, -0.7])
    v = np.array([0.3, 0.2])
    phi = X @ a + 0.25
    phidot = np.zeros(m.N)
    vel = np.tile(v, (m.N, 1))
    Q = 9
    kap = np.full((m.E, Q), 1e-6)
    src = np.zeros((m.E, Q))
    sup = S.SUPG(m, np.zeros(m.N, bool))
    dt = 0.01
    p1, d1 = sup.integrate(phi, phidot, vel, kap, src, dt)
    idx = np.indices((res[1] + 1, res[0] + 1)).reshape(2, -1)[::-1]
    inner = np.all([(idx[k] >= 4) & (idx[k] <= res[k] - 4) for k in range(2)], axis=0)
    va = float(v @ a)
    # first step from phidot = 0: half the translation, phidot becomes -u.grad(phi)
    assert np.allclose(p1[inner], phi[inner] - 0.5 * dt * va, rtol=0, atol=1e-13)
    assert np.allclose(d1[inner], -va, rtol=0, atol=1e-12)
    p2, d2 = sup.integrate(p1, d1, vel, kap, src, dt)
    inner2 = np.all([(idx[k] >= 6) & (idx[k] <= res[k] - 6) for k in range(2)], axis=0)
    # second step: the full translation
    assert np.allclose(p2[inner2], phi[inner2] - 1.5 * dt * va, rtol=0, atol=1e-12)
    assert np.allclose(d2[inner2], -va, rtol=0, atol=1e-11)


def test_pure_diffusion_relaxes_to_parabola():
    res = (8, 2)
    m = S.Q1Mesh(res, (0.0, 0.0), (1.0, 0.25))
    X = m.coords()
    bc = (np.abs(X[:, 0]) < 1e-12) | (np.abs(X[:, 0] - 1.0) < 1e-12)
    phi = np.zeros(m.N)
    phidot = np.zeros(m.N)
    vel = np.zeros((m.N, 2))
    kap = np.ones((m.E, 9))
    src = np.full((m.E, 9), 2.0)
    sup = S.SUPG(m, bc)
    dt = sup.max_dt(vel, kap)
    h = 1.0 / 8
    assert dt == pytest.approx(0.5 * (0.25 / 2) ** 2 / 1.0, rel=1e-12)  # min separation is the y spacing
    for _ in range(1200):
        phi, phidot = sup.integrate(phi, phidot, vel, kap, src, dt)
    exact = X[:, 0] * (1.0 - X[:, 0])
    assert np.allclose(phi, exact, rtol=0, atol=1e-9)
    assert np.abs(phidot).max() < 1e-8
    del h


def test_max_dt_advective_closed_form():
    m = S.Q1Mesh((10, 5), (0.0, 0.0), (2.0, 1.0))
    vel = np.zeros((m.N, 2))
    vel[7] = (3.0, -0.5)
    vel[20] = (0.0, 4.0)
    sup = S.SUPG(m, np.zeros(m.N, bool))
    kap = np.full((m.E, 9), 1e-3)
    # h = 0.2 in both directions; advective min |h / v| = 0.2 / 4
    assert sup.max_dt(vel, kap) == pytest.approx(0.5 * 0.2 / 4.0, rel=1e-12)
