"""The Schur-stencil network oracle (oracle/nn_oracle.py StencilNet, snn_features): its
hand-derived backward pass against central finite differences of its own loss (float64), a
planted mistake that the check catches, and closed forms of the features and of the stencil.
"""
import numpy as np

from oracle import nn_oracle as nno


def _case(grid=(5, 4, 3), C=8, L=2, B=2, seed=0, head_scale=0.3):
    net = nno.StencilNet(grid, C, L)
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, net.nparams)
    N = int(np.prod(grid))
    x = rng.normal(size=(B, 9, N))
    r = rng.normal(size=(B, N))
    t = rng.normal(size=(B, N))
    return net, p, x, r, t


def test_stencil_backward_against_finite_differences():
    net, p, x, r, t = _case()
    L, g, _ = net.loss_grad(p, x, r, t)
    rng = np.random.default_rng(1)
    h = 1e-6
    for k in rng.choice(net.nparams, 24, replace=False):
        e = np.zeros(net.nparams)
        e[k] = h
        fd = (net.loss_grad(p + e, x, r, t)[0] - net.loss_grad(p - e, x, r, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
    # a directional derivative along a random direction
    d = rng.normal(size=net.nparams)
    fd = (net.loss_grad(p + h * d, x, r, t)[0] - net.loss_grad(p - h * d, x, r, t)[0]) / (2 * h)
    assert abs(fd - g @ d) <= 1e-6 * max(1.0, abs(fd))


def test_planted_mistake_is_caught():
    """Dropping the identity from the stencil (C = head instead of I + head) changes the loss
    the finite differences see, so the analytic gradient of the correct loss no longer matches."""
    net, p, x, r, t = _case(seed=3)
    L, g, _ = net.loss_grad(p, x, r, t)

    class Wrong(nno.StencilNet):
        def stencil(self, y, r_):
            return super().stencil(y, r_) - r_  # identity term removed

    w = Wrong(net.grid, net.C, net.L)
    h = 1e-6
    k = net.nparams - 14  # a head bias: its gradient depends on the identity term
    e = np.zeros(net.nparams)
    e[k] = h
    fd = (w.loss_grad(p + e, x, r, t)[0] - w.loss_grad(p - e, x, r, t)[0]) / (2 * h)
    assert abs(fd - g[k]) > 1e-3 * max(1.0, abs(g[k]))


def test_identity_stencil_when_head_is_zero():
    """A zero head gives C = I: the stencil output equals r and the loss is |r - t|^2/|t|^2."""
    net, p, x, r, t = _case(seed=5)
    ci, co, wo, bo, k = net.specs[-1]
    p[wo:bo + co] = 0.0
    y, _ = net.forward(p, x)
    out = net.stencil(y, r)
    np.testing.assert_allclose(out, r, atol=1e-15)
    L, _, _ = net.loss_grad(p, x, r, t)
    np.testing.assert_allclose(L, np.mean(((r - t) ** 2).sum(1) / (t * t).sum(1)), rtol=1e-12)


def test_stencil_closed_form_shift():
    """A head that outputs -1 at the centre tap and +1 at the x+ tap gives (C r)_e = r_(e+x),
    zero on the last x layer."""
    grid = (4, 3, 2)
    net = nno.StencilNet(grid, 8, 1)
    p = np.zeros(net.nparams)
    ci, co, wo, bo, k = net.specs[-1]
    p[bo + 13] = -1.0  # centre tap offset -1 -> identity removed
    p[bo + 14] = 1.0   # tap (dx, dy, dz) = (+1, 0, 0)
    x = np.zeros((1, 9, 24))
    r = np.arange(24, dtype=float)[None, :] + 1.0
    y, _ = net.forward(p, x)
    out = net.stencil(y, r).reshape(2, 3, 4)
    rr = r.reshape(2, 3, 4)
    np.testing.assert_array_equal(out[:, :, :3], rr[:, :, 1:])
    np.testing.assert_array_equal(out[:, :, 3], 0.0)


def test_features_closed_form():
    grid = (3, 2, 2)
    S = np.full(12, 2.0)
    S[0] = 2.0 * np.exp(-2.0)  # element (0,0,0): t = -ln S larger by 2
    f = nno.snn_features(grid, S).reshape(9, 2, 2, 3)
    # element (1,0,0) sees its x- neighbour with t larger by 2: tanh(1)
    assert np.isclose(f[0, 0, 0, 1], np.tanh(1.0))
    # element (0,0,0) sees its x+ neighbour with t smaller by 2: tanh(-1); no x- neighbour
    assert np.isclose(f[1, 0, 0, 0], np.tanh(-1.0)) and f[0, 0, 0, 0] == 0.0
    # uniform elsewhere: zero log-ratio
    assert f[0, 1, 1, 2] == 0.0
    # walls: x in {-1, 0, +1} on 3 elements; y and z two layers each: -1, +1
    np.testing.assert_array_equal(f[6, 0, 0], [-1.0, 0.0, 1.0])
    np.testing.assert_array_equal(f[7, 0, :, 0], [-1.0, 1.0])
    np.testing.assert_array_equal(f[8, :, 0, 0], [-1.0, 1.0])
