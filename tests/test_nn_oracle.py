"""The network oracle's hand-derived backward pass, pinned by central finite differences of its
own forward pass (float64), plus the identities its building blocks must satisfy."""
import numpy as np
import pytest

from oracle import nn_oracle as nno


def _case(grid=(5, 4, 6), C=3, D=3, B=2, seed=0):
    net = nno.Net(grid, C, D)
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.4, net.nparams)
    N = grid[0] * grid[1] * grid[2]
    x = rng.normal(size=(B, 4, N))
    t = rng.normal(size=(B, 3, N))
    m = (rng.random((B, 3, N)) > 0.2).astype(np.float64)
    return net, p, x, t, m


def test_restrict_is_prolong_transpose():
    rng = np.random.default_rng(1)
    for fine in [(5, 5, 5), (4, 6, 7), (3, 8, 5)]:
        coarse = nno.coarser(fine)
        a = rng.normal(size=(fine[2], fine[1], fine[0]))
        b = rng.normal(size=(coarse[2], coarse[1], coarse[0]))
        lhs = np.sum(nno.restrict(a, fine, coarse) * b)
        rhs = np.sum(a * nno.prolong(b, fine, coarse))
        assert abs(lhs - rhs) < 1e-12 * (abs(lhs) + 1)


def test_prolong_reproduces_linear_fields():
    # trilinear prolongation is exact for fields linear in the node index (odd fine grid)
    fine = (9, 7, 5)
    coarse = nno.coarser(fine)
    k, j, i = np.meshgrid(np.arange(coarse[2]), np.arange(coarse[1]), np.arange(coarse[0]), indexing="ij")
    c = 2 * (1.0 + 2 * i - 3 * j + 0.5 * k)
    K, J, I = np.meshgrid(np.arange(fine[2]), np.arange(fine[1]), np.arange(fine[0]), indexing="ij")
    f = 1.0 + I - 1.5 * J + 0.25 * K
    assert np.allclose(nno.prolong(c, fine, coarse), 2 * f)


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1, 2, 3, 4, 5))
    w = rng.normal(size=(3, 2, 27))
    b = rng.normal(size=3)
    y = nno.conv(x, w, b)
    # brute force at every output point
    for o in range(3):
        for z in range(3):
            for yy in range(4):
                for xx in range(5):
                    acc = b[o]
                    for t, (dz, dy, dx) in enumerate(nno.TAPS):
                        Z, Y, X = z + dz, yy + dy, xx + dx
                        if 0 <= Z < 3 and 0 <= Y < 4 and 0 <= X < 5:
                            acc += (w[o, :, t] * x[0, :, Z, Y, X]).sum()
                    assert abs(acc - y[0, o, z, yy, xx]) < 1e-12


def test_loss_of_exact_target_is_zero():
    net, p, x, t, m = _case()
    y, _ = net.forward(p, x)
    L, g = net.loss(y, y.copy(), m)
    assert L == 0.0 and not g.any()


@pytest.mark.parametrize("D", [2, 3])
def test_backward_matches_finite_differences(D):
    net, p, x, t, m = _case(D=D)
    L, g, _ = net.loss_grad(p, x, t, m)
    rng = np.random.default_rng(5)
    idx = rng.choice(net.nparams, 40, replace=False)
    # every bias and the first/last weight of every convolution as well
    for ci, co, wo, bo in net.specs:
        idx = np.concatenate([idx, [wo, bo - 1, bo, bo + co - 1]])
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        Lp = net.loss(net.forward(p + e, x)[0], t, m)[0]
        Lm = net.loss(net.forward(p - e, x)[0], t, m)[0]
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_backward_directional_derivative():
    net, p, x, t, m = _case(D=3, seed=3)
    L, g, _ = net.loss_grad(p, x, t, m)
    d = np.random.default_rng(9).normal(size=p.shape)
    h = 1e-6
    Lp = net.loss(net.forward(p + h * d, x)[0], t, m)[0]
    Lm = net.loss(net.forward(p - h * d, x)[0], t, m)[0]
    fd = (Lp - Lm) / (2 * h)
    assert abs(fd - g @ d) <= 1e-6 * max(1.0, abs(fd))


def test_planted_backward_mistake_is_caught():
    """A dropped skip-connection gradient (a plausible slip) must fail the FD check."""
    net, p, x, t, m = _case(D=3, seed=4)
    y, cache = net.forward(p, x)
    L, gy = net.loss(y, t, m)
    good = net.backward(p, cache, gy)
    saved = nno.prolong

    def bad_prolong(a, fine, coarse, s=1.0):
        return 0.0 * saved(a, fine, coarse, s) if s == 0.125 else saved(a, fine, coarse, s)

    nno.prolong = bad_prolong
    try:
        bad = net.backward(p, cache, gy)
    finally:
        nno.prolong = saved
    # encoder-level parameters at deep levels differ
    assert np.abs(bad - good).max() > 1e-3 * np.abs(good).max()
