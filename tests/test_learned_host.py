"""Host-side pieces of the learned correction's data pipeline (no GPU): the random-field
generator, the nodal-field interpolant and the sample scaling the inference path mirrors."""
import numpy as np

from underworld2_amd import learned


def test_grf_is_standardised_and_seeded():
    rng = np.random.default_rng(0)
    f = learned.gaussian_random_field((17, 17, 17), 0.2, rng)
    assert f.shape == (17, 17, 17)
    assert abs(f.mean()) < 1e-12 and abs(f.std() - 1.0) < 1e-12
    g = learned.gaussian_random_field((17, 17, 17), 0.2, np.random.default_rng(0))
    assert np.array_equal(f, g)


def test_grf_correlation_length_orders_smoothness():
    # mean squared neighbour difference falls as the correlation length grows
    rough = []
    for ell in (0.1, 0.2, 0.3):
        acc = 0.0
        for s in range(4):
            f = learned.gaussian_random_field((33, 33, 33), ell, np.random.default_rng(s))
            acc += np.mean(np.diff(f, axis=2) ** 2)
        rough.append(acc)
    assert rough[0] > rough[1] > rough[2]


def test_nodal_field_reproduces_trilinear_functions():
    nodes = (5, 4, 6)
    z, y, x = np.meshgrid(*[np.linspace(0, 1, n) for n in nodes[::-1]], indexing="ij")
    v = 1.0 + 2 * x - 3 * y + 0.5 * z + x * y * z
    fn = learned.NodalField(v)
    pts = np.random.default_rng(1).random((200, 3))
    exact = 1.0 + 2 * pts[:, 0] - 3 * pts[:, 1] + 0.5 * pts[:, 2]
    # the xyz term is not trilinear across cells; compare the affine part on a pure affine field
    fa = learned.NodalField(1.0 + 2 * x - 3 * y + 0.5 * z)
    assert np.allclose(fa.evaluate(pts)[:, 0], exact, atol=1e-12)
    # nodes are reproduced exactly for any field
    node_pts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    assert np.allclose(fn.evaluate(node_pts)[:, 0], v.ravel(), atol=1e-12)


def test_make_sample_scaling_and_layout():
    rng = np.random.default_rng(2)
    N = 27
    b = rng.normal(size=N * 3) * 7.0
    x = rng.normal(size=N * 3)
    feat = rng.normal(size=N).astype(np.float32)
    mask = np.ones((3, N), np.float32)
    xi, t, m = learned.make_sample(b, x, feat, mask)
    rms = np.sqrt(np.mean(b * b))
    assert np.allclose(xi[:3].T.reshape(-1) * rms, b, rtol=1e-6)
    assert np.allclose(t.T.reshape(-1) * rms, x, rtol=1e-6)
    assert np.array_equal(xi[3], feat)
    assert abs(np.sqrt(np.mean(xi[:3].astype(np.float64) ** 2)) - 1.0) < 1e-6
