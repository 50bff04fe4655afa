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


def test_restart_defaults_follow_method_precision_and_model():
    """FGMRES restart chosen by the solver front end (host logic, no GPU): 30 for the Schur
    method, 60 for the coupled one, 70 with the learned Schur stencil in mixed precision only
    (the FP32 Arnoldi basis's restart rise hands the tail to approxS there), 60 with the stencil
    in FP64 or with the coarse correction; an explicit restart always wins."""
    from underworld2_amd import systems

    class _Stencil(learned.LearnedSchurStencil):
        def __init__(self):
            pass

        def attach(self, stokes):  # the handle is not needed for the options
            pass

    class _Coarse(learned.LearnedCoarseCorrection):
        def __init__(self):
            pass

        def attach(self, stokes, level_res=None):
            pass

    def restart(method, precision, model=None, explicit=None):
        s = systems.StokesSolver(systems.Stokes.__new__(systems.Stokes))
        s.set_method(method)
        s.set_precision(precision)
        s.set_learned(model)
        if explicit is not None:
            s.set_restart(explicit)
        return s._opts().restart

    assert restart("schur", "fp64") == 30
    assert restart("block", "fp64") == 60
    assert restart("block", "mixed") == 60
    assert restart("block", "mixed", _Stencil()) == 70
    assert restart("block", "fp64", _Stencil()) == 60
    assert restart("block", "mixed", _Coarse()) == 60
    assert restart("block", "mixed", _Stencil(), explicit=40) == 40
    s = systems.StokesSolver(systems.Stokes.__new__(systems.Stokes))
    s.set_method("block")
    s.set_precision("mixed")
    s.set_learned(_Stencil())
    s.set_learned(None)  # detaching returns to the plain default
    assert s._opts().restart == 60
