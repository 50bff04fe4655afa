"""Pins of the oracle's PerCellSpaceFillerLayout (oracle/uwg_oracle.c orc_layout_sobol).

* The reference docstring (PY/swarm/layouts.py:160-172): 1 cell on the unit square, 4 particles
  per cell -> [[0.5, 0.5], [0.25, 0.75], [0.75, 0.25], [0.375, 0.625]].
* The x axis (table entry 0: x + 1, m = {1}) against scipy's unscrambled Sobol sequence, whose
  second dimension uses the same polynomial and initial number (Gray-code order, first point 0
  dropped), for 4095 draws.
* Every axis (entries 0, 1, 2) is a proper Sobol sequence: with the initial 0, the first 2^k
  values are exactly the multiples of 2^-k (k <= 12), which fails for an even or oversized
  direction number or a wrong recursion tap.
* The Park-Miller skip: a particle whose draw exceeds 0.9 advances every axis' generator by
  one extra number; counted against the raw Park-Miller stream.
"""
import numpy as np

from oracle import oracle as O


def _axis_values(dim, axis, n):
    """The generator values in [0, 1) that n particles used on one axis: a row of n unit cells
    with one particle each, so particle k's coordinate minus the cell's corner is
    (xi + 1) / 2 = the value drawn."""
    res = (n, 1) if dim == 2 else (n, 1, 1)
    m = O.Mesh(res, (0.0,) * dim, (float(n),) + (1.0,) * (dim - 1))
    c = O.layout_sobol(m, 1)
    c[:, 0] -= np.arange(n)
    return c[:, axis]


def test_sobol_docstring():
    m = O.Mesh((1, 1), (0.0, 0.0), (1.0, 1.0))
    np.testing.assert_array_equal(O.layout_sobol(m, 4), [[0.5, 0.5], [0.25, 0.75], [0.75, 0.25], [0.375, 0.625]])


def _skips(n, seed=13):
    r = O.rng_stream(seed, n) / 2147483647.0
    return r > 0.9


def test_sobol_x_axis_against_scipy_and_skips():
    from scipy.stats import qmc
    n = 4095
    skip = _skips(n)
    assert skip.any() and not skip.all()
    seq = qmc.Sobol(d=2, scramble=False).random(2 * n + 1)[1:, 1]
    # draw index of particle k (1-based): k + 1 + skips before it + its own skip
    idx = np.arange(n) + 1 + np.concatenate([[0], np.cumsum(skip)[:-1]]) + skip
    vals = _axis_values(2, 0, n)
    np.testing.assert_array_equal(vals, seq[idx - 1])


def test_sobol_axes_are_nets():
    """Each axis' raw generator is a proper Sobol sequence (with 0 first, the first 2^k values
    are exactly the multiples of 2^-k), and the layout takes its values at the draw indices the
    Park-Miller skips give."""
    n = 4095
    skip = _skips(n)
    idx = np.arange(n) + 1 + np.concatenate([[0], np.cumsum(skip)[:-1]]) + skip
    for axis in (0, 1, 2):
        raw = O.sobol_stream(axis, 2 * n + 1)
        for k in range(1, 13):
            N = 2 ** k
            np.testing.assert_array_equal(np.sort(np.concatenate([[0.0], raw[:N - 1]])), np.arange(N) / N,
                                          err_msg=f"axis {axis} k {k}")
        for dim in (2, 3):
            if axis < dim:
                np.testing.assert_array_equal(_axis_values(dim, axis, n), raw[idx - 1], err_msg=f"axis {axis} dim {dim}")
