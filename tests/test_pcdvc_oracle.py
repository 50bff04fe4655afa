"""Pins of the population-control oracle (oracle/pcdvc_oracle.py) -- no GPU.

* the reference's doctest (PY/swarm/_population_control.py:60-68): PerCellGaussLayout with 4
  points per direction on the default 4x4 mesh, thresholds 0 and maxSplits 9999: 256 -> 512;
* a one-particle element splits at the element centre (its Voronoi cell is the element);
* the swap-with-last deletion sequence worked by hand on one cell.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import pcdvc_oracle as PC


def _gauss_layout(mesh, n):
    g, _ = np.polynomial.legendre.leggauss(n)
    d = mesh.dim
    h = (mesh.hi - mesh.lo) / mesh.res
    idx = np.indices(tuple(mesh.res[::-1])).reshape(d, -1)[::-1]
    cen = np.stack([mesh.lo[k] + (idx[k] + 0.5) * h[k] for k in range(d)], -1)
    q = np.stack(np.meshgrid(*([g] * d), indexing="ij"), -1).reshape(-1, d)[:, ::-1]
    return (cen[:, None, :] + q[None, :, :] * (0.5 * h)[None, None, :]).reshape(-1, d)


def tables_of(mesh, coords):
    cells = O.cell_of(mesh, coords)
    order = np.argsort(cells, kind="stable")
    ptr = np.zeros(mesh.E + 1, dtype=np.int64)
    ptr[1:] = np.cumsum(np.bincount(cells, minlength=mesh.E))
    return ptr, order.astype(np.int32)


def test_reference_doctest_splits_every_particle():
    mesh = O.Mesh((4, 4), (0.0, 0.0), (1.0, 1.0))
    coords = _gauss_layout(mesh, 4)
    assert coords.shape[0] == 256
    ptr, pid = tables_of(mesh, coords)
    out, src, nptr, npid = PC.repopulate(mesh, coords, ptr, pid, 0.0, 0.0, 0, 9999)
    assert out.shape[0] == 512
    assert np.array_equal(src[:256], np.arange(256))
    assert np.all(np.diff(nptr) == 32)


def test_single_particle_element_splits_at_centre():
    mesh = O.Mesh((2, 1), (0.0, 0.0), (2.0, 1.0))
    gx, gy = np.meshgrid([1.25, 1.75], [0.125, 0.375, 0.625, 0.875])
    c1 = np.stack([gx.ravel(), gy.ravel()], 1)  # eight even Voronoi cells: w = 0.5 each
    coords = np.concatenate([[[0.1, 0.3]], c1])
    ptr, pid = tables_of(mesh, coords)
    out, src, nptr, npid = PC.repopulate(mesh, coords, ptr, pid)
    assert out.shape[0] == 10 and src[9] == 0
    assert np.allclose(out[9], [0.5, 0.5], rtol=0, atol=1e-14)
    assert list(npid[nptr[0]:nptr[1]]) == [0, 9]


def test_deletion_swap_with_last_by_hand():
    mesh = O.Mesh((1, 1), (0.0, 0.0), (1.0, 1.0))
    coords = np.linspace(0.1, 0.9, 12).reshape(6, 2)
    ptr = np.array([0, 6])
    pid = np.arange(6, dtype=np.int32)
    w = np.array([0.5, 0.01, 0.5, 0.5, 0.02, 0.5])  # positions 1 and 4 below 0.05 * 4
    cen = np.zeros((6, 2))
    out, src, nptr, npid = PC.population_control(mesh, coords, ptr, pid, cen, w, 0.05, 0.9, 2, 0)
    # delete 4 (largest index first): table [0,1,2,3,5] -> last particle 5 moves to 4 -> [0,1,2,3,4]
    # delete 1: table [0,4,2,3] -> last particle 4 (data of 5) moves to 1 -> [0,1,2,3]
    assert list(src) == [0, 5, 2, 3]
    assert list(npid) == [0, 1, 2, 3]
    assert np.array_equal(out, coords[[0, 5, 2, 3]])
    with pytest.raises(AssertionError):
        assert list(src) == [0, 1, 2, 3]
