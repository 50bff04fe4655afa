"""Swarm population control on the GPU against the CPU oracle (oracle/pcdvc_oracle.py).

An uneven swarm (seeded random counts of 1-12 particles per element, some elements with a
tight cluster) is repopulated with splits and deletions enabled; the new coordinates, the cell
tables and the particle each SwarmVariable row came from must equal the oracle's exactly.  The
reference's doctest (PerCellGaussLayout 4, thresholds 0, maxSplits 9999: 256 -> 512) runs
through the Python API.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import pcdvc_oracle as PC

pytestmark = pytest.mark.gpu


def _uneven_coords(mesh_o, seed):
    rng = np.random.default_rng(seed)
    d = mesh_o.dim
    h = (mesh_o.hi - mesh_o.lo) / mesh_o.res
    idx = np.indices(tuple(mesh_o.res[::-1])).reshape(d, -1)[::-1].T
    pts = []
    for e, ijk in enumerate(idx):
        lo = mesh_o.lo + ijk * h
        n = int(rng.integers(1, 13))
        x = lo + rng.uniform(0.02, 0.98, (n, d)) * h
        if e % 5 == 0:  # a tight cluster: tiny Voronoi cells to delete
            x = np.concatenate([x, lo + (0.5 + 0.01 * rng.standard_normal((4, d))) * h])
        pts.append(x)
    return np.concatenate(pts)


@pytest.mark.parametrize("res", [(6, 5), (3, 3, 2)])
def test_population_control_matches_oracle(res):
    import underworld2_amd as uw
    from test_pcdvc_oracle import tables_of
    d = len(res)
    print(f"population control {res}", flush=True)
    mesh = uw.mesh.FeMesh_Cartesian("Q1/dQ0", res, (0.0,) * d, (1.0,) * d)
    mo = O.Mesh(res, (0.0,) * d, (1.0,) * d)
    coords = _uneven_coords(mo, 5)
    swarm = uw.swarm.Swarm(mesh)
    swarm.add_particles_with_coordinates(coords)
    tag = swarm.add_variable("long", 1)
    tag.data[:, 0] = np.arange(coords.shape[0])
    ptr, pid = tables_of(mo, coords)
    gptr, gpid = swarm.cell_tables()
    assert np.array_equal(gptr, ptr) and np.array_equal(gpid, pid)
    args = dict(deleteThreshold=0.02, splitThreshold=0.2, maxDeletions=2, maxSplits=3)
    out, src, nptr, npid = PC.repopulate(mo, coords, ptr, pid, 0.02, 0.2, 2, 3)
    uw.swarm.PopulationControl(swarm, **args).repopulate()
    print(f"{coords.shape[0]} -> {out.shape[0]} particles", flush=True)
    assert out.shape[0] != coords.shape[0]
    assert swarm.particleLocalCount == out.shape[0]
    assert np.array_equal(swarm.particleCoordinates.data, out)
    assert np.array_equal(tag.data[:, 0], src)
    gptr, gpid = swarm.cell_tables()
    assert np.array_equal(gptr, nptr) and np.array_equal(gpid, npid)
    owners = swarm.owningCell.data[:, 0]
    assert np.array_equal(owners[npid], np.repeat(np.arange(mo.E), np.diff(nptr)))
    # the repopulated swarm integrates: DVC weights sum to the element volume everywhere
    xi, w = swarm.integration_points()
    sums = np.add.reduceat(w, gptr[:-1])
    assert np.allclose(sums, 2.0 ** d, rtol=1e-12)


def test_reference_doctest_gauss_layout():
    import underworld2_amd as uw
    mesh = uw.mesh.FeMesh_Cartesian()
    swarm = uw.swarm.Swarm(mesh)
    swarm.populate_using_layout(uw.swarm.layouts.PerCellGaussLayout(swarm, 4))
    assert swarm.particleLocalCount == 256
    uw.swarm.PopulationControl(swarm, deleteThreshold=0., splitThreshold=0., maxDeletions=0,
                               maxSplits=9999).repopulate()
    assert swarm.particleGlobalCount == 512
