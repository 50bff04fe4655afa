"""The allocation-free interpolation entry point (uwg_field_interpolate_noalloc) against the
allocating one (uwg_field_interpolate), through the C ABI.

Both run the same interpolation kernel on the same device field, so values and flags must be
bit-identical.  Two calls in a row on each field: the largest accepted batch
(UWG_NOALLOC_MAX_POINTS points), then a smaller one, so the second call reads a scratch array
that still holds the first call's data.  Points outside the mesh are included (flag 0).  A batch
above the maximum is refused with an argument error before any device work.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_POINTS = 65536


def _field(uw, elem, res, ncomp, seed):
    dim = len(res)
    mesh = uw.mesh.FeMesh_Cartesian(elem, res, (0.0,) * dim, (1.0,) + (0.75,) * (dim - 1))
    f = uw.mesh.MeshVariable(mesh, ncomp)
    f.data[:] = np.random.default_rng(seed).standard_normal(f.data.shape)
    f._push()
    return mesh, f


def _points(rng, n, dim):
    hi = np.array((1.0,) + (0.75,) * (dim - 1))
    x = rng.uniform(-0.05, 1.05, (n, dim)) * hi
    return np.ascontiguousarray(x)


def _call(L, fn, f, x, nc):
    n = x.shape[0]
    v = np.full((n, nc), np.nan)
    ok = np.full(n, -1, dtype=np.int32)
    rc = fn(f._h, n, L.dptr(x), L.dptr(v), L.iptr(ok))
    return rc, v, ok


@pytest.mark.parametrize("elem,res,ncomp", [("Q1/dQ0", (40, 30), 2), ("Q1/dQ0", (12, 10, 9), 3),
                                            ("Q2/dPc1", (6, 5, 4), 3)])
def test_noalloc_interpolate_matches_allocating_entry_point(elem, res, ncomp):
    import underworld2_amd as uw
    from underworld2_amd import _lib as L
    print(f"noalloc interpolation: {elem} {res}, {ncomp} components", flush=True)
    lib = L.lib()
    mesh, f = _field(uw, elem, res, ncomp, 3)
    rng = np.random.default_rng(7)
    for n in (MAX_POINTS, 777):
        x = _points(rng, n, len(res))
        rc0, v0, ok0 = _call(L, lib.uwg_field_interpolate, f, x, ncomp)
        rc1, v1, ok1 = _call(L, lib.uwg_field_interpolate_noalloc, f, x, ncomp)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(ok0, ok1)
        assert 0 < ok1.sum() < n  # inside and outside points both present
        inside = ok1 == 1
        assert np.array_equal(v0[inside], v1[inside])
        print(f"  n = {n}: {int(inside.sum())} inside, identical", flush=True)


def test_noalloc_interpolate_rejects_oversized_batch():
    import underworld2_amd as uw
    from underworld2_amd import _lib as L
    lib = L.lib()
    mesh, f = _field(uw, "Q1/dQ0", (8, 8), 1, 1)
    x = np.zeros((MAX_POINTS + 1, 2))
    rc, v, ok = _call(L, lib.uwg_field_interpolate_noalloc, f, x, 1)
    assert rc < 0
    assert "at most 65536 points" in lib.uwg_last_error().decode()
    assert np.all(ok == -1)  # nothing was written
    rc, v, ok = _call(L, lib.uwg_field_interpolate_noalloc, f, np.zeros((0, 2)), 1)
    assert rc == 0
