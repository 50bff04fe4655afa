"""save/load of meshes, mesh variables, swarms and swarm variables through HDF5, on the GPU
(the objects live in the HIP library; the files are written on the host).

* Mesh.save layout (PY/mesh/_mesh.py:560-600): attributes and 'vertices'; 'en_map' rows list each
  element's nodes x-fastest (CartesianGenerator.c:743-800), checked against the node coordinates.
* MeshVariable / Swarm / SwarmVariable round trips are exact; a swarm file with rows outside the
  mesh loads the others and SwarmVariable.load maps rows through the same selection.
* MeshVariable.load(interpolate=True) through the file's 'mesh' link reproduces a bilinear field
  exactly on a finer mesh.
* The reference Blankenbach notebook's LoadFromFile path (docs/examples/03_BlankenbachBenchmark
  .ipynb, case a): the reference's 64^2 steady-state field (tests/golden/tempfield_inp_64_Ra1e4.h5)
  is loaded, interpolated onto the notebook's 128^2 mesh and run for its two steps; Nu 4.884,
  v_rms 42.865, q1 8.020 and q2 0.589 at the notebook's rtol 1e-2.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("et,res", [("Q1/dQ0", (5, 3)), ("Q2/dPc1", (2, 3, 2))])
def test_mesh_save_layout_and_load(tmp_path, et, res):
    import underworld2_amd as uw
    from underworld2_amd import _h5
    d = len(res)
    lo, hi = (0.0, -1.0, 0.5)[:d], (2.0, 1.0, 1.5)[:d]
    print(f"mesh save/load {et} {res}", flush=True)
    mesh = uw.mesh.FeMesh_Cartesian(et, res, lo, hi)
    fn = str(tmp_path / "mesh.h5")
    h = mesh.save(fn, note=3)
    assert h.filename == fn and h.pyobj is mesh
    o = 2 if et.startswith("Q2") else 1
    with _h5.File(fn, "r") as f:
        a = f.attrs
        assert a["dimensions"] == d
        assert np.array_equal(a["mesh resolution"], res) and a["mesh resolution"].dtype == np.int64
        assert np.array_equal(a["max"], hi) and np.array_equal(a["min"], lo)
        assert bool(a["regular"]) and a["elementType"] == ("Q2" if o == 2 else "Q1") and a["note"] == "3"
        V = f.read_dataset("vertices")
        en = f.read_dataset("en_map")
    assert np.array_equal(V, mesh.data)
    assert en.dtype == np.uint32 and en.shape == (np.prod(res), (o + 1) ** d)
    h_el = (np.array(hi) - np.array(lo)) / np.array(res)
    eidx = np.indices(res[::-1]).reshape(d, -1)[::-1].T  # element (i, j, k), x fastest
    lidx = np.indices((o + 1,) * d).reshape(d, -1)[::-1].T  # local node offsets, x fastest
    want = np.array(lo) + (eidx[:, None, :] + lidx[None, :, :] / o) * h_el
    assert np.allclose(V[en.astype(np.int64)], want, rtol=0, atol=1e-12)
    mesh.load(fn)
    other = uw.mesh.FeMesh_Cartesian(et, tuple(r + 1 for r in res), lo, hi)
    with pytest.raises(RuntimeError, match="resolution"):
        other.load(fn)


def test_meshvariable_round_trip_link_and_interpolation(tmp_path):
    import underworld2_amd as uw
    from underworld2_amd import _h5
    print("meshvariable save/load 2D Q1 8x6 + dQ0, interpolation 16^2 -> 32^2", flush=True)
    mesh = uw.mesh.FeMesh_Cartesian("Q1/dQ0", (8, 6), (0.0, 0.0), (1.0, 2.0))
    v = uw.mesh.MeshVariable(mesh, 2)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    rng = np.random.default_rng(4)
    v.data[:] = rng.standard_normal(v.data.shape)
    p.data[:] = rng.standard_normal(p.data.shape)
    mh = mesh.save(str(tmp_path / "m.h5"))
    v.save(str(tmp_path / "v.h5"), mh)
    p.save(str(tmp_path / "p.h5"), step=5)
    with _h5.File(str(tmp_path / "v.h5"), "r") as f:
        assert f.attrs["elementType"] == b"Q1" and f.attrs["units"] == "None"
        assert np.array_equal(f.open_linked("mesh").attrs["mesh resolution"], [8, 6])
    with _h5.File(str(tmp_path / "p.h5"), "r") as f:
        assert f.attrs["elementType"] == b"DQ0" and f.attrs["step"] == "5"
    v2 = uw.mesh.MeshVariable(mesh, 2)
    p2 = uw.mesh.MeshVariable(mesh.subMesh, 1)
    v2.load(str(tmp_path / "v.h5"))
    p2.load(str(tmp_path / "p.h5"))
    assert np.array_equal(v2.data, v.data) and np.array_equal(p2.data, p.data)
    with pytest.raises(RuntimeError, match="incompatible data shape"):
        uw.mesh.MeshVariable(mesh, 1).load(str(tmp_path / "v.h5"))

    coarse = uw.mesh.FeMesh_Cartesian("Q1/dQ0", (16, 16), (0.0, 0.0), (1.0, 1.0))
    fine = uw.mesh.FeMesh_Cartesian("Q1/dQ0", (32, 32), (0.0, 0.0), (1.0, 1.0))
    X = coarse.data
    T = uw.mesh.MeshVariable(coarse, 1)
    T.data[:, 0] = 0.3 + 1.5 * X[:, 0] - 0.7 * X[:, 1] + 2.0 * X[:, 0] * X[:, 1]
    ch = coarse.save(str(tmp_path / "c.h5"))
    T.save(str(tmp_path / "T.h5"), ch)
    T.save(str(tmp_path / "T_nolink.h5"))
    Tf = uw.mesh.MeshVariable(fine, 1)
    with pytest.raises(RuntimeError, match="different resolution"):
        Tf.load(str(tmp_path / "T.h5"))
    with pytest.raises(RuntimeError, match="associated"):
        Tf.load(str(tmp_path / "T_nolink.h5"), interpolate=True)
    Tf.load(str(tmp_path / "T.h5"), interpolate=True)
    Y = fine.data
    assert np.allclose(Tf.data[:, 0], 0.3 + 1.5 * Y[:, 0] - 0.7 * Y[:, 1] + 2.0 * Y[:, 0] * Y[:, 1],
                       rtol=0, atol=1e-13)


def test_swarm_and_swarmvariable_round_trip(tmp_path):
    import underworld2_amd as uw
    from underworld2_amd.swarm import _save_particle_data
    print("swarm save/load 2D 12x10, 7 particles per cell", flush=True)
    mesh = uw.mesh.FeMesh_Cartesian("Q1/dQ0", (12, 10), (0.0, 0.0), (1.2, 1.0))
    sw = uw.swarm.Swarm(mesh)
    sw.populate_using_layout(uw.swarm.layouts.PerCellRandomLayout(sw, 7, seed=3))
    mat = sw.add_variable("int", 1)
    strain = sw.add_variable("double", 2)
    P = sw.particleLocalCount
    mat.data[:, 0] = np.arange(P) % 5
    strain.data[:] = np.random.default_rng(2).random((P, 2))
    sh = sw.save(str(tmp_path / "s.h5"))
    mat.save(str(tmp_path / "mat.h5"), swarmHandle=sh)
    strain.save(str(tmp_path / "strain.h5"))

    sw2 = uw.swarm.Swarm(mesh)
    mat2 = sw2.add_variable("int", 1)
    strain2 = sw2.add_variable("double", 2)
    with pytest.raises(RuntimeError, match="correct state"):
        mat2.load(str(tmp_path / "mat.h5"))
    sw2.load(str(tmp_path / "s.h5"))
    mat2.load(str(tmp_path / "mat.h5"))
    strain2.load(str(tmp_path / "strain.h5"))
    assert sw2.particleLocalCount == P
    assert np.array_equal(sw2.particleCoordinates.data, sw.particleCoordinates.data)
    assert mat2.data.dtype == np.int32 and np.array_equal(mat2.data, mat.data)
    assert np.array_equal(strain2.data, strain.data)

    # rows outside the mesh are skipped by Swarm.load and by the variable mapping after it
    xy = np.array([[0.1, 0.2], [1.5, 0.5], [0.6, 0.9], [-0.1, 0.3], [1.2, 1.0]])
    val = np.array([[10], [11], [12], [13], [14]], dtype=np.int32)
    _save_particle_data(str(tmp_path / "s3.h5"), xy)
    _save_particle_data(str(tmp_path / "v3.h5"), val)
    sw3 = uw.swarm.Swarm(mesh)
    v3 = sw3.add_variable("int", 1)
    sw3.load(str(tmp_path / "s3.h5"))
    assert sw3.particleLocalCount == 3
    assert np.array_equal(sw3.particleCoordinates.data, xy[[0, 2, 4]])
    with pytest.raises(RuntimeError, match="has 3 particles"):
        v3.load(str(tmp_path / "v3.h5"))


def _nusselt(T, h):
    dTdy = (T[-1] - T[-2]) / h
    top = np.sum(h * 0.5 * (dTdy[1:] + dTdy[:-1]))
    bot = np.sum(h * 0.5 * (T[0, 1:] + T[0, :-1]))
    return -top / bot


def test_blankenbach_case_a_load_from_file():
    import underworld2_amd as uw
    res = 128
    print(f"Blankenbach case a from the reference's 64^2 h5 field, {res}^2, two steps", flush=True)
    mesh = uw.mesh.FeMesh_Cartesian("Q1/dQ0", (res, res), (0.0, 0.0), (1.0, 1.0))
    vel = uw.mesh.MeshVariable(mesh, 2)
    p = uw.mesh.MeshVariable(mesh.subMesh, 1)
    T = uw.mesh.MeshVariable(mesh, 1)
    Tdot = uw.mesh.MeshVariable(mesh, 1)
    vel.data[:] = 0.0
    p.data[:] = 0.0
    Tdot.data[:] = 0.0
    mesh64 = uw.mesh.FeMesh_Cartesian("Q1/dQ0", (64, 64), (0.0, 0.0), (1.0, 1.0))
    T64 = uw.mesh.MeshVariable(mesh64, 1)
    T64.load(os.path.join(HERE, "golden", "tempfield_inp_64_Ra1e4.h5"))
    T.data[:] = T64.evaluate(mesh)
    iw = mesh.specialSets["MinI_VertexSet"] + mesh.specialSets["MaxI_VertexSet"]
    jw = mesh.specialSets["MinJ_VertexSet"] + mesh.specialSets["MaxJ_VertexSet"]
    stokes = uw.systems.Stokes(vel, p, fn_viscosity=1.0, fn_bodyforce=1.0e4 * T * (0.0, 1.0),
                               conditions=[uw.conditions.DirichletCondition(vel, (iw, jw))])
    solver = uw.systems.Solver(stokes)
    solver.set_method("block")
    solver.set_outer_rtol(1e-7)
    ad = uw.systems.AdvectionDiffusion(phiField=T, phiDotField=Tdot, velocityField=vel, fn_diffusivity=1.0,
                                       conditions=[uw.conditions.DirichletCondition(T, (jw,))])
    step, nu_last = 0, -1.0
    while step <= 1:
        solver.solve()
        vrms = stokes.velocity_rms()
        nu = _nusselt(T.data[:, 0].reshape(res + 1, res + 1), 1.0 / res)
        print(f"step {step}: v_rms {vrms:.4f} Nu {nu:.5f}", flush=True)
        if abs((nu - nu_last) / nu) < 1e-8:
            break
        nu_last = nu
        ad.integrate(ad.get_max_dt())
        step += 1
    Tg = T.data[:, 0].reshape(res + 1, res + 1)
    q1 = -(Tg[res, 0] - Tg[res - 1, 0]) * res
    q2 = -(Tg[res, res] - Tg[res - 1, res]) * res
    print(f"Nu {nu:.4f} v_rms {vrms:.4f} q1 {q1:.4f} q2 {q2:.4f}", flush=True)
    assert np.isclose(nu, 4.884, rtol=1e-2)
    assert np.isclose(vrms, 42.865, rtol=1e-2)
    assert np.isclose(q1, 8.020, rtol=1e-2)
    assert np.isclose(q2, 0.589, rtol=1e-2)
