"""HDF5 I/O (underworld2_amd/_h5.py) on the CPU.

* The reference's own steady-state Blankenbach field (tests/golden/tempfield_inp_64_Ra1e4.h5,
  written by the reference's MeshVariable.save through h5py) is read through the library and
  checked against the notebook's documented values for it
  (docs/examples/03_BlankenbachBenchmark.ipynb): Nusselt number 4.884 and corner gradients
  q1 = 8.020, q2 = 0.589, each at the notebook's rtol 1e-2.  A transposed or mis-typed read
  fails all three (q1 of the transposed field is 0).
* Every attribute and dataset kind the save methods write round-trips, and the header of a
  written file has the HDF5 types the reference's h5py files carry (checked with h5dump when
  it is present).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from underworld2_amd import _h5

HERE = os.path.dirname(os.path.abspath(__file__))
FIELD = os.path.join(HERE, "golden", "tempfield_inp_64_Ra1e4.h5")


def _nusselt(T, h):
    """-int_top dT/dy dx / int_bottom T dx of a Q1 field on a uniform grid (exact for Q1)."""
    dTdy = (T[-1] - T[-2]) / h
    top = np.sum(h * 0.5 * (dTdy[1:] + dTdy[:-1]))
    bot = np.sum(h * 0.5 * (T[0, 1:] + T[0, :-1]))
    return -top / bot


def test_reference_blankenbach_field_reads_to_documented_values():
    with _h5.File(FIELD, "r") as f:
        assert f.dataset_shape("data") == (65 * 65, 1)
        raw = f.read_dataset("data")
        assert raw.dtype == np.float32  # the file type, read in its natural dtype
        T = f.read_dataset("data", np.float64)
        assert np.array_equal(T, raw.astype(np.float64))
        assert f.attrs["elementType"] == b"Q1"  # np.string_ attribute (fixed length, NULLPAD)
        assert f.attrs.keys() == ["elementType"]
        assert "mesh" in f  # an external link whose target the reference did not ship
        assert f.open_linked("mesh") is None
    T = T[:, 0].reshape(65, 65)  # node id = i + 65 j, x fastest
    h = 1.0 / 64
    assert np.isclose(_nusselt(T, h), 4.884, rtol=1e-2)
    q1 = -(T[64, 0] - T[63, 0]) / h
    q2 = -(T[64, 64] - T[63, 64]) / h
    assert np.isclose(q1, 8.020, rtol=1e-2)
    assert np.isclose(q2, 0.589, rtol=1e-2)
    assert np.all(T[0] == 1.0) and np.all(T[64] == 0.0)
    assert not np.isclose(_nusselt(T.T, h), 4.884, rtol=1e-2) and -(T.T[64, 0] - T.T[63, 0]) / h == 0.0


def test_attribute_and_dataset_round_trip(tmp_path):
    fn = str(tmp_path / "a.h5")
    target = str(tmp_path / "b.h5")
    with _h5.File(target, "w") as g:
        g.attrs["mesh resolution"] = np.array([4, 3], dtype=np.int64)
        g.attrs["regular"] = np.bool_(True)
    data = {
        "f64": np.arange(12.0).reshape(6, 2) * 0.5,
        "f32": (np.arange(5, dtype=np.float32) / 3).reshape(5, 1),
        "i32": np.arange(-6, 6, dtype=np.int32).reshape(4, 3),
        "u32": np.arange(8, dtype=np.uint32).reshape(2, 4),
        "i8": np.array([[1], [0], [-3]], dtype=np.int8),
        "empty": np.zeros((0, 3)),
    }
    with _h5.File(fn, "w") as f:
        for k, v in data.items():
            f.write_dataset(k, v)
        f.write_dataset("f64", data["f64"] + 1.0)  # replacing an existing dataset
        f.attrs["dimensions"] = 2
        f.attrs["max"] = (1.0, 2.5)
        f.attrs["regular"] = np.bool_(False)
        f.attrs["elementType"] = np.bytes_(b"DQ0")
        f.attrs["units"] = "None"
        f.attrs["note"] = "ünïcode"
        f.attrs["proc_offset"] = np.array([7], dtype=np.int64)
        f.attrs["units"] = "None"  # rewriting an attribute
        f.external_link("mesh", target, ".")
    with _h5.File(fn, "r") as f:
        assert sorted(f.attrs.keys()) == sorted(["dimensions", "max", "regular", "elementType", "units", "note",
                                                 "proc_offset"])
        a = f.attrs
        assert a["dimensions"] == 2 and np.asarray(a["dimensions"]).shape == ()
        assert np.array_equal(a["max"], [1.0, 2.5])
        assert a["regular"] is np.False_ or a["regular"] == False  # noqa: E712
        assert a["elementType"] == b"DQ0"
        assert a["units"] == "None" and a["note"] == "ünïcode"
        assert np.array_equal(a["proc_offset"], [7]) and a["proc_offset"].dtype == np.int64
        assert a.get("absent") is None
        for k, v in data.items():
            got = f.read_dataset(k)
            want = v + 1.0 if k == "f64" else v
            assert got.dtype == v.dtype and np.array_equal(got, want), k
        assert np.array_equal(f.read_dataset("i32", np.float64), data["i32"].astype(np.float64))
        m = f.open_linked("mesh")
        assert m is not None
        assert np.array_equal(m.attrs["mesh resolution"], [4, 3])
        assert bool(m.attrs["regular"])
    with pytest.raises(KeyError):
        with _h5.File(fn, "r") as f:
            f.read_dataset("nope")
    with pytest.raises(FileNotFoundError):
        _h5.File(str(tmp_path / "missing.h5"), "r")


def test_particle_file_layout_matches_reference_headers(tmp_path):
    """The swarm layout: 'data' float64 [P][dim] and 'proc_offset' int64[1] (the header of the
    reference's docs/examples/input/07_ShearBandsPureShear/particle_coords.h5), plus 'units'."""
    from underworld2_amd.swarm import _save_particle_data
    fn = str(tmp_path / "swarm.h5")
    xy = np.random.default_rng(1).random((100, 2))
    _save_particle_data(fn, xy)
    with _h5.File(fn, "r") as f:
        assert f.dataset_shape("data") == (100, 2)
        assert f.read_dataset("data").dtype == np.float64
        assert np.array_equal(f.read_dataset("data"), xy)
        po = f.attrs["proc_offset"]
        assert po.dtype == np.int64 and po.shape == (1,) and po[0] == 100
        assert f.attrs["units"] == "None"
    h5dump = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
    if h5dump is None:
        pytest.skip("h5dump not present")
    hdr = subprocess.run([h5dump, "-H", fn], capture_output=True, text=True, check=True).stdout
    assert 'ATTRIBUTE "proc_offset" {\n      DATATYPE  H5T_STD_I64LE\n      DATASPACE  SIMPLE { ( 1 ) / ( 1 ) }' in hdr
    assert 'DATASET "data" {\n      DATATYPE  H5T_IEEE_F64LE\n      DATASPACE  SIMPLE { ( 100, 2 ) / ( 100, 2 ) }' in hdr
    assert "STRSIZE H5T_VARIABLE" in hdr and "H5T_CSET_UTF8" in hdr  # h5py's str attribute
