"""Meshes and mesh variables, mirroring ``underworld.mesh`` (PY/mesh/_mesh.py).

``FeMesh_Cartesian("Q1/dQ0", elementRes, minCoord, maxCoord)`` is a regular box mesh of
Q1 velocity elements with a piecewise-constant (dQ0) pressure sub-mesh.  Node numbering is
lexicographic with x fastest (CartesianGenerator.c:2001-2039); element-local nodes are
ordered (-,-,-), (+,-,-), (-,+,-), (+,+,-), then +z.

``FeMesh_Cartesian("Q2/dPc1", ...)`` has (2 res + 1)^dim vertices on the same element grid
(C2Generator.c:147-260) and a dPc1 sub-mesh of dim+1 pressure vertices per element, stored
element-major, at the element-local points of Inner2DGenerator.c:286-353.

``MeshVariable.data`` is a host numpy array of shape [nodes][nodeDofCount] (or
[elements][count] on the dQ0 sub-mesh).  The authoritative copy lives on the GPU; the host
view is refreshed on access and pushed back before the next device operation.
"""
import ctypes
import os as _os

import numpy as np

from . import _h5
from . import _lib as L
from . import function as _fn


def _check_no_units(attrs, filename):
    """Files whose 'units' attribute names a unit hold dimensional data (PY/scaling); this engine
    works in model units only, so those are refused rather than misread."""
    u = attrs.get("units")
    if u is not None:
        u = u.decode() if isinstance(u, bytes) else str(u)
        if u not in ("None", ""):
            raise NotImplementedError(f"'{filename}' holds data in units '{u}'; dimensional (pint) data are "
                                      "not provided by this engine")


class IndexSet:
    """Node index set (PY/container/_indexset.py); ``a + b`` is the union."""

    def __init__(self, indices, size):
        self._idx = np.unique(np.asarray(indices, dtype=np.int64))
        self.size = size

    def __add__(self, other):
        return IndexSet(np.union1d(self._idx, other._idx), self.size)

    def __len__(self):
        return len(self._idx)

    @property
    def data(self):
        return self._idx

    def __contains__(self, i):
        return int(i) in set(self._idx.tolist())


_DPC1_LOCAL = {2: ((-0.5, -0.5), (0.5, -0.5), (0.0, 0.5)),
               3: ((-0.5, -0.5, -0.5), (0.25, 0.25, 0.25), (0.5, -0.25, -0.5), (-0.25, 0.5, 0.5))}


class _SubMesh:
    """The pressure mesh: one value per element (dQ0) or dim+1 per element (dPc1)."""

    def __init__(self, parent, kind):
        self.parent = parent
        self.elementType = kind
        self.elementRes = parent.elementRes
        self.dim = parent.dim
        self._per_element = 1 if kind == "DQ0" else parent.dim + 1

    @property
    def nodesGlobal(self):
        return self.parent.elementsGlobal * self._per_element

    def add_variable(self, nodeDofCount, dataType="double", **kwargs):
        """MeshVariable on this sub-mesh (PY/mesh/_mesh.py FeMesh.add_variable)."""
        return MeshVariable(self, nodeDofCount, dataType, **kwargs)

    @property
    def data(self):
        """Pressure vertex coordinates: element centres (dQ0) or the dPc1 inner vertices."""
        p = self.parent
        res = np.array(p.elementRes)
        lo, hi = np.array(p.minCoord), np.array(p.maxCoord)
        h = (hi - lo) / res
        idx = np.indices(tuple(res[::-1])).reshape(p.dim, -1)[::-1]
        cen = np.stack([lo[d] + (idx[d] + 0.5) * h[d] for d in range(p.dim)], -1)
        if self._per_element == 1:
            return cen
        loc = np.array(_DPC1_LOCAL[p.dim])
        return (cen[:, None, :] + loc[None, :, :] * (0.5 * h)[None, None, :]).reshape(-1, p.dim)


class FeMesh_Cartesian:
    """Regular Cartesian finite-element mesh (PY/mesh/_mesh.py:1046-1190).

    The element pairs ``"Q1/dQ0"`` and ``"Q2/dPc1"`` are provided by this engine.
    """

    def __init__(self, elementType="Q1/dQ0", elementRes=(4, 4), minCoord=(0.0, 0.0), maxCoord=(1.0, 1.0),
                 periodic=None, partitioned=True, **kwargs):
        et = elementType.upper().replace("-", "/").replace(",", "/") if isinstance(elementType, str) else None
        if et not in ("Q1/DQ0", "Q1", "Q2/DPC1", "Q2"):
            raise ValueError("Only the 'Q1/dQ0' and 'Q2/dPc1' element types are provided by this engine.")
        if periodic is not None and any(periodic):
            raise ValueError("Periodic meshes are not supported by this engine.")
        if len(elementRes) not in (2, 3) or len(minCoord) != len(elementRes) or len(maxCoord) != len(elementRes):
            raise ValueError("elementRes, minCoord and maxCoord must have the same length (2 or 3).")
        self._order = 2 if et.startswith("Q2") else 1
        self.elementType = "Q2" if self._order == 2 else "Q1"
        self.dim = len(elementRes)
        self.elementRes = tuple(int(r) for r in elementRes)
        self.minCoord = tuple(float(c) for c in minCoord)
        self.maxCoord = tuple(float(c) for c in maxCoord)
        h = ctypes.c_void_p()
        res = np.array(self.elementRes, dtype=np.int32)
        lo = np.array(self.minCoord)
        hi = np.array(self.maxCoord)
        create = L.lib().uwg_mesh_create_q2 if self._order == 2 else L.lib().uwg_mesh_create
        L.check(create(self.dim, L.iptr(res), L.dptr(lo), L.dptr(hi), ctypes.byref(h)), "FeMesh_Cartesian")
        self._h = h
        self._nn = tuple(self._order * r + 1 for r in self.elementRes)
        self.nodesGlobal = int(np.prod(self._nn))
        self.elementsGlobal = int(np.prod(self.elementRes))
        self.subMesh = _SubMesh(self, "DPC1" if self._order == 2 else "DQ0")
        self._data = None
        self._special = None

    def add_variable(self, nodeDofCount, dataType="double", **kwargs):
        """MeshVariable on this mesh's nodes (PY/mesh/_mesh.py FeMesh.add_variable)."""
        return MeshVariable(self, nodeDofCount, dataType, **kwargs)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                L.lib().uwg_mesh_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def data(self):
        """Node coordinates [nodes][dim]."""
        if self._data is None:
            out = np.zeros((self.nodesGlobal, self.dim))
            L.check(L.lib().uwg_mesh_coords(self._h, L.dptr(out)), "mesh.data")
            out.flags.writeable = False
            self._data = out
        return self._data

    @property
    def specialSets(self):
        """Wall node sets: Min/Max I/J/K _VertexSet, plus AllWalls_VertexSet."""
        if self._special is None:
            n = list(self._nn)
            idx = np.indices(n[::-1]).reshape(self.dim, -1)[::-1]
            names = "IJK"
            sets = {}
            allw = np.zeros(self.nodesGlobal, dtype=bool)
            for d in range(self.dim):
                lo = np.flatnonzero(idx[d] == 0)
                hi = np.flatnonzero(idx[d] == n[d] - 1)
                sets[f"Min{names[d]}_VertexSet"] = IndexSet(lo, self.nodesGlobal)
                sets[f"Max{names[d]}_VertexSet"] = IndexSet(hi, self.nodesGlobal)
                allw[lo] = True
                allw[hi] = True
            sets["AllWalls_VertexSet"] = IndexSet(np.flatnonzero(allw), self.nodesGlobal)
            self._special = sets
        return self._special

    def _element_nodes(self):
        """[elements][(order+1)^dim] global node ids; each element's nodes x-fastest over its node
        grid (the Q1 order of CartesianGenerator.c:743-800, the Q2 order of C2Generator.c)."""
        d, o = self.dim, self._order
        nn = np.array(self._nn, dtype=np.int64)
        stride = np.concatenate([[1], np.cumprod(nn[:-1])])
        eidx = np.indices(self.elementRes[::-1]).reshape(d, -1)[::-1].astype(np.int64)  # x fastest
        lidx = np.indices((o + 1,) * d).reshape(d, -1)[::-1].astype(np.int64)
        ids = sum((o * eidx[k][:, None] + lidx[k][None, :]) * stride[k] for k in range(d))
        return ids.astype(np.uint32)

    def save(self, filename, units=None, **kwargs):
        """Mesh.save (PY/mesh/_mesh.py:560-600): root attributes 'dimensions', 'mesh resolution',
        'max', 'min', 'regular', 'elementType' and each keyword as a string attribute; datasets
        'vertices' [nodes][dim] (float64) and 'en_map' [elements][nodes per element] (uint32)."""
        if not isinstance(filename, str):
            raise TypeError("'filename', must be of type 'str'")
        if units:
            raise NotImplementedError("dimensional (pint) output is not provided by this engine")
        with _h5.File(filename, "w") as f:
            f.attrs["dimensions"] = self.dim
            f.attrs["mesh resolution"] = np.array(self.elementRes, dtype=np.int64)
            f.attrs["max"] = np.array(self.maxCoord)
            f.attrs["min"] = np.array(self.minCoord)
            f.attrs["regular"] = np.bool_(True)
            f.attrs["elementType"] = self.elementType
            for k, v in kwargs.items():
                f.attrs[str(k)] = str(v)
            f.write_dataset("vertices", self.data)
            f.write_dataset("en_map", self._element_nodes())
        return _h5.SavedFileData(self, filename)

    def load(self, filename):
        """Mesh.load (PY/mesh/_mesh.py:602-680): the file's resolution must equal this mesh's and
        its vertex array must have this mesh's shape.  The engine's meshes are regular boxes, so
        vertices that differ from this mesh's own (a deformed mesh) are refused."""
        if not isinstance(filename, str):
            raise TypeError("Expected filename to be provided as a string")
        with _h5.File(filename, "r") as f:
            res = f.attrs.get("mesh resolution")
            if res is None:
                raise RuntimeError("Can't read the 'mesh resolution' for the field hdf5 file, was it created correctly?")
            if not np.array_equal(np.asarray(res).ravel(), np.array(self.elementRes)):
                raise RuntimeError("Provided file mesh resolution does not appear to correspond to\n"
                                   "resolution of mesh object.")
            _check_no_units(f.attrs, filename)
            shape = f.dataset_shape("vertices")
            if len(shape) != 2 or shape[1] != self.dim:
                raise RuntimeError("Can't load hdf5 '{0}', incompatible data shape".format(filename))
            if shape[0] != self.nodesGlobal:
                raise RuntimeError("Provided data file appears to be for a different resolution mesh.")
            verts = f.read_dataset("vertices", np.float64)
        span = np.array(self.maxCoord) - np.array(self.minCoord)
        if np.abs(verts - self.data).max(initial=0.0) > 1e-12 * span.max():
            raise NotImplementedError("the file holds a deformed mesh; this engine's meshes are regular boxes")

    def _min_separation(self):
        v = ctypes.c_double()
        L.check(L.lib().uwg_mesh_min_separation(self._h, ctypes.byref(v)))
        return v.value


class MeshVariable(_fn.Function):
    """Nodal (or dQ0 element) field, ``MeshVariable(mesh, nodeDofCount)``; as a function it
    is interpolated at the evaluation points (FeVariableFn.cpp)."""

    def __init__(self, mesh, nodeDofCount, dataType="double", **kwargs):
        if dataType != "double":
            raise ValueError("Only 'double' mesh variables are provided by this engine.")
        on_elements = isinstance(mesh, _SubMesh)
        base = mesh.parent if on_elements else mesh
        self.mesh = mesh
        self._base = base
        self.nodeDofCount = int(nodeDofCount)
        self._on_elements = on_elements
        n = mesh.nodesGlobal
        kind = 0 if not on_elements else (1 if mesh.elementType == "DQ0" else 2)
        h = ctypes.c_void_p()
        L.check(L.lib().uwg_field_create(base._h, self.nodeDofCount, kind, ctypes.byref(h)), "MeshVariable")
        self._h = h
        self._host = np.zeros((n, self.nodeDofCount))
        self._dev_newer = False
        self._host_touched = False
        self.version = 0  # bumps whenever the values may have changed

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                L.lib().uwg_field_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def data(self):
        if self._dev_newer:
            L.check(L.lib().uwg_field_get(self._h, L.dptr(self._host)), "MeshVariable.data")
            self._dev_newer = False
        self._host_touched = True
        self.version += 1
        return self._host

    @data.setter
    def data(self, value):
        self.data[...] = value

    def save(self, filename, meshHandle=None, units=None, **kwargs):
        """MeshVariable.save (PY/mesh/_meshvariable.py:330-470): dataset 'data' [nodes][dof]
        (float64, global node order), attributes 'units' (the string 'None'), 'elementType'
        (fixed-length ASCII) and each keyword as a string; with ``meshHandle`` (what
        ``mesh.save`` returned, or a file name) an external link 'mesh' to that file's root."""
        if not isinstance(filename, str):
            raise TypeError("Expected 'filename' to be provided as a string")
        if units:
            raise NotImplementedError("dimensional (pint) output is not provided by this engine")
        mesh_file = None
        if meshHandle:
            if not isinstance(meshHandle, (str, _h5.SavedFileData)):
                raise TypeError("Expected 'meshHandle' to be of type 'uw.utils.SavedFileData'")
            mesh_file = meshHandle if isinstance(meshHandle, str) else meshHandle.filename
            if not _os.path.exists(mesh_file):
                raise ValueError("You are trying to link against the mesh file '{}'\n"
                                 "that does not appear to exist. If you need to link \n"
                                 "against a mesh file, please make sure it is created first.".format(mesh_file))
        data = np.array(self.data, dtype=np.float64)
        with _h5.File(filename, "w") as f:
            f.write_dataset("data", data)
            for k, v in kwargs.items():
                f.attrs[str(k)] = str(v)
            f.attrs["units"] = str(None)
            f.attrs["elementType"] = np.bytes_(self.mesh.elementType.encode())
            if mesh_file is not None:
                f.external_link("mesh", mesh_file, ".")
        return _h5.SavedFileData(self, filename)

    def load(self, filename, interpolate=False):
        """MeshVariable.load (PY/mesh/_meshvariable.py:472-600).  A file with this variable's node
        count is copied (converted to float64); another resolution needs ``interpolate=True`` and
        the file's 'mesh' link, from whose attributes the file's regular mesh is rebuilt and the
        file field interpolated at this variable's nodes."""
        if not isinstance(filename, str):
            raise TypeError("Expected filename to be provided as a string")
        with _h5.File(filename, "r") as f:
            shape = f.dataset_shape("data")
            if len(shape) != 2 or shape[1] != self.nodeDofCount:
                raise RuntimeError("Can't load hdf5 '{0}', incompatible data shape".format(filename))
            _check_no_units(f.attrs, filename)
            if shape[0] == self.mesh.nodesGlobal:
                self.data[:] = f.read_dataset("data", np.float64)
                return
            if not interpolate:
                raise RuntimeError("Provided data file appears to be for a different resolution MeshVariable.\n"
                                   "If you would like to interpolate the data to the current variable, please set\n"
                                   "the 'interpolate' parameter. Check docstring for important caveats of "
                                   "interpolation method.")
            m = f.open_linked("mesh")
            if m is None:
                raise RuntimeError("The hdf5 field to be loaded with interpolation must have an associated "
                                   "'mesh' hdf5 file. Resave the field with its associated mesh."
                                   "i.e. myField.save(\"filename.h5\", meshFilename)")
            res, hi, lo = (m.attrs.get(k) for k in ("mesh resolution", "max", "min"))
            for val, k in ((res, "mesh resolution"), (hi, "max"), (lo, "min")):
                if val is None:
                    raise RuntimeError(f"Can't read the '{k}' for the field hdf5 file, was it created correctly?")
            regular = m.attrs.get("regular")
            if regular is not None and not bool(np.asarray(regular).ravel()[0]):
                raise RuntimeError("Saved mesh file appears to correspond to a irregular mesh.\n"
                                   "Interpolating from irregular mesh not currently supported.")
            et = m.attrs.get("elementType")
            et = "Q1" if et is None else (et.decode() if isinstance(et, bytes) else str(et))
            src_mesh = FeMesh_Cartesian(et, tuple(int(r) for r in np.ravel(res)), tuple(np.ravel(lo)),
                                        tuple(np.ravel(hi)))
            if shape[0] != src_mesh.nodesGlobal:
                raise RuntimeError("The saved mesh file can't be read onto the interpolation grid.\n"
                                   "Note: only nodal (vertex) variables can be interpolated by this engine")
            src = MeshVariable(src_mesh, self.nodeDofCount)
            src.data[:] = f.read_dataset("data", np.float64)
        self.data[:] = src.evaluate(self.mesh.data)

    # device synchronisation helpers used by the systems
    def _push(self):
        if self._host_touched:
            L.check(L.lib().uwg_field_set(self._h, L.dptr(self._host)), "MeshVariable push")
            self._host_touched = False

    def _mark_device_written(self):
        self._dev_newer = True
        self._host_touched = False
        self.version += 1

    def _interpolate(self, coords):
        """Interpolate the nodal field at points (FeVariable_InterpolateValueAt)."""
        if self._on_elements:
            raise NotImplementedError("evaluation of sub-mesh (element) variables is not provided by this engine")
        self._push()
        x = L.f64(coords).reshape(-1, self._base.dim)
        v = np.zeros((x.shape[0], self.nodeDofCount))
        ok = np.zeros(x.shape[0], dtype=np.int32)
        L.check(L.lib().uwg_field_interpolate(self._h, x.shape[0], L.dptr(x), L.dptr(v), L.iptr(ok)), "evaluate")
        if not ok.all():
            raise RuntimeError("Some coordinates lie outside the mesh")
        return v

    # function interface (a MeshVariable used as a function is evaluated by interpolation)
    def _fn_inputs(self):
        return {("meshvar", id(self))}

    def _fn_eval(self, ctx):
        return self._interpolate(ctx.coords)
