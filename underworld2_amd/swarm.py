"""Particle swarms, mirroring ``underworld.swarm`` (PY/swarm/_swarm.py, layouts.py).

Particle coordinates and per-cell particle tables live on the GPU.  ``particleCoordinates
.data`` is a host view refreshed after device updates; assigning into it and calling
``update_particle_owners()`` pushes the new coordinates back.
"""
import ctypes
import os as _os

import numpy as np

from . import _h5
from . import _lib as L
from . import function as _fn
from .mesh import _check_no_units


def _save_particle_data(filename, data, swarmHandle=None, units=None, **kwargs):
    """SwarmVariable.save's layout (PY/swarm/_swarmvariable.py:419-500): dataset 'data'
    [particles][components] in the variable's dtype, attributes 'proc_offset' (the one process's
    particle count, int64[1]), 'units' (the string 'None') and each keyword as a string, and with
    ``swarmHandle`` an external link 'swarm' to that swarm file."""
    if not isinstance(filename, str):
        raise TypeError("'filename' parameter must be of type 'str'")
    if units:
        raise NotImplementedError("dimensional (pint) output is not provided by this engine")
    swarm_file = None
    if swarmHandle is not None:
        if not isinstance(swarmHandle, (str, _h5.SavedFileData)):
            raise TypeError("Expected 'swarmHandle' to be of type 'uw.utils.SavedFileData'")
        swarm_file = swarmHandle if isinstance(swarmHandle, str) else swarmHandle.filename
        if not _os.path.exists(swarm_file):
            raise ValueError("You are trying to link against the swarm file '{}'\n"
                             "that does not appear to exist.".format(swarm_file))
    with _h5.File(filename, "w") as f:
        f.write_dataset("data", data)
        if swarm_file is not None:
            f.external_link("swarm", swarm_file, "./")
        f.attrs["proc_offset"] = np.array([data.shape[0]], dtype=np.int64)
        f.attrs["units"] = str(None)
        for k, v in kwargs.items():
            f.attrs[str(k)] = str(v)


class _CoordVariable:
    def __init__(self, swarm):
        self._swarm = swarm

    @property
    def data(self):
        return self._swarm._coords_host()

    def save(self, filename, collective=False, swarmHandle=None, units=None, **kwargs):
        _save_particle_data(filename, np.array(self.data, dtype=np.float64), swarmHandle, units, **kwargs)
        return _h5.SavedFileData(self, filename)


class SwarmVariable(_fn.Function):
    """Per-particle data ``swarm.add_variable(dataType, count)`` (PY/swarm/_swarmvariable.py).

    Data are held on the host; the integer variable used as a material index is uploaded
    to the device by the Stokes system when it assembles."""

    _dtypes = {"double": np.float64, "float": np.float32, "int": np.int32, "char": np.int8, "short": np.int16,
               "long": np.int64, "bool": np.bool_}

    def __init__(self, swarm, dataType, count):
        if dataType not in self._dtypes:
            raise ValueError(f"unsupported dataType {dataType}")
        self.swarm = swarm
        self.dataType = dataType
        self.count = int(count)
        self._data = np.zeros((swarm.particleLocalCount, self.count), dtype=self._dtypes[dataType])
        self.version = 0

    @property
    def data(self):
        if self._data.shape[0] != self.swarm.particleLocalCount:
            new = np.zeros((self.swarm.particleLocalCount, self.count), dtype=self._data.dtype)
            k = min(self._data.shape[0], new.shape[0])
            new[:k] = self._data[:k]
            self._data = new
        self.version += 1
        return self._data

    def save(self, filename, collective=False, swarmHandle=None, units=None, **kwargs):
        """SwarmVariable.save (PY/swarm/_swarmvariable.py:364-500); ``collective`` has no effect
        in one process."""
        _save_particle_data(filename, np.array(self.data), swarmHandle, units, **kwargs)
        return _h5.SavedFileData(self, filename)

    def load(self, filename, collective=False):
        """SwarmVariable.load (PY/swarm/_swarmvariable.py:244-362): after ``swarm.load`` of the
        matching swarm file, row ``gid`` of the file's 'data' goes to the particle that swarm.load
        created from row ``gid`` of the swarm file."""
        if not isinstance(filename, str):
            raise TypeError("'filename' parameter must be of type 'str'")
        sw = self.swarm
        if getattr(sw, "_checkpointMapsToState", None) != sw.stateId:
            raise RuntimeError("'Swarm' associate with this 'SwarmVariable' does not appear to be in the correct "
                               "state.\nPlease ensure that you have loaded the swarm prior to loading any swarm "
                               "variables.")
        gids = sw._local2globalMap
        with _h5.File(filename, "r") as f:
            shape = f.dataset_shape("data")
            if len(shape) != 2 or shape[1] != self.count:
                raise RuntimeError("Cannot load file data on current swarm. Data in file '{0}', has {1} components "
                                   "-the particlesCoords has {2} components".format(filename, shape[-1], self.count))
            if shape[0] != sw.particleGlobalCount:
                raise RuntimeError("It appears that the swarm has {} particles, but provided h5 file has {} data "
                                   "points. Please check that both the Swarm and the SwarmVariable were saved at "
                                   "the same time, and that you have reloaded using the correct "
                                   "files.".format(sw.particleGlobalCount, shape[0]))
            vals = f.read_dataset("data", self._data.dtype)
            _check_no_units(f.attrs, filename)
        self.data[:len(gids)] = vals[gids]

    def _fn_inputs(self):
        return {("swarmvar", id(self))}

    def _fn_eval(self, ctx):
        if id(self) in ctx.override:
            return ctx.override[id(self)]
        if ctx.swarm is not self.swarm:
            raise RuntimeError("SwarmVariable evaluated outside its own swarm")
        return self.data


class Swarm:
    """``Swarm(mesh, particleEscape=False)``."""

    def __init__(self, mesh, particleEscape=False, **kwargs):
        if particleEscape:
            raise ValueError("particleEscape=True is not supported by this engine")
        self.mesh = mesh
        self.particleEscape = False
        h = ctypes.c_void_p()
        L.check(L.lib().uwg_swarm_create(mesh._h, ctypes.byref(h)), "Swarm")
        self._h = h
        self._variables = []
        self._host = np.zeros((0, mesh.dim))
        self._host_valid = True
        self._host_touched = False
        self.particleCoordinates = _CoordVariable(self)
        self.stateId = 0

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                L.lib().uwg_swarm_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def particleLocalCount(self):
        P = ctypes.c_long()
        L.check(L.lib().uwg_swarm_count(self._h, ctypes.byref(P)))
        return P.value

    @property
    def particleGlobalCount(self):
        return self.particleLocalCount

    def _coords_host(self):
        if not self._host_valid:
            P = self.particleLocalCount
            self._host = np.zeros((P, self.mesh.dim))
            if P:
                L.check(L.lib().uwg_swarm_get_coords(self._h, L.dptr(self._host)), "particleCoordinates")
            self._host_valid = True
        self._host_touched = True
        return self._host

    def _push(self):
        if self._host_touched and self._host_valid and self._host.shape[0] == self.particleLocalCount:
            out = ctypes.c_long()
            rc = L.lib().uwg_swarm_set_coords(self._h, L.dptr(np.ascontiguousarray(self._host)), ctypes.byref(out))
            self._host_touched = False
            self.stateId += 1
            if rc < 0:
                raise RuntimeError("Particles appear to have left the domain, but swarm flag `particleEscape` "
                                   "is False. Check your velocity field or your particle relocation routines, "
                                   "or set the `particleEscape` swarm constructor parameter to True to allow escape.")

    def _device_moved(self):
        self._host_valid = False
        self._host_touched = False
        self.stateId += 1

    def add_variable(self, dataType, count):
        v = SwarmVariable(self, dataType, count)
        self._variables.append(v)
        return v

    def populate_using_layout(self, layout):
        layout._populate(self)
        self._device_moved()

    def add_particles_with_coordinates(self, coordinatesArray):
        c = L.f64(coordinatesArray).reshape(-1, self.mesh.dim)
        P0 = self.particleLocalCount
        rc = L.lib().uwg_swarm_add_particles(self._h, c.shape[0], L.dptr(c))
        self._device_moved()
        if rc < 0:
            raise RuntimeError(L.last_error())
        return np.arange(P0, P0 + c.shape[0], dtype=np.int32)

    def save(self, filename, collective=False, units=None, **kwargs):
        """Swarm.save (PY/swarm/_swarm.py:221-268): the particle coordinates in the layout of
        ``SwarmVariable.save``."""
        if not isinstance(filename, str):
            raise TypeError("Expected filename to be provided as a string")
        self.particleCoordinates.save(filename, collective, units=units, **kwargs)
        return _h5.SavedFileData(self, filename)

    def load(self, filename, collective=False, try_optimise=True):
        """Swarm.load (PY/swarm/_swarm.py:270-420): *adds* the file's particles to the swarm.  Rows
        whose coordinates lie outside the mesh are skipped, as the reference skips particles no
        process owns; the file rows that were added are kept, in order, for ``SwarmVariable.load``."""
        if not isinstance(filename, str):
            raise TypeError("Expected 'filename' to be provided as a string")
        dim = self.mesh.dim
        with _h5.File(filename, "r") as f:
            _check_no_units(f.attrs, filename)
            shape = f.dataset_shape("data")
            if len(shape) != 2 or shape[1] != dim:
                raise RuntimeError("Cannot load file data on current swarm. Data in file '{0}', has {1} components "
                                   "-the particlesCoords has {2} components".format(filename, shape[-1], dim))
            xyz = f.read_dataset("data", np.float64)
        lo = np.array(self.mesh.minCoord)
        hi = np.array(self.mesh.maxCoord)
        inside = np.all((xyz >= lo) & (xyz <= hi), axis=1)
        valid = np.flatnonzero(inside).astype(np.int64)
        if valid.size:
            self.add_particles_with_coordinates(xyz[valid])
        self._local2globalMap = valid
        self._checkpointMapsToState = self.stateId

    def update_particle_owners(self):
        """Swarm.update_particle_owners (PY/swarm/_swarm.py:613-654)."""
        if self._host_touched:
            self._push()
            return
        out = ctypes.c_long()
        rc = L.lib().uwg_swarm_update_owners(self._h, ctypes.byref(out))
        self.stateId += 1
        if rc < 0:
            raise RuntimeError(L.last_error())

    @property
    def owningCell(self):
        class _V:
            pass

        v = _V()
        P = self.particleLocalCount
        o = np.zeros(P, dtype=np.int32)
        if P:
            L.check(L.lib().uwg_swarm_get_owners(self._h, L.iptr(o)))
        v.data = o.reshape(-1, 1)
        return v

    def cell_tables(self):
        """(cell_ptr[E+1], cell_pid[P]): each element's particles in table order."""
        E = self.mesh.elementsGlobal
        ptr = np.zeros(E + 1, dtype=np.int32)
        pid = np.zeros(max(self.particleLocalCount, 1), dtype=np.int32)
        L.check(L.lib().uwg_swarm_get_cell_tables(self._h, L.iptr(ptr), L.iptr(pid)))
        return ptr, pid[:self.particleLocalCount]

    def integration_points(self, dvc=True, resolution=15):
        """Voronoi integration swarm (coincident map + DVC weights), slot order."""
        self._push()
        L.check(L.lib().uwg_swarm_integration_points(self._h, 1 if dvc else 0, resolution), "repopulate")
        P = self.particleLocalCount
        xi = np.zeros((P, self.mesh.dim))
        w = np.zeros(P)
        L.check(L.lib().uwg_swarm_get_integration(self._h, L.dptr(xi), L.dptr(w)))
        return xi, w


class PopulationControl:
    """``PopulationControl(swarm, deleteThreshold=0.006, splitThreshold=0.25, maxDeletions=0,
    maxSplits=3)`` (PY/swarm/_population_control.py:13-140): per element, particles whose DVC
    volume fraction exceeds ``splitThreshold`` are split (a copy at the centroid of their Voronoi
    cell), those below ``deleteThreshold`` deleted.  Every SwarmVariable follows its particles."""

    def __init__(self, swarm, deleteThreshold=0.006, splitThreshold=0.25, maxDeletions=0, maxSplits=3,
                 aggressive=False, aggressiveThreshold=0.8, particlesPerCell=None, **kwargs):
        if not isinstance(swarm, Swarm):
            raise ValueError("Provided 'swarm' must be of class 'Swarm'.")
        for name, v in (("deleteThreshold", deleteThreshold), ("splitThreshold", splitThreshold)):
            if not isinstance(v, (int, float)):
                raise TypeError(f"'{name}' parameter must of type 'float'.")
            if v < 0.0 or v > 1.0:
                raise ValueError(f"'{name}' parameter must take values in [0.,1.].")
        if deleteThreshold > splitThreshold:
            raise ValueError("'deleteThreshold' must be less than 'splitThreshold'.")
        for name, v in (("maxDeletions", maxDeletions), ("maxSplits", maxSplits)):
            if not isinstance(v, int):
                raise TypeError(f"'{name}' parameter must of type 'int'.")
            if v < 0:
                raise ValueError(f"'{name}' parameter must take non-negative values.")
        if not isinstance(aggressive, bool):
            raise TypeError("'aggressive' parameter must of type 'bool'.")
        if aggressive:
            raise NotImplementedError("aggressive population control is not provided by this engine")
        self._swarm = swarm
        self._args = (float(deleteThreshold), float(splitThreshold), int(maxDeletions), int(maxSplits))

    def repopulate(self):
        """Split / delete particles, then move every swarm variable's rows with its particles."""
        sw = self._swarm
        sw._push()
        P0 = sw.particleLocalCount
        for v in sw._variables:  # bring every variable to the current particle count first
            v.data
        cap = P0 + sw.mesh.elementsGlobal * self._args[3]
        src = np.zeros(max(cap, 1), dtype=np.int64)
        pn = ctypes.c_long()
        L.check(L.lib().uwg_swarm_population_control(sw._h, *self._args, cap,
                                                      src.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                                                      ctypes.byref(pn)), "PopulationControl.repopulate")
        src = src[:pn.value]
        if pn.value != P0 or np.any(src != np.arange(pn.value)):
            for v in sw._variables:
                v._data = v._data[src].copy()
                v.version += 1
            sw._device_moved()


class _Layouts:
    class PerCellGaussLayout:
        """``PerCellGaussLayout(swarm, gaussPointCount)`` (PY/swarm/layouts.py:53-100): particles
        at the tensor Gauss-Legendre points of every element, elements in order, x fastest."""

        def __init__(self, swarm, gaussPointCount, **kwargs):
            if not isinstance(gaussPointCount, int):
                raise TypeError("'gaussPointCount' object passed in must be of type 'int'")
            if gaussPointCount not in (1, 2, 3, 4, 5):
                raise ValueError("'gaussPointCount' object passed in must take a value between \n between 1 "
                                 "and 5 inclusive.")
            self.swarm = swarm
            self.gaussPointCount = gaussPointCount

        def _populate(self, swarm):
            m = swarm.mesh
            d = m.dim
            g, _ = np.polynomial.legendre.leggauss(self.gaussPointCount)
            res = np.array(m.elementRes)
            lo, hi = np.array(m.minCoord), np.array(m.maxCoord)
            h = (hi - lo) / res
            idx = np.indices(tuple(res[::-1])).reshape(d, -1)[::-1]          # element (i, j, k), x fastest
            cen = np.stack([lo[k] + (idx[k] + 0.5) * h[k] for k in range(d)], -1)
            q = np.stack(np.meshgrid(*([g] * d), indexing="ij"), -1).reshape(-1, d)[:, ::-1]  # x fastest
            pts = (cen[:, None, :] + q[None, :, :] * (0.5 * h)[None, None, :]).reshape(-1, d)
            swarm.add_particles_with_coordinates(pts)

    class PerCellRandomLayout:
        """``PerCellRandomLayout(swarm, particlesPerCell, seed=13)`` (PY/swarm/layouts.py:186-229)."""

        def __init__(self, swarm, particlesPerCell, seed=13, **kwargs):
            self.swarm = swarm
            self.particlesPerCell = int(particlesPerCell)
            self.seed = int(seed)

        def _populate(self, swarm):
            L.check(L.lib().uwg_swarm_populate_random(swarm._h, self.particlesPerCell, self.seed), "populate")

    class PerCellSpaceFillerLayout:
        """``PerCellSpaceFillerLayout(swarm, particlesPerCell)`` (PY/swarm/layouts.py:146-181):
        per-axis Sobol sequences, with the reference's random skip (seed 13)."""

        def __init__(self, swarm, particlesPerCell, **kwargs):
            if not isinstance(particlesPerCell, int):
                raise TypeError("'particlesPerCell' object passed in must be of type 'int'.")
            if particlesPerCell < 1:
                raise ValueError("'particlesPerCell' object passed in must take a value greater than zero.")
            self.swarm = swarm
            self.particlesPerCell = particlesPerCell
            self.seed = 13

        def _populate(self, swarm):
            L.check(L.lib().uwg_swarm_populate_sobol(swarm._h, self.particlesPerCell, self.seed), "populate")


layouts = _Layouts()
