"""HDF5 files in the layouts the reference's ``save``/``load`` methods use, through the HDF5 C
library (``libhdf5``, 1.10 ABI) over ctypes.  The reference writes them with h5py
(PY/utils/_io.py:25-100); this module reproduces the subset those methods touch:

* whole-array datasets of float/integer type (read with conversion to the requested dtype);
* attributes: numeric scalars and 1-D arrays, h5py's bool (an int8 enum FALSE/TRUE),
  fixed-length NULLPAD ASCII strings (``np.string_``) and variable-length UTF-8 strings
  (h5py's ``str``);
* external links (``h5py.ExternalLink``) and the attributes of the object they point at.

The library is located at import of the first file operation (``UWG_HDF5_LIB`` overrides the
search); its absence raises ``RuntimeError`` naming what was searched.
"""
import ctypes
import ctypes.util
import os

import numpy as np

hid_t = ctypes.c_int64
herr_t = ctypes.c_int
htri_t = ctypes.c_int
hsize_t = ctypes.c_ulonglong

_ACC_RDONLY, _ACC_RDWR, _ACC_TRUNC = 0x0000, 0x0001, 0x0002
_P_DEFAULT = 0
_S_ALL = 0
_S_SCALAR = 0
_T_INTEGER, _T_FLOAT, _T_STRING, _T_ENUM = 0, 1, 3, 8
_T_STR_NULLTERM, _T_STR_NULLPAD = 0, 1
_T_CSET_ASCII, _T_CSET_UTF8 = 0, 1
_T_SGN_NONE = 0
_T_VARIABLE = ctypes.c_size_t(-1).value

_lib = None
_types = {}

_SIGNATURES = {
    "H5open": ([], herr_t),
    "H5Eset_auto2": ([hid_t, ctypes.c_void_p, ctypes.c_void_p], herr_t),
    "H5Fcreate": ([ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t], hid_t),
    "H5Fopen": ([ctypes.c_char_p, ctypes.c_uint, hid_t], hid_t),
    "H5Fclose": ([hid_t], herr_t),
    "H5Screate": ([ctypes.c_int], hid_t),
    "H5Screate_simple": ([ctypes.c_int, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)], hid_t),
    "H5Sget_simple_extent_ndims": ([hid_t], ctypes.c_int),
    "H5Sget_simple_extent_dims": ([hid_t, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)], ctypes.c_int),
    "H5Sclose": ([hid_t], herr_t),
    "H5Dcreate2": ([hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t], hid_t),
    "H5Dopen2": ([hid_t, ctypes.c_char_p, hid_t], hid_t),
    "H5Dwrite": ([hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p], herr_t),
    "H5Dread": ([hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p], herr_t),
    "H5Dget_space": ([hid_t], hid_t),
    "H5Dget_type": ([hid_t], hid_t),
    "H5Dclose": ([hid_t], herr_t),
    "H5Dvlen_reclaim": ([hid_t, hid_t, hid_t, ctypes.c_void_p], herr_t),
    "H5Acreate2": ([hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t], hid_t),
    "H5Aopen": ([hid_t, ctypes.c_char_p, hid_t], hid_t),
    "H5Aexists": ([hid_t, ctypes.c_char_p], htri_t),
    "H5Adelete": ([hid_t, ctypes.c_char_p], herr_t),
    "H5Awrite": ([hid_t, hid_t, ctypes.c_void_p], herr_t),
    "H5Aread": ([hid_t, hid_t, ctypes.c_void_p], herr_t),
    "H5Aget_space": ([hid_t], hid_t),
    "H5Aget_type": ([hid_t], hid_t),
    "H5Aclose": ([hid_t], herr_t),
    "H5Aget_name_by_idx": ([hid_t, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, hsize_t, ctypes.c_char_p,
                            ctypes.c_size_t, hid_t], ctypes.c_ssize_t),
    "H5Tcopy": ([hid_t], hid_t),
    "H5Tclose": ([hid_t], herr_t),
    "H5Tset_size": ([hid_t, ctypes.c_size_t], herr_t),
    "H5Tset_strpad": ([hid_t, ctypes.c_int], herr_t),
    "H5Tset_cset": ([hid_t, ctypes.c_int], herr_t),
    "H5Tget_class": ([hid_t], ctypes.c_int),
    "H5Tget_size": ([hid_t], ctypes.c_size_t),
    "H5Tget_sign": ([hid_t], ctypes.c_int),
    "H5Tget_super": ([hid_t], hid_t),
    "H5Tis_variable_str": ([hid_t], htri_t),
    "H5Tenum_create": ([hid_t], hid_t),
    "H5Tenum_insert": ([hid_t, ctypes.c_char_p, ctypes.c_void_p], herr_t),
    "H5Lexists": ([hid_t, ctypes.c_char_p, hid_t], htri_t),
    "H5Ldelete": ([hid_t, ctypes.c_char_p, hid_t], herr_t),
    "H5Lcreate_external": ([ctypes.c_char_p, ctypes.c_char_p, hid_t, ctypes.c_char_p, hid_t, hid_t], herr_t),
    "H5Oopen": ([hid_t, ctypes.c_char_p, hid_t], hid_t),
    "H5Oclose": ([hid_t], herr_t),
}

_NATIVE = {  # numpy dtype -> name of the library's native type global
    np.dtype(np.float64): "H5T_NATIVE_DOUBLE_g", np.dtype(np.float32): "H5T_NATIVE_FLOAT_g",
    np.dtype(np.int8): "H5T_NATIVE_SCHAR_g", np.dtype(np.uint8): "H5T_NATIVE_UCHAR_g",
    np.dtype(np.int16): "H5T_NATIVE_SHORT_g", np.dtype(np.uint16): "H5T_NATIVE_USHORT_g",
    np.dtype(np.int32): "H5T_NATIVE_INT_g", np.dtype(np.uint32): "H5T_NATIVE_UINT_g",
    np.dtype(np.int64): "H5T_NATIVE_LLONG_g", np.dtype(np.uint64): "H5T_NATIVE_ULLONG_g",
}


def _load():
    global _lib
    if _lib is not None:
        return _lib
    tried = []
    cands = [os.environ.get("UWG_HDF5_LIB"), "libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103",
             ctypes.util.find_library("hdf5")]
    for c in cands:
        if not c:
            continue
        tried.append(c)
        try:
            lib = ctypes.CDLL(c)
        except OSError:
            continue
        for name, (args, res) in _SIGNATURES.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        if lib.H5open() < 0:
            continue
        lib.H5Eset_auto2(0, None, None)  # errors are reported through return codes
        for dt, g in _NATIVE.items():
            _types[dt] = hid_t.in_dll(lib, g).value
        _types["C_S1"] = hid_t.in_dll(lib, "H5T_C_S1_g").value
        _lib = lib
        return lib
    raise RuntimeError("HDF5 I/O needs the HDF5 C library (libhdf5.so.103, the 1.10 ABI); tried: "
                       + ", ".join(tried) + ". Set UWG_HDF5_LIB to its path.")


def _ok(rc, what):
    if rc < 0:
        raise RuntimeError(f"HDF5: {what} failed")
    return rc


def _native(dtype):
    dt = np.dtype(dtype)
    if dt == np.bool_:
        dt = np.dtype(np.int8)
    if dt not in _types:
        raise TypeError(f"HDF5: no native type for numpy dtype {dt}")
    return _types[dt]


def _space(shape):
    lib = _lib
    if shape == ():
        return _ok(lib.H5Screate(_S_SCALAR), "H5Screate")
    dims = (hsize_t * len(shape))(*shape)
    return _ok(lib.H5Screate_simple(len(shape), dims, None), "H5Screate_simple")


def _shape_of(space):
    lib = _lib
    nd = _ok(lib.H5Sget_simple_extent_ndims(space), "extent ndims")
    dims = (hsize_t * max(nd, 1))()
    _ok(lib.H5Sget_simple_extent_dims(space, dims, None), "extent dims")
    return tuple(int(dims[i]) for i in range(nd))


def _dtype_of(ftype):
    """numpy dtype a numeric HDF5 type is read into (its natural memory type)."""
    lib = _lib
    cls = lib.H5Tget_class(ftype)
    size = lib.H5Tget_size(ftype)
    if cls == _T_FLOAT:
        return np.dtype(np.float32 if size == 4 else np.float64)
    if cls == _T_INTEGER:
        signed = lib.H5Tget_sign(ftype) != _T_SGN_NONE
        return np.dtype(("i" if signed else "u") + str(size))
    if cls == _T_ENUM:  # h5py's bool: an integer enum read through its base type
        base = _ok(lib.H5Tget_super(ftype), "H5Tget_super")
        try:
            dt = _dtype_of(base)
        finally:
            lib.H5Tclose(base)
        return np.dtype(np.bool_) if size == 1 else dt
    raise TypeError(f"HDF5: unsupported type class {cls}")


def _bool_enum():
    lib = _lib
    t = _ok(lib.H5Tenum_create(_types[np.dtype(np.int8)]), "H5Tenum_create")
    for name, v in ((b"FALSE", 0), (b"TRUE", 1)):
        c = ctypes.c_int8(v)
        _ok(lib.H5Tenum_insert(t, name, ctypes.byref(c)), "H5Tenum_insert")
    return t


class _Attrs:
    """Attribute access on an open object (``f.attrs[...]`` in h5py terms)."""

    def __init__(self, oid):
        self._oid = oid

    def __contains__(self, name):
        return _lib.H5Aexists(self._oid, name.encode()) > 0

    def keys(self):
        out = []
        buf = ctypes.create_string_buffer(1024)
        n = 0
        while True:
            r = _lib.H5Aget_name_by_idx(self._oid, b".", 0, 0, n, buf, 1024, _P_DEFAULT)
            if r < 0:
                return out
            out.append(buf.value.decode())
            n += 1

    def get(self, name, default=None):
        return self[name] if name in self else default

    def __setitem__(self, name, value):
        lib = _lib
        key = name.encode()
        if lib.H5Aexists(self._oid, key) > 0:
            _ok(lib.H5Adelete(self._oid, key), "H5Adelete")
        if isinstance(value, (bytes, np.bytes_)):  # fixed-length NULLPAD ASCII, as np.string_
            data = bytes(value)
            t = _ok(lib.H5Tcopy(_types["C_S1"]), "H5Tcopy")
            lib.H5Tset_size(t, max(len(data), 1))
            lib.H5Tset_strpad(t, _T_STR_NULLPAD)
            sp = _space(())
            buf = ctypes.create_string_buffer(data, max(len(data), 1))
            self._write(key, t, sp, buf, t)
            return
        if isinstance(value, str):  # variable-length UTF-8, as h5py writes str
            t = _ok(lib.H5Tcopy(_types["C_S1"]), "H5Tcopy")
            lib.H5Tset_size(t, _T_VARIABLE)
            lib.H5Tset_cset(t, _T_CSET_UTF8)
            sp = _space(())
            ptr = ctypes.c_char_p(value.encode())
            self._write(key, t, sp, ctypes.byref(ptr), t)
            return
        arr = np.asarray(value)
        if arr.dtype == np.bool_:
            t = _bool_enum()
            arr = np.array(arr, dtype=np.int8, order="C")
            self._write(key, t, _space(arr.shape), arr.ctypes.data_as(ctypes.c_void_p), t)
            return
        if arr.dtype.kind in "iu":
            arr = arr.astype(arr.dtype.newbyteorder("="))
        elif arr.dtype.kind == "f":
            arr = arr.astype(np.float64 if arr.dtype.itemsize > 4 else np.float32)
        else:
            raise TypeError(f"HDF5: unsupported attribute value {value!r}")
        arr = np.array(arr, order="C")  # keeps 0-d values scalar (a SCALAR dataspace, as h5py writes)
        t = _native(arr.dtype)
        self._write(key, t, _space(arr.shape), arr.ctypes.data_as(ctypes.c_void_p), None)

    def _write(self, key, ftype, space, buf, close_type):
        lib = _lib
        try:
            a = _ok(lib.H5Acreate2(self._oid, key, ftype, space, _P_DEFAULT, _P_DEFAULT), "H5Acreate2")
            try:
                _ok(lib.H5Awrite(a, ftype, buf), "H5Awrite")
            finally:
                lib.H5Aclose(a)
        finally:
            lib.H5Sclose(space)
            if close_type is not None:
                lib.H5Tclose(close_type)

    def __getitem__(self, name):
        lib = _lib
        a = lib.H5Aopen(self._oid, name.encode(), _P_DEFAULT)
        if a < 0:
            raise KeyError(name)
        try:
            ft = _ok(lib.H5Aget_type(a), "H5Aget_type")
            sp = _ok(lib.H5Aget_space(a), "H5Aget_space")
            try:
                shape = _shape_of(sp)
                if lib.H5Tget_class(ft) == _T_STRING:
                    return self._read_str(a, ft, sp, shape)
                dt = _dtype_of(ft)
                out = np.zeros(shape, dtype=dt)
                if lib.H5Tget_class(ft) == _T_ENUM:
                    raw = np.zeros(shape, dtype=np.int8 if dt == np.bool_ else dt)
                    _ok(lib.H5Aread(a, ft, raw.ctypes.data_as(ctypes.c_void_p)), "H5Aread")
                    out[...] = raw
                else:
                    _ok(lib.H5Aread(a, _native(dt), out.ctypes.data_as(ctypes.c_void_p)), "H5Aread")
                return out[()] if shape == () else out
            finally:
                lib.H5Sclose(sp)
                lib.H5Tclose(ft)
        finally:
            lib.H5Aclose(a)

    @staticmethod
    def _read_str(a, ft, sp, shape):
        lib = _lib
        n = int(np.prod(shape)) if shape else 1
        if lib.H5Tis_variable_str(ft) > 0:
            mt = _ok(lib.H5Tcopy(_types["C_S1"]), "H5Tcopy")
            lib.H5Tset_size(mt, _T_VARIABLE)
            lib.H5Tset_cset(mt, _T_CSET_UTF8)
            ptrs = (ctypes.c_char_p * n)()
            try:
                _ok(lib.H5Aread(a, mt, ptrs), "H5Aread")
                vals = [p.decode() if p is not None else "" for p in ptrs]
                lib.H5Dvlen_reclaim(mt, sp, _P_DEFAULT, ptrs)
            finally:
                lib.H5Tclose(mt)
            return vals[0] if shape == () else np.array(vals, dtype=object).reshape(shape)
        size = lib.H5Tget_size(ft)
        buf = ctypes.create_string_buffer(size * n)
        _ok(lib.H5Aread(a, ft, buf), "H5Aread")
        vals = [np.bytes_(buf.raw[i * size:(i + 1) * size].rstrip(b"\0")) for i in range(n)]
        return vals[0] if shape == () else np.array(vals).reshape(shape)


class _Linked:
    """An object reached through a link (e.g. an external link to another file's root)."""

    def __init__(self, oid):
        self._oid = oid
        self.attrs = _Attrs(oid)

    def close(self):
        if self._oid is not None:
            _lib.H5Oclose(self._oid)
            self._oid = None


class File:
    """An HDF5 file opened with ``mode`` 'r' (read), 'r+' (read/write) or 'w' (create, truncate)."""

    def __init__(self, name, mode="r"):
        lib = _load()
        self.filename = str(name)
        fn = os.fsencode(self.filename)
        if mode == "w":
            fid = lib.H5Fcreate(fn, _ACC_TRUNC, _P_DEFAULT, _P_DEFAULT)
        elif mode in ("r", "r+"):
            if not os.path.exists(self.filename):
                raise FileNotFoundError(self.filename)
            fid = lib.H5Fopen(fn, _ACC_RDONLY if mode == "r" else _ACC_RDWR, _P_DEFAULT)
        else:
            raise ValueError(f"unsupported mode {mode!r}")
        if fid < 0:
            raise OSError(f"HDF5: cannot open '{self.filename}' (mode {mode!r})")
        self._fid = fid
        self.attrs = _Attrs(fid)
        self._opened = []

    def close(self):
        for o in self._opened:
            o.close()
        self._opened = []
        if self._fid is not None:
            _lib.H5Fclose(self._fid)
            self._fid = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __contains__(self, name):
        return _lib.H5Lexists(self._fid, name.encode(), _P_DEFAULT) > 0

    def write_dataset(self, name, array):
        """Create (replacing) dataset ``name`` holding ``array`` in its own dtype."""
        lib = _lib
        arr = np.asarray(array)
        if arr.dtype == np.bool_:
            arr = arr.astype(np.int8)
        arr = np.ascontiguousarray(arr.astype(arr.dtype.newbyteorder("=")))
        if name in self:
            _ok(lib.H5Ldelete(self._fid, name.encode(), _P_DEFAULT), "H5Ldelete")
        t = _native(arr.dtype)
        sp = _space(arr.shape)
        try:
            d = _ok(lib.H5Dcreate2(self._fid, name.encode(), t, sp, _P_DEFAULT, _P_DEFAULT, _P_DEFAULT),
                    "H5Dcreate2")
            try:
                if arr.size:
                    _ok(lib.H5Dwrite(d, t, _S_ALL, _S_ALL, _P_DEFAULT, arr.ctypes.data_as(ctypes.c_void_p)),
                        "H5Dwrite")
            finally:
                lib.H5Dclose(d)
        finally:
            lib.H5Sclose(sp)

    def dataset_shape(self, name):
        lib = _lib
        d = lib.H5Dopen2(self._fid, name.encode(), _P_DEFAULT)
        if d < 0:
            raise KeyError(f"no dataset '{name}' in '{self.filename}'")
        try:
            sp = _ok(lib.H5Dget_space(d), "H5Dget_space")
            try:
                return _shape_of(sp)
            finally:
                lib.H5Sclose(sp)
        finally:
            lib.H5Dclose(d)

    def read_dataset(self, name, dtype=None):
        """The whole dataset, converted to ``dtype`` (default: the file type's natural dtype)."""
        lib = _lib
        d = lib.H5Dopen2(self._fid, name.encode(), _P_DEFAULT)
        if d < 0:
            raise KeyError(f"no dataset '{name}' in '{self.filename}'")
        try:
            sp = _ok(lib.H5Dget_space(d), "H5Dget_space")
            ft = _ok(lib.H5Dget_type(d), "H5Dget_type")
            try:
                shape = _shape_of(sp)
                dt = np.dtype(dtype) if dtype is not None else _dtype_of(ft)
                out = np.zeros(shape, dtype=np.int8 if dt == np.bool_ else dt)
                if out.size:
                    _ok(lib.H5Dread(d, _native(out.dtype), _S_ALL, _S_ALL, _P_DEFAULT,
                                    out.ctypes.data_as(ctypes.c_void_p)), "H5Dread")
                return out.astype(np.bool_) if dt == np.bool_ else out
            finally:
                lib.H5Tclose(ft)
                lib.H5Sclose(sp)
        finally:
            lib.H5Dclose(d)

    def external_link(self, name, target_file, target_path="."):
        """``f[name] = h5py.ExternalLink(target_file, target_path)``, replacing an existing link."""
        lib = _lib
        if name in self:
            _ok(lib.H5Ldelete(self._fid, name.encode(), _P_DEFAULT), "H5Ldelete")
        _ok(lib.H5Lcreate_external(os.fsencode(str(target_file)), target_path.encode(), self._fid, name.encode(),
                                   _P_DEFAULT, _P_DEFAULT), "H5Lcreate_external")

    def open_linked(self, name):
        """The object ``name`` resolves to (following an external link); ``None`` if it cannot be reached."""
        if name not in self:
            return None
        oid = _lib.H5Oopen(self._fid, name.encode(), _P_DEFAULT)
        if oid < 0:
            return None
        o = _Linked(oid)
        self._opened.append(o)
        return o


class SavedFileData:
    """What ``save`` returns (PY/utils/_utils.py ``SavedFileData``): the saved object and its file."""

    def __init__(self, pyobj, filename):
        self.pyobj = pyobj
        self.filename = filename
