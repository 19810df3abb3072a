"""HDF5 files through the HDF5 C library (ctypes), without h5py.

The run directory of the reference (python/bdsim/storage.py) is written with
h5py, which is not installed here; the HDF5 library itself is (libhdf5 1.10,
found under the interpreter's prefix or the loader path). This module writes
and reads the subset the run directory needs, with h5py's conventions so that
either side reads the other's files:

* datasets of int64 / float64 (contiguous, or chunked with gzip compression);
* attributes of int64, float64, bool (h5py's enum of int8, FALSE = 0,
  TRUE = 1) and str (variable-length UTF-8), one value each.

Errors raise OSError; the library's own error stack printing is switched off.
"""
import ctypes
import ctypes.util
import os
import sys
from ctypes import byref, c_char_p, c_int, c_int64, c_uint, c_void_p, c_size_t

import numpy as np

hid_t = c_int64
hsize_t = ctypes.c_ulonglong

H5F_ACC_RDONLY, H5F_ACC_TRUNC = 0, 2
H5P_DEFAULT = H5S_ALL = 0
H5S_SCALAR = 0
H5T_INTEGER, H5T_FLOAT, H5T_STRING, H5T_ENUM = 0, 1, 3, 8
H5T_CSET_UTF8 = 1
H5T_VARIABLE = ctypes.c_size_t(-1).value
H5_INDEX_NAME = H5_ITER_INC = 0

_lib = None
_T = {}


def _candidates():
    found = ctypes.util.find_library("hdf5")
    if found:
        yield found
    for prefix in {sys.prefix, sys.base_prefix, "/opt/conda", "/usr"}:
        for sub in ("lib", "lib/x86_64-linux-gnu", "lib64"):
            d = os.path.join(prefix, sub)
            if os.path.isdir(d):
                for f in sorted(os.listdir(d)):
                    if f.startswith("libhdf5.so"):
                        yield os.path.join(d, f)


def available():
    try:
        lib()
        return True
    except OSError:
        return False


def lib():
    """The HDF5 library, loaded and initialised once."""
    global _lib
    if _lib is not None:
        return _lib
    err = None
    for path in _candidates():
        try:
            L = ctypes.CDLL(path)
            break
        except OSError as e:
            err = e
    else:
        raise OSError(f"the HDF5 library (libhdf5) was not found: {err}")
    sig = {
        "H5open": ([], c_int), "H5Eset_auto2": ([hid_t, c_void_p, c_void_p], c_int),
        "H5Fcreate": ([c_char_p, c_uint, hid_t, hid_t], hid_t),
        "H5Fopen": ([c_char_p, c_uint, hid_t], hid_t), "H5Fclose": ([hid_t], c_int),
        "H5Screate": ([c_int], hid_t), "H5Sclose": ([hid_t], c_int),
        "H5Screate_simple": ([c_int, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)], hid_t),
        "H5Sget_simple_extent_ndims": ([hid_t], c_int),
        "H5Sget_simple_extent_dims": ([hid_t, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)], c_int),
        "H5Pcreate": ([hid_t], hid_t), "H5Pclose": ([hid_t], c_int),
        "H5Pset_chunk": ([hid_t, c_int, ctypes.POINTER(hsize_t)], c_int),
        "H5Pset_deflate": ([hid_t, c_uint], c_int),
        "H5Dcreate2": ([hid_t, c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t], hid_t),
        "H5Dopen2": ([hid_t, c_char_p, hid_t], hid_t), "H5Dclose": ([hid_t], c_int),
        "H5Dwrite": ([hid_t, hid_t, hid_t, hid_t, hid_t, c_void_p], c_int),
        "H5Dread": ([hid_t, hid_t, hid_t, hid_t, hid_t, c_void_p], c_int),
        "H5Dget_space": ([hid_t], hid_t), "H5Dget_type": ([hid_t], hid_t),
        "H5Lexists": ([hid_t, c_char_p, hid_t], c_int),
        "H5Acreate2": ([hid_t, c_char_p, hid_t, hid_t, hid_t, hid_t], hid_t),
        "H5Awrite": ([hid_t, hid_t, c_void_p], c_int), "H5Aread": ([hid_t, hid_t, c_void_p], c_int),
        "H5Aclose": ([hid_t], c_int), "H5Aget_num_attrs": ([hid_t], c_int),
        "H5Aopen_by_idx": ([hid_t, c_char_p, c_int, c_int, hsize_t, hid_t, hid_t], hid_t),
        "H5Aget_name": ([hid_t, c_size_t, c_char_p], ctypes.c_ssize_t),
        "H5Aget_type": ([hid_t], hid_t), "H5Aget_space": ([hid_t], hid_t),
        "H5Tcopy": ([hid_t], hid_t), "H5Tclose": ([hid_t], c_int),
        "H5Tset_size": ([hid_t, c_size_t], c_int), "H5Tset_cset": ([hid_t, c_int], c_int),
        "H5Tget_class": ([hid_t], c_int), "H5Tget_size": ([hid_t], c_size_t),
        "H5Tis_variable_str": ([hid_t], c_int),
        "H5Tenum_create": ([hid_t], hid_t), "H5Tenum_insert": ([hid_t, c_char_p, c_void_p], c_int),
        "H5free_memory": ([c_void_p], c_int),
        "H5Gopen2": ([hid_t, c_char_p, hid_t], hid_t), "H5Gclose": ([hid_t], c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes, f.restype = args, res
    if L.H5open() < 0:
        raise OSError("H5open failed")
    L.H5Eset_auto2(0, None, None)                   # no error stack printing
    for name in ("H5T_NATIVE_DOUBLE_g", "H5T_NATIVE_LLONG_g", "H5T_NATIVE_INT8_g",
                 "H5T_C_S1_g", "H5P_CLS_DATASET_CREATE_ID_g"):
        _T[name[:-2]] = hid_t.in_dll(L, name).value
    _lib = L
    return L


def _ok(rc, what):
    if rc < 0:
        raise OSError(f"HDF5: {what} failed")
    return rc


def _bool_type():
    L = lib()
    t = _ok(L.H5Tenum_create(_T["H5T_NATIVE_INT8"]), "enum type")
    for name, v in ((b"FALSE", 0), (b"TRUE", 1)):
        val = ctypes.c_int8(v)
        _ok(L.H5Tenum_insert(t, name, byref(val)), "enum insert")
    return t


def _str_type():
    L = lib()
    t = _ok(L.H5Tcopy(_T["H5T_C_S1"]), "string type")
    _ok(L.H5Tset_size(t, H5T_VARIABLE), "string size")
    _ok(L.H5Tset_cset(t, H5T_CSET_UTF8), "string charset")
    return t


class File:
    """An HDF5 file opened for writing (mode "w": created or truncated) or
    reading (mode "r"). Use as a context manager."""

    def __init__(self, path, mode="r"):
        L = lib()
        p = os.fsencode(path)
        if mode == "w":
            self.id = _ok(L.H5Fcreate(p, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), f"create {path}")
        elif mode == "r":
            self.id = _ok(L.H5Fopen(p, H5F_ACC_RDONLY, H5P_DEFAULT), f"open {path}")
        else:
            raise ValueError("mode must be 'r' or 'w'")

    def close(self):
        if self.id is not None:
            lib().H5Fclose(self.id)
            self.id = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- datasets -------------------------------------------------------
    def create_dataset(self, name, data, compression=None, chunks=None):
        """int64 or float64 array; compression "gzip" (level 4) or None."""
        L = lib()
        a = np.ascontiguousarray(data)
        if a.dtype.kind in "iub":
            a, mem_t = np.ascontiguousarray(a, dtype=np.int64), _T["H5T_NATIVE_LLONG"]
        elif a.dtype.kind == "f":
            a, mem_t = np.ascontiguousarray(a, dtype=np.float64), _T["H5T_NATIVE_DOUBLE"]
        else:
            raise TypeError(f"dataset {name}: unsupported dtype {a.dtype}")
        dims = (hsize_t * max(a.ndim, 1))(*(a.shape or (1,)))
        space = _ok(L.H5Screate_simple(a.ndim, dims, None) if a.ndim else L.H5Screate(H5S_SCALAR),
                    "dataspace")
        dcpl = H5P_DEFAULT
        try:
            if compression is not None and a.ndim > 0 and a.size > 0:
                if compression not in ("gzip", 4):
                    raise ValueError(f"compression {compression!r}: only gzip")
                dcpl = _ok(L.H5Pcreate(_T["H5P_CLS_DATASET_CREATE_ID"]), "dcpl")
                ch = chunks or tuple([min(a.shape[0], 64)] + list(a.shape[1:]))
                _ok(L.H5Pset_chunk(dcpl, a.ndim, (hsize_t * a.ndim)(*ch)), "chunk")
                _ok(L.H5Pset_deflate(dcpl, 4), "deflate")
            d = _ok(L.H5Dcreate2(self.id, name.encode(), mem_t, space, H5P_DEFAULT, dcpl,
                                 H5P_DEFAULT), f"create dataset {name}")
            try:
                if a.size:
                    _ok(L.H5Dwrite(d, mem_t, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                   a.ctypes.data_as(c_void_p)), f"write {name}")
            finally:
                L.H5Dclose(d)
        finally:
            if dcpl != H5P_DEFAULT:
                L.H5Pclose(dcpl)
            L.H5Sclose(space)

    def __contains__(self, name):
        return lib().H5Lexists(self.id, name.encode(), H5P_DEFAULT) > 0

    def read(self, name):
        """The whole dataset as int64 or float64."""
        L = lib()
        d = _ok(L.H5Dopen2(self.id, name.encode(), H5P_DEFAULT), f"open dataset {name}")
        try:
            t = L.H5Dget_type(d)
            cls = L.H5Tget_class(t)
            L.H5Tclose(t)
            if cls == H5T_INTEGER:
                dt, mem_t = np.int64, _T["H5T_NATIVE_LLONG"]
            elif cls == H5T_FLOAT:
                dt, mem_t = np.float64, _T["H5T_NATIVE_DOUBLE"]
            else:
                raise OSError(f"dataset {name}: type class {cls} not supported")
            sp = L.H5Dget_space(d)
            nd = L.H5Sget_simple_extent_ndims(sp)
            dims = (hsize_t * max(nd, 1))()
            L.H5Sget_simple_extent_dims(sp, dims, None)
            L.H5Sclose(sp)
            out = np.empty(tuple(int(dims[i]) for i in range(nd)), dtype=dt)
            if out.size:
                _ok(L.H5Dread(d, mem_t, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                              out.ctypes.data_as(c_void_p)), f"read {name}")
            return out
        finally:
            L.H5Dclose(d)

    # ---- attributes ------------------------------------------------------
    def set_attr(self, name, value):
        L = lib()
        space = _ok(L.H5Screate(H5S_SCALAR), "scalar dataspace")
        own_t = None
        if isinstance(value, (bool, np.bool_)):
            own_t = _bool_type()
            t, buf = own_t, ctypes.c_int8(1 if value else 0)
        elif isinstance(value, (int, np.integer)):
            t, buf = _T["H5T_NATIVE_LLONG"], c_int64(int(value))
        elif isinstance(value, (float, np.floating)):
            t, buf = _T["H5T_NATIVE_DOUBLE"], ctypes.c_double(float(value))
        elif isinstance(value, str):
            own_t = _str_type()
            t, buf = own_t, c_char_p(value.encode("utf-8"))
        else:
            L.H5Sclose(space)
            raise TypeError(f"attribute {name}: unsupported type {type(value).__name__}")
        try:
            a = _ok(L.H5Acreate2(self.id, name.encode(), t, space, H5P_DEFAULT, H5P_DEFAULT),
                    f"create attribute {name}")
            try:
                _ok(L.H5Awrite(a, t, byref(buf)), f"write attribute {name}")
            finally:
                L.H5Aclose(a)
        finally:
            if own_t is not None:
                L.H5Tclose(own_t)
            L.H5Sclose(space)

    def attrs(self):
        """{name: value} of the file's (root group's) scalar attributes."""
        L = lib()
        out = {}
        g = _ok(L.H5Gopen2(self.id, b"/", H5P_DEFAULT), "open root group")
        try:
            count = _ok(L.H5Aget_num_attrs(g), "count attributes")
        finally:
            L.H5Gclose(g)
        for i in range(count):
            a = _ok(L.H5Aopen_by_idx(self.id, b".", H5_INDEX_NAME, H5_ITER_INC, i, H5P_DEFAULT,
                                     H5P_DEFAULT), "open attribute")
            try:
                n = L.H5Aget_name(a, 0, None)
                nb = ctypes.create_string_buffer(n + 1)
                L.H5Aget_name(a, n + 1, nb)
                name = nb.value.decode()
                t = L.H5Aget_type(a)
                try:
                    out[name] = self._read_attr(a, t, name)
                finally:
                    L.H5Tclose(t)
            finally:
                L.H5Aclose(a)
        return out

    @staticmethod
    def _read_attr(a, t, name):
        L = lib()
        cls = L.H5Tget_class(t)
        if cls == H5T_ENUM:                          # h5py's bool
            v = ctypes.c_int8()
            mt = _bool_type()
            try:
                _ok(L.H5Aread(a, mt, byref(v)), f"read {name}")
            finally:
                L.H5Tclose(mt)
            return bool(v.value)
        if cls == H5T_INTEGER:
            v = c_int64()
            _ok(L.H5Aread(a, _T["H5T_NATIVE_LLONG"], byref(v)), f"read {name}")
            return int(v.value)
        if cls == H5T_FLOAT:
            v = ctypes.c_double()
            _ok(L.H5Aread(a, _T["H5T_NATIVE_DOUBLE"], byref(v)), f"read {name}")
            return float(v.value)
        if cls == H5T_STRING:
            if L.H5Tis_variable_str(t) > 0:
                mt = _str_type()
                p = c_char_p()
                try:
                    _ok(L.H5Aread(a, mt, byref(p)), f"read {name}")
                    s = p.value.decode("utf-8")
                finally:
                    L.H5Tclose(mt)
                    if p:
                        L.H5free_memory(ctypes.cast(p, c_void_p))
                return s
            size = L.H5Tget_size(t)
            buf = ctypes.create_string_buffer(size + 1)
            _ok(L.H5Aread(a, t, buf), f"read {name}")
            return buf.value.decode("utf-8")
        raise OSError(f"attribute {name}: type class {cls} not supported")
