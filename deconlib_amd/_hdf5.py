"""Minimal ctypes binding of the HDF5 C library (1.10 API) for the ``.psf.h5`` format.

Only what ``io.py`` needs: files, groups, scalar / 1-D attributes (numbers, number
arrays, UTF-8 strings), n-D datasets (optionally chunked and deflate-compressed), link
queries.  The library is the system's ``libhdf5`` (found through ``ctypes.util`` or the
conda prefix); when it is absent, using this module raises ``RuntimeError``.
"""

from __future__ import annotations

import ctypes as C
import ctypes.util
import glob
import os
from typing import Any, List, Optional, Tuple

import numpy as np

hid_t = C.c_int64
hsize_t = C.c_ulonglong
herr_t = C.c_int
htri_t = C.c_int

H5F_ACC_RDONLY = 0x0000
H5F_ACC_TRUNC = 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5T_SGN_NONE = 0
H5T_CSET_UTF8 = 1
H5T_VARIABLE = C.c_size_t(-1).value
H5_INDEX_NAME, H5_ITER_INC = 0, 0


class _GInfo(C.Structure):
    _fields_ = [("storage_type", C.c_int), ("nlinks", hsize_t), ("max_corder", C.c_int64), ("mounted", C.c_bool)]


_LIB = None


def _candidates() -> List[str]:
    out = []
    found = ctypes.util.find_library("hdf5")
    if found:
        out.append(found)
    for prefix in (os.environ.get("CONDA_PREFIX"), "/opt/conda", "/usr"):
        if prefix:
            out += sorted(glob.glob(os.path.join(prefix, "lib", "libhdf5.so*")))
            out += sorted(glob.glob(os.path.join(prefix, "lib", "x86_64-linux-gnu", "libhdf5*.so*")))
    return [p for p in out if "hl" not in os.path.basename(p) and "fortran" not in os.path.basename(p)]


def lib():
    """The loaded HDF5 library (opened once; its error printing switched off)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    err = None
    for path in _candidates():
        try:
            L = C.CDLL(path)
            break
        except OSError as e:
            err = e
    else:
        raise RuntimeError(f"the HDF5 C library (libhdf5) was not found: .psf.h5 I/O is unavailable ({err})")
    sigs = {
        "H5open": (herr_t, []),
        "H5Eset_auto2": (herr_t, [hid_t, C.c_void_p, C.c_void_p]),
        "H5Fcreate": (hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]),
        "H5Fopen": (hid_t, [C.c_char_p, C.c_uint, hid_t]),
        "H5Fclose": (herr_t, [hid_t]),
        "H5Gcreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t]),
        "H5Gopen2": (hid_t, [hid_t, C.c_char_p, hid_t]),
        "H5Gclose": (herr_t, [hid_t]),
        "H5Gget_info": (herr_t, [hid_t, C.POINTER(_GInfo)]),
        "H5Lexists": (htri_t, [hid_t, C.c_char_p, hid_t]),
        "H5Lget_name_by_idx": (C.c_ssize_t, [hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, C.c_char_p, C.c_size_t,
                                             hid_t]),
        "H5Screate": (hid_t, [C.c_int]),
        "H5Screate_simple": (hid_t, [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
        "H5Sget_simple_extent_ndims": (C.c_int, [hid_t]),
        "H5Sget_simple_extent_dims": (C.c_int, [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
        "H5Sclose": (herr_t, [hid_t]),
        "H5Tcopy": (hid_t, [hid_t]),
        "H5Tset_size": (herr_t, [hid_t, C.c_size_t]),
        "H5Tset_cset": (herr_t, [hid_t, C.c_int]),
        "H5Tget_class": (C.c_int, [hid_t]),
        "H5Tget_size": (C.c_size_t, [hid_t]),
        "H5Tget_sign": (C.c_int, [hid_t]),
        "H5Tis_variable_str": (htri_t, [hid_t]),
        "H5Tclose": (herr_t, [hid_t]),
        "H5Pcreate": (hid_t, [hid_t]),
        "H5Pset_chunk": (herr_t, [hid_t, C.c_int, C.POINTER(hsize_t)]),
        "H5Pset_deflate": (herr_t, [hid_t, C.c_uint]),
        "H5Pclose": (herr_t, [hid_t]),
        "H5Dcreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        "H5Dopen2": (hid_t, [hid_t, C.c_char_p, hid_t]),
        "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
        "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
        "H5Dget_space": (hid_t, [hid_t]),
        "H5Dget_type": (hid_t, [hid_t]),
        "H5Dclose": (herr_t, [hid_t]),
        "H5Acreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t]),
        "H5Aopen": (hid_t, [hid_t, C.c_char_p, hid_t]),
        "H5Aexists": (htri_t, [hid_t, C.c_char_p]),
        "H5Awrite": (herr_t, [hid_t, hid_t, C.c_void_p]),
        "H5Aread": (herr_t, [hid_t, hid_t, C.c_void_p]),
        "H5Aget_space": (hid_t, [hid_t]),
        "H5Aget_type": (hid_t, [hid_t]),
        "H5Aclose": (herr_t, [hid_t]),
        "H5free_memory": (herr_t, [C.c_void_p]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.H5open() < 0:
        raise RuntimeError("H5open failed")
    L.H5Eset_auto2(0, None, None)  # errors come back as return codes, not stderr dumps
    _LIB = L
    return L


def _global(name: str) -> int:
    return hid_t.in_dll(lib(), name).value


def native(dtype: np.dtype) -> int:
    """Memory (native) type of a numpy dtype."""
    dt = np.dtype(dtype)
    table = {
        np.dtype(np.float32): "H5T_NATIVE_FLOAT_g", np.dtype(np.float64): "H5T_NATIVE_DOUBLE_g",
        np.dtype(np.int8): "H5T_NATIVE_INT8_g", np.dtype(np.uint8): "H5T_NATIVE_UINT8_g",
        np.dtype(np.int16): "H5T_NATIVE_INT16_g", np.dtype(np.uint16): "H5T_NATIVE_UINT16_g",
        np.dtype(np.int32): "H5T_NATIVE_INT32_g", np.dtype(np.uint32): "H5T_NATIVE_UINT32_g",
        np.dtype(np.int64): "H5T_NATIVE_INT64_g", np.dtype(np.uint64): "H5T_NATIVE_UINT64_g",
    }
    if dt not in table:
        raise TypeError(f"no HDF5 type for {dt}")
    return _global(table[dt])


def file_type(dtype: np.dtype) -> int:
    """Little-endian standard type stored in the file for a numpy dtype (what h5py writes)."""
    dt = np.dtype(dtype)
    if dt.kind == "f":
        return _global({4: "H5T_IEEE_F32LE_g", 8: "H5T_IEEE_F64LE_g"}[dt.itemsize])
    if dt.kind in "iu":
        return _global(f"H5T_STD_{'I' if dt.kind == 'i' else 'U'}{8 * dt.itemsize}LE_g")
    raise TypeError(f"no HDF5 type for {dt}")


def _check(v: int, what: str) -> int:
    if v < 0:
        raise OSError(f"HDF5: {what} failed")
    return v


def _dims(shape) -> Tuple[Any, int]:
    n = len(shape)
    return (hsize_t * max(n, 1))(*[int(s) for s in shape]), n


def _space(shape) -> int:
    L = lib()
    if shape == ():
        return _check(L.H5Screate(H5S_SCALAR), "H5Screate")
    d, n = _dims(shape)
    return _check(L.H5Screate_simple(n, d, None), "H5Screate_simple")


def _str_type() -> int:
    L = lib()
    t = _check(L.H5Tcopy(_global("H5T_C_S1_g")), "H5Tcopy")
    L.H5Tset_size(t, H5T_VARIABLE)
    L.H5Tset_cset(t, H5T_CSET_UTF8)
    return t


class File:
    """An HDF5 file opened for writing (``mode="w"``, truncating) or reading (``"r"``)."""

    def __init__(self, path: str, mode: str = "r"):
        L = lib()
        p = os.fsencode(str(path))
        if mode == "w":
            self.id = L.H5Fcreate(p, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
        elif mode == "r":
            self.id = L.H5Fopen(p, H5F_ACC_RDONLY, H5P_DEFAULT)
        else:
            raise ValueError(f"mode must be 'r' or 'w', got {mode!r}")
        if self.id < 0:
            raise OSError(f"HDF5: cannot open {path!r} (mode {mode!r})")
        self._groups: List[int] = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self) -> None:
        L = lib()
        for g in self._groups:
            L.H5Gclose(g)
        self._groups = []
        if self.id >= 0:
            L.H5Fclose(self.id)
            self.id = -1

    def create_group(self, loc: int, name: str) -> int:
        g = _check(lib().H5Gcreate2(loc, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2")
        self._groups.append(g)
        return g

    def open_group(self, loc: int, name: str) -> int:
        g = _check(lib().H5Gopen2(loc, name.encode(), H5P_DEFAULT), f"opening group {name!r}")
        self._groups.append(g)
        return g


def exists(loc: int, name: str) -> bool:
    return lib().H5Lexists(loc, name.encode(), H5P_DEFAULT) > 0


def members(loc: int) -> List[str]:
    """Link names of a group in name order."""
    L = lib()
    info = _GInfo()
    _check(L.H5Gget_info(loc, C.byref(info)), "H5Gget_info")
    out = []
    for i in range(int(info.nlinks)):
        n = _check(L.H5Lget_name_by_idx(loc, b".", H5_INDEX_NAME, H5_ITER_INC, i, None, 0, H5P_DEFAULT), "name")
        buf = C.create_string_buffer(n + 1)
        L.H5Lget_name_by_idx(loc, b".", H5_INDEX_NAME, H5_ITER_INC, i, buf, n + 1, H5P_DEFAULT)
        out.append(buf.value.decode())
    return out


def write_attr(loc: int, name: str, value: Any) -> None:
    """Scalar number, 1-D number array or string attribute (h5py's encodings: numbers as
    little-endian standard types, strings as variable-length UTF-8)."""
    L = lib()
    if isinstance(value, str):
        t = _str_type()
        s = _space(())
        a = _check(L.H5Acreate2(loc, name.encode(), t, s, H5P_DEFAULT, H5P_DEFAULT), f"attribute {name!r}")
        buf = C.c_char_p(value.encode("utf-8"))
        err = L.H5Awrite(a, t, C.byref(buf))
        L.H5Aclose(a)
        L.H5Sclose(s)
        L.H5Tclose(t)
        _check(err, f"writing attribute {name!r}")
        return
    arr = np.asarray(value)
    if arr.dtype == np.bool_:
        arr = arr.astype(np.uint8)
    if arr.dtype.kind not in "iuf" or arr.ndim > 1:
        raise TypeError(f"attribute {name!r}: unsupported value of type {type(value).__name__}")
    arr = np.array(arr, order="C")  # keeps 0-d values 0-d: a scalar dataspace
    s = _space(arr.shape)
    a = _check(L.H5Acreate2(loc, name.encode(), file_type(arr.dtype), s, H5P_DEFAULT, H5P_DEFAULT),
               f"attribute {name!r}")
    err = L.H5Awrite(a, native(arr.dtype), arr.ctypes.data_as(C.c_void_p))
    L.H5Aclose(a)
    L.H5Sclose(s)
    _check(err, f"writing attribute {name!r}")


def has_attr(loc: int, name: str) -> bool:
    return lib().H5Aexists(loc, name.encode()) > 0


def _shape_of(space: int) -> Tuple[int, ...]:
    L = lib()
    n = L.H5Sget_simple_extent_ndims(space)
    if n <= 0:
        return ()
    d = (hsize_t * n)()
    L.H5Sget_simple_extent_dims(space, d, None)
    return tuple(int(v) for v in d)


def _numpy_of(t: int) -> np.dtype:
    L = lib()
    cls, size = L.H5Tget_class(t), int(L.H5Tget_size(t))
    if cls == H5T_FLOAT:
        return np.dtype({4: np.float32, 8: np.float64}[size])
    if cls == H5T_INTEGER:
        unsigned = L.H5Tget_sign(t) == H5T_SGN_NONE
        return np.dtype(("u" if unsigned else "i") + str(size))
    raise TypeError("unsupported HDF5 type class")


def read_attr(loc: int, name: str) -> Any:
    """A string, a number or a 1-D number array."""
    L = lib()
    a = _check(L.H5Aopen(loc, name.encode(), H5P_DEFAULT), f"attribute {name!r}")
    t = L.H5Aget_type(a)
    s = L.H5Aget_space(a)
    try:
        shape = _shape_of(s)
        if L.H5Tget_class(t) == H5T_STRING:
            if L.H5Tis_variable_str(t) > 0:
                mt = _str_type()
                ptr = C.c_char_p()
                err = L.H5Aread(a, mt, C.byref(ptr))
                L.H5Tclose(mt)
                _check(err, f"reading attribute {name!r}")
                v = (ptr.value or b"").decode("utf-8")
                L.H5free_memory(C.cast(ptr, C.c_void_p))
                return v
            size = int(L.H5Tget_size(t))
            buf = C.create_string_buffer(size + 1)
            _check(L.H5Aread(a, t, buf), f"reading attribute {name!r}")
            return buf.raw[:size].split(b"\0", 1)[0].decode("utf-8")
        dt = _numpy_of(t)
        out = np.empty(shape, dtype=dt)
        _check(L.H5Aread(a, native(dt), out.ctypes.data_as(C.c_void_p)), f"reading attribute {name!r}")
        return out[()] if shape == () else out
    finally:
        L.H5Sclose(s)
        L.H5Tclose(t)
        L.H5Aclose(a)


def write_dataset(loc: int, name: str, data: np.ndarray, gzip: Optional[int] = None) -> None:
    """n-D number array; with ``gzip`` a chunked layout with the deflate filter at that level."""
    L = lib()
    arr = np.array(data, order="C")
    s = _space(arr.shape)
    dcpl = H5P_DEFAULT
    if gzip is not None and arr.ndim > 0 and arr.size > 0:
        dcpl = _check(L.H5Pcreate(_global("H5P_CLS_DATASET_CREATE_ID_g")), "H5Pcreate")
        # chunks of at most about 1 MiB: whole trailing axes, the leading ones split
        chunk = list(arr.shape)
        for ax in range(arr.ndim - 1):
            while int(np.prod(chunk)) * arr.itemsize > (1 << 20) and chunk[ax] > 1:
                chunk[ax] = (chunk[ax] + 1) // 2
        cd, n = _dims(chunk)
        _check(L.H5Pset_chunk(dcpl, n, cd), "H5Pset_chunk")
        _check(L.H5Pset_deflate(dcpl, int(gzip)), "H5Pset_deflate")
    d = L.H5Dcreate2(loc, name.encode(), file_type(arr.dtype), s, H5P_DEFAULT, dcpl, H5P_DEFAULT)
    if dcpl != H5P_DEFAULT:
        L.H5Pclose(dcpl)
    if d < 0:
        L.H5Sclose(s)
        raise OSError(f"HDF5: creating dataset {name!r} failed")
    err = L.H5Dwrite(d, native(arr.dtype), H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data_as(C.c_void_p))
    L.H5Dclose(d)
    L.H5Sclose(s)
    _check(err, f"writing dataset {name!r}")


def read_dataset(loc: int, name: str) -> np.ndarray:
    """The whole dataset in its stored number type."""
    L = lib()
    d = _check(L.H5Dopen2(loc, name.encode(), H5P_DEFAULT), f"dataset {name!r}")
    t = L.H5Dget_type(d)
    s = L.H5Dget_space(d)
    try:
        out = np.empty(_shape_of(s), dtype=_numpy_of(t))
        if out.size:
            _check(L.H5Dread(d, native(out.dtype), H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data_as(C.c_void_p)),
                   f"reading dataset {name!r}")
        return out
    finally:
        L.H5Sclose(s)
        L.H5Tclose(t)
        L.H5Dclose(d)
