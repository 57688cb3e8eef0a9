"""HDF5 output of the production and gold-standard probe runs.

Restates the reference's HDF5 package extension (ext/MagnetoPhotonicHDF5Ext.jl) on the
HDF5 C library through ctypes: the production file (`write_production_h5`, :417-454), the
gold-standard probe file (`write_goldstd_h5`, :528-549) and per-shot checkpoints
(`append_probe_shot_h5`, :515-526), the streaming movie writers of the pump and probe
phases (`production_frame_writer` / `probe_frame_writer`, :551-697), the magnetisation
snapshot (`save_magnetization` / `load_magnetization` / `load_production_magnetization`,
:699-772), the state dump (`save_hdf5_state` / `load_hdf5_state`, :14-93) and the schema
check (`h5_schema_signature` / `assert_h5_schema_compatible`, :774-814).

Layout.  The reference writes Julia arrays through HDF5.jl, which stores a column-major
array of size (d1, d2, ..., dn) as an HDF5 dataset of dims (dn, ..., d2, d1) over the same
bytes.  This module takes and returns numpy arrays in the Julia dims, first index fastest
(as `FDTDState.field` returns them), and writes them the same way, so a file written here
reads back in HDF5.jl with the reference's shapes.  Scalars are scalar dataspaces, strings
variable-length UTF-8, `Int` is Int64, the run flags Int8; zero-size arrays get a null
dataspace, as HDF5.jl gives them.  Indices the reference stores (active cells, material
cells) are 1-based in the file.

Movie frames are staged on the device: `mp_plane_abs` (pump |Ez| planes) and
`mp_probe_frames` (probe |Ey|, |Ez| planes) write float32 planes into device buffers, which
are downloaded and written as one frame slab of a chunked (one slice per frame) dataset.
"""
from __future__ import annotations

import ctypes as C
import datetime
import math
import os
from typing import Optional

import numpy as np

from . import dispersion
from .grid import FDTDParams
from .sources import source_index


class H5Error(RuntimeError):
    """The HDF5 library is missing or an HDF5 call failed."""


# ---- the HDF5 C library ------------------------------------------------------------------

_LIB_NAMES = ("libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so",
              "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so")

hid_t = C.c_int64
hsize_t = C.c_uint64
H5P_DEFAULT = 0
H5S_ALL = 0
H5F_ACC_RDONLY, H5F_ACC_RDWR, H5F_ACC_TRUNC = 0, 1, 2
H5S_SCALAR, H5S_SIMPLE, H5S_NULL = 0, 1, 2
H5S_SELECT_SET = 0
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5T_SGN_NONE = 0
H5T_CSET_UTF8 = 1
H5I_GROUP, H5I_DATASET = 2, 5
H5_INDEX_NAME, H5_ITER_INC = 0, 0
_VARIABLE = C.c_size_t(-1).value

_ITER_CB = C.CFUNCTYPE(C.c_int, hid_t, C.c_char_p, C.c_void_p, C.c_void_p)


class _Lib:
    def __init__(self):
        err = []
        self.h = None
        for name in _LIB_NAMES:
            try:
                self.h = C.CDLL(name)
                break
            except OSError as e:
                err.append(f"{name}: {e}")
        if self.h is None:
            raise H5Error("the HDF5 C library (libhdf5) was not found: " + "; ".join(err))
        h = self.h
        sig = {
            "H5open": (C.c_int, []),
            "H5Eset_auto2": (C.c_int, [hid_t, C.c_void_p, C.c_void_p]),
            "H5Fcreate": (hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]),
            "H5Fopen": (hid_t, [C.c_char_p, C.c_uint, hid_t]),
            "H5Fclose": (C.c_int, [hid_t]),
            "H5Fflush": (C.c_int, [hid_t, C.c_int]),
            "H5Gcreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t]),
            "H5Gopen2": (hid_t, [hid_t, C.c_char_p, hid_t]),
            "H5Gclose": (C.c_int, [hid_t]),
            "H5Lexists": (C.c_int, [hid_t, C.c_char_p, hid_t]),
            "H5Literate": (C.c_int, [hid_t, C.c_int, C.c_int, C.POINTER(hsize_t), _ITER_CB, C.c_void_p]),
            "H5Oopen": (hid_t, [hid_t, C.c_char_p, hid_t]),
            "H5Oclose": (C.c_int, [hid_t]),
            "H5Iget_type": (C.c_int, [hid_t]),
            "H5Screate": (hid_t, [C.c_int]),
            "H5Screate_simple": (hid_t, [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
            "H5Sclose": (C.c_int, [hid_t]),
            "H5Sget_simple_extent_type": (C.c_int, [hid_t]),
            "H5Sget_simple_extent_ndims": (C.c_int, [hid_t]),
            "H5Sget_simple_extent_dims": (C.c_int, [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
            "H5Sselect_hyperslab": (C.c_int, [hid_t, C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t),
                                              C.POINTER(hsize_t), C.POINTER(hsize_t)]),
            "H5Pcreate": (hid_t, [hid_t]),
            "H5Pset_chunk": (C.c_int, [hid_t, C.c_int, C.POINTER(hsize_t)]),
            "H5Pclose": (C.c_int, [hid_t]),
            "H5Dcreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
            "H5Dopen2": (hid_t, [hid_t, C.c_char_p, hid_t]),
            "H5Dclose": (C.c_int, [hid_t]),
            "H5Dget_space": (hid_t, [hid_t]),
            "H5Dget_type": (hid_t, [hid_t]),
            "H5Dwrite": (C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
            "H5Dread": (C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
            "H5Dvlen_reclaim": (C.c_int, [hid_t, hid_t, hid_t, C.c_void_p]),
            "H5Tcopy": (hid_t, [hid_t]),
            "H5Tclose": (C.c_int, [hid_t]),
            "H5Tset_size": (C.c_int, [hid_t, C.c_size_t]),
            "H5Tset_cset": (C.c_int, [hid_t, C.c_int]),
            "H5Tget_class": (C.c_int, [hid_t]),
            "H5Tget_size": (C.c_size_t, [hid_t]),
            "H5Tget_sign": (C.c_int, [hid_t]),
            "H5Tis_variable_str": (C.c_int, [hid_t]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        if h.H5open() < 0:
            raise H5Error("H5open failed")
        h.H5Eset_auto2(0, None, None)   # errors come back as status codes, raised here

        def g(name):
            return hid_t.in_dll(h, name).value
        self.native = {np.dtype(np.float64): g("H5T_NATIVE_DOUBLE_g"), np.dtype(np.float32): g("H5T_NATIVE_FLOAT_g"),
                       np.dtype(np.int64): g("H5T_NATIVE_LLONG_g"), np.dtype(np.int32): g("H5T_NATIVE_INT_g"),
                       np.dtype(np.int16): g("H5T_NATIVE_SHORT_g"), np.dtype(np.int8): g("H5T_NATIVE_SCHAR_g"),
                       np.dtype(np.uint8): g("H5T_NATIVE_UCHAR_g"), np.dtype(np.uint16): g("H5T_NATIVE_USHORT_g"),
                       np.dtype(np.uint32): g("H5T_NATIVE_UINT_g"), np.dtype(np.uint64): g("H5T_NATIVE_ULLONG_g")}
        self.c_s1 = g("H5T_C_S1_g")
        self.dcpl_cls = g("H5P_CLS_DATASET_CREATE_ID_g")

    def check(self, rc, what):
        if rc < 0:
            raise H5Error(f"HDF5: {what} failed")
        return rc

    def vlen_str(self):
        t = self.check(self.h.H5Tcopy(self.c_s1), "H5Tcopy")
        self.h.H5Tset_size(t, _VARIABLE)
        self.h.H5Tset_cset(t, H5T_CSET_UTF8)
        return t


_lib: Optional[_Lib] = None


def _h5() -> _Lib:
    global _lib
    if _lib is None:
        _lib = _Lib()
    return _lib


def available() -> bool:
    """Whether the HDF5 C library can be loaded."""
    try:
        _h5()
        return True
    except H5Error:
        return False


def _dims(seq):
    arr = (hsize_t * max(len(seq), 1))(*[int(v) for v in seq])
    return arr


_JULIA_INT = {1: "Int8", 2: "Int16", 4: "Int32", 8: "Int64"}
_JULIA_UINT = {1: "UInt8", 2: "UInt16", 4: "UInt32", 8: "UInt64"}


class H5File:
    """An open HDF5 file.  mode: 'w' (create, truncate), 'cw' (open read-write, create if
    missing), 'r+' (read-write), 'r' (read-only), the modes of HDF5.h5open."""

    def __init__(self, path: str, mode: str = "r"):
        L = _h5()
        self.L = L
        self.path = str(path)
        p = self.path.encode()
        if mode == "w":
            fid = L.h.H5Fcreate(p, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
        elif mode == "cw":
            fid = (L.h.H5Fopen(p, H5F_ACC_RDWR, H5P_DEFAULT) if os.path.exists(self.path)
                   else L.h.H5Fcreate(p, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT))
        elif mode == "r+":
            fid = L.h.H5Fopen(p, H5F_ACC_RDWR, H5P_DEFAULT)
        elif mode == "r":
            fid = L.h.H5Fopen(p, H5F_ACC_RDONLY, H5P_DEFAULT)
        else:
            raise ValueError(f"unknown HDF5 file mode {mode!r}")
        if fid < 0:
            raise H5Error(f"cannot open {self.path!r} (mode {mode})")
        self.fid = fid

    def close(self):
        if self.fid is not None:
            self.L.h.H5Fclose(self.fid)
            self.fid = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- paths
    def exists(self, path: str) -> bool:
        """_h5_dataset_exists (:149-157): every component of the path is present."""
        h = self.L.h
        cur = ""
        for part in [q for q in path.split("/") if q]:
            cur = f"{cur}/{part}" if cur else part
            if h.H5Lexists(self.fid, cur.encode(), H5P_DEFAULT) <= 0:
                return False
        return True

    def ensure_group(self, path: str) -> None:
        """_ensure_group (:95-106)."""
        h = self.L.h
        cur = ""
        for part in [q for q in path.split("/") if q]:
            cur = f"{cur}/{part}" if cur else part
            if h.H5Lexists(self.fid, cur.encode(), H5P_DEFAULT) <= 0:
                gid = self.L.check(h.H5Gcreate2(self.fid, cur.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                                   f"create group /{cur}")
                h.H5Gclose(gid)

    def _parent(self, path: str):
        parts = [q for q in path.split("/") if q]
        if not parts:
            raise ValueError("empty HDF5 path")
        self.ensure_group("/".join(parts[:-1]))
        return "/".join(parts)

    # -- writing
    def put(self, path: str, value) -> None:
        """_put! (:108-115) with _h5_value (:9-13): a scalar, string, tuple/list of numbers or
        array (Julia dims) as a dataset at path, groups created on the way."""
        name = self._parent(path)
        h, L = self.L.h, self.L
        if value is None:
            value = "nothing"
        if isinstance(value, (tuple, list)):
            value = _h5_tuple(value)
        if isinstance(value, str):
            t = L.vlen_str()
            sp = L.check(h.H5Screate(H5S_SCALAR), "H5Screate")
            try:
                d = L.check(h.H5Dcreate2(self.fid, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                            f"create /{name}")
                buf = (C.c_char_p * 1)(value.encode("utf-8"))
                L.check(h.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, C.cast(buf, C.c_void_p)), f"write /{name}")
                h.H5Dclose(d)
            finally:
                h.H5Sclose(sp)
                h.H5Tclose(t)
            return
        arr = np.asarray(_h5_number(value))
        if arr.dtype not in L.native:
            raise H5Error(f"/{name}: unsupported element type {arr.dtype}")
        t = L.native[arr.dtype]
        if arr.ndim == 0:
            sp = h.H5Screate(H5S_SCALAR)
            data = np.ascontiguousarray(arr)
        elif arr.size == 0:
            sp = h.H5Screate(H5S_NULL)
            data = None
        else:
            data = np.ascontiguousarray(arr.T)          # Julia dims -> HDF5 dims (reversed)
            sp = h.H5Screate_simple(data.ndim, _dims(data.shape), None)
        L.check(sp, "create dataspace")
        try:
            d = L.check(h.H5Dcreate2(self.fid, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                        f"create /{name}")
            if data is not None:
                L.check(h.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, data.ctypes.data_as(C.c_void_p)),
                        f"write /{name}")
            h.H5Dclose(d)
        finally:
            h.H5Sclose(sp)

    def create_empty(self, path: str, dtype, dims, chunk=None) -> None:
        """_put_empty_dataset! (:132-145): a dataset of Julia dims `dims`, unwritten
        (reads as the fill value 0), optionally chunked with Julia chunk dims."""
        name = self._parent(path)
        h, L = self.L.h, self.L
        t = L.native[np.dtype(dtype)]
        hd = list(reversed([int(v) for v in dims]))
        sp = L.check(h.H5Screate_simple(len(hd), _dims(hd), None), "create dataspace")
        dcpl = H5P_DEFAULT
        try:
            if chunk is not None:
                dcpl = L.check(h.H5Pcreate(L.dcpl_cls), "H5Pcreate")
                hc = list(reversed([int(v) for v in chunk]))
                L.check(h.H5Pset_chunk(dcpl, len(hc), _dims(hc)), "H5Pset_chunk")
            d = L.check(h.H5Dcreate2(self.fid, name.encode(), t, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT),
                        f"create /{name}")
            h.H5Dclose(d)
        finally:
            if dcpl != H5P_DEFAULT:
                h.H5Pclose(dcpl)
            h.H5Sclose(sp)

    def write_frame(self, path: str, frame: int, plane: np.ndarray) -> None:
        """dset[:, :, frame] = plane (frame 1-based, plane in Julia dims (d1, d2)) on a
        (d1, d2, nf) dataset: one slab of the chunked movie."""
        h, L = self.L.h, self.L
        d = L.check(h.H5Dopen2(self.fid, path.strip("/").encode(), H5P_DEFAULT), f"open /{path}")
        try:
            fs = h.H5Dget_space(d)
            nd = h.H5Sget_simple_extent_ndims(fs)
            hd = (hsize_t * 3)()
            h.H5Sget_simple_extent_dims(fs, hd, None)
            if nd != 3 or not (1 <= frame <= hd[0]) or (hd[1], hd[2]) != tuple(reversed(plane.shape)):
                h.H5Sclose(fs)
                raise H5Error(f"/{path}: frame {frame} of shape {plane.shape} does not fit the movie")
            data = np.ascontiguousarray(np.asarray(plane, np.float32).T)
            L.check(h.H5Sselect_hyperslab(fs, H5S_SELECT_SET, _dims([frame - 1, 0, 0]), None,
                                          _dims([1, hd[1], hd[2]]), None), "select frame")
            ms = h.H5Screate_simple(3, _dims([1, hd[1], hd[2]]), None)
            try:
                L.check(h.H5Dwrite(d, L.native[np.dtype(np.float32)], ms, fs, H5P_DEFAULT,
                                   data.ctypes.data_as(C.c_void_p)), f"write frame {frame} of /{path}")
            finally:
                h.H5Sclose(ms)
                h.H5Sclose(fs)
        finally:
            h.H5Dclose(d)

    # -- reading
    def _info(self, d):
        h = self.L.h
        t = h.H5Dget_type(d)
        sp = h.H5Dget_space(d)
        cls, size, sign = h.H5Tget_class(t), h.H5Tget_size(t), h.H5Tget_sign(t)
        kind = h.H5Sget_simple_extent_type(sp)
        nd = h.H5Sget_simple_extent_ndims(sp) if kind == H5S_SIMPLE else 0
        hd = (hsize_t * max(nd, 1))()
        if nd:
            h.H5Sget_simple_extent_dims(sp, hd, None)
        h.H5Sclose(sp)
        return t, cls, size, sign, kind, [int(hd[i]) for i in range(nd)]

    def read(self, path: str):
        """HDF5.read of one dataset: a str, a numpy scalar, or an array in Julia dims."""
        h, L = self.L.h, self.L
        d = L.check(h.H5Dopen2(self.fid, path.strip("/").encode(), H5P_DEFAULT), f"open /{path}")
        try:
            t, cls, size, sign, kind, hd = self._info(d)
            try:
                if cls == H5T_STRING:
                    if h.H5Tis_variable_str(t) > 0:
                        mt = L.vlen_str()
                        buf = (C.c_char_p * 1)()
                        L.check(h.H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, C.cast(buf, C.c_void_p)),
                                f"read /{path}")
                        out = buf[0].decode("utf-8") if buf[0] is not None else ""
                        sp = h.H5Screate(H5S_SCALAR)
                        h.H5Dvlen_reclaim(mt, sp, H5P_DEFAULT, C.cast(buf, C.c_void_p))
                        h.H5Sclose(sp)
                        h.H5Tclose(mt)
                        return out
                    raw = C.create_string_buffer(size)
                    L.check(h.H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, C.cast(raw, C.c_void_p)), f"read /{path}")
                    return raw.raw.rstrip(b"\0").decode("utf-8")
                if cls == H5T_FLOAT:
                    dt = {4: np.float32, 8: np.float64}[size]
                elif cls == H5T_INTEGER:
                    dt = np.dtype(("u" if sign == H5T_SGN_NONE else "i") + str(size))
                else:
                    raise H5Error(f"/{path}: unsupported HDF5 type class {cls}")
                dt = np.dtype(dt)
                if kind == H5S_NULL:
                    return np.zeros(0, dt)
                out = np.zeros(tuple(hd), dt)
                L.check(h.H5Dread(d, L.native[dt], H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data_as(C.c_void_p)),
                        f"read /{path}")
                return out[()] if out.ndim == 0 else out.T
            finally:
                h.H5Tclose(t)
        finally:
            h.H5Dclose(d)

    def keys(self, group: str = "") -> list:
        """Names of the links of a group, in name order."""
        h, L = self.L.h, self.L
        gid = L.check(h.H5Gopen2(self.fid, (group.strip("/") or "/").encode(), H5P_DEFAULT), f"open group /{group}")
        names = []

        def cb(_g, name, _info, _data):
            names.append(name.decode())
            return 0
        fn = _ITER_CB(cb)
        try:
            L.check(h.H5Literate(gid, H5_INDEX_NAME, H5_ITER_INC, None, fn, None), f"iterate /{group}")
        finally:
            h.H5Gclose(gid)
        return names

    def kind(self, path: str) -> str:
        h = self.L.h
        o = self.L.check(h.H5Oopen(self.fid, path.strip("/").encode(), H5P_DEFAULT), f"open /{path}")
        try:
            t = h.H5Iget_type(o)
        finally:
            h.H5Oclose(o)
        return "group" if t == H5I_GROUP else ("dataset" if t == H5I_DATASET else "other")

    def signature_row(self, path: str) -> dict:
        h = self.L.h
        d = self.L.check(h.H5Dopen2(self.fid, path.strip("/").encode(), H5P_DEFAULT), f"open /{path}")
        try:
            t, cls, size, sign, kind, hd = self._info(d)
            h.H5Tclose(t)
        finally:
            h.H5Dclose(d)
        if cls == H5T_STRING:
            dtype = "String"
        elif cls == H5T_FLOAT:
            dtype = {2: "Float16", 4: "Float32", 8: "Float64"}[size]
        elif cls == H5T_INTEGER:
            dtype = (_JULIA_UINT if sign == H5T_SGN_NONE else _JULIA_INT)[size]
        else:
            dtype = f"class{cls}"
        return dict(path=path.strip("/"), shape=tuple(reversed(hd)), dtype=dtype)


def _h5_number(x):
    """The element types HDF5.jl writes for Julia numbers: Int -> Int64, Float64, and the
    fixed-width numpy kinds as they are; a Python bool becomes an Int8 flag."""
    if isinstance(x, (bool, np.bool_)):
        return np.int8(1 if x else 0)
    if isinstance(x, int):
        return np.int64(x)
    if isinstance(x, float):
        return np.float64(x)
    return x


def _h5_tuple(x):
    """_h5_value(::Tuple) (:13): a tuple of numbers becomes a vector, anything else its string."""
    if all(isinstance(v, (int, float, bool, np.integer, np.floating)) for v in x):
        if all(isinstance(v, (int, bool, np.integer)) for v in x):
            return np.asarray([int(v) for v in x], np.int64)
        return np.asarray([float(v) for v in x], np.float64)
    return str(tuple(x))


# ---- payload helpers (:117-130) ----------------------------------------------------------

def _get(nt, key, default=None):
    if nt is None:
        return default
    if isinstance(nt, dict):
        return nt.get(key, default)
    return getattr(nt, key, default)


def _vec(nt, key, n, fill_value=math.nan):
    v = _get(nt, key)
    return np.full(n, float(fill_value)) if v is None else np.asarray(v, np.float64).reshape(-1)


def _ivec(nt, key, n, fill_value=0):
    v = _get(nt, key)
    return np.full(n, int(fill_value), np.int64) if v is None else np.asarray(v, np.int64).reshape(-1)


# ---- state accessors ---------------------------------------------------------------------

def _spec(state):
    return state.spec


def _dims_from_state(state):
    """_dims_from_state (:177-181): (Nx, Ny, Nz, n_material)."""
    nx, ny, nz = _spec(state).shape
    return nx, ny, nz, int(_spec(state).n_film)


def _ct(state):
    return np.dtype(_spec(state).compute_dtype)


def _gd_value(state, key):
    """A field of the state's (compute-type) ferrimagnet parameters, in that type, as the
    reference's state.model.params holds it; the default parameters when the state has none."""
    gd = _spec(state).gd
    if gd is None:
        from . import models
        return float(models.ferrimagnet_parameters()[key])
    return _ct(state).type(gd[key])


def _ferri(state):
    gd = _spec(state).gd
    if gd is None:
        from . import models
        return models.ferrimagnet_parameters()
    return gd


def _state_n(state):
    return int(getattr(state, "n", 0))


def _center_index(n):
    """cld(n, 2), 0-based."""
    return (n + 1) // 2 - 1


def _eps_slices(state, x_um=(10.0, 20.0, 30.0, 40.0), slice_z_pos=None, slice_y_pos=0.0):
    """_eps_slices (:183-201): the xy slice at kz, the xz slice at jy and yz slices at the
    x positions (micrometres), Float64, in the reference's orientation."""
    sp = _spec(state)
    if sp.epsr is None:
        raise ValueError("the state keeps no permittivity volume")
    eps = np.asarray(sp.epsr, np.float64).reshape(sp.shape, order="F")
    g = sp.grid
    kz = _center_index(g.z.n) if slice_z_pos is None else source_index(g.z, float(slice_z_pos))
    jy = source_index(g.y, float(slice_y_pos))
    xy = eps[:, :, kz]
    xz = eps[:, jy, :]
    yz = {}
    for x in x_um:
        yz[int(round(x))] = eps[source_index(g.x, float(x) * 1e-6), :, :]
    return xy, xz, yz


# ---- production file ---------------------------------------------------------------------

_METADATA_STRINGS = {
    "Hsw_role": "Hsw0 is a parameter record; no applied switching field during optical run",
    "Q_spin_formula": "Q_spin reservoirs follow the coupled four-temperature spin-bath energy exchange",
    "Q_spin_units": "J/m^3",
    "absorption_ade_work_formula": "positive local E dot J_ADE work accumulated in active magneto-optic cells",
    "absorption_ade_work_note": "matches the production :ade_work fluence path; nonpositive local work is clipped for heating",
    "absorption_cycle_average_formula": "0.5*omega*eps0*eps_imag*|E|^2",
    "ade_update_model": "diagonal ADE polarization current",
    "magnetization_internal_units": "reduced sublattice moments",
    "magnetization_model": "two-sublattice ferrimagnet LLB with 4TM reservoirs",
    "magnetization_model_source": "MagnetoPhotonic generalized ferrimagnet model",
    "magnetization_physical_units": "A/m",
    "maxwell_update_model": "Yee + CPML + ADE + MO",
    "pml_launch_model": "KernelAbstractions",
    "precision": "storage Float32 / selectable compute precision",
    "production_curve_cleanup": "package production replication writer",
    "pump_coupling": "mode source into NOT-gate arm",
    "pump_frame_hdf5_chunking": "frame-major float32 movies",
    "pump_frame_storage_precision": "Float32",
    "use_split_maxwell_pml_requested": "false",
    "initial_magnetization_source": "ferrimagnet equilibrium at T0",
}

_JULIA_FLOAT = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64"}


def _write_metadata(f: H5File, state, metadata=None, frame_count: int = 0):
    """_write_metadata! (:203-303)."""
    md = metadata or {}
    sp = _spec(state)
    Nx, Ny, Nz, _ = _dims_from_state(state)
    xy, xz, yz = _eps_slices(state, slice_z_pos=_get(md, "eps_slice_z"), slice_y_pos=_get(md, "eps_slice_y", 0.0))
    p = sp.params
    ct = _JULIA_FLOAT[_ct(state)]
    fld = _JULIA_FLOAT[np.dtype(sp.field_dtype)]
    ade_work = sp.absorption_model == "ade_work"
    scalars = {
        "Nx": Nx, "Ny": Ny, "Nz": Nz, "dt": float(sp.dt),
        "eps_r_vac": float(p.epsr_vac), "eps_r_sio2": float(p.epsr_sio2),
        "Q_voigt_TM": _gd_value(state, "Q_voigt_TM"), "Q_voigt_RE": _gd_value(state, "Q_voigt_RE"),
        "Hsw0_T": _gd_value(state, "Hsw0"), "pump_E_scale": _get(md, "pump_E_scale", 1.0),
        "target_F_abs_mJcm2": _get(md, "target_F_abs_mJcm2", math.nan),
        "E_amplitude_base": _get(md, "E_amplitude_base", math.nan),
        "E_amplitude_scale_to_target_F_abs": _get(md, "E_amplitude_scale_to_target_F_abs", 1.0),
        "E_amplitude": _get(md, "E_amplitude", math.nan),
        "absorbed_fluence_avg_mJcm2": _get(md, "absorbed_fluence_avg_mJcm2", math.nan),
        "absorbed_fluence_local_max_mJcm2": _get(md, "absorbed_fluence_local_max_mJcm2", math.nan),
        "fluence_scale_from_base": _get(md, "fluence_scale_from_base", _get(md, "E_amplitude_scale_to_target_F_abs", 1.0)),
        "U_abs_local_max_Jm3": _get(md, "U_abs_local_max_Jm3", math.nan),
        "absorption_model_id": 2 if ade_work else 1,
        "brillouin_iters": int(sp.brillouin_iters),
        "debug_max_probe_steps": _get(md, "debug_max_probe_steps", 0),
        "debug_max_pump_steps": _get(md, "debug_max_pump_steps", 0),
        "debug_nan_trace_dense_steps": _get(md, "debug_nan_trace_dense_steps", 0),
        "debug_nan_trace_interval": _get(md, "debug_nan_trace_interval", 0),
        "debug_nan_trace_steps": _get(md, "debug_nan_trace_steps", 0),
        "debug_progress_interval": _get(md, "debug_progress_interval", 0),
        "frame_skip": _get(md, "frame_skip", 0),
        "initial_m_RE_x_reduced": _get(md, "initial_m_RE_x_reduced", -1.0),
        "initial_m_TM_x_reduced": _get(md, "initial_m_TM_x_reduced", 1.0),
        "n_relaxation_trajectories": _get(md, "n_relaxation_trajectories", 1),
        "probe_amplitude_V_m": _get(md, "probe_amplitude_V_m", 1.0e6),
        "probe_duration_s": _get(md, "probe_duration_s", 0.6e-12),
        "probe_lambda_nm": _get(md, "probe_lambda_nm", 532.0),
        "probe_post_x_um_requested": _get(md, "probe_post_x_um_requested", 48.0),
        "probe_pre_x_um_requested": _get(md, "probe_pre_x_um_requested", 33.0),
        "probe_spectrum_bins": _get(md, "probe_spectrum_bins", 5),
        "probe_trace_stride": _get(md, "probe_trace_stride", 10),
        "pump_MLUPS": _get(md, "pump_MLUPS", math.nan),
        "pump_elapsed_s": _get(md, "pump_elapsed_s", math.nan),
        "pump_ns_per_cell_step": _get(md, "pump_ns_per_cell_step", math.nan),
        "steps_pump_nominal": _get(md, "steps_pump_nominal", _state_n(state)),
        "steps_pump_run": _get(md, "steps_pump_run", _state_n(state)),
    }
    # run flags are Int8 in the reference files (a strict type check compares them)
    flags = {"debug_disable_ade": 0, "debug_disable_multiphysics": 0, "debug_disable_source": 0,
             "debug_nan_trace": 0, "debug_skip_equilibration": 0, "debug_stop_after_pump": 0,
             "debug_stop_after_pump_effective": 0, "deterministic_relaxation": 1, "async_frame_io": 0,
             "run_probe": 1, "save_probe_frames": 1, "save_pump_frames": 1 if frame_count > 0 else 0,
             "use_split_maxwell_pml_effective": 0}
    strings = dict(_METADATA_STRINGS)
    strings.update({
        "absorption_model": sp.absorption_model,
        "ade_precision": ct, "cpml_precision": ct, "em_field_storage_precision": fld, "field_precision": fld,
        "magnetization_precision": ct, "maxwell_coef_precision": ct, "thermal_precision": ct,
    })
    for k, v in scalars.items():
        f.put(f"metadata/{k}", v)
    for k, v in flags.items():
        f.put(f"metadata/{k}", np.int8(v))
    for k, v in strings.items():
        f.put(f"metadata/{k}", v)
    f.put("metadata/bulk_grid_shape", np.asarray([Nx, Ny, Nz], np.int64))
    f.put("metadata/pml_cells_xyz", np.asarray([_get(md, "pml_x", 0), _get(md, "pml_y", 0), _get(md, "pml_z", 0)], np.int64))
    f.put("metadata/pml_thickness_nm_xyz", np.asarray([_get(md, "pml_x_nm", math.nan), _get(md, "pml_y_nm", math.nan),
                                                      _get(md, "pml_z_nm", math.nan)], np.float64))
    f.put("metadata/threads_3d", np.asarray([8, 8, 4], np.int64))
    g = sp.grid
    f.put("metadata/x_coords", np.asarray(g.x.centers, np.float64) * 1e6)
    f.put("metadata/y_coords", np.asarray(g.y.centers, np.float64) * 1e6)
    f.put("metadata/z_coords", np.asarray(g.z.centers, np.float64) * 1e6)
    f.put("metadata/eps_z_xy", xy)
    f.put("metadata/eps_z_xz", xz)
    for key in (10, 20, 30, 40):
        f.put(f"metadata/eps_z_yz_{key}", yz[key])


FILM_TS_KEYS = (
    "mag_time", "mx_TM", "mx_RE",
    "m_TM_x_reduced", "m_TM_y_reduced", "m_TM_z_reduced",
    "m_RE_x_reduced", "m_RE_y_reduced", "m_RE_z_reduced",
    "M_TM_x_Apm", "M_TM_y_Apm", "M_TM_z_Apm",
    "M_RE_x_Apm", "M_RE_y_Apm", "M_RE_z_Apm",
    "M_net_x_Apm", "M_net_y_Apm", "M_net_z_Apm",
    "Te_K", "Tl_K", "Ts_TM_K", "Ts_RE_K", "Te_avg", "Tl_avg", "Ts_avg",
    "mag_TM_norm_avg", "mag_RE_norm_avg",
)

PUMP_ACTIVE_KEYS = (
    "m_TM_x_reduced_active_cells", "m_TM_y_reduced_active_cells", "m_TM_z_reduced_active_cells",
    "m_RE_x_reduced_active_cells", "m_RE_y_reduced_active_cells", "m_RE_z_reduced_active_cells",
    "M_TM_x_Apm_active_cells", "M_TM_y_Apm_active_cells", "M_TM_z_Apm_active_cells",
    "M_RE_x_Apm_active_cells", "M_RE_y_Apm_active_cells", "M_RE_z_Apm_active_cells",
    "M_net_x_Apm_active_cells", "M_net_y_Apm_active_cells", "M_net_z_Apm_active_cells",
    "mag_TM_norm_active_cells", "mag_RE_norm_active_cells",
    "Te_active_cells", "Tl_active_cells", "Ts_TM_active_cells", "Ts_RE_active_cells",
)


def _write_film_timeseries(f: H5File, group: str, data, n: int = 0):
    """_write_film_timeseries! (:316-321)."""
    n = n if n > 0 else len(_vec(data, "mag_time", 0))
    for key in FILM_TS_KEYS:
        f.put(f"{group}/{key}", _vec(data, key, n))


def _put_movie(f: H5File, path: str, source, key: str, dims):
    """_put_movie! (:160-175): a movie already streamed into the file stays; otherwise the
    payload's cube, or an unwritten chunked (d1, d2, 1) dataset of the given dims."""
    if f.exists(path):
        return
    chunk = (max(1, dims[0]), max(1, dims[1]), 1)
    v = _get(source, key)
    if v is None:
        if any(int(d) == 0 for d in dims):
            f.put(path, np.zeros(dims, np.float32))
            return
        f.create_empty(path, np.float32, dims, chunk=chunk)
        return
    vh = np.asarray(v)
    if vh.size == 0 and all(int(d) > 0 for d in dims):
        f.create_empty(path, np.float32, dims, chunk=chunk)
        return
    f.put(path, vh.astype(np.float32))


def _write_pump_group(f: H5File, state, data=None, active=None, movies=None):
    """_write_pump_group! (:323-347)."""
    Nx, Ny, Nz, Na = _dims_from_state(state)
    n = len(_vec(data, "mag_time", 0))
    _write_film_timeseries(f, "pump", data, n)
    for key in PUMP_ACTIVE_KEYS:
        f.put(f"pump/{key}", _vec(active, key, Na))
    for key in ("U_abs_J_m3", "U_abs_active_cells", "active_cell_volume_m3", "pabs_peak_W_m3", "pabs_peak_active_cells"):
        f.put(f"pump/{key}", _vec(active, key, Na, fill_value=0.0))
    f.put("pump/active_linear_index", _ivec(active, "active_linear_index", Na))
    f.put("pump/hot_cell_index", int(_get(active, "hot_cell_index", 1)))
    nf = int(_get(movies, "frame_count", 0))
    _put_movie(f, "pump/Ez_xy", movies, "Ez_xy", (Nx, Ny, nf))
    _put_movie(f, "pump/Ez_xz", movies, "Ez_xz", (Nx, Nz, nf))
    for key in ("Ez_yz_10", "Ez_yz_20", "Ez_yz_30", "Ez_yz_40"):
        _put_movie(f, f"pump/{key}", movies, key, (Ny, Nz, nf))


def _write_relax_group(f: H5File, group: str, data, elapsed_s=math.nan):
    """_write_relax_group! (:349-352)."""
    _write_film_timeseries(f, group, data)
    f.put(f"{group}/elapsed_s", float(elapsed_s))


_SHOT_FINAL_KEYS = ("final_Te_K_active", "final_Tl_K_active", "final_Ts_RE_K_active", "final_Ts_TM_K_active",
                    "final_U_abs_J_m3_active", "final_m_RE_x_reduced_active", "final_m_RE_y_reduced_active",
                    "final_m_RE_z_reduced_active", "final_m_TM_x_reduced_active", "final_m_TM_y_reduced_active",
                    "final_m_TM_z_reduced_active")
_SHOT_SCALAR_KEYS = ("active_cell_MLUPS", "core_Uabs_fraction_of_max", "final_core_switch_fraction",
                     "final_switch_fraction", "final_core_mean_m_RE_x", "final_core_mean_m_TM_x",
                     "final_mixture_check_m_RE_x", "final_mixture_check_m_TM_x",
                     "final_re_reversed_fraction", "final_switched_cell_mean_m_RE_x",
                     "final_switched_cell_mean_m_TM_x", "final_tm_reversed_fraction",
                     "final_unswitched_cell_mean_m_RE_x", "final_unswitched_cell_mean_m_TM_x",
                     "ns_per_active_cell_step")


def _write_shot_group(f: H5File, state, data=None, active=None, elapsed_s=math.nan):
    """_write_shot_group! (:354-389)."""
    Nx, Ny, Nz, Na = _dims_from_state(state)
    _write_relax_group(f, "shot_1", data, elapsed_s)
    for key in _SHOT_FINAL_KEYS:
        fb = np.zeros(Na, np.float64)
        for sub in ("m_RE_x", "m_RE_y", "m_RE_z", "m_TM_x", "m_TM_y", "m_TM_z"):
            if key.startswith(f"final_{sub}"):
                fb = _vec(active, f"{sub}_reduced_active_cells", Na)
        v = _get(active, key)
        f.put(f"shot_1/{key}", fb if v is None else np.asarray(v))
    for key in _SHOT_SCALAR_KEYS:
        f.put(f"shot_1/{key}", float(_get(active, key, math.nan)))
    f.put("shot_1/final_core_cell_count", int(_get(active, "final_core_cell_count", Na)))
    f.put("shot_1/hot_cell_index", int(_get(active, "hot_cell_index", 1)))
    n = len(_vec(data, "mag_time", 0))
    f.put("shot_1/hot_cell_Te_K", _vec(active, "hot_cell_Te_K", n))
    f.put("shot_1/hot_cell_mx_RE", _vec(active, "hot_cell_mx_RE", n))
    f.put("shot_1/hot_cell_mx_TM", _vec(active, "hot_cell_mx_TM", n))
    f.put("shot_1/switched_fraction_TM_time", _vec(active, "switched_fraction_TM_time", n))


_PROBE_POLE_POSITIONS_ROWS = np.asarray([[0.0, 3.0e14], [4.0e15, 8.0e14]], np.float64)


def _write_probe_metadata_only(f: H5File, state, metadata=None):
    """_write_probe_metadata_only! (:391-415)."""
    md = metadata or {}
    gd = _ferri(state)
    eps0 = FDTDParams().eps0
    dt = float(_spec(state).dt)
    dp = dispersion.build_probe_poles(dt, eps0, gd)
    mp_ = dispersion.build_probe_mo_poles(dt, eps0, gd)
    ade_work = _spec(state).absorption_model == "ade_work"
    f.put("probe/Hsw0_T_for_parameter_record", _gd_value(state, "Hsw0"))
    f.put("probe/Q_voigt_RE", _gd_value(state, "Q_voigt_RE"))
    f.put("probe/Q_voigt_TM", _gd_value(state, "Q_voigt_TM"))
    f.put("probe/absorption_model", _spec(state).absorption_model)
    f.put("probe/absorption_model_id", 2 if ade_work else 1)
    ct = _ct(state) if _spec(state).gd is not None else np.dtype(np.float64)
    f.put("probe/eps_probe_target_re_im", np.asarray([gd["eps_real_probe"], gd["eps_imag_probe"]], np.float64).astype(ct))
    f.put("probe/final_core_switch_fraction_from_shot_1", _get(md, "final_core_switch_fraction", math.nan))
    f.put("probe/final_switch_fraction_from_shot_1", _get(md, "final_switch_fraction", math.nan))
    f.put("probe/geometry_note", "60um NOT-gate Y-branch waveguide with active GdFeCo film")
    f.put("probe/mo_coupling", "off-diagonal magneto-optic ADE coupling from frozen sublattice magnetization")
    f.put("probe/output_h5", _get(md, "output_h5", ""))
    f.put("probe/source_h5", _get(md, "source_h5", ""))
    f.put("probe/probe_run_mode", "gold_standard_probe_only")
    f.put("probe/probe_tau_note", "probe_tau is the Gaussian envelope tau used by the reference probe")
    f.put("probe/probe_tau_s", _get(md, "probe_tau_s", 24e-15))
    f.put("probe/readout_scope", "mode-overlap DFT, net flux energies, Jones Faraday/Kerr readout")
    f.put("probe/tra_method", "bare-waveguide reference netflux normalization with initial/switched contrast")
    # one row per pole, columns (C1, C2, C3): Julia (n_poles, 3)
    f.put("probe/probe_diag_poles_C1_C2_C3", np.asarray([[p[0], p[1], p[2]] for p in dp], np.float64))
    f.put("probe/probe_mo_poles_C1_C2_C3", np.asarray([[p[0], p[1], p[2]] for p in mp_], np.float64))
    f.put("probe/probe_pole_positions_omega0_gamma", _PROBE_POLE_POSITIONS_ROWS)


def write_production_h5(path: str, state, pump=None, relaxation=None, shot_1=None, active=None,
                        pump_active=None, shot_active=None, movies=None, metadata=None, overwrite: bool = True):
    """write_production_h5 (:417-454): metadata, the pump / relaxation / shot_1 groups, the
    probe parameter record, probe/reference and phase2_resume_finalization.  With
    overwrite=False an existing file opens read-write, so movies streamed into it during the
    run stay."""
    md = metadata or {}
    shot_1 = relaxation if shot_1 is None else shot_1
    pump_active = active if pump_active is None else pump_active
    shot_active = active if shot_active is None else shot_active
    c0 = FDTDParams().c0
    with H5File(path, "w" if overwrite else "cw") as f:
        _write_metadata(f, state, md, frame_count=int(_get(movies, "frame_count", 0)))
        _write_pump_group(f, state, pump, pump_active, movies)
        _write_relax_group(f, "relaxation", relaxation, _get(md, "relaxation_elapsed_s", math.nan))
        _write_shot_group(f, state, shot_1, shot_active, _get(md, "shot_elapsed_s", math.nan))
        _write_probe_metadata_only(f, state, md)
        f.put("probe/reference/Q_voigt_RE", _gd_value(state, "Q_voigt_RE"))
        f.put("probe/reference/Q_voigt_TM", _gd_value(state, "Q_voigt_TM"))
        f.put("probe/reference/dt", float(_spec(state).dt))
        f.put("probe/reference/lambda_nm", 532.0)
        f.put("probe/reference/omega_bins_rad_s", (2 * math.pi * c0 / 532e-9) * (1.0 + np.linspace(-0.04, 0.04, 5)))
        f.put("probe/reference/pole_fit_note", "probe poles fitted at lambda_probe")
        f.put("probe/reference/probe_amplitude_V_m", _get(md, "probe_amplitude_V_m", 1.0e6))
        f.put("probe/reference/probe_frame_count", 0)
        f.put("probe/reference/probe_frame_skip", 0)
        f.put("probe/reference/probe_frame_target", _get(md, "probe_frame_target", 903))
        f.put("probe/reference/probe_tau_fs", 24.0)
        f.put("probe/reference/state_label", "reference")
        f.put("probe/reference/steps_probe", _get(md, "steps_probe", 0))
        f.put("probe/reference/x_post_um", 48.0)
        f.put("probe/reference/x_pre_um", 33.0)
        f.put("phase2_resume_finalization/Hsw0_T", _gd_value(state, "Hsw0"))
        f.put("phase2_resume_finalization/dt_relax_s", _get(md, "dt_relax_s", math.nan))
        f.put("phase2_resume_finalization/elapsed_s", _get(md, "phase2_elapsed_s", math.nan))
        f.put("phase2_resume_finalization/final_core_switch_fraction", _get(shot_active, "final_core_switch_fraction", math.nan))
        f.put("phase2_resume_finalization/final_switch_fraction", _get(shot_active, "final_switch_fraction", math.nan))
        f.put("phase2_resume_finalization/source_h5", _get(md, "source_h5", ""))
        f.put("phase2_resume_finalization/status", "complete")
        f.put("phase2_resume_finalization/steps_relax", _get(md, "steps_relax", 0))
    return path


# ---- probe shots and the gold-standard file ----------------------------------------------

_SHOT_NUM_KEYS = ("A", "R", "T", "T_plus_R_plus_A", "absorbed_energy_J", "dt",
                  "ellipticity_faraday_deg", "ellipticity_kerr_deg", "group_delay_fs",
                  "incident_energy_J", "input_pulse_rms_fs", "lambda_nm",
                  "net_energy_post_J", "net_energy_pre_J", "probe_amplitude_V_m",
                  "probe_frame_count", "probe_frame_skip", "probe_frame_target",
                  "probe_tau_fs", "probe_trace_stride", "pulse_broadening_fs",
                  "reflected_energy_J", "steps_probe", "theta_faraday_deg", "theta_kerr_deg",
                  "throughput_loss", "transmitted_energy_J", "transmitted_pulse_rms_fs",
                  "x_post_um", "x_pre_um")
_SHOT_VEC_KEYS = ("T_omega", "R_omega", "Sx_incident_im", "Sx_incident_re",
                  "Sx_post_net_im", "Sx_post_net_re", "Sx_pre_net_im", "Sx_pre_net_re",
                  "jones_faraday_Ey_re_im", "jones_faraday_Ez_re_im",
                  "jones_kerr_Ey_re_im", "jones_kerr_Ez_re_im",
                  "omega_bins_rad_s", "Ey_trace_post", "Ey_trace_pre",
                  "Ez_trace_post", "Ez_trace_pre", "trace_time_s")
PROBE_MOVIE_KEYS = ("Ey_xy", "Ez_xy", "Ey_yz_post", "Ey_yz_pre", "Ez_yz_post", "Ez_yz_pre")


def _shot_value(v, default):
    if v is None:
        return default
    if isinstance(v, (list, tuple)) and all(isinstance(x, (int, float, np.integer, np.floating)) for x in v):
        return np.asarray(v, np.float64)
    return v


def _write_probe_shot(f: H5File, group: str, shot, state=None, write_movies: bool = False, eps_slice_z=None):
    """_write_probe_shot! (:456-507)."""
    for key in _SHOT_NUM_KEYS:
        f.put(f"{group}/{key}", _shot_value(_get(shot, key), math.nan))
    f.put(f"{group}/is_reference", np.int8(int(_get(shot, "is_reference", 0))))
    f.put(f"{group}/Q_voigt_RE", math.nan if state is None else _gd_value(state, "Q_voigt_RE"))
    f.put(f"{group}/Q_voigt_TM", math.nan if state is None else _gd_value(state, "Q_voigt_TM"))
    for key in _SHOT_VEC_KEYS:
        f.put(f"{group}/{key}", _shot_value(_get(shot, key), np.zeros(0, np.float64)))
    f.put(f"{group}/normalization_method", _get(shot, "normalization_method", "gold_standard_bare_waveguide_reference_netflux"))
    f.put(f"{group}/pole_fit_note", _get(shot, "pole_fit_note", "probe ADE poles fitted at lambda_probe"))
    f.put(f"{group}/state_label", _get(shot, "state_label", group))
    if write_movies and _get(shot, "probe_frame_indices") is not None:
        f.put(f"{group}/probe_frame_indices", np.asarray(_get(shot, "probe_frame_indices"), np.int64))
    if write_movies and _get(shot, "probe_frame_times_s") is not None:
        f.put(f"{group}/probe_frame_times_s", np.asarray(_get(shot, "probe_frame_times_s"), np.float64))
    if write_movies:
        frames = _get(shot, "frames", {}) or {}
        fc = int(_get(shot, "probe_frame_count", 0))
        if state is None:
            for key in PROBE_MOVIE_KEYS:
                _put_movie(f, f"{group}/{key}", frames, key, (fc, 0, 0))
        else:
            Nx, Ny, Nz, _ = _dims_from_state(state)
            _put_movie(f, f"{group}/Ey_xy", frames, "Ey_xy", (Nx, Ny, fc))
            _put_movie(f, f"{group}/Ez_xy", frames, "Ez_xy", (Nx, Ny, fc))
            for key in ("Ey_yz_post", "Ey_yz_pre", "Ez_yz_post", "Ez_yz_pre"):
                _put_movie(f, f"{group}/{key}", frames, key, (Ny, Nz, fc))
        if state is not None:
            xpre = float(_get(shot, "x_pre_um", 33.0))
            xpost = float(_get(shot, "x_post_um", 48.0))
            xy, _, yz = _eps_slices(state, x_um=(xpre, xpost), slice_z_pos=eps_slice_z)
            f.put(f"{group}/eps_xy", xy)
            f.put(f"{group}/eps_yz_post", yz[int(round(xpost))])
            f.put(f"{group}/eps_yz_pre", yz[int(round(xpre))])


def append_probe_shot_h5(path: str, label: str, shot, state=None):
    """append_probe_shot_h5! (:509-526): one completed shot below /probe/<label>, the file
    closed before returning; an existing group is an error."""
    group = label if label.startswith("probe/") else f"probe/{label}"
    with H5File(path, "r+") as f:
        if f.exists(group):
            raise H5Error(f"probe shot group already exists: /{group}")
        _write_probe_shot(f, group, shot, state=state, write_movies=False)
    return path


CONTRAST_KEYS = ("delta_A", "delta_R", "delta_T", "delta_ellipticity_faraday_deg",
                 "delta_ellipticity_kerr_deg", "delta_group_delay_fs",
                 "delta_pulse_broadening_fs", "delta_theta_faraday_deg",
                 "delta_theta_kerr_deg", "initial_T_R_A", "initial_T_plus_R_plus_A",
                 "switched_T_R_A", "switched_T_plus_R_plus_A")


def write_goldstd_h5(path: str, shots, state=None, reference=None, initial=None, switched=None,
                     contrast=None, metadata=None, overwrite: bool = True):
    """write_goldstd_h5 (:528-549): the probe parameter record (with a state), the reference,
    initial and switched shots (movies for the last two) and their contrast."""
    from .probe import probe_contrast
    md = dict(metadata or {})
    reference = _get(shots, "reference") if reference is None else reference
    initial = _get(shots, "initial") if initial is None else initial
    switched = _get(shots, "switched") if switched is None else switched
    contrast = probe_contrast(initial, switched) if contrast is None else contrast
    with H5File(path, "w" if overwrite else "cw") as f:
        if state is not None:
            _write_probe_metadata_only(f, state, {"output_h5": path, **md})
        slice_z = _get(md, "eps_slice_z")
        _write_probe_shot(f, "probe/reference", reference, state=state, write_movies=False)
        _write_probe_shot(f, "probe/initial", initial, state=state, write_movies=True, eps_slice_z=slice_z)
        _write_probe_shot(f, "probe/switched", switched, state=state, write_movies=True, eps_slice_z=slice_z)
        for key in CONTRAST_KEYS:
            v = _get(contrast, key, math.nan)
            f.put(f"probe/contrast/{key}", np.asarray(v, np.float64) if isinstance(v, (list, tuple)) else v)
    return path


# ---- streaming movie writers -------------------------------------------------------------

def _device_staging(nbytes: int):
    from .state import DeviceBuffer
    return DeviceBuffer.zeros(nbytes)


class ProductionFrameWriter:
    """production_frame_writer / record_production_frame! / close_frame_writer!
    (:551-638): the six pump |Ez| movies, Ez_xy (Nx, Ny, nf) at z = kz, Ez_xz (Nx, Nz, nf)
    at y = jy and Ez_yz_10..40 (Ny, Nz, nf) at x = 10, 20, 30, 40 um, chunked one slice per
    frame; each `record` stages the six planes on the device (`mp_plane_abs`), downloads
    them and writes frame `frame_idx`.  Frames past frame_count are ignored."""

    KEYS = ("Ez_xy", "Ez_xz", "Ez_yz_10", "Ez_yz_20", "Ez_yz_30", "Ez_yz_40")

    def __init__(self, path: str, state, frame_count: int, slice_z_pos=math.nan, slice_y_pos=0.0,
                 x_um=(10.0, 20.0, 30.0, 40.0)):
        sp = _spec(state)
        g = sp.grid
        self.nx, self.ny, self.nz = sp.shape
        self.kz = _center_index(g.z.n) if math.isnan(float(slice_z_pos)) else source_index(g.z, float(slice_z_pos))
        self.jy = source_index(g.y, float(slice_y_pos))
        self.ix = tuple(source_index(g.x, float(x) * 1e-6) for x in x_um)
        self.frame_idx = 1
        self.frame_count = int(frame_count)
        self.f = H5File(path, "cw")
        nf = self.frame_count
        self.dims = {"Ez_xy": (self.nx, self.ny), "Ez_xz": (self.nx, self.nz)}
        for k in self.KEYS[2:]:
            self.dims[k] = (self.ny, self.nz)
        for k in self.KEYS:
            d1, d2 = self.dims[k]
            self.f.create_empty(f"pump/{k}", np.float32, (d1, d2, nf), chunk=(d1, d2, 1))
        self.dev = None

    def planes(self):
        """(key, axis, index) of each movie plane (0-based)."""
        out = [("Ez_xy", 2, self.kz), ("Ez_xz", 1, self.jy)]
        out += [(k, 0, i) for k, i in zip(self.KEYS[2:], self.ix)]
        return out

    def write_frame(self, planes: dict) -> "ProductionFrameWriter":
        """Write one frame from host planes (Julia dims) and advance."""
        if self.frame_idx > self.frame_count:
            return self
        for k in self.KEYS:
            self.f.write_frame(f"pump/{k}", self.frame_idx, planes[k])
        self.frame_idx += 1
        return self

    def record(self, state) -> "ProductionFrameWriter":
        """record_production_frame! (:626-636) on a device state."""
        if self.frame_idx > self.frame_count:
            return self
        from . import _lib
        from .abi import MP_F32, MP_F64
        sp = _spec(state)
        code = MP_F32 if np.dtype(sp.field_dtype) == np.float32 else MP_F64
        if self.dev is None:
            self.dev = _device_staging(4 * max(d1 * d2 for d1, d2 in self.dims.values()))
        out = {}
        ez = state.f["ez"]
        for key, axis, idx in self.planes():
            d1, d2 = self.dims[key]
            _lib.call("plane_abs", ez.vp, code, self.nx, self.ny, self.nz, axis, idx, self.dev.vp)
            out[key] = self.dev.download(np.float32, d1 * d2).reshape((d1, d2), order="F")
        return self.write_frame(out)

    def close(self):
        """close_frame_writer! (:638)."""
        if self.dev is not None:
            self.dev.free()
            self.dev = None
        self.f.close()
        return self


def production_frame_writer(path: str, state, frame_count: int, slice_z_pos=math.nan, slice_y_pos=0.0,
                            x_um=(10.0, 20.0, 30.0, 40.0)) -> ProductionFrameWriter:
    return ProductionFrameWriter(path, state, frame_count, slice_z_pos, slice_y_pos, x_um)


def record_production_frame(w: ProductionFrameWriter, state) -> ProductionFrameWriter:
    return w.record(state)


class ProbeFrameWriter:
    """probe_frame_writer / record_probe_frame! (:640-697): the six probe movies of one shot
    below `group`, Ey_xy / Ez_xy (Nx, Ny, nf) at z = kz and Ey/Ez_yz_pre/post (Ny, Nz, nf) at
    the pre and post x-planes, staged on the device in one `mp_probe_frames` launch."""

    KEYS = ("Ey_xy", "Ez_xy", "Ey_yz_pre", "Ez_yz_pre", "Ey_yz_post", "Ez_yz_post")

    def __init__(self, path: str, state, group: str, frame_count: int, pre_x, post_x, slice_z_pos=math.nan):
        sp = _spec(state)
        g = sp.grid
        self.nx, self.ny, self.nz = sp.shape
        self.kz = _center_index(g.z.n) if math.isnan(float(slice_z_pos)) else source_index(g.z, float(slice_z_pos))
        self.ipre = source_index(g.x, pre_x)
        self.ipost = source_index(g.x, post_x)
        self.group = group.strip("/")
        self.frame_idx = 1
        self.frame_count = int(frame_count)
        self.f = H5File(path, "cw")
        nf = self.frame_count
        self.dims = {k: ((self.nx, self.ny) if k.endswith("_xy") else (self.ny, self.nz)) for k in self.KEYS}
        for k in ("Ey_xy", "Ez_xy", "Ey_yz_pre", "Ez_yz_pre", "Ey_yz_post", "Ez_yz_post"):
            d1, d2 = self.dims[k]
            self.f.create_empty(f"{self.group}/{k}", np.float32, (d1, d2, nf), chunk=(d1, d2, 1))
        self.dev = None

    def write_frame(self, planes: dict) -> "ProbeFrameWriter":
        if self.frame_idx > self.frame_count:
            return self
        for k in self.KEYS:
            self.f.write_frame(f"{self.group}/{k}", self.frame_idx, planes[k])
        self.frame_idx += 1
        return self

    def record(self, state) -> "ProbeFrameWriter":
        """record_probe_frame! (:685-697) on a device state."""
        if self.frame_idx > self.frame_count:
            return self
        from . import _lib
        if self.dev is None:
            self.dev = {k: _device_staging(4 * d1 * d2) for k, (d1, d2) in self.dims.items()}
        ptrs = [self.dev[k].vp for k in self.KEYS]
        _lib.call("probe_frames", C.byref(state.yee), self.kz, self.ipre, self.ipost, *ptrs)
        out = {}
        for k in self.KEYS:
            d1, d2 = self.dims[k]
            out[k] = self.dev[k].download(np.float32, d1 * d2).reshape((d1, d2), order="F")
        return self.write_frame(out)

    def close(self):
        if self.dev is not None:
            for b in self.dev.values():
                b.free()
            self.dev = None
        self.f.close()
        return self


def probe_frame_writer(path: str, state, group: str, frame_count: int, pre_x, post_x,
                       slice_z_pos=math.nan) -> ProbeFrameWriter:
    return ProbeFrameWriter(path, state, group, frame_count, pre_x, post_x, slice_z_pos)


def record_probe_frame(w: ProbeFrameWriter, state) -> ProbeFrameWriter:
    return w.record(state)


def close_frame_writer(w):
    return w.close()


# ---- magnetisation snapshot and state dump -----------------------------------------------

MAG_KEYS = ("m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z")


def magnetization_snapshot(state) -> dict:
    """magnetization_snapshot (src/fdtd/Solver.jl:448-467) as the file holds it: the six
    moments in the compute type, material_cells 1-based (Int64), material_fill."""
    sp = _spec(state)
    ct = _ct(state)
    fs = state.film_state() if sp.n_film else {}
    out = {k: np.asarray(fs[k]).astype(ct) if k in fs else np.zeros(0, np.float64) for k in MAG_KEYS}
    out["material_cells"] = np.asarray(sp.film_cells, np.int64) + 1
    out["material_fill"] = (np.asarray(sp.film_fill).astype(ct) if sp.film_fill is not None
                            else np.zeros(0, ct))
    return out


def save_magnetization(path: str, state, overwrite: bool = True):
    """save_magnetization (:699-707)."""
    if not _spec(state).n_film:
        raise ValueError("state has no magnetization")
    snap = magnetization_snapshot(state)
    with H5File(path, "w" if overwrite else "cw") as f:
        for k in MAG_KEYS + ("material_cells", "material_fill"):
            f.put(f"magnetization/{k}", snap[k])
    return path


def _read_node(f: H5File, path: str):
    """_read_node (:76-83): a dataset's value, or a dict of a group's children."""
    if path and f.kind(path) == "dataset":
        return f.read(path)
    return {k: _read_node(f, f"{path}/{k}" if path else k) for k in f.keys(path)}


def load_magnetization(path: str, group: str = "magnetization") -> dict:
    """load_magnetization (:709-713)."""
    with H5File(path, "r") as f:
        return _read_node(f, group)


def load_production_magnetization(path: str, state, group: str = "shot_1") -> dict:
    """load_production_magnetization (:715-772): the post-relax moments of a production file
    (`<group>/final_m_*_reduced_active`) mapped onto the state's film cells by the full-grid
    linear index (`pump/active_linear_index`, 1-based); cells the file lacks keep the
    as-grown equilibrium of its metadata (m_x = initial_m_*_x_reduced, m_y = m_z = 0)."""
    sp = _spec(state)
    if not sp.n_film:
        raise ValueError("state has no magnetization to populate (n_material == 0)")
    cells = np.asarray(sp.film_cells, np.int64) + 1
    n = len(cells)
    with H5File(path, "r") as f:
        dims = tuple(int(f.read(f"metadata/{k}")) for k in ("Nx", "Ny", "Nz"))
        if dims != tuple(sp.shape):
            raise ValueError(f"production h5 grid ({dims[0]}x{dims[1]}x{dims[2]}) != state grid "
                             f"({sp.shape[0]}x{sp.shape[1]}x{sp.shape[2]}); frozen magnetization can only be "
                             "mapped by linear index on an identical mesh")
        idx = np.asarray(f.read("pump/active_linear_index"), np.int64)
        src = {k: np.asarray(f.read(f"{group}/final_{k}_reduced_active"), np.float64) for k in MAG_KEYS}
        im_tm = float(f.read("metadata/initial_m_TM_x_reduced")) if f.exists("metadata/initial_m_TM_x_reduced") else 1.0
        im_re = float(f.read("metadata/initial_m_RE_x_reduced")) if f.exists("metadata/initial_m_RE_x_reduced") else -1.0
    out = {"m_TM_x": np.full(n, im_tm), "m_TM_y": np.zeros(n), "m_TM_z": np.zeros(n),
           "m_RE_x": np.full(n, im_re), "m_RE_y": np.zeros(n), "m_RE_z": np.zeros(n)}
    lut = {int(v): i for i, v in enumerate(idx)}
    matched = 0
    for p in range(n):
        j = lut.get(int(cells[p]))
        if j is not None:
            for k in MAG_KEYS:
                out[k][p] = src[k][j]
            matched += 1
    out["matched"] = matched
    return out


def save_hdf5_state(path: str, state, overwrite: bool = True, backend: str = "HIPBackend"):
    """save_hdf5_state (:48-74) for an FDTDState: provenance, then state/n, dt, backend,
    compute_T, the nine field volumes, the thermal baths and the magnetisation (film cells)
    and the region's material lists (1-based cells)."""
    sp = _spec(state)
    ct = _JULIA_FLOAT[_ct(state)]
    with H5File(path, "w" if overwrite else "cw") as f:
        f.put("provenance/package", "MagnetoPhotonic")
        f.put("provenance/version", "mphip")
        f.put("provenance/saved_at", datetime.datetime.now().isoformat(timespec="seconds"))
        f.put("state/n", _state_n(state))
        f.put("state/dt", float(sp.dt))
        f.put("state/backend", backend)
        f.put("state/compute_T", ct)
        for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Dx", "Dy", "Dz"):
            f.put(f"state/fields/{comp}", state.field(comp.lower()))
        if sp.n_film:
            fs = state.film_state()
            for k in ("Te", "Tl", "Ts_TM", "Ts_RE"):
                f.put(f"state/thermal/{k}", np.asarray(fs[k]))
            for k in MAG_KEYS:
                f.put(f"state/magnetization/{k}", np.asarray(fs[k]))
        snap = magnetization_snapshot(state)
        f.put("state/region/material_cells", snap["material_cells"])
        f.put("state/region/material_fill", snap["material_fill"])
        f.put("state/region/n_material", int(sp.n_film))
    return path


def load_hdf5_state(path: str) -> dict:
    """load_hdf5_state (:85-89): the whole file as nested dicts."""
    with H5File(path, "r") as f:
        return _read_node(f, "")


# ---- schema ------------------------------------------------------------------------------

def h5_schema_signature(path: str) -> list:
    """h5_schema_signature (:774-791): (path, shape, dtype) of every dataset, shapes in
    Julia dims, dtypes as Julia element type names."""
    rows = []
    with H5File(path, "r") as f:
        def visit(prefix):
            for k in f.keys(prefix):
                p = f"{prefix}/{k}" if prefix else k
                if f.kind(p) == "dataset":
                    rows.append(f.signature_row(p))
                else:
                    visit(p)
        visit("")
    return rows


def _normalize_dtype(s: str) -> str:
    """_normalize_dtype (:793-799): every string flavour is "String"."""
    return "String" if ("String" in s or "Cstring" in s or "Char" in s) else s


def assert_h5_schema_compatible(candidate: str, reference: str, require_shapes: bool = True,
                                require_types: bool = True) -> bool:
    """assert_h5_schema_compatible (:801-812): every dataset of `reference` exists in
    `candidate` with the same shape and element type."""
    cand = {r["path"]: r for r in h5_schema_signature(candidate)}
    ref = {r["path"]: r for r in h5_schema_signature(reference)}
    missing = sorted(set(ref) - set(cand))
    if missing:
        raise H5Error(f"HDF5 schema missing datasets: {missing}")
    bad = []
    for path, rr in ref.items():
        cr = cand[path]
        if require_shapes and cr["shape"] != rr["shape"]:
            bad.append(f"{path} shape {cr['shape']} != {rr['shape']}")
        if require_types and _normalize_dtype(cr["dtype"]) != _normalize_dtype(rr["dtype"]):
            bad.append(f"{path} dtype {cr['dtype']} != {rr['dtype']}")
    if bad:
        raise H5Error(f"HDF5 schema mismatches: {bad}")
    return True
