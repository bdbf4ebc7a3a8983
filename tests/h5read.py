"""Minimal NetCDF-4 (HDF5) reader for the tests, through libhdf5 and ctypes.

Reads root-group datasets (whole, converted to numpy), global attributes
(text or numeric) and the h5dump header of a file.  Only what the tests of
the host's NetCDF-4 output need; nothing of the product imports it.
"""
from __future__ import annotations

import ctypes
import os
import re
import shutil
import subprocess
from typing import Dict, Optional

import numpy as np

CANDIDATES = [os.environ.get("CENTRE_HDF5_LIB", ""), "libhdf5.so", "/opt/conda/lib/libhdf5.so"]
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"

_lib = None


def lib():
    """libhdf5, or None when it cannot be loaded here."""
    global _lib
    if _lib is None:
        for c in CANDIDATES:
            if not c:
                continue
            try:
                _lib = ctypes.CDLL(c)
                break
            except OSError:
                continue
        if _lib is None:
            return None
        L = _lib
        for name in ("H5Fopen", "H5Dopen2", "H5Dget_space", "H5Dget_type", "H5Aopen",
                     "H5Aget_space", "H5Aget_type", "H5Tget_native_type", "H5Tcopy"):
            getattr(L, name).restype = ctypes.c_int64
        L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, ctypes.c_int64]
        L.H5Dopen2.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64]
        L.H5Aopen.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64]
        L.H5Aexists.argtypes = [ctypes.c_int64, ctypes.c_char_p]
        for name in ("H5Dget_space", "H5Dget_type", "H5Aget_space", "H5Aget_type", "H5Tcopy",
                     "H5Fclose", "H5Dclose", "H5Aclose", "H5Sclose", "H5Tclose",
                     "H5Tget_class", "H5Tget_size", "H5Tget_sign", "H5Sget_simple_extent_ndims"):
            getattr(L, name).argtypes = [ctypes.c_int64]
        L.H5Tget_native_type.argtypes = [ctypes.c_int64, ctypes.c_int]
        L.H5Tget_size.restype = ctypes.c_size_t
        L.H5Sget_simple_extent_dims.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.H5Dread.argtypes = [ctypes.c_int64] * 5 + [ctypes.c_void_p]
        L.H5Aread.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p]
        L.H5open()
        L.H5Eset_auto2(ctypes.c_int64(0), None, None)
    return _lib


T_INTEGER, T_FLOAT, T_STRING = 0, 1, 3


def _dtype(L, tid) -> np.dtype:
    cls, size = L.H5Tget_class(tid), L.H5Tget_size(tid)
    if cls == T_INTEGER:
        signed = L.H5Tget_sign(tid) == 1
        return np.dtype(f"{'i' if signed else 'u'}{size}")
    if cls == T_FLOAT:
        return np.dtype(f"f{size}")
    raise ValueError(f"type class {cls} not supported")


def _dims(L, sid):
    nd = L.H5Sget_simple_extent_ndims(sid)
    if nd <= 0:
        return ()
    d = (ctypes.c_uint64 * nd)()
    L.H5Sget_simple_extent_dims(sid, d, None)
    return tuple(int(x) for x in d)


class File:
    def __init__(self, path: str):
        L = lib()
        if L is None:
            raise OSError("libhdf5 not found")
        self.L = L
        self.f = L.H5Fopen(path.encode(), 0, 0)
        if self.f < 0:
            raise OSError(f"{path}: not an HDF5 file")

    def close(self):
        if self.f >= 0:
            self.L.H5Fclose(self.f)
            self.f = -1

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def read(self, name: str) -> np.ndarray:
        L = self.L
        d = L.H5Dopen2(self.f, name.encode(), 0)
        if d < 0:
            raise KeyError(name)
        try:
            s, t = L.H5Dget_space(d), L.H5Dget_type(d)
            shape, dt = _dims(L, s), _dtype(L, t)
            nt = L.H5Tget_native_type(t, 0)
            out = np.empty(shape, dt)
            if out.size and L.H5Dread(d, nt, 0, 0, 0, out.ctypes.data) < 0:
                raise OSError(f"read of {name} failed")
            for h, c in ((nt, L.H5Tclose), (t, L.H5Tclose), (s, L.H5Sclose)):
                c(h)
            return out
        finally:
            L.H5Dclose(d)

    def attr(self, name: str):
        """A global attribute: str for text, numpy array otherwise."""
        L = self.L
        if L.H5Aexists(self.f, name.encode()) <= 0:
            raise KeyError(name)
        a = L.H5Aopen(self.f, name.encode(), 0)
        try:
            t, s = L.H5Aget_type(a), L.H5Aget_space(a)
            if L.H5Tget_class(t) == T_STRING:
                n = L.H5Tget_size(t)
                buf = ctypes.create_string_buffer(n + 1)
                L.H5Aread(a, t, buf)
                v = buf.raw[:n].split(b"\0", 1)[0].decode()
            else:
                shape, dt = _dims(L, s), _dtype(L, t)
                nt = L.H5Tget_native_type(t, 0)
                v = np.empty(shape or (1,), dt)
                L.H5Aread(a, nt, v.ctypes.data)
                L.H5Tclose(nt)
            L.H5Tclose(t)
            L.H5Sclose(s)
            return v
        finally:
            L.H5Aclose(a)


def is_hdf5(path: str) -> bool:
    with open(path, "rb") as f:
        return f.read(8) == b"\x89HDF\r\n\x1a\n"


def header(path: str) -> Optional[str]:
    """h5dump's header of a file (creation order, properties, attribute
    values), with object addresses, the file name and the _NCProperties
    provenance text taken out; None when h5dump is absent."""
    if not os.path.exists(H5DUMP):
        return None
    out = subprocess.run([H5DUMP, "-q", "creation_order", "-H", "-p", "-A", path],
                         check=True, capture_output=True, text=True).stdout
    lines = out.splitlines()[1:]
    out = "\n".join(lines)
    out = re.sub(r"DATASET [0-9]+ ", "DATASET ", out)
    out = re.sub(r'(ATTRIBUTE "_NCProperties" \{.*?STRSIZE )[0-9]+(;.*?\(0\): )"[^"]*"',
                 r'\1N\2"PROVENANCE"', out, flags=re.S)
    return out


def attrs(path: str, names) -> Dict[str, object]:
    with File(path) as f:
        return {n: f.attr(n) for n in names}


class _Attrs:
    def __init__(self, path: str):
        self.path = path

    def __getitem__(self, name: str):
        with File(self.path) as f:
            return f.attr(name)


class Dataset:
    """The read interface of ``centre_amd.cdf``'s classic datasets
    (``read(name)``, ``attrs[name]``) over a NetCDF-4 file."""

    def __init__(self, path: str):
        self.path = path
        self.attrs = _Attrs(path)

    def read(self, name: str) -> np.ndarray:
        with File(self.path) as f:
            return f.read(name)


def open_any(path: str):
    """A NetCDF file of either family, opened by content."""
    if is_hdf5(path):
        return Dataset(path)
    from centre_amd import cdf
    return cdf.open_dataset(path)


# ---- writing (test fixtures in the reference's NetCDF-4 BAT layout) ---------

_FILE_TYPES = {np.dtype("f4"): "H5T_IEEE_F32LE_g", np.dtype("f8"): "H5T_IEEE_F64LE_g",
               np.dtype("i4"): "H5T_STD_I32LE_g", np.dtype("u1"): "H5T_STD_U8LE_g",
               np.dtype("u4"): "H5T_STD_U32LE_g"}


def write_netcdf4(path: str, dims, variables, attrs: Dict[str, object]) -> None:
    """A NetCDF-4 file as netcdf-c lays it out: dims = [(name, length,
    unlimited)], variables = [(name, dim names, array, chunks or None)],
    global attrs (str or numpy arrays).  Dimensions become dimension scales
    (libhdf5_hl) with _Netcdf4Dimid; creation order is tracked."""
    L = lib()
    hl = ctypes.CDLL(os.path.join(os.path.dirname(L._name), "libhdf5_hl.so")
                     if os.path.isabs(L._name) else "libhdf5_hl.so")
    g = lambda n: ctypes.c_int64.in_dll(L, n).value  # noqa: E731
    i64 = ctypes.c_int64
    for n in ("H5Pcreate", "H5Fcreate", "H5Screate_simple", "H5Screate", "H5Dcreate2",
              "H5Acreate2", "H5Tcopy"):
        getattr(L, n).restype = i64
    L.H5Pcreate.argtypes = [i64]
    L.H5Fcreate.argtypes = [ctypes.c_char_p, ctypes.c_uint, i64, i64]
    L.H5Dcreate2.argtypes = [i64, ctypes.c_char_p, i64, i64, i64, i64, i64]
    L.H5Acreate2.argtypes = [i64, ctypes.c_char_p, i64, i64, i64, i64]
    L.H5Dwrite.argtypes = [i64] * 5 + [ctypes.c_void_p]
    L.H5Awrite.argtypes = [i64, i64, ctypes.c_void_p]
    L.H5Pset_chunk.argtypes = [i64, ctypes.c_int, ctypes.c_void_p]
    L.H5Screate_simple.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    L.H5Pset_link_creation_order.argtypes = [i64, ctypes.c_uint]
    L.H5Pset_attr_creation_order.argtypes = [i64, ctypes.c_uint]
    L.H5Tset_size.argtypes = [i64, ctypes.c_size_t]
    for n in ("H5Pclose", "H5Tset_strpad", "H5Tset_cset"):
        getattr(L, n).argtypes = [i64] if n == "H5Pclose" else [i64, ctypes.c_int]
    hl.H5DSset_scale.argtypes = [i64, ctypes.c_char_p]
    hl.H5DSattach_scale.argtypes = [i64, i64, ctypes.c_uint]
    unl = ctypes.c_uint64(2 ** 64 - 1).value

    def put_attr(loc, name, v):
        if isinstance(v, str):
            t = L.H5Tcopy(g("H5T_C_S1_g"))
            L.H5Tset_size(t, max(1, len(v)))
            L.H5Tset_strpad(t, 0)
            sp = L.H5Screate(0)
            a = L.H5Acreate2(loc, name.encode(), t, sp, 0, 0)
            L.H5Awrite(a, t, ctypes.c_char_p(v.encode() or b"\0"))
            L.H5Tclose(t)
        else:
            arr = np.ascontiguousarray(v)
            t = g(_FILE_TYPES[arr.dtype])
            n = (ctypes.c_uint64 * 1)(arr.size)
            sp = L.H5Screate_simple(1, n, None) if arr.ndim else L.H5Screate(0)
            a = L.H5Acreate2(loc, name.encode(), t, sp, 0, 0)
            L.H5Awrite(a, t, arr.ctypes.data)
        L.H5Aclose(a)
        L.H5Sclose(sp)

    fcpl = L.H5Pcreate(g("H5P_CLS_FILE_CREATE_ID_g"))
    L.H5Pset_link_creation_order(fcpl, 3)
    L.H5Pset_attr_creation_order(fcpl, 3)
    f = L.H5Fcreate(path.encode(), 2, fcpl, 0)
    L.H5Pclose(fcpl)
    if f < 0:
        raise OSError(f"cannot create {path}")
    for k, v in attrs.items():
        put_attr(f, k, v)
    scales, length = {}, {}
    for i, (name, n, is_unl) in enumerate(dims):
        dcpl = L.H5Pcreate(g("H5P_CLS_DATASET_CREATE_ID_g"))
        L.H5Pset_attr_creation_order(dcpl, 3)
        cur = (ctypes.c_uint64 * 1)(0 if is_unl else n)
        mx = (ctypes.c_uint64 * 1)(unl if is_unl else n)
        if is_unl:
            L.H5Pset_chunk(dcpl, 1, (ctypes.c_uint64 * 1)(1))
        sp = L.H5Screate_simple(1, cur, mx)
        # a variable named like the dimension but not its coordinate
        # variable (the BAT file's bond[frame][bond]): netcdf-c names the
        # dimension's scale _nc4_non_coord_<name>
        clash = any(v[0] == name for v in variables)
        sname = ("_nc4_non_coord_" + name) if clash else name
        d = L.H5Dcreate2(f, sname.encode(), g("H5T_IEEE_F32BE_g"), sp, 0, dcpl, 0)
        hl.H5DSset_scale(d, b"This is a netCDF dimension but not a netCDF variable.%10d" % cur[0])
        put_attr(d, "_Netcdf4Dimid", np.int32(i))
        L.H5Sclose(sp)
        L.H5Pclose(dcpl)
        scales[name], length[name] = d, (n, is_unl)
    for name, dn, arr, chunks in variables:
        arr = np.ascontiguousarray(arr)
        nd = len(dn)
        dcpl = L.H5Pcreate(g("H5P_CLS_DATASET_CREATE_ID_g"))
        L.H5Pset_attr_creation_order(dcpl, 3)
        has_unl = any(length[x][1] for x in dn)
        if chunks is None and has_unl:
            chunks = [1 if length[x][1] else length[x][0] for x in dn]
        if chunks is not None:
            L.H5Pset_chunk(dcpl, nd, (ctypes.c_uint64 * nd)(*chunks))
        cur = (ctypes.c_uint64 * nd)(*arr.shape)
        mx = (ctypes.c_uint64 * nd)(*[unl if length[x][1] else length[x][0] for x in dn])
        sp = L.H5Screate_simple(nd, cur, mx)
        t = g(_FILE_TYPES[arr.dtype])
        d = L.H5Dcreate2(f, name.encode(), t, sp, 0, dcpl, 0)
        if d < 0:
            raise OSError(f"cannot create variable {name}")
        if nd > 1:
            put_attr(d, "_Netcdf4Coordinates", np.array([[x for x, _, _ in dims].index(y) for y in dn],
                                                        np.int32))
        for k, x in enumerate(dn):
            hl.H5DSattach_scale(d, scales[x], k)
        if arr.size:
            L.H5Dwrite(d, t, 0, 0, 0, arr.ctypes.data)
        L.H5Dclose(d)
        L.H5Sclose(sp)
        L.H5Pclose(dcpl)
    for d in scales.values():
        L.H5Dclose(d)
    L.H5Fclose(f)
