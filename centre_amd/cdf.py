"""NetCDF classic-family (CDF-1, CDF-2, CDF-5) reader and writer in numpy.

Written from the published NetCDF classic / 64-bit-offset / 64-bit-data
format specification (Unidata, "NetCDF File Format Specification").  The
reference links netcdf-c for all of its file I/O; this module provides the
subset CENTRE uses without that dependency:

* the BAT trajectory input (``Netcdf_BAT.cpp:24-132``: dims ``frame``,
  ``bond``, ``angle``, ``dihedral``; global ``convention == "CENTRE"``),
* the bin-code hand-off files ``bin_{bonds,angles,torsions}.nc``
  (``Netcdf_TrjInt.cpp:542-755``, created with ``NC_64BIT_DATA`` = CDF-5),
* the optional count files ``hist_*.nc`` (``Netcdf_HistUtils.cpp:808-1029``,
  CDF-5), and the entropy-contribution layout when written as CDF-5.

Everything is big-endian.  Record variables (leading dimension unlimited)
are interleaved record by record; each per-record slab is padded to a
4-byte boundary unless the file has exactly one record variable.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

NC_DIMENSION = 0x0A
NC_VARIABLE = 0x0B
NC_ATTRIBUTE = 0x0C

# nc_type -> numpy big-endian dtype
_NC2NP = {
    1: np.dtype(">i1"),   # NC_BYTE
    2: np.dtype("S1"),    # NC_CHAR
    3: np.dtype(">i2"),   # NC_SHORT
    4: np.dtype(">i4"),   # NC_INT
    5: np.dtype(">f4"),   # NC_FLOAT
    6: np.dtype(">f8"),   # NC_DOUBLE
    7: np.dtype(">u1"),   # NC_UBYTE
    8: np.dtype(">u2"),   # NC_USHORT
    9: np.dtype(">u4"),   # NC_UINT
    10: np.dtype(">i8"),  # NC_INT64
    11: np.dtype(">u8"),  # NC_UINT64
}
_NP2NC = {
    np.dtype("i1"): 1, np.dtype("i2"): 3, np.dtype("i4"): 4,
    np.dtype("f4"): 5, np.dtype("f8"): 6, np.dtype("u1"): 7,
    np.dtype("u2"): 8, np.dtype("u4"): 9, np.dtype("i8"): 10,
    np.dtype("u8"): 11,
}
NC_TYPE_BY_NAME = {
    "byte": 1, "char": 2, "short": 3, "int": 4, "float": 5, "double": 6,
    "ubyte": 7, "ushort": 8, "uint": 9, "int64": 10, "uint64": 11,
}


class CDFError(ValueError):
    """Malformed or unsupported classic-format file."""


def _pad4(n: int) -> int:
    return (n + 3) & ~3


@dataclass
class Variable:
    name: str
    dims: Tuple[str, ...]
    nc_type: int
    attrs: Dict[str, object]
    vsize: int
    begin: int
    is_record: bool
    shape: Tuple[int, ...] = ()


@dataclass
class Dataset:
    """An opened classic-format file (read-only view over its bytes)."""

    path: str
    version: int
    numrecs: int
    dims: Dict[str, int]
    unlimited: Optional[str]
    attrs: Dict[str, object]
    variables: Dict[str, Variable]
    _buf: memoryview = field(repr=False, default=None)
    _recsize: int = 0

    # -- data access -------------------------------------------------------
    def read(self, name: str, start: Optional[Sequence[int]] = None,
             count: Optional[Sequence[int]] = None) -> np.ndarray:
        """Read a variable (or a hyperslab of it) as a native-endian array."""
        v = self.variables[name]
        dt = _NC2NP[v.nc_type]
        shape = list(v.shape)
        if start is None:
            start = [0] * len(shape)
        if count is None:
            count = [s - o for s, o in zip(shape, start)]
        start = list(start)
        count = list(count)
        for s, c, n in zip(start, count, shape):
            if s < 0 or c < 0 or s + c > n:
                raise CDFError(f"hyperslab out of range for {name}: "
                               f"start={start} count={count} shape={shape}")
        if not v.is_record:
            arr = np.frombuffer(self._buf, dtype=dt,
                                count=int(np.prod(shape)) if shape else 1,
                                offset=v.begin).reshape(shape)
            sl = tuple(slice(s, s + c) for s, c in zip(start, count))
            out = arr[sl]
        else:
            inner = shape[1:]
            n_inner = int(np.prod(inner)) if inner else 1
            recs = []
            for r in range(start[0], start[0] + count[0]):
                off = v.begin + r * self._recsize
                a = np.frombuffer(self._buf, dtype=dt, count=n_inner,
                                  offset=off).reshape(inner)
                sl = tuple(slice(s, s + c) for s, c in zip(start[1:], count[1:]))
                recs.append(a[sl])
            out = (np.stack(recs) if recs
                   else np.zeros([0] + list(count[1:]), dtype=dt))
        if dt.kind == "S":
            return out.copy()
        return out.astype(dt.newbyteorder("="))


class _Reader:
    def __init__(self, buf: memoryview):
        self.b = buf
        self.p = 0
        self.version = 1

    def i32(self) -> int:
        v = struct.unpack_from(">i", self.b, self.p)[0]
        self.p += 4
        return v

    def nonneg(self) -> int:
        if self.version == 5:
            v = struct.unpack_from(">q", self.b, self.p)[0]
            self.p += 8
        else:
            v = struct.unpack_from(">i", self.b, self.p)[0]
            self.p += 4
        return v

    def offset(self) -> int:
        if self.version == 1:
            return self.i32()
        v = struct.unpack_from(">q", self.b, self.p)[0]
        self.p += 8
        return v

    def name(self) -> str:
        n = self.nonneg()
        s = bytes(self.b[self.p:self.p + n]).decode("utf-8")
        self.p += _pad4(n)
        return s

    def values(self, nc_type: int, n: int):
        dt = _NC2NP.get(nc_type)
        if dt is None:
            raise CDFError(f"unknown nc_type {nc_type}")
        nbytes = dt.itemsize * n
        raw = bytes(self.b[self.p:self.p + nbytes])
        self.p += _pad4(nbytes)
        if nc_type == 2:
            return raw.decode("latin-1")
        return np.frombuffer(raw, dtype=dt).astype(dt.newbyteorder("="))

    def att_list(self) -> Dict[str, object]:
        tag = self.i32()
        n = self.nonneg()
        if tag == 0 and n == 0:
            return {}
        if tag != NC_ATTRIBUTE:
            raise CDFError(f"expected NC_ATTRIBUTE tag, got {tag:#x}")
        out = {}
        for _ in range(n):
            nm = self.name()
            t = self.i32()
            k = self.nonneg()
            out[nm] = self.values(t, k)
        return out


def open_dataset(path: str) -> Dataset:
    """Parse the header of a CDF-1/2/5 file and map its bytes."""
    with open(path, "rb") as f:
        data = f.read()
    buf = memoryview(data)
    if len(data) < 4 or data[:3] != b"CDF":
        raise CDFError(f"{path}: not a NetCDF classic-format file "
                       f"(HDF5-based NetCDF-4 is not handled here)")
    r = _Reader(buf)
    r.version = data[3]
    if r.version not in (1, 2, 5):
        raise CDFError(f"{path}: unsupported CDF version {r.version}")
    r.p = 4
    numrecs = r.nonneg()
    if r.version != 5 and numrecs == -1:
        numrecs = -1  # streaming; resolved below
    # dims
    tag = r.i32()
    ndims = r.nonneg()
    dims: Dict[str, int] = {}
    dim_names: List[str] = []
    unlimited = None
    if not (tag == 0 and ndims == 0):
        if tag != NC_DIMENSION:
            raise CDFError(f"expected NC_DIMENSION tag, got {tag:#x}")
        for _ in range(ndims):
            nm = r.name()
            ln = r.nonneg()
            if ln == 0:
                unlimited = nm
            dims[nm] = ln
            dim_names.append(nm)
    gatts = r.att_list()
    tag = r.i32()
    nvars = r.nonneg()
    variables: Dict[str, Variable] = {}
    if not (tag == 0 and nvars == 0):
        if tag != NC_VARIABLE:
            raise CDFError(f"expected NC_VARIABLE tag, got {tag:#x}")
        for _ in range(nvars):
            nm = r.name()
            nd = r.nonneg()
            dimids = [r.nonneg() for _ in range(nd)]
            vatts = r.att_list()
            t = r.i32()
            vsize = r.nonneg()
            begin = r.offset()
            vd = tuple(dim_names[i] for i in dimids)
            is_rec = bool(vd) and vd[0] == unlimited
            variables[nm] = Variable(nm, vd, t, vatts, vsize, begin, is_rec)
    rec_vars = [v for v in variables.values() if v.is_record]
    if len(rec_vars) == 1:
        v = rec_vars[0]
        inner = [dims[d] for d in v.dims[1:]]
        recsize = _NC2NP[v.nc_type].itemsize * (int(np.prod(inner)) if inner else 1)
    else:
        recsize = sum(v.vsize for v in rec_vars)
    if numrecs < 0 and rec_vars:
        first = min(v.begin for v in rec_vars)
        numrecs = (len(data) - first) // max(recsize, 1)
    if unlimited is not None:
        dims[unlimited] = numrecs
    for v in variables.values():
        v.shape = tuple(dims[d] for d in v.dims)
    return Dataset(path, r.version, numrecs, dims, unlimited, gatts,
                   variables, buf, recsize)


# ---------------------------------------------------------------------------
# writer
# ---------------------------------------------------------------------------
class _Writer:
    def __init__(self, version: int):
        self.version = version
        self.parts: List[bytes] = []

    def i32(self, v: int):
        self.parts.append(struct.pack(">i", v))

    def nonneg(self, v: int):
        self.parts.append(struct.pack(">q" if self.version == 5 else ">i", v))

    def offset(self, v: int):
        self.parts.append(struct.pack(">i" if self.version == 1 else ">q", v))

    def name(self, s: str):
        b = s.encode("utf-8")
        self.nonneg(len(b))
        self.parts.append(b + b"\0" * (_pad4(len(b)) - len(b)))

    def att(self, name: str, value):
        self.name(name)
        if isinstance(value, (str, bytes)):
            b = value.encode("latin-1") if isinstance(value, str) else value
            self.i32(2)
            self.nonneg(len(b))
            self.parts.append(b + b"\0" * (_pad4(len(b)) - len(b)))
            return
        a = np.atleast_1d(np.asarray(value))
        t = _NP2NC[a.dtype.newbyteorder("=") if a.dtype.byteorder not in "=|" else a.dtype]
        raw = a.astype(_NC2NP[t]).tobytes()
        self.i32(t)
        self.nonneg(a.size)
        self.parts.append(raw + b"\0" * (_pad4(len(raw)) - len(raw)))

    def att_list(self, attrs: Dict[str, object]):
        if not attrs:
            self.i32(0)
            self.nonneg(0)
            return
        self.i32(NC_ATTRIBUTE)
        self.nonneg(len(attrs))
        for k, v in attrs.items():
            self.att(k, v)

    def data(self) -> bytes:
        return b"".join(self.parts)


@dataclass
class VarSpec:
    name: str
    dims: Tuple[str, ...]
    data: np.ndarray
    attrs: Dict[str, object] = field(default_factory=dict)


def write_dataset(path: str, dims: Sequence[Tuple[str, int]],
                  variables: Sequence[VarSpec],
                  attrs: Optional[Dict[str, object]] = None,
                  version: int = 5, unlimited: Optional[str] = None) -> None:
    """Write a complete classic-format file.

    ``dims`` is an ordered list of ``(name, length)``; the unlimited
    dimension is named by ``unlimited`` and its length is the record count.
    Each ``VarSpec.data`` must already have the full shape of its dims.
    """
    attrs = attrs or {}
    dim_index = {n: i for i, (n, _) in enumerate(dims)}
    dim_len = dict(dims)
    numrecs = dim_len[unlimited] if unlimited else 0
    recvars = [v for v in variables if v.dims and v.dims[0] == unlimited]
    fixvars = [v for v in variables if not (v.dims and v.dims[0] == unlimited)]
    for v in variables:
        exp = tuple(dim_len[d] for d in v.dims)
        if tuple(v.data.shape) != exp:
            raise CDFError(f"{v.name}: data shape {v.data.shape} != dims {exp}")

    def vsize_of(v: VarSpec) -> int:
        dt = _NC2NP[_NP2NC[np.dtype(v.data.dtype.str.lstrip("<>=|"))]]
        shp = [dim_len[d] for d in v.dims]
        if v.dims and v.dims[0] == unlimited:
            shp = shp[1:]
        n = int(np.prod(shp)) if shp else 1
        return _pad4(n * dt.itemsize)

    def header(begins: Dict[str, int]) -> bytes:
        w = _Writer(version)
        w.parts.append(b"CDF" + bytes([version]))
        w.nonneg(numrecs)
        if dims:
            w.i32(NC_DIMENSION)
            w.nonneg(len(dims))
            for n, ln in dims:
                w.name(n)
                w.nonneg(0 if n == unlimited else ln)
        else:
            w.i32(0)
            w.nonneg(0)
        w.att_list(attrs)
        if variables:
            w.i32(NC_VARIABLE)
            w.nonneg(len(variables))
            for v in fixvars + recvars:
                w.name(v.name)
                w.nonneg(len(v.dims))
                for d in v.dims:
                    w.nonneg(dim_index[d])
                w.att_list(v.attrs)
                w.i32(_NP2NC[np.dtype(v.data.dtype.str.lstrip("<>=|"))])
                w.nonneg(vsize_of(v))
                w.offset(begins.get(v.name, 0))
        else:
            w.i32(0)
            w.nonneg(0)
        return w.data()

    hlen = len(header({}))
    begins: Dict[str, int] = {}
    pos = hlen
    for v in fixvars:
        begins[v.name] = pos
        pos += vsize_of(v)
    for v in recvars:
        begins[v.name] = pos
        pos += vsize_of(v)
    hdr = header(begins)
    assert len(hdr) == hlen
    body: List[bytes] = []
    for v in fixvars:
        dt = _NC2NP[_NP2NC[np.dtype(v.data.dtype.str.lstrip("<>=|"))]]
        raw = np.ascontiguousarray(v.data).astype(dt).tobytes()
        body.append(raw + b"\0" * (vsize_of(v) - len(raw)))
    single = len(recvars) == 1
    for r in range(numrecs):
        for v in recvars:
            dt = _NC2NP[_NP2NC[np.dtype(v.data.dtype.str.lstrip("<>=|"))]]
            raw = np.ascontiguousarray(v.data[r]).astype(dt).tobytes()
            pad = 0 if single else vsize_of(v) - len(raw)
            body.append(raw + b"\0" * pad)
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(b"".join(body))
