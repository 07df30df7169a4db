"""netCDF classic (CDF-1) and 64-bit-offset (CDF-2) files: encoder and decoder.

Exodus II, the mesh and result format Goma reads (rd_mesh.c:152, rd_exo.c)
and writes (wr_exo.c:89, wr_soln.c:40) through SEACAS, is a set of naming
conventions over netCDF.  This module implements the published netCDF
classic-format grammar directly (header = magic, numrecs, dimension list,
global attribute list, variable list; big-endian data; every name, attribute
and variable padded to 4 bytes; record variables interleaved record by
record), so that the mesh/solution I/O needs no netCDF library.

Variables are given as ``Var(name, dims, data, attrs)``: ``dims`` names
dimensions of the file; a variable whose first dimension is the unlimited one
is a record variable and ``data`` then has the number of records as its first
axis.  Types follow ``data``'s dtype: int8 -> NC_BYTE, 'S1' -> NC_CHAR,
int16 -> NC_SHORT, int32 -> NC_INT, float32 -> NC_FLOAT, float64 -> NC_DOUBLE.
Attributes: ``str`` -> NC_CHAR, Python ``int`` -> NC_INT, Python ``float`` ->
NC_DOUBLE, numpy scalars and arrays by their dtype.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

NC_BYTE, NC_CHAR, NC_SHORT, NC_INT, NC_FLOAT, NC_DOUBLE = 1, 2, 3, 4, 5, 6
_TAG_DIM, _TAG_VAR, _TAG_ATT = 10, 11, 12
_DTYPE = {NC_BYTE: np.dtype(">i1"), NC_CHAR: np.dtype("S1"), NC_SHORT: np.dtype(">i2"),
          NC_INT: np.dtype(">i4"), NC_FLOAT: np.dtype(">f4"), NC_DOUBLE: np.dtype(">f8")}
_KIND = {("i", 1): NC_BYTE, ("S", 1): NC_CHAR, ("i", 2): NC_SHORT, ("i", 4): NC_INT,
         ("f", 4): NC_FLOAT, ("f", 8): NC_DOUBLE}


class NetCDFError(ValueError):
    pass


@dataclass
class Var:
    name: str
    dims: tuple
    data: np.ndarray
    attrs: dict = field(default_factory=dict)


@dataclass
class Dataset:
    dims: dict  # name -> length (the unlimited one holds the record count)
    unlimited: str | None
    attrs: dict
    variables: dict  # name -> Var


def _pad4(n: int) -> int:
    return (n + 3) & ~3


def _nc_type(dt: np.dtype) -> int:
    key = (dt.kind if dt.kind != "u" else "?", dt.itemsize)
    if key not in _KIND:
        raise NetCDFError(f"no netCDF classic type for dtype {dt}")
    return _KIND[key]


def _name(s: str) -> bytes:
    b = s.encode("utf-8")
    return struct.pack(">i", len(b)) + b + b"\0" * (_pad4(len(b)) - len(b))


def _att_values(v):
    if isinstance(v, str):
        return NC_CHAR, np.frombuffer(v.encode("utf-8"), dtype="S1")
    if isinstance(v, bool):
        raise NetCDFError("boolean attributes have no netCDF type")
    if isinstance(v, int):
        return NC_INT, np.array([v], dtype=">i4")
    if isinstance(v, float):
        return NC_DOUBLE, np.array([v], dtype=">f8")
    a = np.atleast_1d(np.asarray(v))
    t = _nc_type(a.dtype)
    return t, a.astype(_DTYPE[t])


def _att_list(attrs: dict) -> bytes:
    if not attrs:
        return b"\0" * 8
    out = [struct.pack(">ii", _TAG_ATT, len(attrs))]
    for k, v in attrs.items():
        t, vals = _att_values(v)
        raw = vals.tobytes()
        out.append(_name(k) + struct.pack(">ii", t, vals.size) + raw + b"\0" * (_pad4(len(raw)) - len(raw)))
    return b"".join(out)


def write(path, dims: dict, variables: list, attrs: dict | None = None, version: int = 2) -> None:
    """Write a netCDF file.  ``dims`` maps names to lengths; ``None`` marks the
    unlimited (record) dimension, whose length is taken from the record variables."""
    if version not in (1, 2):
        raise NetCDFError("version must be 1 (classic) or 2 (64-bit offset)")
    names = list(dims)
    unlimited = [k for k, n in dims.items() if n is None]
    if len(unlimited) > 1:
        raise NetCDFError("at most one unlimited dimension")
    ulim = unlimited[0] if unlimited else None
    seen = set()
    recs = None
    plan = []  # (var, nc_type, is_record, slab bytes)
    for v in variables:
        if v.name in seen:
            raise NetCDFError(f"duplicate variable {v.name}")
        seen.add(v.name)
        for d in v.dims:
            if d not in dims:
                raise NetCDFError(f"variable {v.name}: unknown dimension {d}")
        if ulim in v.dims[1:]:
            raise NetCDFError(f"variable {v.name}: the unlimited dimension must come first")
        a = np.asarray(v.data)
        t = _nc_type(a.dtype)
        is_rec = bool(v.dims) and v.dims[0] == ulim
        shape = tuple(dims[d] for d in v.dims[is_rec:])
        if is_rec:
            if recs is None:
                recs = a.shape[0]
            elif a.shape[0] != recs:
                raise NetCDFError(f"variable {v.name}: {a.shape[0]} records, others have {recs}")
            if a.shape[1:] != shape:
                raise NetCDFError(f"variable {v.name}: shape {a.shape} does not fit {v.dims}")
        elif a.shape != shape:
            raise NetCDFError(f"variable {v.name}: shape {a.shape} does not fit {v.dims}")
        slab = int(np.prod(shape, dtype=np.int64)) * _DTYPE[t].itemsize
        plan.append((v, t, is_rec, slab))
    recs = recs or 0
    nrec_vars = sum(1 for p in plan if p[2])
    off_fmt = ">i" if version == 1 else ">q"

    def header(begins):
        out = [b"CDF" + bytes([version]), struct.pack(">i", recs)]
        if names:
            out.append(struct.pack(">ii", _TAG_DIM, len(names)))
            for k in names:
                out.append(_name(k) + struct.pack(">i", 0 if dims[k] is None else dims[k]))
        else:
            out.append(b"\0" * 8)
        out.append(_att_list(attrs or {}))
        if plan:
            out.append(struct.pack(">ii", _TAG_VAR, len(plan)))
            for (v, t, is_rec, slab), b in zip(plan, begins):
                vsize = slab if (is_rec and nrec_vars == 1) else _pad4(slab)
                out.append(_name(v.name) + struct.pack(">i", len(v.dims))
                           + b"".join(struct.pack(">i", names.index(d)) for d in v.dims)
                           + _att_list(v.attrs) + struct.pack(">iI", t, vsize if vsize <= 2**32 - 4 else 2**32 - 1)
                           + struct.pack(off_fmt, b))
        else:
            out.append(b"\0" * 8)
        return b"".join(out)

    hlen = len(header([0] * len(plan)))
    begins, pos = [], hlen
    for v, t, is_rec, slab in plan:
        if not is_rec:
            begins.append(pos)
            pos += _pad4(slab)
        else:
            begins.append(None)
    rec_start = pos
    recsize = 0
    for i, (v, t, is_rec, slab) in enumerate(plan):
        if is_rec:
            begins[i] = rec_start + recsize
            recsize += slab if nrec_vars == 1 else _pad4(slab)
    if version == 1 and rec_start + recsize * max(recs, 1) >= 2**31:
        raise NetCDFError("file too large for the classic format; use version 2")
    with open(path, "wb") as f:
        f.write(header(begins))
        for v, t, is_rec, slab in plan:
            if not is_rec:
                raw = np.ascontiguousarray(v.data, dtype=_DTYPE[t]).tobytes()
                f.write(raw + b"\0" * (_pad4(slab) - slab))
        rv = [(np.ascontiguousarray(v.data, dtype=_DTYPE[t]), slab) for v, t, r, slab in plan if r]
        for r in range(recs):
            for a, slab in rv:
                raw = a[r:r + 1].tobytes()  # a slice keeps the big-endian layout
                f.write(raw + (b"" if nrec_vars == 1 else b"\0" * (_pad4(slab) - slab)))


class _Reader:
    def __init__(self, buf: memoryview):
        self.buf, self.pos = buf, 0

    def take(self, n: int) -> bytes:
        if self.pos + n > len(self.buf):
            raise NetCDFError("truncated netCDF header")
        b = bytes(self.buf[self.pos:self.pos + n])
        self.pos += n
        return b

    def i32(self) -> int:
        return struct.unpack(">i", self.take(4))[0]

    def name(self) -> str:
        n = self.i32()
        s = self.take(_pad4(n))[:n]
        return s.decode("utf-8")

    def atts(self) -> dict:
        tag, n = self.i32(), self.i32()
        if tag == 0 and n == 0:
            return {}
        if tag != _TAG_ATT:
            raise NetCDFError("bad attribute list tag")
        out = {}
        for _ in range(n):
            k = self.name()
            t, m = self.i32(), self.i32()
            if t not in _DTYPE:
                raise NetCDFError(f"attribute {k}: bad type {t}")
            nb = m * _DTYPE[t].itemsize
            raw = self.take(_pad4(nb))[:nb]
            if t == NC_CHAR:
                out[k] = raw.rstrip(b"\0").decode("utf-8", errors="replace")
            else:
                a = np.frombuffer(raw, dtype=_DTYPE[t]).astype(_DTYPE[t].newbyteorder("="))
                out[k] = a[0].item() if a.size == 1 else a
        return out


def _parse_header(buf):
    """Header of a classic / 64-bit-offset file: (version, recs, dnames,
    dlens, unlimited, global attributes, [(name, dims, attrs, type, vsize,
    begin)], recsize)."""
    r = _Reader(buf)
    magic = r.take(4)
    if magic[:3] != b"CDF" or magic[3] not in (1, 2):
        raise NetCDFError("not a netCDF classic or 64-bit-offset file (netCDF-4/HDF5 is not read)")
    version = magic[3]
    recs = r.i32()
    if recs < 0:
        raise NetCDFError("streaming record count is not supported")
    tag, n = r.i32(), r.i32()
    dnames, dlens = [], []
    if not (tag == 0 and n == 0):
        if tag != _TAG_DIM:
            raise NetCDFError("bad dimension list tag")
        for _ in range(n):
            dnames.append(r.name())
            dlens.append(r.i32())
            if dlens[-1] < 0:
                raise NetCDFError(f"dimension {dnames[-1]}: negative length")
    unlimited = None
    for k, ln in zip(dnames, dlens):
        if ln == 0:
            unlimited = k
    gatts = r.atts()
    tag, n = r.i32(), r.i32()
    raw_vars = []
    if not (tag == 0 and n == 0):
        if tag != _TAG_VAR:
            raise NetCDFError("bad variable list tag")
        for _ in range(n):
            name = r.name()
            nd = r.i32()
            ids = [r.i32() for _ in range(nd)]
            vatts = r.atts()
            t, vsize = r.i32(), r.i32()
            begin = struct.unpack(">i" if version == 1 else ">q", r.take(4 if version == 1 else 8))[0]
            if t not in _DTYPE:
                raise NetCDFError(f"variable {name}: bad type {t}")
            if any(i < 0 or i >= len(dnames) for i in ids):
                raise NetCDFError(f"variable {name}: dimension id out of range")
            if begin < 0:
                raise NetCDFError(f"variable {name}: negative data offset")
            raw_vars.append((name, [dnames[i] for i in ids], vatts, t, vsize, begin))
    nrec_vars = sum(1 for v in raw_vars if v[1] and v[1][0] == unlimited)
    recsize = 0
    for name, vd, vatts, t, vsize, begin in raw_vars:
        if vd and vd[0] == unlimited:
            slab = int(np.prod([dlens[dnames.index(d)] for d in vd[1:]], dtype=np.int64)) * _DTYPE[t].itemsize
            recsize += slab if nrec_vars == 1 else _pad4(slab)
    return version, recs, dnames, dlens, unlimited, gatts, raw_vars, recsize


def read(path, only=None) -> Dataset:
    """Read a netCDF classic or 64-bit-offset file: every variable, or those
    named in ``only``."""
    with open(path, "rb") as f:
        buf = memoryview(f.read())
    version, recs, dnames, dlens, unlimited, gatts, raw_vars, recsize = _parse_header(buf)
    dims = dict(zip(dnames, dlens))
    if unlimited is not None:
        dims[unlimited] = recs
    out = {}
    for name, vd, vatts, t, vsize, begin in raw_vars:
        if only is not None and name not in only:
            continue
        dt = _DTYPE[t]
        is_rec = bool(vd) and vd[0] == unlimited
        shape = [dims[d] for d in vd[is_rec:]]
        count = int(np.prod(shape, dtype=np.int64))
        if not is_rec:
            end = begin + count * dt.itemsize
            if end > len(buf):
                raise NetCDFError(f"variable {name}: data past the end of the file")
            a = np.frombuffer(buf[begin:end], dtype=dt).reshape(shape)
        else:
            a = np.empty([recs] + shape, dtype=dt)
            for k in range(recs):
                s = begin + k * recsize
                e = s + count * dt.itemsize
                if e > len(buf):
                    raise NetCDFError(f"variable {name}: record {k} past the end of the file")
                a[k] = np.frombuffer(buf[s:e], dtype=dt).reshape(shape)
        if t != NC_CHAR:
            a = a.astype(dt.newbyteorder("="))
        else:
            a = a.copy()
        out[name] = Var(name, tuple(vd), a, vatts)
    return Dataset(dims=dims, unlimited=unlimited, attrs=gatts, variables=out)


def append_record(path, values: dict) -> int:
    """Append one record to an existing file: ``values`` maps every record
    variable to its data for the new record (shape = the variable's shape
    without the record axis).  The record is written after the last one and
    the header's record count updated in place.  Returns the new count."""
    with open(path, "rb") as f:
        buf = memoryview(f.read())
    version, recs, dnames, dlens, unlimited, gatts, raw_vars, recsize = _parse_header(buf)
    size = len(buf)
    buf.release()
    rvars = [v for v in raw_vars if v[1] and v[1][0] == unlimited]
    if not rvars:
        raise NetCDFError("the file has no record variables")
    names = {v[0] for v in rvars}
    if set(values) != names:
        raise NetCDFError(f"a record needs exactly the record variables {sorted(names)}")
    rec_start = min(v[5] for v in rvars)
    if size < rec_start + recs * recsize:
        raise NetCDFError("the file is shorter than its records")
    pos0 = rec_start + recs * recsize
    single = len(rvars) == 1
    slabs = []  # every shape is checked before anything is written
    for name, vd, vatts, t, vsize, begin in rvars:
        shape = tuple(dlens[dnames.index(d)] for d in vd[1:])
        a = np.array(values[name], dtype=_DTYPE[t], order="C")  # keeps 0-d for scalar records
        if a.shape != shape:
            raise NetCDFError(f"variable {name}: shape {a.shape}, the record needs {shape}")
        raw = a.tobytes()
        slabs.append((pos0 + (begin - rec_start), raw + (b"" if single else b"\0" * (_pad4(len(raw)) - len(raw)))))
    with open(path, "r+b") as f:
        for off, raw in slabs:
            f.seek(off)
            f.write(raw)
        f.truncate(pos0 + recsize)
        f.seek(4)
        f.write(struct.pack(">i", recs + 1))
    return recs + 1


def chars(strings, width: int) -> np.ndarray:
    """Fixed-width NC_CHAR array [len(strings), width], NUL padded."""
    a = np.zeros((len(strings), width), dtype="S1")
    for i, s in enumerate(strings):
        b = s.encode("utf-8")[:width]
        a[i, :len(b)] = np.frombuffer(b, dtype="S1")
    return a


def strings(a: np.ndarray) -> list:
    """Inverse of ``chars``: rows of an NC_CHAR array as Python strings."""
    return [b"".join(row.tolist()).split(b"\0", 1)[0].decode("utf-8", errors="replace")
            for row in np.atleast_2d(a)]
