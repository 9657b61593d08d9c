"""JLD2 checkpoints (rcmps_store schema 1, src/storage.jl:4-38) read and written without h5py.

JLD2 files are HDF5 files written by Julia's JLD2 package (here "HDF5-based Julia Data Format,
version 0.2.0").  This module reads the subset of the HDF5 file format that such a checkpoint
uses, following the published HDF5 File Format Specification (version 3.0):

- superblock version 2 or 3 (after JLD2's 512-byte text preamble), 8-byte offsets and lengths;
- version-2 object headers ("OHDR") with continuation blocks ("OCHK");
- hard links held as link messages in the root group's header;
- dataspace messages (versions 1 and 2), layout messages (versions 3 and 4: compact and
  contiguous storage), datatype messages of classes fixed-point, floating-point, fixed-length
  string and compound (version 3 member encoding), inline or shared (committed in the file).

read_jld2 returns a StateRecord like the reference GPU package's reader (gpu/src/rcmps_gpu/io.py:
261-298): R, K, gradient as row-major numpy arrays (HDF5 stores Julia's column-major arrays with
reversed dimensions), e0 (None when not finite), model, coupling, integration and manifold.

write_jld2 writes a checkpoint in the layout JLD2 0.2 gives it (object order, committed Julia
types with their julia_type / field_names / field_types attributes, global heap, checksums), so
that Julia's rcmps_load reads it: rewriting the reference's three checkpoints from their own
contents reproduces each file byte for byte after the preamble (which names the Julia version).
"""
from __future__ import annotations

import math
import struct
from pathlib import Path

import numpy as np

from .io import Integration, Manifold, StateRecord

SIGNATURE = b"\x89HDF\r\n\x1a\n"


class JLD2Error(ValueError):
    pass


class _H5:
    def __init__(self, data: bytes):
        self.data = data
        start = data.find(SIGNATURE)
        if start < 0:
            raise JLD2Error("no HDF5 superblock")
        version = data[start + 8]
        if version not in (2, 3):
            raise JLD2Error(f"superblock version {version} is not supported")
        if data[start + 9] != 8 or data[start + 10] != 8:
            raise JLD2Error("only 8-byte offsets and lengths are supported")
        self.base, _ext, _eof, self.root = struct.unpack_from("<QQQQ", data, start + 12)

    def u(self, fmt, off):
        return struct.unpack_from("<" + fmt, self.data, off)

    def header(self, addr):
        """Messages (type, flags, body) of the version-2 object header at `addr`."""
        d = self.data
        off = self.base + addr
        if d[off:off + 4] != b"OHDR" or d[off + 4] != 2:
            raise JLD2Error(f"object header at {addr} is not a version-2 header")
        flags = d[off + 5]
        p = off + 6 + (16 if flags & 0x20 else 0) + (4 if flags & 0x10 else 0)
        nb = 1 << (flags & 3)
        size = int.from_bytes(d[p:p + nb], "little")
        p += nb
        msgs = []
        blocks = [(p, p + size)]
        while blocks:
            p, end = blocks.pop(0)
            while p + 4 <= end:
                mtype, msize, mflags = d[p], self.u("H", p + 1)[0], d[p + 3]
                p += 4 + (2 if flags & 0x04 else 0)
                body = d[p:p + msize]
                if mtype == 0x10:   # continuation: address, length of an "OCHK" block
                    caddr, clen = struct.unpack_from("<QQ", body, 0)
                    cp = self.base + caddr
                    if d[cp:cp + 4] != b"OCHK":
                        raise JLD2Error("bad continuation block")
                    blocks.append((cp + 4, cp + clen - 4))   # the last 4 bytes are a checksum
                msgs.append((mtype, mflags, body))
                p += msize
        return msgs

    def links(self, addr):
        out = {}
        for mtype, _f, b in self.header(addr):
            if mtype != 0x06:
                continue
            fl = b[1]
            p = 2
            if fl & 0x08:
                if b[p] != 0:
                    raise JLD2Error("only hard links are supported")
                p += 1
            if fl & 0x04:
                p += 8
            if fl & 0x10:
                p += 1
            ln = 1 << (fl & 3)
            n = int.from_bytes(b[p:p + ln], "little")
            p += ln
            name = b[p:p + n].decode()
            p += n
            out[name] = struct.unpack_from("<Q", b, p)[0]
        return out

    # ---- datatypes -------------------------------------------------------------------------
    def datatype(self, b, off=0):
        """(descriptor, bytes consumed) of the datatype encoded at b[off:]."""
        cls, ver = b[off] & 0x0F, b[off] >> 4
        f0, f1, f2 = b[off + 1], b[off + 2], b[off + 3]
        size = struct.unpack_from("<I", b, off + 4)[0]
        p = off + 8
        if cls == 0:
            if f0 & 0x01:
                raise JLD2Error("big-endian integers are not supported")
            return ("int", size, bool(f0 & 0x08)), p + 4 - off
        if cls == 1:
            if f0 & 0x01:
                raise JLD2Error("big-endian floats are not supported")
            return ("float", size), p + 12 - off
        if cls == 3:
            return ("str", size), p - off
        if cls == 6:
            if ver != 3:
                raise JLD2Error(f"compound datatype version {ver} is not supported")
            nmem = f0 | (f1 << 8)
            width = 1 if size < 1 << 8 else 2 if size < 1 << 16 else 3 if size < 1 << 24 else 4
            members = []
            for _ in range(nmem):
                e = b.index(b"\x00", p)
                name = b[p:e].decode()
                p = e + 1
                moff = int.from_bytes(b[p:p + width], "little")
                p += width
                mt, used = self.datatype(b, p)
                p += used
                members.append((name, moff, mt))
            return ("compound", size, members), p - off
        raise JLD2Error(f"datatype class {cls} is not supported")

    def dataset_type(self, mflags, body):
        if mflags & 0x02:   # shared: a datatype committed in the file
            if body[0] != 3 or body[1] != 2:
                raise JLD2Error("unsupported shared-message encoding")
            addr = struct.unpack_from("<Q", body, 2)[0]
            for t, _f, b in self.header(addr):
                if t == 0x03:
                    return self.datatype(b)[0]
            raise JLD2Error("committed datatype without a datatype message")
        return self.datatype(body)[0]

    # ---- datasets --------------------------------------------------------------------------
    def dataset(self, addr):
        dims, dtype, raw = None, None, None
        for t, f, b in self.header(addr):
            if t == 0x01:
                ver, rank = b[0], b[1]
                if ver == 1:
                    dims = struct.unpack_from(f"<{rank}Q", b, 8)
                elif ver == 2:
                    dims = () if b[3] == 0 else struct.unpack_from(f"<{rank}Q", b, 4)
                else:
                    raise JLD2Error(f"dataspace version {ver} is not supported")
            elif t == 0x03:
                dtype = self.dataset_type(f, b)
            elif t == 0x08:
                ver, cls = b[0], b[1]
                if ver not in (3, 4):
                    raise JLD2Error(f"layout version {ver} is not supported")
                if cls == 0:
                    n = struct.unpack_from("<H", b, 2)[0]
                    raw = b[4:4 + n]
                elif cls == 1:
                    a, n = struct.unpack_from("<QQ", b, 2)
                    raw = self.data[self.base + a:self.base + a + n] if a != (1 << 64) - 1 else b""
                else:
                    raise JLD2Error("chunked or virtual storage is not supported")
        if dims is None or dtype is None or raw is None:
            raise JLD2Error(f"object at {addr} is not a dataset")
        return _decode(raw, dtype, dims)


def _np_dtype(dt):
    kind = dt[0]
    if kind == "int":
        return np.dtype(f"<{'i' if dt[2] else 'u'}{dt[1]}")
    if kind == "float":
        return np.dtype(f"<f{dt[1]}")
    if kind == "str":
        return np.dtype(f"S{dt[1]}")
    names = [m[0] for m in dt[2]]
    return np.dtype({"names": names, "formats": [_np_dtype(m[2]) for m in dt[2]],
                     "offsets": [m[1] for m in dt[2]], "itemsize": dt[1]})


def _is_complex(dt):
    return dt[0] == "compound" and [m[0] for m in dt[2]] == ["re", "im"] and all(m[2][0] == "float" for m in dt[2])


def _decode(raw, dt, dims):
    count = int(np.prod(dims)) if dims else 1
    nd = _np_dtype(dt)
    arr = np.frombuffer(raw, dtype=nd, count=count)
    if _is_complex(dt):
        arr = arr["re"].astype(np.float64) + 1j * arr["im"].astype(np.float64)
    if dt[0] == "str":
        return arr[0].rstrip(b"\x00").decode()
    if not dims:
        return arr[0]
    # HDF5 dims are Julia's reversed: the C-order array is the transpose of the Julia matrix
    return np.asarray(arr).reshape(dims).T


def read_jld2_datasets(path) -> dict:
    """Top-level datasets of a JLD2 file by name (groups such as _types are skipped)."""
    h = _H5(Path(path).read_bytes())
    out = {}
    for name, addr in h.links(h.root).items():
        if name.startswith("_"):
            continue
        out[name] = h.dataset(addr)
    return out


def read_jld2(path) -> StateRecord:
    """An rcmps_store checkpoint (schema 1) as a StateRecord; gradient and energy are None when e0
    is not finite (a state never evaluated)."""
    f = read_jld2_datasets(path)
    if int(f.get("schema_version", 0)) != 1:
        raise ValueError("unsupported schema; migrate old JLD with examples/migrate_legacy.jl")
    R = np.ascontiguousarray(f["R"], dtype=np.complex128)
    K = np.ascontiguousarray(f["K"], dtype=np.complex128)
    if R.ndim != 2 or R.shape[0] != R.shape[1] or R.shape != K.shape or R.shape[0] == 0:
        raise ValueError("R, K must be nonempty, square and of equal size")
    coupling = complex(f["coupling"])
    if coupling.imag != 0:
        raise ValueError("model coupling must be real")
    e0 = float(f["e0"])
    evaluated = math.isfinite(e0)
    g = f.get("grad")
    row = f["integration"]
    integration = Integration(int(row["n_quad"]), float(row["max_point"]), float(row["atol"]), float(row["rtol"]),
                              int(row["maxiters"]) if "maxiters" in row.dtype.names else None)
    row = f.get("manifold")
    manifold = None if row is None else Manifold(
        float(row["metric_reg"]), float(row["precond_reg"]),
        float(row["precond_reg_prefactor"]) if "precond_reg_prefactor" in row.dtype.names else None)
    return StateRecord(D=R.shape[0], model=str(f["model"]), coupling=coupling.real, R=R, K=K,
                       mass=float(f.get("mass", 1.0)),
                       gradient=np.ascontiguousarray(g, dtype=np.complex128) if evaluated and g is not None else None,
                       energy=e0 if evaluated else None, integration=integration, manifold=manifold,
                       info=str(f["info"]) if "info" in f else None, source=str(path))


# =============================================================================================
# Writer: the rcmps_store schema-1 checkpoint (src/storage.jl:4-24) in the layout JLD2 0.2 writes
# for it, so that Julia's rcmps_load reads it back.  Objects follow one another in the order JLD2
# creates them (its 512-byte preamble, a version-2 superblock, then version-2 object headers with
# lookup3 checksums, one 4096-byte global heap collection for variable-length data, contiguous data
# blocks 8-byte aligned):
#   schema_version (Int64), R and K (Matrix{ComplexF64}: the committed DataType and Complex types,
#   the Float64 parameter stored as a DataType value), model (String), coupling (ComplexF64), e0
#   (Float64), grad, info (String), integration and manifold (NamedTuples: committed NamedTuple,
#   Symbol, Tuple{Symbol...} and Tuple{...} types with their julia_type / field_names /
#   field_types attributes), the _types group of the nine committed types, the root group.
# Addresses are relative to the superblock at byte 512; the superblock's end-of-file field holds the
# file length.
# =============================================================================================
PREAMBLE = b"HDF5-based Julia Data Format, version 0.2.0\x00 (Julia 1.12.5 64-bit LE)"
_BASE = 512
_GCOL_SIZE = 4096
_UNDEF = 0xFFFFFFFFFFFFFFFF


def lookup3(data: bytes, initval: int = 0) -> int:
    """Bob Jenkins' lookup3 hashlittle (2006), the checksum of HDF5 version-2 metadata."""
    mask = 0xFFFFFFFF

    def rot(x, k):
        return ((x << k) | (x >> (32 - k))) & mask

    n = len(data)
    a = b = c = (0xDEADBEEF + n + initval) & mask
    i = 0
    while n > 12:
        a = (a + int.from_bytes(data[i:i + 4], "little")) & mask
        b = (b + int.from_bytes(data[i + 4:i + 8], "little")) & mask
        c = (c + int.from_bytes(data[i + 8:i + 12], "little")) & mask
        a = (a - c) & mask; a ^= rot(c, 4); c = (c + b) & mask
        b = (b - a) & mask; b ^= rot(a, 6); a = (a + c) & mask
        c = (c - b) & mask; c ^= rot(b, 8); b = (b + a) & mask
        a = (a - c) & mask; a ^= rot(c, 16); c = (c + b) & mask
        b = (b - a) & mask; b ^= rot(a, 19); a = (a + c) & mask
        c = (c - b) & mask; c ^= rot(b, 4); b = (b + a) & mask
        n -= 12
        i += 12
    if n == 0:
        return c
    tail = data[i:] + bytes(12 - n)
    a = (a + int.from_bytes(tail[0:4], "little")) & mask
    b = (b + int.from_bytes(tail[4:8], "little")) & mask
    c = (c + int.from_bytes(tail[8:12], "little")) & mask
    c ^= b; c = (c - rot(b, 14)) & mask
    a ^= c; a = (a - rot(c, 11)) & mask
    b ^= a; b = (b - rot(a, 25)) & mask
    c ^= b; c = (c - rot(b, 16)) & mask
    a ^= c; a = (a - rot(c, 4)) & mask
    b ^= a; b = (b - rot(a, 14)) & mask
    c ^= b; c = (c - rot(b, 24)) & mask
    return c


# ---- datatype encodings (HDF5 datatype message, version 3) ----------------------------------
_T_I64 = bytes.fromhex("3008000008000000" "00004000")
_T_F64 = bytes.fromhex("31203f0008000000" "000040" "00340b0034ff030000")
_T_SYMBOL = bytes.fromhex("3911010010000000" "300000000100000000000800")   # vlen UTF-8 string
_T_REF = bytes.fromhex("3700000008000000")                                  # object reference


def _t_string(n: int) -> bytes:
    return bytes.fromhex("33110000") + struct.pack("<I", n)


def _t_compound(size: int, members) -> bytes:
    """Version-3 compound: member name (null-terminated), offset in the smallest width, type."""
    width = 1 if size < 1 << 8 else 2 if size < 1 << 16 else 3 if size < 1 << 24 else 4
    out = bytes([0x36]) + struct.pack("<H", len(members)) + b"\x00" + struct.pack("<I", size)
    for name, off, t in members:
        out += name.encode() + b"\x00" + off.to_bytes(width, "little") + t
    return out


_T_DATATYPE = _t_compound(32, [("name", 0, _T_SYMBOL),
                               ("parameters", 16, bytes.fromhex("3900000010000000") + _T_REF)])
_T_COMPLEX = _t_compound(16, [("re", 0, _T_F64), ("im", 8, _T_F64)])


def _shared(addr: int) -> bytes:
    return bytes([3, 2]) + struct.pack("<Q", addr)


_SCALAR = bytes.fromhex("02000000")


def _simple(*dims) -> bytes:
    return bytes([2, len(dims), 0, 1]) + b"".join(struct.pack("<Q", d) for d in dims)


class _Writer:
    def __init__(self):
        self.buf = bytearray()          # content at relative addresses 0..
        self.gcol = None                # relative address of the global heap collection
        self.gobjs = []                 # its objects (bytes)
        self.types = []                 # (name, address) of committed datatypes

    # -- space ---------------------------------------------------------------------------
    def at(self, addr: int, data: bytes):
        end = addr + len(data)
        if len(self.buf) < end:
            self.buf.extend(bytes(end - len(self.buf)))
        self.buf[addr:end] = data

    def cursor(self) -> int:
        return len(self.buf)

    def reserve(self, n: int) -> int:
        a = self.cursor()
        self.buf.extend(bytes(n))
        return a

    def align8(self):
        self.reserve((-self.cursor()) % 8)

    # -- global heap -----------------------------------------------------------------------
    def heap(self, payload: bytes) -> bytes:
        """Store payload as the next global heap object; its 12-byte heap ID (collection, index)."""
        if self.gcol is None:
            self.align8()
            self.gcol = self.reserve(_GCOL_SIZE + 16)
        self.gobjs.append(payload)
        return struct.pack("<QI", self.gcol, len(self.gobjs))

    def vlen(self, payload: bytes, count: int) -> bytes:
        return struct.pack("<I", count) + self.heap(payload)

    def symbol(self, text: str) -> bytes:
        b = text.encode()
        return self.vlen(b, len(b))

    def refs(self, addrs) -> bytes:
        if not addrs:
            return bytes(16)
        return self.vlen(b"".join(struct.pack("<Q", a) for a in addrs), len(addrs))

    def flush_heap(self):
        if self.gcol is None:
            return
        body = bytearray()
        for i, obj in enumerate(self.gobjs, 1):
            body += struct.pack("<HH4xQ", i, 1, len(obj)) + obj + bytes((-len(obj)) % 8)
        free = _GCOL_SIZE - 16 - len(body)
        body += struct.pack("<HH4xQ", 0, 0, free)
        self.at(self.gcol, b"GCOL\x01\x00\x00\x00" + struct.pack("<Q", _GCOL_SIZE) + bytes(body))

    # -- object headers -------------------------------------------------------------------
    @staticmethod
    def header(messages) -> bytes:
        """Version-2 object header of (type, flags, body) messages, with its checksum."""
        body = b"".join(bytes([t]) + struct.pack("<H", len(b)) + bytes([f]) + b for t, f, b in messages)
        flags = 0 if len(body) < 256 else 1
        head = b"OHDR\x02" + bytes([flags]) + len(body).to_bytes(1 << flags, "little") + body
        return head + struct.pack("<I", lookup3(head))

    @staticmethod
    def attribute(name: str, dtype: bytes, space: bytes, data: bytes, shared: bool) -> tuple:
        nm = name.encode() + b"\x00"
        body = bytes([2, 1 if shared else 0]) + struct.pack("<HHH", len(nm), len(dtype), len(space)) + nm + dtype + space
        return (0x0C, 0, body + data)

    def dataset(self, space: bytes, dtype: bytes, shared: bool, layout: bytes) -> int:
        msgs = [(0x05, 0, b"\x03\x09"), (0x01, 0, space), (0x03, 3 if shared else 1, dtype), (0x08, 0, layout),
                (0x00, 0, bytes(16))]
        h = self.header(msgs)
        a = self.reserve(len(h))
        self.at(a, h)
        return a

    def compact(self, space: bytes, dtype: bytes, shared: bool, data: bytes) -> int:
        return self.dataset(space, dtype, shared, b"\x04\x00" + struct.pack("<H", len(data)) + data)

    def contiguous(self, space: bytes, dtype: bytes, shared: bool, data: bytes) -> int:
        size = len(self.header([(0x05, 0, b""), (0x01, 0, space), (0x03, 0, dtype), (0x08, 0, bytes(18)),
                                (0x00, 0, bytes(16))])) + 2
        hdr = self.cursor()
        start = hdr + size
        start += (-start) % 8
        a = self.dataset(space, dtype, shared, b"\x04\x01" + struct.pack("<QQ", start, len(data)))
        assert a == hdr
        self.align8()
        assert self.cursor() == start
        self.at(start, data)
        return a

    def committed(self, dtype: bytes, build) -> int:
        """Reserve the header of a committed datatype at the cursor, run build(address) -> list of
        attribute messages (which may write further objects after the header), then fill it in."""
        a = self.cursor()
        # the header's size is known before its attribute values are: they are fixed-size records
        probe = build(None)
        h_len = len(self.header([(0x03, 0x40, dtype)] + probe))
        self.reserve(h_len)
        msgs = build(a)
        h = self.header([(0x03, 0x40, dtype)] + msgs)
        assert len(h) == h_len
        self.at(a, h)
        return a

    def group(self, links) -> int:
        msgs = [(0x02, 0, b"\x00\x00" + struct.pack("<QQ", _UNDEF, _UNDEF)), (0x0A, 0, b"\x00\x00")]
        for name, addr in links:
            nm = name.encode()
            msgs.append((0x06, 0, bytes([1, 0x10, 1, len(nm)]) + nm + struct.pack("<Q", addr)))
        msgs.append((0x00, 0, bytes(16)))
        h = self.header(msgs)
        a = self.reserve(len(h))
        self.at(a, h)
        return a


class _Types:
    """The committed datatypes of a schema-1 checkpoint, created on first use in JLD2's order."""

    def __init__(self, w: _Writer):
        self.w = w
        self.datatype = None
        self.complex = None
        self.symbol = None

    def _julia_type(self, name: str, params, target):
        """julia_type attribute: a DataType record (name, parameter references) whose type is the
        committed DataType type (itself, for the DataType type)."""
        w = self.w
        if target is None:   # size probe: the record is 32 bytes whatever it holds
            return self.w.attribute("julia_type", _shared(0), _SCALAR, bytes(32), True)
        rec = w.symbol(name) + w.refs(params() if callable(params) else params)
        return w.attribute("julia_type", _shared(self.datatype if self.datatype is not None else target),
                           _SCALAR, rec, True)

    def _datatype_value(self, name: str) -> int:
        """A DataType value (e.g. Core.Float64 as a type parameter) stored as a scalar dataset."""
        w = self.w
        return w.compact(_SCALAR, _shared(self.datatype), True, w.symbol(name) + bytes(16))

    def ensure_datatype(self):
        if self.datatype is None:
            w = self.w

            def build(a):
                return [self._julia_type("Core.DataType", [], a)]
            self.datatype = w.committed(_T_DATATYPE, build)
            w.types.append(self.datatype)

    def ensure_complex(self) -> int:
        self.ensure_datatype()
        if self.complex is None:
            w = self.w

            def build(a):
                if a is None:
                    return [self._julia_type("", [], None)]
                name = w.symbol("Base.Complex")
                param = self._datatype_value("Core.Float64")
                rec = name + w.refs([param])
                return [w.attribute("julia_type", _shared(self.datatype), _SCALAR, rec, True)]
            self.complex = w.committed(_T_COMPLEX, build)
            w.types.append(self.complex)
        return self.complex

    def ensure_symbol(self) -> int:
        if self.symbol is None:
            w = self.w
            self.symbol = w.committed(_T_SYMBOL, lambda a: [self._julia_type("Core.Symbol", [], a)])
            w.types.append(self.symbol)
        return self.symbol

    def symbol_tuple(self, n: int) -> int:
        """Tuple{Symbol, ..., Symbol} (n fields "1".."n")."""
        w = self.w
        sym = self.ensure_symbol()
        dtype = _t_compound(16 * n, [(str(i + 1), 16 * i, _T_SYMBOL) for i in range(n)])

        def build(a):
            if a is None:
                return [self._julia_type("", [], None),
                        w.attribute("field_names", _T_SYMBOL, _simple(n), bytes(16 * n), False),
                        w.attribute("field_types", _T_REF, _simple(n), bytes(8 * n), False)]
            jt = self._julia_type("Tuple", [sym] * n, a)
            names = b"".join(w.symbol(str(i + 1)) for i in range(n))
            return [jt, w.attribute("field_names", _T_SYMBOL, _simple(n), names, False),
                    w.attribute("field_types", _T_REF, _simple(n), struct.pack(f"<{n}Q", *([sym] * n)), False)]
        a = w.committed(dtype, build)
        w.types.append(a)
        return a

    def value_tuple(self, kinds, julia_name: str) -> int:
        """Tuple{kinds...} of Int64 ("i") / Float64 ("f") fields "1".."n", each parameter a DataType
        value written after the type (one per distinct kind, in first-use order)."""
        w = self.w
        members = [(str(i + 1), 8 * i, _T_I64 if k == "i" else _T_F64) for i, k in enumerate(kinds)]
        dtype = _t_compound(8 * len(kinds), members)

        def build(a):
            if a is None:
                return [self._julia_type("", [], None)]
            name = w.symbol(julia_name)
            seen = {}
            for k in kinds:
                if k not in seen:
                    seen[k] = self._datatype_value("Core.Int64" if k == "i" else "Core.Float64")
            rec = name + w.refs([seen[k] for k in kinds])
            return [w.attribute("julia_type", _shared(self.datatype), _SCALAR, rec, True)]
        a = w.committed(dtype, build)
        w.types.append(a)
        return a

    def named_tuple(self, fields_, tuple_name: str) -> int:
        """NamedTuple{names, Tuple{types}} for fields (name, kind); returns the committed type."""
        w = self.w
        n = len(fields_)
        members = [(nm, 8 * i, _T_I64 if k == "i" else _T_F64) for i, (nm, k) in enumerate(fields_)]
        dtype = _t_compound(8 * n, members)

        def build(a):
            if a is None:
                return [self._julia_type("", [], None)]
            name = w.symbol("Core.NamedTuple")
            names_type = self.symbol_tuple(n)
            names_value = w.compact(_SCALAR, _shared(names_type), True, b"".join(w.symbol(nm) for nm, _ in fields_))
            types_type = self.value_tuple([k for _, k in fields_], tuple_name)
            rec = name + w.refs([names_value, types_type])
            return [w.attribute("julia_type", _shared(self.datatype), _SCALAR, rec, True)]
        a = w.committed(dtype, build)
        w.types.append(a)
        return a


def write_jld2(path, fields_: dict) -> Path:
    """Write a schema-1 checkpoint.  fields_: R, K, grad (Julia matrices as row-major numpy arrays,
    complex128), model (str), coupling (complex), e0 (float; NaN for a state never evaluated), info
    (non-empty str), integration (n_quad, max_point, atol, rtol, maxiters), manifold (metric_reg,
    precond_reg_prefactor, precond_reg).  The file is written through a temporary file and an
    atomic rename."""
    import os
    import tempfile

    R = np.asarray(fields_["R"], dtype=np.complex128)
    D = R.shape[0]
    info = str(fields_["info"])
    if not info:
        raise ValueError("info must be a non-empty string")
    w = _Writer()
    t = _Types(w)
    w.reserve(48)   # superblock

    def matrix(a):
        a = np.asarray(a, dtype=np.complex128)
        if a.shape != (D, D):
            raise ValueError("R, K and grad must be D x D")
        # Julia's column-major D x D matrix = the row-major transpose's bytes
        return np.ascontiguousarray(a.T).astype("<c16").tobytes()

    root = []
    root.append(("schema_version", w.compact(_SCALAR, _T_I64, False, struct.pack("<q", 1))))
    cplx = None
    for key, name in (("R", "R"), ("K", "K")):
        if cplx is None:
            cplx = t.ensure_complex()
        root.append((name, w.contiguous(_simple(D, D), _shared(cplx), True, matrix(fields_[key]))))
    model = str(fields_["model"]).encode()
    root.append(("model", w.compact(_SCALAR, _t_string(len(model)), False, model)))
    c = complex(fields_["coupling"])
    root.append(("coupling", w.compact(_SCALAR, _shared(cplx), True, struct.pack("<dd", c.real, c.imag))))
    root.append(("e0", w.compact(_SCALAR, _T_F64, False, struct.pack("<d", float(fields_["e0"])))))
    root.append(("grad", w.contiguous(_simple(D, D), _shared(cplx), True, matrix(fields_["grad"]))))
    ib = info.encode()
    root.append(("info", w.compact(_SCALAR, _t_string(len(ib)), False, ib)))
    n_quad, max_point, atol, rtol, maxiters = fields_["integration"]
    nt = t.named_tuple([("n_quad", "i"), ("max_point", "f"), ("atol", "f"), ("rtol", "f"), ("maxiters", "i")],
                       "Core.Tuple")
    root.append(("integration", w.compact(_SCALAR, _shared(nt), True,
                                          struct.pack("<qdddq", int(n_quad), float(max_point), float(atol),
                                                      float(rtol), int(maxiters)))))
    mt = t.named_tuple([("metric_reg", "f"), ("precond_reg_prefactor", "f"), ("precond_reg", "f")], "Tuple")
    root.append(("manifold", w.compact(_SCALAR, _shared(mt), True, struct.pack("<ddd", *map(float, fields_["manifold"])))))
    types = sorted(w.types)   # _types/00000001.. in the order the types were placed
    root.append(("_types", w.group([(f"{i + 1:08d}", a) for i, a in enumerate(types)])))
    root_addr = w.group(root)
    w.flush_heap()
    eof = _BASE + len(w.buf)
    sb = b"\x89HDF\r\n\x1a\n\x02\x08\x08\x00" + struct.pack("<QQQQ", _BASE, _UNDEF, eof, root_addr)
    w.at(0, sb + struct.pack("<I", lookup3(sb)))
    content = PREAMBLE + bytes(_BASE - len(PREAMBLE)) + bytes(w.buf)
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=path.parent, suffix=".jld2.tmp")
    try:
        with os.fdopen(fd, "wb") as f:
            f.write(content)
        os.replace(tmp, path)
    finally:
        if os.path.exists(tmp):
            os.remove(tmp)
    return path
