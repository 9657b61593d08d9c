"""JLD2 checkpoints (rcmps_store schema 1, src/storage.jl:4-38) read without h5py.

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
