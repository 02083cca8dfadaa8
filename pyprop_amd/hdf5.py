"""A self-contained HDF5 reader and writer for the files PyProp's serialization uses.

PyProp stores wavefunctions and tensor potentials with PyTables
(pyprop/serialization/CommonHDF.py:66-189, WavefunctionHDF.py:2-148,
TensorPotentialHDF.py:10-168).  Neither h5py, PyTables nor libhdf5 is part of this
build, so the subset of the HDF5 file format those files use is implemented here from the
published format specification:

* superblock version 0 or 1 (8-byte offsets and lengths);
* groups as symbol tables: a version-1 B-tree (node type 0) over symbol-table nodes and a
  local heap of link names;
* version-1 object headers with continuation blocks;
* messages: dataspace (versions 1, 2), datatype (fixed-point, IEEE floating point,
  fixed-length string, compound), data layout (versions 1-3: contiguous, chunked with a
  type-1 version-1 B-tree, compact), attribute (versions 1-3), symbol table, fill value
  and modification time (the last two are skipped when reading);
* no filters: a chunked dataset with a filter pipeline is rejected.

The writer produces superblock version 0, symbol-table groups, contiguous datasets and
version-1 attributes, the layout PyTables itself writes apart from chunking.  Complex
arrays are the compound {r: float64, i: float64} of PyTables' ComplexAtom(itemsize=16).
"""
from __future__ import annotations

import mmap
import os
import struct
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class HDF5Error(RuntimeError):
    pass


# --------------------------------------------------------------------------------------
# datatypes
# --------------------------------------------------------------------------------------

def _parse_datatype(buf: bytes, pos: int) -> Tuple[np.dtype, int]:
    """Datatype message at pos -> (numpy dtype, bytes used)."""
    cv = buf[pos]
    cls, ver = cv & 0x0F, cv >> 4
    b0, b1, b2 = buf[pos + 1], buf[pos + 2], buf[pos + 3]
    size = struct.unpack_from("<I", buf, pos + 4)[0]
    p = pos + 8
    if cls == 0:  # fixed point
        order = ">" if b0 & 1 else "<"
        signed = bool(b0 & 0x08)
        return np.dtype("%s%s%d" % (order, "i" if signed else "u", size)), 8 + 4
    if cls == 1:  # floating point
        order = ">" if b0 & 1 else "<"
        if size not in (2, 4, 8):
            raise HDF5Error("unsupported float size %d" % size)
        return np.dtype("%sf%d" % (order, size)), 8 + 12
    if cls == 3:  # fixed-length string
        return np.dtype("S%d" % size), 8
    if cls == 6:  # compound
        nmemb = b0 | (b1 << 8)
        names, formats, offsets = [], [], []
        for _ in range(nmemb):
            end = buf.find(b"\0", p)
            name = buf[p:end].decode()
            if ver < 3:
                p += ((end - p) // 8 + 1) * 8  # name padded to a multiple of 8 with its NUL
            else:
                p = end + 1
            if ver == 1:
                off = struct.unpack_from("<I", buf, p)[0]
                ndims = buf[p + 4]
                if ndims:
                    raise HDF5Error("array members of compound types are not supported")
                p += 4 + 1 + 3 + 4 + 4 + 16
            elif ver == 2:
                off = struct.unpack_from("<I", buf, p)[0]
                p += 4
            else:
                nb = max(1, (size.bit_length() + 7) // 8)
                off = int.from_bytes(buf[p:p + nb], "little")
                p += nb
            dt, used = _parse_datatype(buf, p)
            p += used
            names.append(name)
            formats.append(dt)
            offsets.append(off)
        return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": size}), p - pos
    raise HDF5Error("unsupported datatype class %d" % cls)


def _float_message(size: int) -> bytes:
    if size == 8:
        return struct.pack("<BBBBI", 0x11, 0x20, 63, 0, 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
    if size == 4:
        return struct.pack("<BBBBI", 0x11, 0x20, 31, 0, 4) + struct.pack("<HHBBBBI", 0, 32, 23, 8, 0, 23, 127)
    raise HDF5Error("unsupported float size %d" % size)


def _datatype_message(dt: np.dtype) -> bytes:
    """Version-1 datatype message for a little-endian numpy dtype."""
    dt = np.dtype(dt)
    if dt.kind == "c":
        half = dt.itemsize // 2
        members = b""
        for name, off in (("r", 0), ("i", half)):
            nm = name.encode() + b"\0"
            nm += b"\0" * ((8 - len(nm) % 8) % 8)
            members += nm + struct.pack("<IB3xI4x4I", off, 0, 0, 0, 0, 0, 0) + _float_message(half)
        return struct.pack("<BBBBI", 0x16, 2, 0, 0, dt.itemsize) + members
    if dt.kind == "f":
        return _float_message(dt.itemsize)
    if dt.kind in "iu":
        bits = 0x08 if dt.kind == "i" else 0
        return struct.pack("<BBBBI", 0x10, bits, 0, 0, dt.itemsize) + struct.pack("<HH", 0, 8 * dt.itemsize)
    if dt.kind == "S":
        return struct.pack("<BBBBI", 0x13, 0, 0, 0, dt.itemsize)  # null-terminated ASCII
    if dt.kind == "b":
        return _datatype_message(np.dtype("u1"))
    raise HDF5Error("unsupported dtype %s" % dt)


def _storage_dtype(dt: np.dtype) -> np.dtype:
    dt = np.dtype(dt)
    if dt.kind == "c":
        half = np.dtype("<f%d" % (dt.itemsize // 2))
        return np.dtype({"names": ["r", "i"], "formats": [half, half], "offsets": [0, dt.itemsize // 2],
                         "itemsize": dt.itemsize})
    if dt.kind == "b":
        return np.dtype("u1")
    return dt.newbyteorder("<") if dt.byteorder == ">" else dt


def _to_numpy(raw: np.ndarray) -> np.ndarray:
    """A compound {r, i} of two floats becomes a complex array."""
    dt = raw.dtype
    if dt.names and len(dt.names) == 2 and set(dt.names) == {"r", "i"}:
        r, i = dt.fields["r"][0], dt.fields["i"][0]
        if r == i and r.kind == "f":
            out = np.empty(raw.shape, dtype=np.dtype("c%d" % (2 * r.itemsize)))
            out.real = raw["r"]
            out.imag = raw["i"]
            return out
    return raw


# --------------------------------------------------------------------------------------
# reader
# --------------------------------------------------------------------------------------

@dataclass
class Dataset:
    name: str
    shape: tuple
    dtype: np.dtype
    attrs: dict
    layout: tuple  # ("contiguous", addr, size) | ("chunked", btree, chunk dims) | ("compact", bytes)


@dataclass
class Group:
    name: str
    attrs: dict = field(default_factory=dict)
    members: Dict[str, object] = field(default_factory=dict)


class H5File:
    """Read-only view of an HDF5 file: H5File(path)["/group/dataset"] -> numpy array."""

    def __init__(self, path: str):
        self.path = path
        with open(path, "rb") as f:
            # mapped, not read: datasets are copied out only when asked for
            self.buf = mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ)
        buf = self.buf
        if buf[:8] != SIGNATURE:
            raise HDF5Error("not an HDF5 file (signature)")
        ver = buf[8]
        if ver not in (0, 1):
            raise HDF5Error("superblock version %d not supported" % ver)
        so, sl = buf[13], buf[14]
        if so != 8 or sl != 8:
            raise HDF5Error("only 8-byte offsets and lengths are supported")
        p = 24 + (4 if ver == 1 else 0)
        self.base, _free, self.eof, _drv = struct.unpack_from("<QQQQ", buf, p)
        p += 32
        _name_off, root_hdr = struct.unpack_from("<QQ", buf, p)
        self.root = self._read_object("/", root_hdr)

    # -- low level ---------------------------------------------------------------------
    def _u(self, fmt, pos):
        return struct.unpack_from(fmt, self.buf, self.base + pos if pos != UNDEF else pos)

    def _messages(self, addr: int) -> List[Tuple[int, bytes]]:
        buf = self.buf
        a = self.base + addr
        if buf[a] != 1:
            raise HDF5Error("object header version %d not supported" % buf[a])
        nmsg = struct.unpack_from("<H", buf, a + 2)[0]
        size = struct.unpack_from("<I", buf, a + 8)[0]
        blocks = [(a + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            p = start
            while p + 8 <= start + length and len(out) < nmsg:
                mtype, msize, _flags = struct.unpack_from("<HHB", buf, p)
                data = buf[p + 8:p + 8 + msize]
                p += 8 + msize
                if mtype == 0x10:  # continuation
                    off, ln = struct.unpack_from("<QQ", data, 0)
                    blocks.append((self.base + off, ln))
                out.append((mtype, data))
        return out

    def _heap(self, addr: int) -> bytes:
        a = self.base + addr
        if self.buf[a:a + 4] != b"HEAP":
            raise HDF5Error("bad local heap signature")
        size, _free, data = struct.unpack_from("<QQQ", self.buf, a + 8)
        return self.buf[self.base + data:self.base + data + size]

    def _group_entries(self, btree: int, heap: bytes) -> List[Tuple[str, int]]:
        buf = self.buf
        a = self.base + btree
        if buf[a:a + 4] != b"TREE" or buf[a + 4] != 0:
            raise HDF5Error("bad group B-tree node")
        level = buf[a + 5]
        used = struct.unpack_from("<H", buf, a + 6)[0]
        p = a + 24
        children = []
        for i in range(used):
            p += 8  # key
            children.append(struct.unpack_from("<Q", buf, p)[0])
            p += 8
        out = []
        for c in children:
            if level > 0:
                out.extend(self._group_entries(c, heap))
                continue
            s = self.base + c
            if buf[s:s + 4] != b"SNOD":
                raise HDF5Error("bad symbol table node")
            n = struct.unpack_from("<H", buf, s + 6)[0]
            for k in range(n):
                e = s + 8 + 40 * k
                name_off, hdr = struct.unpack_from("<QQ", buf, e)
                name = heap[name_off:heap.find(b"\0", name_off)].decode()
                out.append((name, hdr))
        return out

    def _attribute(self, data: bytes) -> Tuple[str, object]:
        ver = data[0]
        nsize, tsize, ssize = struct.unpack_from("<HHH", data, 2)
        p = 8 + (1 if ver == 3 else 0)
        pad = (lambda n: (n + 7) // 8 * 8) if ver == 1 else (lambda n: n)
        name = data[p:p + nsize].split(b"\0")[0].decode()
        p += pad(nsize)
        dt, _ = _parse_datatype(data, p)
        p += pad(tsize)
        shape = self._dataspace(data[p:p + ssize])
        p += pad(ssize)
        count = int(np.prod(shape)) if shape else 1
        raw = np.frombuffer(data, dtype=dt, count=count, offset=p).reshape(shape)
        val = _to_numpy(raw).copy()
        if val.shape == () and dt.kind == "S":
            return name, val.item().decode(errors="replace")
        return name, val if val.shape else val[()]

    @staticmethod
    def _dataspace(data: bytes) -> tuple:
        ver, rank, flags = data[0], data[1], data[2]
        if ver == 1:
            p = 8
        elif ver == 2:
            if data[3] == 2:
                return (0,)
            p = 4
        else:
            raise HDF5Error("dataspace version %d not supported" % ver)
        return tuple(struct.unpack_from("<%dQ" % rank, data, p)) if rank else ()

    def _read_object(self, name: str, addr: int):
        msgs = self._messages(addr)
        attrs = {}
        shape = dtype = layout = None
        sym = None
        filtered = False
        for mtype, data in msgs:
            if mtype == 0x0C:
                k, v = self._attribute(data)
                attrs[k] = v
            elif mtype == 0x11:
                sym = struct.unpack_from("<QQ", data, 0)
            elif mtype == 0x01:
                shape = self._dataspace(data)
            elif mtype == 0x03:
                dtype, _ = _parse_datatype(data, 0)
            elif mtype == 0x0B:
                filtered = True
            elif mtype == 0x08:
                layout = self._layout(data, shape)
        if sym is not None:
            g = Group(name, attrs)
            heap = self._heap(sym[1])
            for child, hdr in self._group_entries(sym[0], heap):
                g.members[child] = self._read_object(child, hdr)
            return g
        if layout is None or dtype is None or shape is None:
            raise HDF5Error("object %r is neither a group nor a dataset" % name)
        if filtered and layout[0] == "chunked":
            raise HDF5Error("dataset %r uses a filter pipeline (compression), not supported" % name)
        return Dataset(name, shape, dtype, attrs, layout)

    def _layout(self, data: bytes, shape) -> tuple:
        ver = data[0]
        if ver == 3:
            cls = data[1]
            if cls == 1:
                return ("contiguous",) + struct.unpack_from("<QQ", data, 2)
            if cls == 2:
                nd = data[2]
                btree = struct.unpack_from("<Q", data, 3)[0]
                dims = struct.unpack_from("<%dI" % nd, data, 11)
                return ("chunked", btree, dims)
            if cls == 0:
                n = struct.unpack_from("<H", data, 2)[0]
                return ("compact", data[4:4 + n])
        elif ver in (1, 2):
            nd, cls = data[1], data[2]
            p = 8
            if cls == 0:
                dims = struct.unpack_from("<%dI" % nd, data, p)
                p += 4 * nd
                n = struct.unpack_from("<I", data, p)[0]
                return ("compact", data[p + 4:p + 4 + n])
            addr = struct.unpack_from("<Q", data, p)[0]
            p += 8
            dims = struct.unpack_from("<%dI" % nd, data, p)
            if cls == 1:
                size = int(np.prod(dims))
                return ("contiguous", addr, size)
            p += 4 * nd
            esize = struct.unpack_from("<I", data, p)[0]
            return ("chunked", addr, tuple(dims) + (esize,))
        raise HDF5Error("data layout version %d not supported" % ver)

    def _chunks(self, addr: int, rank: int):
        buf = self.buf
        a = self.base + addr
        if buf[a:a + 4] != b"TREE" or buf[a + 4] != 1:
            raise HDF5Error("bad chunk B-tree node")
        level = buf[a + 5]
        used = struct.unpack_from("<H", buf, a + 6)[0]
        ksize = 8 + 8 * (rank + 1)
        p = a + 24
        for _ in range(used):
            csize, fmask = struct.unpack_from("<II", buf, p)
            offs = struct.unpack_from("<%dQ" % (rank + 1), buf, p + 8)
            child = struct.unpack_from("<Q", buf, p + ksize)[0]
            p += ksize + 8
            if level > 0:
                yield from self._chunks(child, rank)
            else:
                if fmask != 0:
                    raise HDF5Error("filtered chunk")
                yield offs[:rank], child, csize

    # -- public ------------------------------------------------------------------------
    def node(self, path: str):
        cur = self.root
        for part in [p for p in path.split("/") if p]:
            if not isinstance(cur, Group) or part not in cur.members:
                raise KeyError(path)
            cur = cur.members[part]
        return cur

    def __contains__(self, path: str) -> bool:
        try:
            self.node(path)
            return True
        except KeyError:
            return False

    def attrs(self, path: str) -> dict:
        return self.node(path).attrs

    def read(self, path: str) -> np.ndarray:
        ds = self.node(path)
        if not isinstance(ds, Dataset):
            raise HDF5Error("%s is a group" % path)
        count = int(np.prod(ds.shape)) if ds.shape else 1
        kind = ds.layout[0]
        if kind == "contiguous":
            addr = ds.layout[1]
            if addr == UNDEF:
                raw = np.zeros(ds.shape, dtype=ds.dtype)
            else:
                raw = np.frombuffer(self.buf, dtype=ds.dtype, count=count, offset=self.base + addr).reshape(ds.shape)
        elif kind == "compact":
            raw = np.frombuffer(ds.layout[1], dtype=ds.dtype, count=count).reshape(ds.shape)
        else:
            rank = len(ds.shape)
            cdims = tuple(ds.layout[2][:rank])
            raw = np.zeros(ds.shape, dtype=ds.dtype)
            if ds.layout[1] != UNDEF:
                for offs, addr, _ in self._chunks(ds.layout[1], rank):
                    chunk = np.frombuffer(self.buf, dtype=ds.dtype, count=int(np.prod(cdims)),
                                          offset=self.base + addr).reshape(cdims)
                    sl = tuple(slice(o, min(o + c, n)) for o, c, n in zip(offs, cdims, ds.shape))
                    raw[sl] = chunk[tuple(slice(0, s.stop - s.start) for s in sl)]
        return _to_numpy(np.array(raw))

    def walk(self, group: Optional[Group] = None, prefix: str = ""):
        """(path, node) for every object below the root."""
        g = group or self.root
        for name, obj in g.members.items():
            path = prefix + "/" + name
            yield path, obj
            if isinstance(obj, Group):
                yield from self.walk(obj, path)


# --------------------------------------------------------------------------------------
# writer
# --------------------------------------------------------------------------------------

@dataclass
class _WGroup:
    attrs: dict = field(default_factory=dict)
    members: Dict[str, object] = field(default_factory=dict)


@dataclass
class _WDataset:
    array: np.ndarray
    attrs: dict = field(default_factory=dict)


LEAF_K = 16      # symbol-table node capacity 2K = 32 links per group
INTERNAL_K = 16


class H5Writer:
    """Build a file in memory (groups, arrays, attributes), then write it in one pass."""

    def __init__(self):
        self.root = _WGroup()

    def _group(self, path: str, create: bool = True) -> _WGroup:
        cur = self.root
        for part in [p for p in path.split("/") if p]:
            nxt = cur.members.get(part)
            if nxt is None:
                if not create:
                    raise KeyError(path)
                nxt = cur.members[part] = _WGroup()
            if not isinstance(nxt, _WGroup):
                raise HDF5Error("%s is a dataset, not a group" % part)
            cur = nxt
        return cur

    def create_group(self, path: str) -> None:
        self._group(path)

    def create_dataset(self, path: str, array: np.ndarray, attrs: Optional[dict] = None) -> None:
        parent, _, name = path.rstrip("/").rpartition("/")
        if not name:
            raise HDF5Error("invalid dataset path %r" % path)
        self._group(parent).members[name] = _WDataset(np.ascontiguousarray(array), dict(attrs or {}))

    def set_attr(self, path: str, name: str, value) -> None:
        node = self._node(path)
        node.attrs[name] = value

    def remove(self, path: str) -> None:
        parent, _, name = path.rstrip("/").rpartition("/")
        try:
            self._group(parent, create=False).members.pop(name, None)
        except KeyError:
            pass

    def _node(self, path: str):
        cur = self.root
        for part in [p for p in path.split("/") if p]:
            cur = cur.members[part]
        return cur

    @classmethod
    def from_file(cls, path: str) -> "H5Writer":
        """Everything a file holds, to be modified and written again."""
        w = cls()
        f = H5File(path)
        w.root.attrs.update(f.root.attrs)
        for p, obj in f.walk():
            if isinstance(obj, Group):
                w._group(p).attrs.update(obj.attrs)
            else:
                w.create_dataset(p, f.read(p), obj.attrs)
        return w

    # -- layout --------------------------------------------------------------------------
    def write(self, path: str) -> None:
        out = bytearray()

        def alloc(n: int) -> int:
            addr = len(out)
            out.extend(b"\0" * ((n + 7) // 8 * 8))
            return addr

        def put(addr: int, data: bytes) -> None:
            out[addr:addr + len(data)] = data

        alloc(96)  # superblock + root symbol table entry

        def attr_message(name: str, value) -> bytes:
            if isinstance(value, str):
                value = np.array(value.encode())
            elif isinstance(value, bytes):
                value = np.array(value)
            arr = np.asarray(value)
            if arr.dtype.kind == "U":
                arr = arr.astype("S")
            if arr.dtype.kind == "S" and arr.dtype.itemsize == 0:
                arr = arr.astype("S1")
            sdt = _storage_dtype(arr.dtype)
            dtm = _datatype_message(arr.dtype)
            spm = _dataspace_message(arr.shape)
            data = _storage_bytes(arr, sdt)
            nm = name.encode() + b"\0"
            pad = lambda b: b + b"\0" * ((8 - len(b) % 8) % 8)
            body = struct.pack("<BBHHH", 1, 0, len(nm), len(dtm), len(spm)) + pad(nm) + pad(dtm) + pad(spm) + data
            return body

        def object_header(messages: List[Tuple[int, bytes]]) -> int:
            body = b""
            for mtype, data in messages:
                data = data + b"\0" * ((8 - len(data) % 8) % 8)
                body += struct.pack("<HHB3x", mtype, len(data), 0) + data
            hdr = struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(body))
            addr = alloc(len(hdr) + len(body))
            put(addr, hdr + body)
            return addr

        def write_dataset(ds: _WDataset) -> int:
            arr = ds.array
            sdt = _storage_dtype(arr.dtype)
            data = _storage_bytes(arr, sdt)
            daddr = alloc(len(data)) if len(data) else UNDEF
            if len(data):
                put(daddr, data)
            msgs = [(0x01, _dataspace_message(arr.shape)),
                    (0x03, _datatype_message(arr.dtype)),
                    (0x05, struct.pack("<BBBB", 2, 1, 2, 0)),  # fill value v2: early, never, undefined
                    (0x08, struct.pack("<BBQQ", 3, 1, daddr, len(data)))]
            msgs += [(0x0C, attr_message(k, v)) for k, v in ds.attrs.items()]
            return object_header(msgs)

        def write_group(g: _WGroup) -> Tuple[int, int, int]:
            """-> (object header, B-tree, heap) addresses."""
            names = sorted(g.members)
            if len(names) > 2 * LEAF_K:
                raise HDF5Error("more than %d members in one group" % (2 * LEAF_K))
            children = {}
            for n in names:
                obj = g.members[n]
                children[n] = write_group(obj)[0] if isinstance(obj, _WGroup) else write_dataset(obj)
            # local heap: offset 0 = "", then the names
            heap = bytearray(b"\0" * 8)
            offs = {}
            for n in names:
                offs[n] = len(heap)
                nm = n.encode() + b"\0"
                heap += nm + b"\0" * ((8 - len(nm) % 8) % 8)
            data_addr = alloc(len(heap))
            put(data_addr, bytes(heap))
            heap_addr = alloc(32)
            put(heap_addr, b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), UNDEF, data_addr))
            # one symbol table node
            snod_addr = alloc(8 + 2 * LEAF_K * 40)
            snod = b"SNOD" + struct.pack("<BBH", 1, 0, len(names))
            for n in names:
                snod += struct.pack("<QQII16x", offs[n], children[n], 0, 0)
            put(snod_addr, snod)
            # B-tree root (level 0): key 0, child, key = last name
            bt_addr = alloc(24 + (2 * INTERNAL_K + 1) * 8 + 2 * INTERNAL_K * 8)
            last = offs[names[-1]] if names else 0
            put(bt_addr, b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF) + struct.pack("<QQQ", 0, snod_addr, last))
            msgs = [(0x11, struct.pack("<QQ", bt_addr, heap_addr))]
            msgs += [(0x0C, attr_message(k, v)) for k, v in g.attrs.items()]
            return object_header(msgs), bt_addr, heap_addr

        root_hdr, root_bt, root_heap = write_group(self.root)
        sb = SIGNATURE + struct.pack("<BBBBBBBBHHI", 0, 0, 0, 0, 0, 8, 8, 0, LEAF_K, INTERNAL_K, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, len(out), UNDEF)
        sb += struct.pack("<QQII", 0, root_hdr, 1, 0) + struct.pack("<QQ", root_bt, root_heap)
        put(0, sb)
        tmp = path + ".tmp"
        with open(tmp, "wb") as f:
            f.write(out)
        os.replace(tmp, path)


def _dataspace_message(shape: tuple) -> bytes:
    return struct.pack("<BBB5x", 1, len(shape), 0) + b"".join(struct.pack("<Q", int(n)) for n in shape)


def _storage_bytes(arr: np.ndarray, sdt: np.dtype) -> bytes:
    if arr.dtype.kind == "c":
        out = np.empty(arr.shape, dtype=sdt)
        out["r"] = arr.real
        out["i"] = arr.imag
        return out.tobytes()
    return np.ascontiguousarray(arr, dtype=sdt).tobytes()
