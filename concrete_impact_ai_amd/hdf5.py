"""A self-contained reader and writer of the HDF5 subset that the reference's RVE
data files use (``src/fem/rve/io.py``, written there with h5py).

No HDF5 library is importable on this path, so the file format is handled here
from its published specification (HDF Group, "HDF5 File Format Specification",
version 2.0 of the format document): superblock versions 0-3, version-1 object
headers and their continuation blocks (version-2 "OHDR" headers are read too),
symbol-table groups (version-1 B-trees, symbol-table nodes, local heaps) and
link messages of compact groups, datasets with compact, contiguous or chunked
layout (version-1 B-tree chunk index) and the deflate, shuffle, Fletcher-32 and
LZF (filter 32000, the LZF format of liblzf) filters, fixed-point, floating-point,
fixed-length string and variable-length string datatypes (the latter through
the global heap), and attributes.

The writer produces superblock-0 files with the encodings h5py itself writes
(as found in the reference's sample shard): symbol-table groups (multi-level
B-trees for large groups), contiguous datasets (uncompressed), the bool
enumeration, fixed-length byte strings, and version-1 attributes whose ``str``
values are variable-length UTF-8 strings in a global heap.
"""

from __future__ import annotations

import struct
import zlib
from pathlib import Path

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


# ---------------------------------------------------------------------------
# filters


def lzf_decompress(src: bytes, out_size: int) -> bytes:
    """LZF decompression (the liblzf format of h5py's filter 32000): a control
    byte c < 32 copies c + 1 literal bytes; otherwise a back reference of length
    (c >> 5) + 2 (7 + next byte + 2 when c >> 5 == 7) at distance
    ((c & 31) << 8) + next byte + 1."""
    out = bytearray(out_size)
    ip = op = 0
    n = len(src)
    while ip < n:
        c = src[ip]
        ip += 1
        if c < 32:
            c += 1
            out[op:op + c] = src[ip:ip + c]
            ip += c
            op += c
        else:
            length = c >> 5
            ref = op - ((c & 31) << 8) - 1
            if length == 7:
                length += src[ip]
                ip += 1
            ref -= src[ip]
            ip += 1
            length += 2
            if ref < 0:
                raise ValueError("corrupt LZF stream")
            for k in range(length):  # overlapping copies are byte-serial
                out[op + k] = out[ref + k]
            op += length
    if op != out_size:
        raise ValueError(f"LZF stream decoded to {op} bytes, expected {out_size}")
    return bytes(out)


def _unshuffle(data: bytes, itemsize: int) -> bytes:
    a = np.frombuffer(data, dtype=np.uint8)
    n = a.size // itemsize
    return a[:n * itemsize].reshape(itemsize, n).T.tobytes() + a[n * itemsize:].tobytes()


# ---------------------------------------------------------------------------
# reader


class Dataset:
    def __init__(self, f: "File", name: str, msgs: dict):
        self.file = f
        self.name = name
        self.shape = msgs["shape"]
        self.dtype = msgs["dtype"]
        self._layout = msgs["layout"]
        self._filters = msgs.get("filters", [])
        self.attrs = msgs["attrs"]
        self._vlen = msgs.get("vlen")

    def __getitem__(self, key):
        return self.read()[key]

    def read(self):
        return self.file._read_data(self)

    def __repr__(self):
        return f"<Dataset {self.name} {self.shape} {self.dtype}>"


class Group:
    def __init__(self, f: "File", name: str, links: dict, attrs: dict):
        self.file = f
        self.name = name
        self._links = links
        self.attrs = attrs

    def keys(self):
        return list(self._links)

    def __contains__(self, key):
        try:
            self[key]
            return True
        except KeyError:
            return False

    def __iter__(self):
        return iter(self._links)

    def __getitem__(self, key: str):
        node = self
        for part in [p for p in key.split("/") if p]:
            if not isinstance(node, Group) or part not in node._links:
                raise KeyError(key)
            node = node.file._object(node._links[part], f"{node.name.rstrip('/')}/{part}")
        return node

    def visit(self, fn, prefix=""):
        for k in self.keys():
            obj = self[k]
            path = f"{prefix}{k}"
            fn(path, obj)
            if isinstance(obj, Group):
                obj.visit(fn, path + "/")

    def __repr__(self):
        return f"<Group {self.name} {self.keys()}>"


class File(Group):
    """Read-only HDF5 file: ``File(path)["group/dataset"].read()`` -> numpy array."""

    def __init__(self, path):
        self.path = Path(path)
        self.buf = self.path.read_bytes()
        b = self.buf
        base = b.find(SIGNATURE)
        if base != 0:
            raise ValueError(f"{path}: not an HDF5 file (signature at {base})")
        ver = b[8]
        self._objcache: dict = {}
        if ver in (0, 1):
            self.so, self.sl = b[13], b[14]
            p = 24 + (4 if ver == 1 else 0)
            p += 4 * self.so  # base, free space, end of file, driver info
            root_hdr = self._off(p + self.so)
        elif ver in (2, 3):
            self.so, self.sl = b[9], b[10]
            p = 12 + 3 * self.so
            root_hdr = self._off(p)
        else:
            raise ValueError(f"unsupported superblock version {ver}")
        self.superblock_version = ver
        msgs = self._header(root_hdr)
        Group.__init__(self, self, "/", msgs["links"], msgs["attrs"])

    # -- primitives
    def _u(self, p, n):
        return int.from_bytes(self.buf[p:p + n], "little")

    def _off(self, p):
        return self._u(p, self.so)

    def _len(self, p):
        return self._u(p, self.sl)

    # -- object headers
    def _object(self, addr, name):
        if addr not in self._objcache:
            m = self._header(addr)
            if "layout" in m:
                self._objcache[addr] = Dataset(self, name, m)
            else:
                self._objcache[addr] = Group(self, name, m["links"], m["attrs"])
        return self._objcache[addr]

    def _header(self, addr) -> dict:
        b = self.buf
        msgs = []
        if b[addr:addr + 4] == b"OHDR":
            flags = b[addr + 5]
            p = addr + 6
            if flags & 0x20:
                p += 16
            if flags & 0x10:
                p += 4
            size_len = 1 << (flags & 3)
            chunk_size = self._u(p, size_len)
            p += size_len
            blocks = [(p, p + chunk_size, 2, flags)]
        else:
            if b[addr] != 1:
                raise ValueError(f"unsupported object header version {b[addr]} at {addr}")
            nmsg = self._u(addr + 2, 2)
            size = self._u(addr + 8, 4)
            blocks = [(addr + 16, addr + 16 + size, 1, 0)]
        k = 0
        while k < len(blocks):
            start, end, v, flags = blocks[k]
            p = start
            if v == 2 and b[p:p + 4] == b"OCHK":
                p += 4
            while p + (8 if v == 1 else 4) <= end:
                if v == 1:
                    mtype = self._u(p, 2)
                    msize = self._u(p + 2, 2)
                    p += 8
                else:
                    mtype = b[p]
                    msize = self._u(p + 1, 2)
                    p += 4 + (2 if flags & 0x04 else 0)
                data = b[p:p + msize]
                if mtype == 0x10:  # continuation
                    cont = self._off(p)
                    clen = self._len(p + self.so)
                    blocks.append((cont, cont + clen, v, flags))
                else:
                    msgs.append((mtype, data, p))
                p += msize
                if v == 2 and end - p < 4:
                    break
            k += 1
        return self._interpret(msgs)

    def _interpret(self, msgs) -> dict:
        out: dict = {"attrs": {}, "links": {}}
        for mtype, data, pos in msgs:
            if mtype == 0x01:
                out["shape"] = self._dataspace(data)[0]
            elif mtype == 0x03:
                out["dtype"], out["vlen"] = self._datatype(data)
            elif mtype == 0x08:
                out["layout"] = self._layout(data)
            elif mtype == 0x0B:
                out["filters"] = self._filter_pipeline(data)
            elif mtype == 0x0C:
                name, value = self._attribute(data)
                out["attrs"][name] = value
            elif mtype == 0x11:
                btree = int.from_bytes(data[:self.so], "little")
                heap = int.from_bytes(data[self.so:2 * self.so], "little")
                out["links"].update(self._symbol_table(btree, heap))
            elif mtype == 0x06:
                name, addr = self._link(data)
                if addr is not None:
                    out["links"][name] = addr
        return out

    # -- messages
    def _dataspace(self, d):
        ver, rank, flags = d[0], d[1], d[2]
        p = 8 if ver == 1 else 4
        dims = tuple(int.from_bytes(d[p + i * self.sl:p + (i + 1) * self.sl], "little") for i in range(rank))
        if ver == 2 and d[3] == 0:
            return (), flags
        return dims, flags

    def _datatype(self, d):
        cls = d[0] & 0x0F
        bits = d[1] | (d[2] << 8) | (d[3] << 16)
        size = int.from_bytes(d[4:8], "little")
        order = ">" if bits & 1 else "<"
        if cls == 0:
            signed = bool(bits & 0x08)
            return np.dtype(f"{order}{'i' if signed else 'u'}{size}"), None
        if cls == 1:
            return np.dtype(f"{order}f{size}"), None
        if cls == 3:
            return np.dtype(f"S{size}"), None
        if cls == 9:
            vtype = bits & 0x0F
            if vtype == 1:  # string
                return np.dtype(object), "str"
            base, _ = self._datatype(d[8:])
            return np.dtype(object), base
        if cls == 8:  # enumeration: h5py's FALSE / TRUE over int8 is a bool, others the base type
            base, _ = self._datatype(d[8:])
            if (bits & 0xFFFF) == 2 and size == 1:
                return np.dtype(bool), None
            return base, None
        raise ValueError(f"unsupported datatype class {cls}")

    def _layout(self, d):
        ver = d[0]
        if ver == 3:
            cls = d[1]
            if cls == 0:
                size = int.from_bytes(d[2:4], "little")
                return ("compact", bytes(d[4:4 + size]))
            if cls == 1:
                return ("contiguous", int.from_bytes(d[2:2 + self.so], "little"),
                        int.from_bytes(d[2 + self.so:2 + self.so + self.sl], "little"))
            if cls == 2:
                rank = d[2]
                addr = int.from_bytes(d[3:3 + self.so], "little")
                p = 3 + self.so
                dims = [int.from_bytes(d[p + 4 * i:p + 4 * i + 4], "little") for i in range(rank)]
                return ("chunked", addr, dims)
        if ver in (1, 2):
            rank, cls = d[1], d[2]
            p = 8
            addr = None
            if cls != 0:
                addr = int.from_bytes(d[p:p + self.so], "little")
                p += self.so
            dims = [int.from_bytes(d[p + 4 * i:p + 4 * i + 4], "little") for i in range(rank)]
            if cls == 1:
                return ("contiguous", addr, None)
            if cls == 2:
                return ("chunked", addr, dims)
            size = int.from_bytes(d[p + 4 * rank:p + 4 * rank + 4], "little")
            return ("compact", bytes(d[p + 4 * rank + 4:p + 4 * rank + 4 + size]))
        raise ValueError(f"unsupported layout message version {ver}")

    def _filter_pipeline(self, d):
        ver, n = d[0], d[1]
        p = 8 if ver == 1 else 2
        out = []
        for _ in range(n):
            fid = int.from_bytes(d[p:p + 2], "little")
            p += 2
            if ver == 1 or fid >= 256:
                nlen = int.from_bytes(d[p:p + 2], "little")
                p += 2
            else:
                nlen = 0
            flags = int.from_bytes(d[p:p + 2], "little")
            ncv = int.from_bytes(d[p + 2:p + 4], "little")
            p += 4
            if ver == 1:
                p += (nlen + 7) // 8 * 8
            else:
                p += nlen
            cv = [int.from_bytes(d[p + 4 * i:p + 4 * i + 4], "little") for i in range(ncv)]
            p += 4 * ncv
            if ver == 1 and ncv % 2:
                p += 4
            out.append((fid, flags, cv))
        return out

    def _attribute(self, d):
        ver = d[0]
        nsize = int.from_bytes(d[2:4], "little")
        tsize = int.from_bytes(d[4:6], "little")
        ssize = int.from_bytes(d[6:8], "little")
        if ver == 1:
            p = 8
            pad = lambda n: (n + 7) // 8 * 8  # noqa: E731
        else:
            p = 9 if ver == 3 else 8
            pad = lambda n: n  # noqa: E731
        name = bytes(d[p:p + nsize]).split(b"\0")[0].decode("utf-8")
        p += pad(nsize)
        dtype, vlen = self._datatype(d[p:p + tsize])
        p += pad(tsize)
        shape, _ = self._dataspace(d[p:p + ssize])
        p += pad(ssize)
        count = int(np.prod(shape)) if shape else 1
        if vlen is not None:
            vals = [self._vlen_item(d, p + k * (4 + self.so + 4), vlen) for k in range(count)]
            value = vals[0] if not shape else np.asarray(vals, dtype=object).reshape(shape)
        else:
            arr = np.frombuffer(bytes(d[p:p + count * dtype.itemsize]), dtype=dtype)
            value = arr[0] if not shape else arr.reshape(shape)
            if dtype.kind == "S":
                value = (value.decode("utf-8").rstrip("\0") if not shape
                         else np.char.rstrip(np.char.decode(arr.reshape(shape), "utf-8"), "\0"))
            elif not shape:
                value = value.item()
        return name, value

    def _vlen_item(self, d, p, vlen):
        n = int.from_bytes(d[p:p + 4], "little")
        addr = int.from_bytes(d[p + 4:p + 4 + self.so], "little")
        idx = int.from_bytes(d[p + 4 + self.so:p + 8 + self.so], "little")
        raw = self._global_heap_object(addr, idx)
        if vlen == "str":
            return raw.decode("utf-8")
        return np.frombuffer(raw, dtype=vlen)[:n]

    def _global_heap_object(self, addr, index):
        b = self.buf
        if addr == 0 and index == 0:
            return b""
        if b[addr:addr + 4] != b"GCOL":
            raise ValueError("bad global heap collection")
        size = self._len(addr + 8)
        p = addr + 8 + self.sl
        end = addr + size
        while p + 8 + self.sl <= end:
            idx = self._u(p, 2)
            osize = self._len(p + 8)
            data = b[p + 8 + self.sl:p + 8 + self.sl + osize]
            if idx == index:
                return bytes(data)
            if idx == 0:
                break
            p += 8 + self.sl + (osize + 7) // 8 * 8
        raise KeyError(f"global heap object {index} not found")

    def _link(self, d):
        flags = d[1]
        p = 2
        ltype = 0
        if flags & 0x08:
            ltype = d[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        lsize = 1 << (flags & 3)
        nlen = int.from_bytes(d[p:p + lsize], "little")
        p += lsize
        name = bytes(d[p:p + nlen]).decode("utf-8")
        p += nlen
        if ltype != 0:
            return name, None
        return name, int.from_bytes(d[p:p + self.so], "little")

    def _symbol_table(self, btree, heap):
        b = self.buf
        if b[heap:heap + 4] != b"HEAP":
            raise ValueError("bad local heap")
        data_addr = self._off(heap + 8 + 2 * self.sl)

        def name_at(off):
            end = b.index(b"\0", data_addr + off)
            return b[data_addr + off:end].decode("utf-8")

        links = {}

        def walk(node):
            if b[node:node + 4] != b"TREE":
                raise ValueError("bad B-tree node")
            level = b[node + 5]
            used = self._u(node + 6, 2)
            p = node + 8 + 2 * self.so
            children = []
            for i in range(used):
                p += self.sl  # key
                children.append(self._off(p))
                p += self.so
            for c in children:
                if level > 0:
                    walk(c)
                else:
                    if b[c:c + 4] != b"SNOD":
                        raise ValueError("bad symbol table node")
                    nsym = self._u(c + 6, 2)
                    q = c + 8
                    for _ in range(nsym):
                        links[name_at(self._off(q))] = self._off(q + self.so)
                        q += 2 * self.so + 24

        walk(btree)
        return links

    # -- data
    def _read_data(self, ds: Dataset):
        shape = ds.shape
        count = int(np.prod(shape)) if shape else 1
        kind = ds._layout[0]
        if ds._vlen is not None:
            raw = self._raw_bytes(ds, count * (4 + self.so + 4))
            vals = [self._vlen_item(raw, k * (4 + self.so + 4), ds._vlen) for k in range(count)]
            return np.asarray(vals, dtype=object).reshape(shape) if shape else vals[0]
        itemsize = ds.dtype.itemsize
        if kind == "chunked":
            arr = self._read_chunked(ds, shape, itemsize)
        else:
            raw = self._raw_bytes(ds, count * itemsize)
            arr = np.frombuffer(raw, dtype=ds.dtype, count=count).reshape(shape)
        return arr.copy() if shape else arr.reshape(()).item()

    def _raw_bytes(self, ds, nbytes):
        lay = ds._layout
        if lay[0] == "compact":
            return lay[1]
        if lay[0] == "contiguous":
            if lay[1] == UNDEF:
                return bytes(nbytes)
            return self.buf[lay[1]:lay[1] + nbytes]
        return self._read_chunked(ds, ds.shape, ds.dtype.itemsize).tobytes()

    def _filter(self, data, ds, mask, nbytes):
        for k, (fid, _, cv) in reversed(list(enumerate(ds._filters))):
            if mask & (1 << k):
                continue
            if fid == 1:
                data = zlib.decompress(data)
            elif fid == 2:
                data = _unshuffle(data, ds.dtype.itemsize)
            elif fid == 3:
                data = data[:-4]
            elif fid == 32000:
                data = lzf_decompress(data, cv[2] if len(cv) > 2 and cv[2] else nbytes)
            else:
                raise ValueError(f"unsupported HDF5 filter {fid}")
        return data

    def _read_chunked(self, ds, shape, itemsize):
        _, addr, cdims = ds._layout
        rank = len(shape)
        cshape = tuple(cdims[:rank])
        out = np.zeros(shape, dtype=ds.dtype)
        if addr == UNDEF:
            return out
        b = self.buf
        nbytes = int(np.prod(cshape)) * itemsize

        def walk(node):
            if b[node:node + 4] != b"TREE":
                raise ValueError("bad chunk B-tree node")
            level = b[node + 5]
            used = self._u(node + 6, 2)
            p = node + 8 + 2 * self.so
            ksize = 8 + 8 * (rank + 1)
            for i in range(used):
                csize = self._u(p, 4)
                mask = self._u(p + 4, 4)
                offs = [self._u(p + 8 + 8 * j, 8) for j in range(rank)]
                child = self._off(p + ksize)
                p += ksize + self.so
                if level > 0:
                    walk(child)
                    continue
                data = self._filter(bytes(b[child:child + csize]), ds, mask, nbytes)
                chunk = np.frombuffer(data, dtype=ds.dtype, count=int(np.prod(cshape))).reshape(cshape)
                sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, cshape, shape))
                out[sl] = chunk[tuple(slice(0, s.stop - s.start) for s in sl)]

        walk(addr)
        return out


def read_tree(path) -> dict:
    """Every dataset and attribute of a file as {path: array or value}; attributes
    under "<path>@<name>"."""
    f = File(path)
    out = {f"@{k}": v for k, v in f.attrs.items()}

    def visit(name, obj):
        for k, v in obj.attrs.items():
            out[f"{name}@{k}"] = v
        if isinstance(obj, Dataset):
            out[name] = obj.read()

    f.visit(visit)
    return out


# ---------------------------------------------------------------------------
# writer

_LEAF_K, _INTERNAL_K = 4, 16        # symbol-table node capacity 2 * 4, B-tree fan-out 2 * 16
_SNOD_SIZE = 8 + 2 * _LEAF_K * 40
_TREE_SIZE = 24 + (2 * _INTERNAL_K + 1) * 8 + 2 * _INTERNAL_K * 8


def _pad8(b: bytes) -> bytes:
    return b + bytes((-len(b)) % 8)


class _Node:
    def __init__(self, kind, data=None):
        self.kind = kind
        self.children: dict = {}
        self.attrs: dict = {}
        self.data = data


class Writer:
    """Build a file in memory (``group``, ``dataset``, ``attrs``) and ``save`` it.

    Datatypes: float32/64, signed and unsigned integers, bool (the FALSE/TRUE
    enumeration over int8 that h5py maps to numpy bool), fixed-length byte strings
    (numpy ``S``); attribute values may also be ``str`` (variable-length UTF-8,
    stored in the global heap, read back as ``str`` by h5py) or numbers."""

    def __init__(self):
        self.root = _Node("group")

    @property
    def attrs(self) -> dict:
        return self.root.attrs

    def _walk(self, path, create_last_as=None, data=None):
        parts = [p for p in path.split("/") if p]
        node = self.root
        for i, part in enumerate(parts):
            last = i == len(parts) - 1
            if part not in node.children:
                if last and create_last_as == "dataset":
                    node.children[part] = _Node("dataset", data)
                else:
                    node.children[part] = _Node("group")
            elif last and create_last_as == "dataset":
                raise ValueError(f"{path} exists")
            node = node.children[part]
            if node.kind != "group" and not last:
                raise ValueError(f"{path}: {part} is a dataset")
        return node

    def group(self, path) -> _Node:
        node = self._walk(path)
        if node.kind != "group":
            raise ValueError(f"{path} is a dataset")
        return node

    def dataset(self, path, data) -> _Node:
        arr = np.asarray(data)
        if arr.dtype.kind == "U":
            arr = np.char.encode(arr, "utf-8")
        return self._walk(path, "dataset", np.ascontiguousarray(arr))

    def __contains__(self, path):
        node = self.root
        for part in [p for p in path.split("/") if p]:
            if node.kind != "group" or part not in node.children:
                return False
            node = node.children[part]
        return True

    def get(self, path) -> _Node:
        node = self.root
        for part in [p for p in path.split("/") if p]:
            node = node.children[part]
        return node

    # -- encoding
    @staticmethod
    def _dtype_msg(dt: np.dtype) -> bytes:
        dt = np.dtype(dt)
        if dt.kind == "f":
            prec = dt.itemsize * 8
            exp_loc, exp_size, man_size, bias = {8: (52, 11, 52, 1023), 4: (23, 8, 23, 127)}[dt.itemsize]
            return (bytes([0x11, 0x20, prec - 1, 0]) + struct.pack("<I", dt.itemsize)
                    + struct.pack("<HHBBBBI", 0, prec, exp_loc, exp_size, 0, man_size, bias))
        if dt.kind in "iu":
            return (bytes([0x10, 0x08 if dt.kind == "i" else 0x00, 0, 0]) + struct.pack("<I", dt.itemsize)
                    + struct.pack("<HH", 0, dt.itemsize * 8))
        if dt.kind == "b":
            base = bytes([0x10, 0x08, 0, 0]) + struct.pack("<I", 1) + struct.pack("<HH", 0, 8)
            return (bytes([0x18, 2, 0, 0]) + struct.pack("<I", 1) + base + _pad8(b"FALSE\0")
                    + _pad8(b"TRUE\0") + bytes([0, 1]))
        if dt.kind == "S":
            return bytes([0x13, 0x01, 0, 0]) + struct.pack("<I", max(dt.itemsize, 1))
        raise ValueError(f"unsupported dtype {dt}")

    @staticmethod
    def _space_msg(shape, maxdims=True) -> bytes:
        rank = len(shape)
        out = bytes([1, rank, 1 if (maxdims and rank) else 0, 0]) + bytes(4)
        out += b"".join(struct.pack("<Q", int(d)) for d in shape)
        if maxdims and rank:
            out += b"".join(struct.pack("<Q", int(d)) for d in shape)
        return out

    @staticmethod
    def _msg(mtype, data, flags=0) -> bytes:
        data = _pad8(data)
        return struct.pack("<HHB3x", mtype, len(data), flags) + data

    @staticmethod
    def _header(msgs) -> bytes:
        body = b"".join(msgs)
        return struct.pack("<BBHII4x", 1, 0, len(msgs), 1, len(body)) + body

    def _attr_msg(self, name, value, heap_ids) -> bytes:
        nm = name.encode("utf-8") + b"\0"
        if isinstance(value, str):
            dt = bytes([0x19, 0x01, 0x01, 0x00]) + struct.pack("<I", 16) + \
                bytes([0x10, 0, 0, 0]) + struct.pack("<I", 1) + struct.pack("<HH", 0, 8)
            sp = self._space_msg((), False)
            raw = value.encode("utf-8")
            addr, idx = heap_ids[raw]
            data = struct.pack("<IQI", len(raw), addr, idx)
        else:
            arr = np.asarray(value)
            if arr.dtype.kind == "U":
                arr = np.char.encode(arr, "utf-8")
            if arr.dtype.kind in "f":
                arr = arr.astype("<f8")
            elif arr.dtype.kind in "iu" and arr.dtype.itemsize < 8:
                arr = arr.astype("<i8" if arr.dtype.kind == "i" else "<u8")
            dt = self._dtype_msg(arr.dtype)
            sp = self._space_msg(arr.shape, False)
            data = np.ascontiguousarray(arr).tobytes()
        body = struct.pack("<BBHHH", 1, 0, len(nm), len(dt), len(sp)) + _pad8(nm) + _pad8(dt) + _pad8(sp) + data
        return self._msg(0x0C, body)

    # -- layout
    def save(self, path) -> Path:
        path = Path(path)
        strings = []

        def collect(node):
            for v in node.attrs.values():
                if isinstance(v, str):
                    strings.append(v.encode("utf-8"))
            for c in node.children.values():
                collect(c)

        collect(self.root)
        out = bytearray(96)
        heap_ids = {}
        if strings:
            uniq = list(dict.fromkeys(strings))
            objs = b""
            offs = {}
            for i, raw in enumerate(uniq, start=1):
                offs[raw] = i
                objs += struct.pack("<HH4xQ", i, 0, len(raw)) + _pad8(raw)
            size = max(4096, 16 + len(objs) + 16)
            free = size - 16 - len(objs)
            gaddr = len(out)
            out += b"GCOL" + bytes([1, 0, 0, 0]) + struct.pack("<Q", size) + objs
            out += struct.pack("<HH4xQ", 0, 0, free) + bytes(free - 16)
            heap_ids = {raw: (gaddr, i) for raw, i in offs.items()}
        root_addr, btree, lheap = self._write_group(out, self.root, heap_ids)
        sb = bytearray(SIGNATURE + bytes([0, 0, 0, 0, 0, 8, 8, 0]))
        sb += struct.pack("<HH", _LEAF_K, _INTERNAL_K) + struct.pack("<I", 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, len(out), UNDEF)
        sb += struct.pack("<QQI4xQQ", 0, root_addr, 1, btree, lheap)
        out[:96] = sb
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_bytes(bytes(out))
        return path

    @staticmethod
    def _align(out):
        out += bytes((-len(out)) % 8)

    def _write_dataset(self, out, node, heap_ids) -> int:
        arr = node.data
        dt = arr.dtype
        if dt.kind == "f" and dt.byteorder == ">":
            arr = arr.astype(dt.newbyteorder("<"))
        self._align(out)
        nbytes = arr.nbytes
        if nbytes:
            addr = len(out)
            out += arr.tobytes()
        else:
            addr = UNDEF
        self._align(out)
        msgs = [self._msg(0x01, self._space_msg(arr.shape)),
                self._msg(0x03, self._dtype_msg(arr.dtype), flags=1),
                self._msg(0x05, bytes([2, 2, 2, 1]), flags=1),
                self._msg(0x08, bytes([3, 1]) + struct.pack("<QQ", addr, nbytes))]
        msgs += [self._attr_msg(k, v, heap_ids) for k, v in node.attrs.items()]
        haddr = len(out)
        out += self._header(msgs)
        return haddr

    def _write_group(self, out, node, heap_ids):
        names = sorted(node.children, key=lambda s: s.encode("utf-8"))
        addrs = {}
        for nm in names:
            child = node.children[nm]
            if child.kind == "group":
                addrs[nm] = self._write_group(out, child, heap_ids)[0]
            else:
                addrs[nm] = self._write_dataset(out, child, heap_ids)
        # local heap: "" at offset 0, then the names
        seg = bytearray(8)
        name_off = {}
        for nm in names:
            name_off[nm] = len(seg)
            seg += _pad8(nm.encode("utf-8") + b"\0")
        self._align(out)
        seg_addr = len(out)
        out += seg
        heap_addr = len(out)
        out += b"HEAP" + bytes(4) + struct.pack("<QQQ", len(seg), UNDEF, seg_addr)
        # symbol-table nodes, 2 * leaf K entries each
        leaves = []  # (address, first key offset, last name offset)
        for s in range(0, max(len(names), 1), 2 * _LEAF_K) if names else []:
            chunk = names[s:s + 2 * _LEAF_K]
            addr = len(out)
            snod = bytearray(b"SNOD" + bytes([1, 0]) + struct.pack("<H", len(chunk)))
            for nm in chunk:
                snod += struct.pack("<QQI4x16x", name_off[nm], addrs[nm], 0)
            snod += bytes(_SNOD_SIZE - len(snod))
            out += snod
            leaves.append((addr, name_off[chunk[-1]]))
        # B-tree levels over the symbol-table nodes
        level = 0
        nodes = leaves
        while True:
            parents = []
            groups = [nodes[i:i + 2 * _INTERNAL_K] for i in range(0, max(len(nodes), 1), 2 * _INTERNAL_K)] \
                if nodes else [[]]
            for grp in groups:
                addr = len(out)
                t = bytearray(b"TREE" + bytes([0, level]) + struct.pack("<H", len(grp))
                              + struct.pack("<QQ", UNDEF, UNDEF))
                t += struct.pack("<Q", 0)
                for child_addr, last in grp:
                    t += struct.pack("<QQ", child_addr, last)
                t += bytes(_TREE_SIZE - len(t))
                out += t
                parents.append((addr, grp[-1][1] if grp else 0))
            if len(parents) == 1:
                btree = parents[0][0]
                break
            nodes = parents
            level += 1
        msgs = [self._msg(0x11, struct.pack("<QQ", btree, heap_addr))]
        msgs += [self._attr_msg(k, v, heap_ids) for k, v in node.attrs.items()]
        haddr = len(out)
        out += self._header(msgs)
        return haddr, btree, heap_addr
