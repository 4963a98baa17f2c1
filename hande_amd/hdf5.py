"""A small HDF5 reader and writer for HANDE restart files.

HANDE writes its restart files (src/restart_hdf5.F90:307-537) through the
HDF5 Fortran library with its default file-creation properties: a version-0
superblock, groups stored as symbol tables (a version-1 B-tree of symbol
nodes plus a local heap of link names), version-1 object headers and
contiguous (or compact) datasets of little-endian integers, IEEE doubles or
fixed-length strings.  No HDF5 library is present here, so this module reads
and writes exactly that subset of the published HDF5 file format
specification (version 2.0 of the format document: superblock, object header,
B-tree v1, symbol table node, local heap, and the dataspace, datatype,
layout, symbol-table and continuation messages).  Anything outside the subset
(chunked or filtered layouts, variable-length or compound types, version-2
object headers, dense link storage) raises ``HDF5Error`` rather than being
guessed at.

The tree is returned as nested dicts: a group maps link names to groups or
datasets, a dataset is a numpy array (shape in C order, i.e. the reverse of
the Fortran shape HANDE declares) or, for strings, ``bytes``.
"""
from __future__ import annotations

import struct

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF


class HDF5Error(ValueError):
    """The file is not HDF5 or uses a feature outside the supported subset."""


# --------------------------------------------------------------------- reading


class _Reader:
    def __init__(self, data: bytes):
        self.d = data
        if data[:8] != _SIG:
            raise HDF5Error("no HDF5 signature at offset 0")
        ver = data[8]
        if ver not in (0, 1):
            raise HDF5Error(f"superblock version {ver} is not supported")
        self.so = data[13]
        self.sl = data[14]
        if self.so != 8 or self.sl != 8:
            raise HDF5Error("only 8-byte offsets and lengths are supported")
        pos = 24 if ver == 0 else 28
        self.base, _free, self.eof, _drv = struct.unpack_from("<4Q", data, pos)
        pos += 32
        # root group symbol table entry
        _name, self.root_ohdr, cache, _res = struct.unpack_from("<QQII", data, pos)
        if self.base != 0:
            raise HDF5Error("a non-zero base address is not supported")

    def u(self, fmt, pos):
        return struct.unpack_from("<" + fmt, self.d, pos)

    # local heap: names of the links of one group
    def heap(self, addr):
        if self.d[addr:addr + 4] != b"HEAP":
            raise HDF5Error(f"no local heap at {addr}")
        _size, _free, data_addr = self.u("QQQ", addr + 8)
        return data_addr

    def heap_name(self, data_addr, off):
        start = data_addr + off
        end = self.d.index(b"\0", start)
        return self.d[start:end].decode()

    # v1 B-tree of a group: yields symbol table entries (name offset, object header address)
    def group_entries(self, btree, heap_data):
        out = []
        stack = [btree]
        while stack:
            a = stack.pop()
            if self.d[a:a + 4] != b"TREE":
                raise HDF5Error(f"no B-tree node at {a}")
            ntype, level, used = self.u("BBH", a + 4)
            if ntype != 0:
                raise HDF5Error("expected a group B-tree node")
            p = a + 8 + 16  # siblings
            children = []
            for i in range(used):
                p += 8  # key (heap offset)
                (child,) = self.u("Q", p)
                p += 8
                children.append(child)
            if level > 0:
                stack.extend(reversed(children))
                continue
            for c in children:
                if self.d[c:c + 4] != b"SNOD":
                    raise HDF5Error(f"no symbol table node at {c}")
                (nsym,) = self.u("H", c + 6)
                q = c + 8
                for _ in range(nsym):
                    name_off, ohdr = self.u("QQ", q)
                    out.append((self.heap_name(heap_data, name_off), ohdr))
                    q += 40
        return out

    # v1 object header: list of (type, flags, payload bytes)
    def messages(self, addr):
        ver, _res, nmsg, _refs, size = self.u("BBHII", addr)
        if ver != 1:
            raise HDF5Error(f"object header version {ver} at {addr} is not supported")
        blocks = [(addr + 16, size)]
        msgs = []
        while blocks and len(msgs) < nmsg:
            p, n = blocks.pop(0)
            end = p + n
            while p + 8 <= end and len(msgs) < nmsg:
                mtype, msize, mflags = self.u("HHB", p)
                body = self.d[p + 8:p + 8 + msize]
                if mtype == 0x10:  # continuation
                    off, ln = struct.unpack_from("<QQ", body, 0)
                    blocks.append((off, ln))
                msgs.append((mtype, mflags, body))
                p += 8 + msize
        return msgs

    def obj(self, addr):
        msgs = self.messages(addr)
        types = {m[0] for m in msgs}
        if 0x11 in types:
            body = next(m[2] for m in msgs if m[0] == 0x11)
            btree, heap = struct.unpack_from("<QQ", body, 0)
            hd = self.heap(heap)
            return {name: self.obj(o) for name, o in self.group_entries(btree, hd)}
        if 0x02 in types or 0x06 in types:
            raise HDF5Error("new-style (link message) groups are not supported")
        return self.dataset(msgs)

    def dataset(self, msgs):
        space = dtype = layout = None
        for mtype, _f, body in msgs:
            if mtype == 0x01:
                space = _dataspace(body)
            elif mtype == 0x03:
                dtype = _datatype(body)
            elif mtype == 0x08:
                layout = body
        if space is None or dtype is None or layout is None:
            raise HDF5Error("object is neither a symbol-table group nor a dataset")
        for mtype, _f, _b in msgs:
            if mtype == 0x0B:
                raise HDF5Error("filtered datasets are not supported")
        raw = self.layout_bytes(layout)
        kind, np_dt, size = dtype
        count = int(np.prod(space)) if space else 1
        if kind == "str":
            chunks = [raw[i * size:(i + 1) * size].rstrip(b"\0 ") for i in range(count)]
            return chunks[0] if not space else np.array(chunks, dtype=object).reshape(space)
        arr = np.frombuffer(raw[:count * size], dtype=np_dt).copy()
        return arr.reshape(space) if space else arr[0]

    def layout_bytes(self, body):
        ver = body[0]
        if ver == 3:
            cls = body[1]
            if cls == 0:
                (n,) = struct.unpack_from("<H", body, 2)
                return body[4:4 + n]
            if cls == 1:
                addr, n = struct.unpack_from("<QQ", body, 2)
                if addr == _UNDEF:
                    return b""
                return self.d[addr:addr + n]
            raise HDF5Error("chunked datasets are not supported")
        if ver in (1, 2):
            ndim, cls = body[1], body[2]
            if cls == 1:
                (addr,) = struct.unpack_from("<Q", body, 8)
                dims = struct.unpack_from(f"<{ndim}I", body, 16)
                n = int(np.prod(dims))
                return self.d[addr:addr + n]
            if cls == 0:
                p = 8 + 4 * ndim
                (n,) = struct.unpack_from("<I", body, p)
                return body[p + 4:p + 4 + n]
            raise HDF5Error("chunked datasets are not supported")
        raise HDF5Error(f"layout message version {ver} is not supported")


def _dataspace(body):
    ver, ndim, flags = body[0], body[1], body[2]
    if ver == 1:
        p = 8
    elif ver == 2:
        if body[3] == 2:  # null dataspace
            return ()
        p = 4
    else:
        raise HDF5Error(f"dataspace version {ver} is not supported")
    dims = struct.unpack_from(f"<{ndim}Q", body, p)
    return tuple(int(x) for x in dims)


def _datatype(body):
    cls = body[0] & 0x0F
    bits0 = body[1]
    (size,) = struct.unpack_from("<I", body, 4)
    if cls == 0:  # fixed point
        if bits0 & 1:
            raise HDF5Error("big-endian integers are not supported")
        signed = bool(bits0 & 0x08)
        return ("int", np.dtype(("<i" if signed else "<u") + str(size)), size)
    if cls == 1:  # IEEE floating point
        if bits0 & 1:
            raise HDF5Error("big-endian floats are not supported")
        return ("float", np.dtype("<f" + str(size)), size)
    if cls == 3:  # fixed-length string
        return ("str", None, size)
    raise HDF5Error(f"datatype class {cls} is not supported")


def read(path_or_bytes) -> dict:
    """Read an HDF5 file (path or its bytes) into nested dicts of numpy arrays."""
    if isinstance(path_or_bytes, (bytes, bytearray)):
        data = bytes(path_or_bytes)
    else:
        with open(path_or_bytes, "rb") as fh:
            data = fh.read()
    r = _Reader(data)
    return r.obj(r.root_ohdr)


# --------------------------------------------------------------------- writing


class _Writer:
    """Lays out a file in the same subset the reader accepts: version-0
    superblock, symbol-table groups (one leaf B-tree node, one symbol node per
    group) and contiguous datasets with version-1 object headers."""

    LEAF_K = 4      # group leaf node K: a symbol node holds up to 2K entries
    INTERNAL_K = 16

    def __init__(self):
        self.buf = bytearray()

    def alloc(self, n, align=8):
        pad = (-len(self.buf)) % align
        self.buf += b"\0" * pad
        a = len(self.buf)
        self.buf += b"\0" * n
        return a

    def put(self, addr, data):
        self.buf[addr:addr + len(data)] = data

    @staticmethod
    def _msg(mtype, body, flags=0):
        body = bytes(body)
        body += b"\0" * ((-len(body)) % 8)
        return struct.pack("<HHB3x", mtype, len(body), flags) + body

    def _ohdr(self, msgs):
        blob = b"".join(msgs)
        addr = self.alloc(16 + len(blob))
        self.put(addr, struct.pack("<BBHII4x", 1, 0, len(msgs), 1, len(blob)) + blob)
        return addr

    def dataset(self, value):
        if isinstance(value, (bytes, str)):
            s = value.encode() if isinstance(value, str) else value
            n = max(1, len(s))
            space = struct.pack("<BBB5x", 1, 0, 0)  # scalar
            dtype = struct.pack("<B3sI", 0x13, b"\0\0\0", n)  # class 3, version 1, null-terminated ASCII
            raw = s.ljust(n, b"\0")
        else:
            arr = np.asarray(value)
            if arr.dtype == np.bool_:
                arr = arr.astype(np.int32)
            if arr.dtype.kind == "f":
                arr = arr.astype("<f8") if arr.dtype.itemsize == 8 else arr.astype("<f4")
                size = arr.dtype.itemsize
                if size == 8:
                    props = struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
                else:
                    props = struct.pack("<HHBBBBI", 0, 32, 23, 8, 0, 23, 127)
                dtype = struct.pack("<BBBBI", 0x11, 0x20, 0x3F, 0x00, size) + props
            elif arr.dtype.kind in "iu":
                arr = arr.astype("<" + arr.dtype.kind + str(arr.dtype.itemsize))
                size = arr.dtype.itemsize
                bits0 = 0x08 if arr.dtype.kind == "i" else 0x00
                dtype = struct.pack("<BBBBI", 0x10, bits0, 0, 0, size) + struct.pack("<HH", 0, 8 * size)
            else:
                raise HDF5Error(f"cannot write dtype {arr.dtype}")
            if arr.ndim == 0:
                space = struct.pack("<BBB5x", 1, 0, 0)
            else:
                space = struct.pack("<BBB5x", 1, arr.ndim, 0) + struct.pack(f"<{arr.ndim}Q", *arr.shape)
            raw = np.ascontiguousarray(arr).tobytes()
        daddr = self.alloc(len(raw))
        self.put(daddr, raw)
        layout = struct.pack("<BBQQ", 3, 1, daddr, len(raw))
        return self._ohdr([self._msg(0x01, space), self._msg(0x03, dtype, flags=1), self._msg(0x08, layout)])


def write(path, tree: dict) -> bytes:
    """Write nested dicts (groups) of numpy arrays, scalars and strings as HDF5.

    Returns the file's bytes, and writes them to ``path`` unless it is None."""
    data = _write_tree(_Writer(), tree)
    if path is not None:
        with open(path, "wb") as fh:
            fh.write(data)
    return data


def _write_tree(w: _Writer, tree: dict) -> bytes:
    sb = w.alloc(96)

    def group(t):
        names = sorted(t)
        kids = {n: (group(v)[0] if isinstance(v, dict) else w.dataset(v)) for n, v in t.items()}
        return _group_body(w, names, kids)

    root, btree, heap = group(tree)
    eof = len(w.buf)
    blob = _SIG + struct.pack("<BBBBBBBBHHI", 0, 0, 0, 0, 0, 8, 8, 0, w.LEAF_K, w.INTERNAL_K, 0)
    blob += struct.pack("<QQQQ", 0, _UNDEF, eof, _UNDEF)
    blob += struct.pack("<QQII", 0, root, 1, 0) + struct.pack("<QQ", btree, heap)
    w.put(sb, blob)
    return bytes(w.buf)


def _group_body(w: _Writer, names, kids):
    heap_data = bytearray(b"\0" * 8)
    offs = {}
    for n in names:
        offs[n] = len(heap_data)
        s = n.encode() + b"\0"
        heap_data += s + b"\0" * ((-len(s)) % 8)
    free_off = len(heap_data)
    heap_data += struct.pack("<QQ", 1, 16)  # one free block at the end: next = 1 (none), size 16
    data_addr = w.alloc(len(heap_data))
    w.put(data_addr, heap_data)
    heap = w.alloc(32)
    w.put(heap, b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap_data), free_off, data_addr))
    per = 2 * w.LEAF_K
    groups = [names[i:i + per] for i in range(0, len(names), per)] or [[]]
    if len(groups) > 2 * w.INTERNAL_K:
        raise HDF5Error("too many links in one group")
    snods = []
    for g in groups:
        a = w.alloc(8 + 40 * per)
        blob = b"SNOD" + struct.pack("<BBH", 1, 0, len(g))
        for n in g:
            blob += struct.pack("<QQII16x", offs[n], kids[n], 0, 0)
        w.put(a, blob)
        snods.append(a)
    nkeys = 2 * w.INTERNAL_K
    btree = w.alloc(24 + 8 * (nkeys + 1) + 8 * nkeys)
    blob = b"TREE" + struct.pack("<BBHQQ", 0, 0, len(snods), _UNDEF, _UNDEF)
    blob += struct.pack("<Q", 0)
    for g, a in zip(groups, snods):
        blob += struct.pack("<QQ", a, offs[g[-1]] if g else 0)
    w.put(btree, blob)
    ohdr = w._ohdr([w._msg(0x11, struct.pack("<QQ", btree, heap))])
    return ohdr, btree, heap
