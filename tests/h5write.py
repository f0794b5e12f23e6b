"""A small HDF5 writer for the tests (the library only reads HDF5).

Writes the file form libhdf5 and h5py produce by default: superblock version
0, version-1 object headers, groups as symbol tables (a one-node B-tree, a
symbol node and a local heap), datasets with a version-1 dataspace, an IEEE
float or integer datatype, a version-3 layout (contiguous, or chunked with a
one-node version-1 B-tree, optionally through the shuffle and deflate
filters).  Offsets and lengths are 8 bytes.

    write(path, {"images": array, "coords/alpha": array}, chunks=..., deflate=...)
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF


class _Buf:
    def __init__(self):
        self.b = bytearray()

    def alloc(self, n, align=8):
        while len(self.b) % align:
            self.b.append(0)
        at = len(self.b)
        self.b.extend(b"\0" * n)
        return at

    def put(self, at, data):
        self.b[at:at + len(data)] = data


def _msg(mtype, body):
    body = bytes(body)
    body += b"\0" * ((-len(body)) % 8)
    return struct.pack("<HHB3x", mtype, len(body), 0) + body


def _object_header(buf, msgs):
    blob = b"".join(msgs)
    at = buf.alloc(16 + len(blob))
    buf.put(at, struct.pack("<BBHII4x", 1, 0, len(msgs), 1, len(blob)) + blob)
    return at


def _dtype_msg(dt):
    dt = np.dtype(dt)
    bo = 1 if dt.byteorder == ">" else 0
    if dt.kind == "f":
        # IEEE layouts: (exponent location, exponent size, mantissa size, bias)
        e = {2: (10, 5, 10, 15), 4: (23, 8, 23, 127), 8: (52, 11, 52, 1023)}[dt.itemsize]
        sign = dt.itemsize * 8 - 1
        bits = bo | (0x20)   # mantissa normalisation: implied msb set
        body = struct.pack("<B3BI", 0x11, bits, sign, 0, dt.itemsize)
        body += struct.pack("<HHBBBBI", 0, dt.itemsize * 8, e[0], e[1], 0, e[2], e[3])
    elif dt.kind in "iu":
        bits = bo | (0x08 if dt.kind == "i" else 0)
        body = struct.pack("<B3BI", 0x10, bits, 0, 0, dt.itemsize)
        body += struct.pack("<HH", 0, dt.itemsize * 8)
    else:
        raise ValueError(f"unsupported dtype {dt}")
    return _msg(0x03, body)


def _dataset(buf, arr, chunks=None, deflate=None, shuffle=False):
    arr = np.asarray(arr)
    dt = arr.dtype
    rank = arr.ndim
    space = struct.pack("<BBBB4x", 1, rank, 0, 0) + b"".join(struct.pack("<Q", d) for d in arr.shape)
    msgs = [_msg(0x01, space), _dtype_msg(dt)]
    if chunks is None:
        raw = arr.tobytes(order="C")
        at = buf.alloc(len(raw))
        buf.put(at, raw)
        msgs.append(_msg(0x08, struct.pack("<BBQQ", 3, 1, at, len(raw))))
        return _object_header(buf, msgs)
    chunks = tuple(int(c) for c in chunks)
    es = dt.itemsize
    grid = [(-(-n // c)) for n, c in zip(arr.shape, chunks)]
    entries = []
    for idx in np.ndindex(*grid):
        off = [i * c for i, c in zip(idx, chunks)]
        block = np.zeros(chunks, dt)
        sl = tuple(slice(o, min(o + c, n)) for o, c, n in zip(off, chunks, arr.shape))
        part = arr[sl]
        block[tuple(slice(0, s) for s in part.shape)] = part
        data = block.tobytes(order="C")
        if shuffle:
            data = np.frombuffer(data, np.uint8).reshape(-1, es).T.tobytes()
        if deflate is not None:
            data = zlib.compress(data, deflate)
        at = buf.alloc(len(data))
        buf.put(at, data)
        entries.append((len(data), off, at))
    if len(entries) > 64:
        raise ValueError("the one-node chunk B-tree holds at most 64 chunks")
    # B-tree node type 1, level 0: keys (size, filter mask, offsets + element dim) and children
    key = lambda size, off: struct.pack("<II", size, 0) + b"".join(struct.pack("<Q", o) for o in off + [0])
    node = b"TREE" + struct.pack("<BBHQQ", 1, 0, len(entries), UNDEF, UNDEF)
    for size, off, at in entries:
        node += key(size, off) + struct.pack("<Q", at)
    node += key(0, list(arr.shape))
    bt = buf.alloc(len(node))
    buf.put(bt, node)
    msgs.append(_msg(0x08, struct.pack("<BBBQ", 3, 2, rank + 1, bt) +
                     b"".join(struct.pack("<I", c) for c in chunks + (es,))))
    filters = []
    if shuffle:
        filters.append((2, [es]))
    if deflate is not None:
        filters.append((1, [deflate]))
    if filters:
        body = struct.pack("<BB6x", 1, len(filters))
        for fid, cd in filters:
            body += struct.pack("<HHHH", fid, 0, 0, len(cd)) + b"".join(struct.pack("<I", v) for v in cd)
            if len(cd) % 2:
                body += b"\0" * 4
        msgs.append(_msg(0x0B, body))
    return _object_header(buf, msgs)


def _group(buf, children):
    """children: name -> object header address; returns the group's header address"""
    names = sorted(children)
    heap = bytearray(b"\0" * 8)   # offset 0: the empty name
    offs = {}
    for n in names:
        offs[n] = len(heap)
        heap += n.encode() + b"\0"
        heap += b"\0" * ((-len(heap)) % 8)
    hdata = buf.alloc(len(heap))
    buf.put(hdata, bytes(heap))
    hh = buf.alloc(32)
    buf.put(hh, b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), UNDEF, hdata))
    snod = b"SNOD" + struct.pack("<BBH", 1, 0, len(names))
    for n in names:
        snod += struct.pack("<QQII16x", offs[n], children[n], 0, 0)
    sn = buf.alloc(len(snod))
    buf.put(sn, snod)
    node = b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF)
    node += struct.pack("<QQQ", 0, sn, offs[names[-1]] if names else 0)
    bt = buf.alloc(len(node))
    buf.put(bt, node)
    oh = _object_header(buf, [_msg(0x11, struct.pack("<QQ", bt, hh))])
    return oh, bt, hh


def write(path, datasets: dict, chunks=None, deflate=None, shuffle=False):
    """datasets: absolute or relative paths ("images", "coords/alpha") -> arrays.
    chunks: None (contiguous) or a dict path -> chunk shape."""
    buf = _Buf()
    buf.alloc(96)   # superblock version 0 with 8-byte offsets
    tree: dict = {}
    for name, arr in datasets.items():
        parts = [p for p in name.split("/") if p]
        node = tree
        for p in parts[:-1]:
            node = node.setdefault(p, {})
        ch = None if chunks is None else chunks.get(name.strip("/"))
        node[parts[-1]] = _dataset(buf, arr, ch, deflate if ch else None, shuffle and ch is not None)

    def build(node):
        kids = {k: (build(v)[0] if isinstance(v, dict) else v) for k, v in node.items()}
        return _group(buf, kids)

    root, rbt, rheap = build(tree)
    eof = len(buf.b)
    sb = b"\x89HDF\r\n\x1a\n" + struct.pack("<BBBBBBBBHHI", 0, 0, 0, 0, 0, 8, 8, 0, 4, 16, 0)
    sb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
    sb += struct.pack("<QQII", 0, root, 1, 0) + struct.pack("<QQ", rbt, rheap)
    buf.put(0, sb)
    with open(path, "wb") as f:
        f.write(bytes(buf.b))
