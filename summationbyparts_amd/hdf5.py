"""HDF5 output in the reference's layout, written without an HDF5 library.

The multigrid binary writes one HDF5 file per run (multigrid/src/system.rs:
28-53, 87-117, 591-616):

    /t                      u64, shape (n,), unlimited, chunk (1,)   step index of each output
    /<grid>/t               soft link to /t
    /<grid>/x, /<grid>/y    f64, (ny, nx), chunk (ny, nx), deflate 9
    /<grid>/{rho,rhou,rhov,e}
                            f64, (n, ny, nx), unlimited along n, chunk (1, ny, nx),
                            shuffle + deflate 3

`Output` mirrors that: the constructor is BaseSystem::new/create's file set-up,
`output(ntime, fields)` is System::output (one time slice appended to every
variable). The image has no HDF5 library for Python, so this module writes the
file format itself (HDF5 file format specification, version 0 superblock):
raw chunks are appended as they come, and the metadata (object headers,
version 1 B-trees of the groups and of the chunked datasets, local heaps,
symbol-table nodes) is written at `close()` (or `flush()`), after which the
superblock at offset 0 points at the root group. Compression is zlib's
deflate, the shuffle filter the HDF5 byte shuffle.
"""
import os
import struct
import zlib

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
_SIG = b"\x89HDF\r\n\x1a\n"
_GROUP_LEAF_K = 4       # symbol-table node: up to 2K entries
_GROUP_INTERNAL_K = 16  # group B-tree node: up to 2K children
_CHUNK_K = 32           # chunk B-tree node: up to 2K children (the library default for a version 0 superblock)
_HEAP_FREE_NULL = 1     # end of a local heap's free list on disk
_FILTER_DEFLATE, _FILTER_SHUFFLE = 1, 2
_FLAG_OPTIONAL = 1


def _pad8(b):
    return b + b"\0" * (-len(b) % 8)


def _msg(mtype, data, flags=0):
    data = _pad8(data)
    return struct.pack("<HHB3x", mtype, len(data), flags) + data


def _object_header(messages):
    body = b"".join(messages)
    return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(body)) + body


def _dtype_message(dt):
    dt = np.dtype(dt)
    if dt == np.float64:
        # class 1 (floating point), version 1; little endian, implied msb
        # mantissa normalisation, sign at bit 63; offset 0, precision 64,
        # exponent at 52 (11 bits), mantissa at 0 (52 bits), bias 1023
        return _msg(3, struct.pack("<B3BI", 0x11, 0x20, 63, 0, 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023), 1)
    if dt == np.uint64:
        # class 0 (fixed point), version 1, little endian, unsigned
        return _msg(3, struct.pack("<B3BI", 0x10, 0, 0, 0, 8) + struct.pack("<HH", 0, 64), 1)
    raise ValueError(f"unsupported dtype {dt}")


def _dataspace_message(shape, maxshape):
    flags = 1 if maxshape is not None else 0
    b = struct.pack("<BBBx4x", 1, len(shape), flags) + b"".join(struct.pack("<Q", n) for n in shape)
    if maxshape is not None:
        b += b"".join(struct.pack("<Q", UNDEF if m is None else m) for m in maxshape)
    return _msg(1, b)


def _layout_message(btree_addr, chunk, itemsize):
    dims = list(chunk) + [itemsize]
    return _msg(8, struct.pack("<BBBQ", 3, 2, len(dims), btree_addr) + b"".join(struct.pack("<I", d) for d in dims))


def _pipeline_message(filters):
    b = struct.pack("<BB6x", 1, len(filters))
    for fid, cd in filters:
        b += struct.pack("<HHHH", fid, 0, _FLAG_OPTIONAL, len(cd)) + b"".join(struct.pack("<I", v) for v in cd)
        if len(cd) % 2:
            b += b"\0" * 4
    return _msg(11, b)


def _shuffle(raw, itemsize):
    a = np.frombuffer(raw, dtype=np.uint8)
    return a.reshape(-1, itemsize).T.tobytes()


class _Dataset:
    def __init__(self, dtype, shape, maxshape, chunk, filters):
        self.dtype = np.dtype(dtype)
        self.shape = list(shape)
        self.maxshape = maxshape
        self.chunk = tuple(chunk)
        self.filters = filters  # [(filter id, client data)]
        self.chunks = {}  # chunk offset (elements) -> (address, stored size)


class _Group:
    def __init__(self):
        self.children = {}  # name -> _Group | _Dataset | ("soft", target)


class Writer:
    """A minimal HDF5 writer: groups, soft links and chunked datasets of f64
    or u64 with optional shuffle and deflate filters, extendible along axis 0."""

    def __init__(self, path):
        self.path = path
        self.f = open(path, "w+b")
        self.f.write(b"\0" * 96)  # superblock, written at flush/close
        self.end = 96
        self.meta_start = None  # where the last metadata set begins
        self.root = _Group()

    # ------------------------------------------------------------ building
    def _node(self, path):
        g = self.root
        for part in [p for p in path.split("/") if p]:
            g = g.children[part]
        return g

    def _parent(self, path):
        parts = [p for p in path.split("/") if p]
        return self._node("/".join(parts[:-1])), parts[-1]

    def create_group(self, path):
        parent, name = self._parent(path)
        if name in parent.children:
            raise ValueError(f"{path} exists")
        parent.children[name] = _Group()

    def link_soft(self, target, path):
        parent, name = self._parent(path)
        parent.children[name] = ("soft", target)

    def create_dataset(self, path, dtype, shape, chunk, maxshape=None, shuffle=False, deflate=None):
        parent, name = self._parent(path)
        filters = []
        if shuffle:
            filters.append((_FILTER_SHUFFLE, [np.dtype(dtype).itemsize]))
        if deflate is not None:
            filters.append((_FILTER_DEFLATE, [int(deflate)]))
        d = _Dataset(dtype, shape, maxshape, chunk, filters)
        parent.children[name] = d
        return d

    def _append(self, data):
        addr = self.end
        self.f.seek(addr)
        self.f.write(data)
        self.end = addr + len(data)
        pad = -self.end % 8
        if pad:
            self.f.write(b"\0" * pad)
            self.end += pad
        return addr

    def write_chunk(self, path, offset, values):
        """Write the chunk at element offset `offset` (a multiple of the chunk
        shape) of dataset `path`, growing axis 0 to cover it."""
        d = self._node(path)
        a = np.ascontiguousarray(values, dtype=d.dtype).reshape(d.chunk)
        raw = a.astype(d.dtype.newbyteorder("<"), copy=False).tobytes()
        if self.meta_start is not None:
            # new raw data goes over the last metadata set, which the next
            # flush writes again after it
            self.end = self.meta_start
            self.meta_start = None
        for fid, cd in d.filters:
            raw = _shuffle(raw, cd[0]) if fid == _FILTER_SHUFFLE else zlib.compress(raw, cd[0])
        d.chunks[tuple(offset)] = (self._append(raw), len(raw))
        d.shape[0] = max(d.shape[0], offset[0] + d.chunk[0])

    # ------------------------------------------------------------ metadata
    def _chunk_btree(self, d):
        """Version 1 B-tree (node type 1) over the dataset's chunks."""
        rank = len(d.chunk)
        ksize = 8 + 8 * (rank + 1)
        nsize = 24 + (2 * _CHUNK_K + 1) * ksize + 2 * _CHUNK_K * 8

        def key(size, offset):
            return struct.pack("<II", size, 0) + b"".join(struct.pack("<Q", o) for o in offset) + struct.pack("<Q", 0)

        items = sorted(d.chunks.items())
        if not items:
            return UNDEF  # no chunk written yet: no index
        last = items[-1][0]
        end_key = key(0, [last[i] + d.chunk[i] for i in range(rank)])
        # level 0: (left key, child address); the key after a node's last
        # child is the next node's first key, or end_key
        level = [(key(size, off), addr) for off, (addr, size) in items]
        lev = 0
        while True:
            groups = [level[i:i + 2 * _CHUNK_K] for i in range(0, len(level), 2 * _CHUNK_K)]
            base = self.end
            addrs = [base + i * nsize for i in range(len(groups))]
            nxt = []
            for gi, grp in enumerate(groups):
                right = groups[gi + 1][0][0] if gi + 1 < len(groups) else end_key
                left_sib = addrs[gi - 1] if gi > 0 else UNDEF
                right_sib = addrs[gi + 1] if gi + 1 < len(groups) else UNDEF
                body = b"TREE" + struct.pack("<BBHQQ", 1, lev, len(grp), left_sib, right_sib)
                for k, child in grp:
                    body += k + struct.pack("<Q", child)
                body += right
                body += b"\0" * (nsize - len(body))
                a = self._append(body)
                assert a == addrs[gi]
                nxt.append((grp[0][0], a))
            if len(nxt) == 1:
                return nxt[0][1]
            level = nxt
            lev += 1

    def _write_dataset(self, d):
        bt = self._chunk_btree(d)
        maxshape = None
        if d.maxshape is not None:
            maxshape = list(d.maxshape)
        msgs = [
            _dataspace_message(d.shape, maxshape),
            _dtype_message(d.dtype),
            _layout_message(bt, d.chunk, d.dtype.itemsize),
        ]
        if d.filters:
            msgs.append(_pipeline_message(d.filters))
        return self._append(_object_header(msgs))

    def _write_group(self, g):
        """Old-style group: local heap of names, symbol-table nodes, a version
        1 B-tree (node type 0) over them, and an object header holding the
        symbol-table message. Returns (object header, B-tree, heap)."""
        entries = {}
        for name, child in g.children.items():
            if isinstance(child, _Group):
                oh, bt, hp = self._write_group(child)
                entries[name] = ("group", oh, bt, hp)
            elif isinstance(child, _Dataset):
                entries[name] = ("dataset", self._write_dataset(child))
            else:
                entries[name] = ("soft", child[1])
        # local heap: "" at offset 0, then the names and soft-link values
        heap = bytearray(b"\0" * 8)
        off = {}

        def put(s):
            o = len(heap)
            heap.extend(_pad8(s.encode() + b"\0"))
            return o

        names = sorted(entries, key=lambda s: s.encode())
        for n in names:
            off[n] = put(n)
        soft_off = {n: put(entries[n][1]) for n in names if entries[n][0] == "soft"}
        heap_data = self._append(bytes(heap))
        heap_addr = self._append(b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), _HEAP_FREE_NULL, heap_data))

        def entry(n):
            e = entries[n]
            if e[0] == "group":
                return struct.pack("<QQI4xQQ", off[n], e[1], 1, e[2], e[3])
            if e[0] == "dataset":
                return struct.pack("<QQI4x16x", off[n], e[1], 0)
            return struct.pack("<QQI4xI12x", off[n], UNDEF, 2, soft_off[n])

        cap = 2 * _GROUP_LEAF_K
        snods = [names[i:i + cap] for i in range(0, len(names), cap)]
        level = []
        for sn in snods:
            body = b"SNOD" + struct.pack("<BBH", 1, 0, len(sn)) + b"".join(entry(n) for n in sn)
            body += b"\0" * (8 + cap * 40 - len(body))
            level.append((off[sn[0]], off[sn[-1]], self._append(body)))
        nsize = 24 + (2 * _GROUP_INTERNAL_K + 1) * 8 + 2 * _GROUP_INTERNAL_K * 8
        lev = 0
        while True:
            # children (first name, last name, address); keys: "" then the
            # last name of each child
            groups = [level[i:i + 2 * _GROUP_INTERNAL_K] for i in range(0, len(level), 2 * _GROUP_INTERNAL_K)] or [[]]
            base = self.end
            addrs = [base + i * nsize for i in range(len(groups))]
            nxt = []
            for gi, grp in enumerate(groups):
                left_sib = addrs[gi - 1] if gi > 0 else UNDEF
                right_sib = addrs[gi + 1] if gi + 1 < len(groups) else UNDEF
                first_key = groups[gi - 1][-1][1] if gi > 0 else 0
                body = b"TREE" + struct.pack("<BBHQQ", 0, lev, len(grp), left_sib, right_sib) + struct.pack("<Q", first_key)
                for first, last, child in grp:
                    body += struct.pack("<QQ", child, last)
                body += b"\0" * (nsize - len(body))
                a = self._append(body)
                assert a == addrs[gi]
                nxt.append((grp[0][0] if grp else 0, grp[-1][1] if grp else 0, a))
            if len(nxt) == 1:
                btree = nxt[0][2]
                break
            level = nxt
            lev += 1
        oh = self._append(_object_header([_msg(0x11, struct.pack("<QQ", btree, heap_addr))]))
        return oh, btree, heap_addr

    def flush(self):
        """Write the metadata and the superblock; the file is then complete
        (later chunks and another flush append a new metadata set)."""
        if self.meta_start is not None:
            self.end = self.meta_start
        self.meta_start = self.end
        oh, bt, hp = self._write_group(self.root)
        eof = self.end
        self.f.truncate(eof)
        sb = _SIG + struct.pack("<8B", 0, 0, 0, 0, 0, 8, 8, 0) + struct.pack("<HHI", _GROUP_LEAF_K, _GROUP_INTERNAL_K, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
        sb += struct.pack("<QQI4xQQ", 0, oh, 1, bt, hp)
        assert len(sb) == 96
        self.f.seek(0)
        self.f.write(sb)
        self.f.flush()

    def close(self):
        if self.f is None:
            return
        self.flush()
        self.f.close()
        self.f = None


class Output:
    """The multigrid binary's output file (BaseSystem::new, system.rs:28-53;
    its groups and datasets, :87-117; System::output, :591-616).

    names: the grid names; grids: their (x, y) coordinate arrays of shape
    (ny, nx). Each `output(ntime, fields)` appends the step index to /t and
    one (ny, nx) slice of rho, rhou, rhov, e per grid (fields[i] of shape
    (4, ny, nx), the system's state). The file is complete after every
    call; `close()` ends it.
    """

    def __init__(self, path, names, grids):
        self.w = Writer(os.fspath(path))
        self.names = list(names)
        self.shapes = []
        self.ntimes = 0
        self.w.create_dataset("t", np.uint64, (0,), (1,), maxshape=(None,))
        for name, (x, y) in zip(self.names, grids):
            x = np.asarray(x, dtype=np.float64)
            y = np.asarray(y, dtype=np.float64)
            ny, nx = x.shape
            self.shapes.append((ny, nx))
            self.w.create_group(name)
            self.w.link_soft("/t", f"{name}/t")
            for dim, v in (("x", x), ("y", y)):
                self.w.create_dataset(f"{name}/{dim}", np.float64, (ny, nx), (ny, nx), deflate=9)
                self.w.write_chunk(f"{name}/{dim}", (0, 0), v)
            for var in ("rho", "rhou", "rhov", "e"):
                self.w.create_dataset(f"{name}/{var}", np.float64, (0, ny, nx), (1, ny, nx), maxshape=(None, ny, nx),
                                      shuffle=True, deflate=3)
        self.w.flush()

    def output(self, ntime, fields):
        tpos = self.ntimes
        self.w.write_chunk("t", (tpos,), np.array([ntime], dtype=np.uint64))
        for name, (ny, nx), q in zip(self.names, self.shapes, fields):
            q = np.asarray(q, dtype=np.float64)
            if q.shape != (4, ny, nx):
                raise ValueError(f"field of grid {name}: shape {q.shape}, expected {(4, ny, nx)}")
            for c, var in enumerate(("rho", "rhou", "rhov", "e")):
                self.w.write_chunk(f"{name}/{var}", (tpos, 0, 0), q[c])
        self.ntimes += 1
        self.w.flush()

    def close(self):
        self.w.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
