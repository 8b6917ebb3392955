"""HDF5 datasets of the reference's initial-condition and checkpoint files.

The reference reads an initial state ("IC", or a path such as "fields/IC")
and per-cell permeability/porosity ("Permeability", "PermeabilityX/Y/Z",
"Porosity") from HDF5 with PETSc's VecLoad (src/tdyio.c:176-330), in the
mesh's natural cell order.  This module reads such files without the HDF5
library: superblock versions 0/1, version-1 object headers with
continuation blocks, symbol-table groups (v1 B-trees, local heaps), and
contiguous, compact or chunked (unfiltered) datasets of IEEE floats and
integers of either byte order.  That is the layout h5py and PETSc write by
default; anything else raises ``ValueError``.

``write_checkpoint`` writes the reference's restart file (TDyIOOutputCheckpoint,
src/tdyio.c:339-359): the state in natural cell order as dataset "fields/IC"
and the step count as the root attribute "time step", in the same classic
layout (superblock 0, symbol-table groups, a contiguous dataset), so that
``-ic_file <file> -ic_dataset fields/IC`` restarts from it.
"""
from __future__ import annotations

import struct

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF


class _File:
    def __init__(self, data: bytes):
        self.d = data
        off = data.find(_SIG)
        if off != 0:
            raise ValueError("not an HDF5 file (no signature at offset 0)")
        ver = data[8]
        if ver not in (0, 1):
            raise ValueError(f"HDF5 superblock version {ver} is not supported")
        self.so, self.sl = data[13], data[14]
        if self.so != 8 or self.sl != 8:
            raise ValueError("only 8-byte offsets and lengths are supported")
        p = 24 if ver == 0 else 28
        self.base = self.u(p, 8)
        p += 32  # base, free-space, end-of-file, driver addresses
        # root group symbol table entry
        self.root_hdr = self.u(p + 8, 8)
        ctype = self.u(p + 16, 4)
        self.root_stab = (self.u(p + 24, 8), self.u(p + 32, 8)) if ctype == 1 else None

    def u(self, p, n):
        return int.from_bytes(self.d[p:p + n], "little")

    # ---- object headers -------------------------------------------------
    def messages(self, addr):
        d = self.d
        if d[addr:addr + 4] == b"OHDR":
            raise ValueError("version-2 object headers are not supported")
        if d[addr] != 1:
            raise ValueError(f"object header version {d[addr]} is not supported")
        nmsg = self.u(addr + 2, 2)
        size = self.u(addr + 8, 4)
        blocks = [(addr + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            p, n = blocks.pop(0)
            end = p + n
            while p + 8 <= end and len(out) < nmsg:
                mtype, msize = self.u(p, 2), self.u(p + 2, 2)
                body = p + 8
                if mtype == 0x10:  # continuation
                    blocks.append((self.u(body, 8), self.u(body + 8, 8)))
                out.append((mtype, body, msize))
                p = body + msize
        return out

    # ---- groups -----------------------------------------------------------
    def heap_name(self, heap, off):
        if self.d[heap:heap + 4] != b"HEAP":
            raise ValueError("bad local heap")
        seg = self.u(heap + 24, 8)
        s = seg + off
        e = self.d.index(b"\0", s)
        return self.d[s:e].decode()

    def group_entries(self, btree, heap):
        d = self.d
        out = {}
        stack = [btree]
        while stack:
            node = stack.pop()
            if d[node:node + 4] == b"SNOD":
                n = self.u(node + 6, 2)
                for i in range(n):
                    e = node + 8 + 40 * i
                    name = self.heap_name(heap, self.u(e, 8))
                    ctype = self.u(e + 16, 4)
                    stab = (self.u(e + 24, 8), self.u(e + 32, 8)) if ctype == 1 else None
                    out[name] = (self.u(e + 8, 8), stab)
                continue
            if d[node:node + 4] != b"TREE" or d[node + 4] != 0:
                raise ValueError("bad group B-tree node")
            used = self.u(node + 6, 2)
            p = node + 24 + 8  # first key
            for i in range(used):
                stack.append(self.u(p, 8))
                p += 16
        return out

    def group_of(self, hdr, stab):
        if stab is None:
            for mtype, body, _ in self.messages(hdr):
                if mtype == 0x11:
                    stab = (self.u(body, 8), self.u(body + 8, 8))
            if stab is None:
                raise ValueError("not a symbol-table group")
        return self.group_entries(*stab)

    def lookup(self, path):
        hdr, stab = self.root_hdr, self.root_stab
        parts = [p for p in path.split("/") if p]
        for i, name in enumerate(parts):
            ents = self.group_of(hdr, stab)
            if name not in ents:
                raise KeyError(f"{path}: no member '{name}'")
            hdr, stab = ents[name]
        return hdr

    def names(self, path=""):
        hdr, stab = self.root_hdr, self.root_stab
        if path:
            hdr, stab = self.lookup(path), None
        return sorted(self.group_of(hdr, stab))

    # ---- datasets ---------------------------------------------------------
    def dataset(self, hdr):
        shape = dtype = layout = None
        filtered = False
        for mtype, body, msize in self.messages(hdr):
            if mtype == 0x01:
                shape = self._dataspace(body)
            elif mtype == 0x03:
                dtype = self._datatype(body)
            elif mtype == 0x08:
                layout = self._layout(body)
            elif mtype == 0x0B:
                filtered = True
        if shape is None or dtype is None or layout is None:
            raise ValueError("object is not a dataset")
        if filtered:
            raise ValueError("filtered (compressed) datasets are not supported")
        count = int(np.prod(shape)) if shape else 1
        kind = layout[0]
        if kind == "compact":
            raw = layout[1]
            return np.frombuffer(raw, dtype=dtype, count=count).reshape(shape).astype(dtype.newbyteorder("="))
        if kind == "contiguous":
            addr = layout[1]
            if addr == _UNDEF:
                return np.zeros(shape, dtype=dtype.newbyteorder("="))
            a = np.frombuffer(self.d, dtype=dtype, count=count, offset=self.base + addr)
            return a.reshape(shape).astype(dtype.newbyteorder("="))
        _, btree, cdims = kind, layout[1], layout[2]
        out = np.zeros(shape, dtype=dtype.newbyteorder("="))
        nd = len(shape)
        for offs, addr, nbytes in self._chunks(btree, nd):
            c = np.frombuffer(self.d, dtype=dtype, count=int(np.prod(cdims)), offset=self.base + addr)
            c = c.reshape(cdims)
            sl = tuple(slice(o, min(o + cd, s)) for o, cd, s in zip(offs, cdims, shape))
            out[sl] = c[tuple(slice(0, s.stop - s.start) for s in sl)]
        return out

    def _dataspace(self, p):
        ver, nd, flags = self.d[p], self.d[p + 1], self.d[p + 2]
        q = p + (8 if ver == 1 else 4)
        return tuple(self.u(q + 8 * i, 8) for i in range(nd))

    def _datatype(self, p):
        cls, ver = self.d[p] & 0x0F, self.d[p] >> 4
        bits = self.d[p + 1]
        size = self.u(p + 4, 4)
        order = ">" if bits & 1 else "<"
        if cls == 1 and size in (4, 8):
            return np.dtype(f"{order}f{size}")
        if cls == 0 and size in (1, 2, 4, 8):
            return np.dtype(f"{order}{'i' if bits & 8 else 'u'}{size}")
        raise ValueError(f"datatype class {cls} of size {size} is not supported")

    def _layout(self, p):
        ver = self.d[p]
        if ver == 3:
            cls = self.d[p + 1]
            if cls == 0:
                n = self.u(p + 2, 2)
                return ("compact", self.d[p + 4:p + 4 + n])
            if cls == 1:
                return ("contiguous", self.u(p + 2, 8))
            if cls == 2:
                nd = self.d[p + 2]
                btree = self.u(p + 3, 8)
                dims = tuple(self.u(p + 11 + 4 * i, 4) for i in range(nd - 1))
                return ("chunked", btree, dims)
        elif ver in (1, 2):
            nd, cls = self.d[p + 1], self.d[p + 2]
            q = p + 8
            if cls == 1:
                return ("contiguous", self.u(q, 8))
            if cls == 2:
                btree = self.u(q, 8)
                dims = tuple(self.u(q + 8 + 4 * i, 4) for i in range(nd - 1))
                return ("chunked", btree, dims)
        raise ValueError(f"data layout version {ver} is not supported")

    def _chunks(self, node, nd):
        d = self.d
        if d[node:node + 4] != b"TREE" or d[node + 4] != 1:
            raise ValueError("bad chunk B-tree node")
        level, used = d[node + 5], self.u(node + 6, 2)
        ksz = 8 + 8 * (nd + 1)
        p = node + 24
        for i in range(used):
            nbytes = self.u(p, 4)
            mask = self.u(p + 4, 4)
            offs = tuple(self.u(p + 8 + 8 * k, 8) for k in range(nd))
            child = self.u(p + ksz, 8)
            if level > 0:
                yield from self._chunks(child, nd)
            else:
                if mask:
                    raise ValueError("filtered chunks are not supported")
                yield offs, child, nbytes
            p += ksz + 8


def read_dataset(path: str, name: str) -> np.ndarray:
    """The dataset ``name`` (a path such as "IC" or "fields/IC") of an HDF5 file."""
    f = _File(open(path, "rb").read())
    return f.dataset(f.lookup(name))


def list_members(path: str, group: str = "") -> list:
    """Member names of a group ("" for the root group)."""
    f = _File(open(path, "rb").read())
    return f.names(group)


# ---- writing --------------------------------------------------------------

def checkpoint_filename(time: float) -> str:
    """The reference's checkpoint file name for a simulation time
    (``"%11.5e_chk.h5"``, src/tdyio.c:348)."""
    return "%11.5e_%s.h5" % (time, "chk")


class _Writer:
    """Bump allocator over a bytearray, 8-byte aligned."""

    def __init__(self):
        self.b = bytearray()

    def alloc(self, n):
        a = len(self.b)
        n = (n + 7) & ~7
        self.b += bytes(n)
        return a

    def put(self, addr, fmt, *vals):
        struct.pack_into("<" + fmt, self.b, addr, *vals)


_LEAF_K, _NODE_K = 4, 16
_BTREE_SIZE = 24 + (2 * _NODE_K + 1) * 8 + 2 * _NODE_K * 8
_SNOD_SIZE = 8 + 2 * _LEAF_K * 40


def _pad8(b: bytes) -> bytes:
    return b + bytes((-len(b)) % 8)


def _msg(mtype: int, body: bytes) -> bytes:
    body = _pad8(body)
    return struct.pack("<HHB3x", mtype, len(body), 0) + body


def _object_header(w: _Writer, msgs) -> int:
    body = b"".join(msgs)
    a = w.alloc(16 + len(body))
    w.put(a, "BBHII", 1, 0, len(msgs), 1, len(body))
    w.b[a + 16:a + 16 + len(body)] = body
    return a


def _group(w: _Writer, members) -> tuple:
    """A symbol-table group of ``members`` = [(name, header address,
    (btree, heap) or None)]; returns (B-tree address, heap address)."""
    members = sorted(members)
    if len(members) > 2 * _LEAF_K:
        raise ValueError("too many members for one symbol-table node")
    names = b"\0" * 8
    offs = []
    for nm, _, _ in members:
        offs.append(len(names))
        names += _pad8(nm.encode() + b"\0")
    seg_size = len(names) + 16  # trailing free block, as the HDF5 library leaves one
    bt = w.alloc(_BTREE_SIZE)
    heap = w.alloc(32)
    seg = w.alloc(seg_size)
    snod = w.alloc(_SNOD_SIZE)
    w.b[bt:bt + 8] = b"TREE" + bytes([0, 0]) + struct.pack("<H", 1)
    w.put(bt + 8, "QQ", _UNDEF, _UNDEF)
    w.put(bt + 24, "QQQ", 0, snod, offs[-1] if offs else 0)
    w.b[heap:heap + 8] = b"HEAP" + bytes(4)
    w.put(heap + 8, "QQQ", seg_size, len(names), seg)
    w.b[seg:seg + len(names)] = names
    w.put(seg + len(names), "QQ", 1, 16)  # free block: no next block (1), 16 bytes
    w.b[snod:snod + 8] = b"SNOD" + bytes([1, 0]) + struct.pack("<H", len(members))
    for i, ((nm, hdr, stab), off) in enumerate(zip(members, offs)):
        e = snod + 8 + 40 * i
        w.put(e, "QQII", off, hdr, 1 if stab else 0, 0)
        if stab:
            w.put(e + 24, "QQ", *stab)
    return bt, heap


_F64_LE = bytes.fromhex("11203f000800000000004000340b0034ff030000")
_I32_LE = struct.pack("<BBBBIHH", 0x10, 0x08, 0, 0, 4, 0, 32)


def write_checkpoint(path: str, values, time_step: int) -> None:
    """Write ``values`` (the state in natural cell order) as "fields/IC" and
    ``time_step`` as the root attribute "time step"."""
    v = np.ascontiguousarray(np.asarray(values, dtype="<f8").ravel())
    n = v.shape[0]
    w = _Writer()
    w.alloc(96)  # superblock and root symbol-table entry
    # the dataset
    space = struct.pack("<BBBB4xQQ", 1, 1, 1, 0, n, n)
    fill = bytes([2, 2, 2, 1]) + struct.pack("<I", 0)
    ds = _object_header(w, [_msg(0x01, space), _msg(0x03, _F64_LE), _msg(0x05, fill),
                            _msg(0x08, struct.pack("<BBQQ", 3, 1, 0, n * 8))])
    # group "fields"
    bt, heap = _group(w, [("IC", ds, None)])
    fields = _object_header(w, [_msg(0x11, struct.pack("<QQ", bt, heap))])
    # root group with the attribute
    rbt, rheap = _group(w, [("fields", fields, (bt, heap))])
    name = b"time step\0"
    scalar = struct.pack("<BBBB4x", 1, 0, 0, 0)
    attr = (struct.pack("<BBHHH", 1, 0, len(name), len(_I32_LE), len(scalar)) + _pad8(name) + _pad8(_I32_LE)
            + _pad8(scalar) + struct.pack("<i", int(time_step)))
    root = _object_header(w, [_msg(0x11, struct.pack("<QQ", rbt, rheap)), _msg(0x0C, attr)])
    data_addr = w.alloc(n * 8)
    w.b[data_addr:data_addr + n * 8] = v.tobytes()
    # the dataset's layout message points at the data
    lay = ds + 16 + (8 + 24) + (8 + 24) + (8 + 8) + 8
    w.put(lay + 2, "Q", data_addr)
    # superblock
    w.b[0:8] = _SIG
    w.put(8, "BBBBBBBB", 0, 0, 0, 0, 0, 8, 8, 0)
    w.put(16, "HHI", _LEAF_K, _NODE_K, 0)
    w.put(24, "QQQQ", 0, _UNDEF, len(w.b), _UNDEF)
    w.put(56, "QQII", 0, root, 1, 0)
    w.put(80, "QQ", rbt, rheap)
    with open(path, "wb") as fh:
        fh.write(bytes(w.b))


def read_attribute(path: str, name: str, obj: str = ""):
    """A scalar integer or float attribute of an object ("" = root group)."""
    f = _File(open(path, "rb").read())
    hdr = f.root_hdr if not obj else f.lookup(obj)
    for mtype, body, _ in f.messages(hdr):
        if mtype != 0x0C:
            continue
        ver = f.d[body]
        if ver != 1:
            raise ValueError(f"attribute message version {ver} is not supported")
        nsz, tsz, ssz = f.u(body + 2, 2), f.u(body + 4, 2), f.u(body + 6, 2)
        p = body + 8
        nm = f.d[p:p + nsz].split(b"\0")[0].decode()
        p += (nsz + 7) & ~7
        dt = f._datatype(p)
        p += (tsz + 7) & ~7
        p += (ssz + 7) & ~7
        if nm == name:
            return np.frombuffer(f.d, dtype=dt, count=1, offset=p)[0].item()
    raise KeyError(f"{obj or '/'}: no attribute '{name}'")
