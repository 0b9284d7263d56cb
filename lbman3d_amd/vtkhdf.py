"""VTK-HDF ImageData frames of a simulation, in the reference's layout.

The reference writes one file per save, ``<dir>/lbm_<step>.vtkhdf``, through
libhdf5 (SimIO::ExportVTKHDF, src/sim_io.cc:168-211; ImageDataWriter,
src/vtkhdf_writer.h:115-160):

    /VTKHDF                    attributes Type = "ImageData" (fixed-length,
                               NULL-padded ASCII), Version = int64 {2, 0},
                               WholeExtent = int64 {0, nx-1, 0, ny-1, 0, nz-1},
                               Origin = {0, 0, 0}, Spacing = {1, 1, 1},
                               Direction = identity (9 doubles)
    /VTKHDF/PointData/rho      float64 [nz, ny, nx]
    /VTKHDF/PointData/f0..f14  float64 [nz, ny, nx]   populations, debug frames
                                                      only (Params::kDebugLogging)
    /VTKHDF/PointData/order    float64 [nz, ny, nx]   largest eigenvalue of Q
    /VTKHDF/PointData/ux|uy|uz float64 [nz, ny, nx]
    /VTKHDF/PointData/director float64 [nz, ny, nx, 3]

libhdf5 is not part of this build's environment, so the file is written by
``H5File`` below: a self-contained writer of the subset of the HDF5 file
format (HDF Group, "HDF5 File Format Specification", version 0 superblock)
that this layout needs: a version-0 superblock, symbol-table groups (local
heap, version-1 B-tree of one leaf, one symbol-table node), version-1 object
headers, contiguous datasets and attributes of fixed-point, IEEE double and
fixed-length string type. ``read_h5`` is an independent reader of the same
subset used by the tests, which also read every file with the HDF Group's
h5dump where the machine has it.
"""
from __future__ import annotations

import os
import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
SIG = b"\x89HDF\r\n\x1a\n"
# A symbol-table node holds 2*LEAF_K entries, and every group here is one
# node: K = 16 (libhdf5's default is 4; the superblock records the value)
# leaves room for the 21 members of a debug frame's PointData group.
LEAF_K, INTERNAL_K = 16, 16
FREE_NULL = 1                       # end of a local heap's free list on disk


def _pad8(b: bytes) -> bytes:
    return b + b"\0" * (-len(b) % 8)


# ---- datatype and dataspace messages ---------------------------------------

def _dtype_msg(kind: str, size: int) -> bytes:
    """Datatype message, version 1 (class in the low nibble)."""
    if kind == "f8":
        bits = bytes([0x20, 63, 0])                 # LE, implied-MSB mantissa, sign at bit 63
        props = struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
        return bytes([0x10 | 1]) + bits + struct.pack("<I", 8) + props
    if kind in ("i8", "u1"):
        signed = kind == "i8"
        bits = bytes([0x08 if signed else 0x00, 0, 0])
        return bytes([0x10 | 0]) + bits + struct.pack("<I", size) + struct.pack("<HH", 0, 8 * size)
    if kind == "str":
        return bytes([0x10 | 3]) + bytes([0x01, 0, 0]) + struct.pack("<I", size)   # NULL-pad, ASCII
    raise ValueError(kind)


def _dspace_msg(shape) -> bytes:
    """Dataspace message, version 1 (rank 0 is a scalar)."""
    return struct.pack("<BBBB4x", 1, len(shape), 0, 0) + b"".join(struct.pack("<Q", d) for d in shape)


def _kind_of(arr: np.ndarray) -> str:
    if arr.dtype == np.float64:
        return "f8"
    if arr.dtype == np.int64:
        return "i8"
    if arr.dtype == np.uint8:
        return "u1"
    raise TypeError(f"unsupported dtype {arr.dtype}")


def _attr_msg(name: str, value) -> bytes:
    """Attribute message, version 1."""
    nm = name.encode() + b"\0"
    if isinstance(value, str):
        raw = value.encode("ascii")
        dt, ds, data = _dtype_msg("str", len(raw)), _dspace_msg(()), raw
    else:
        arr = np.ascontiguousarray(value)
        arr = arr.astype(arr.dtype.newbyteorder("<"))
        dt, ds, data = _dtype_msg(_kind_of(arr), arr.itemsize), _dspace_msg(arr.shape), arr.tobytes()
    head = struct.pack("<BBHHH", 1, 0, len(nm), len(dt), len(ds))
    return head + _pad8(nm) + _pad8(dt) + _pad8(ds) + data


def _message(mtype: int, body: bytes, flags: int = 0) -> bytes:
    body = _pad8(body)
    return struct.pack("<HHB3x", mtype, len(body), flags) + body


def _object_header(messages) -> bytes:
    data = b"".join(messages)
    return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(data)) + data


class Node:
    def __init__(self, attrs=None):
        self.attrs = dict(attrs or {})


class Group(Node):
    def __init__(self, attrs=None):
        super().__init__(attrs)
        self.children: dict[str, Node] = {}

    def group(self, name, attrs=None) -> "Group":
        g = Group(attrs)
        self.children[name] = g
        return g

    def dataset(self, name, data, attrs=None) -> "Dataset":
        d = Dataset(data, attrs)
        self.children[name] = d
        return d


class Dataset(Node):
    def __init__(self, data, attrs=None):
        super().__init__(attrs)
        arr = np.ascontiguousarray(data)
        self.data = arr.astype(arr.dtype.newbyteorder("<"))


class H5File:
    """Build a tree of groups, datasets and attributes in memory, then
    ``save(path)`` writes it as one HDF5 file."""

    def __init__(self):
        self.root = Group()

    def save(self, path):
        buf = bytearray()

        def alloc(n):
            off = len(buf)
            buf.extend(b"\0" * (n + (-n % 8)))
            return off

        alloc(96)                                   # superblock with root symbol-table entry
        root_hdr, root_btree, root_heap = self._write_group(self.root, buf, alloc)
        eof = len(buf)
        sb = bytearray()
        sb += SIG + bytes([0, 0, 0, 0, 0, 8, 8, 0])
        sb += struct.pack("<HHI", LEAF_K, INTERNAL_K, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
        sb += struct.pack("<QQII", 0, root_hdr, 1, 0) + struct.pack("<QQ", root_btree, root_heap)
        assert len(sb) == 96
        buf[0:96] = sb
        tmp = str(path) + ".tmp"
        with open(tmp, "wb") as fp:
            fp.write(bytes(buf))
        os.replace(tmp, path)

    # Returns (object header address, B-tree address, heap address).
    def _write_group(self, g: Group, buf, alloc):
        names = sorted(g.children)
        if len(names) > 2 * LEAF_K:
            raise ValueError(f"a group holds at most {2 * LEAF_K} members in this writer")
        # children first (their object headers must exist before the table)
        child = {}
        for nm in names:
            node = g.children[nm]
            if isinstance(node, Group):
                child[nm] = self._write_group(node, buf, alloc)
            else:
                child[nm] = (self._write_dataset(node, buf, alloc), None, None)
        # local heap: "" at offset 0, then the names, then one free block
        seg = bytearray(b"\0" * 8)
        offs = {}
        for nm in names:
            offs[nm] = len(seg)
            seg += _pad8(nm.encode() + b"\0")
        free_off = len(seg)
        seg += struct.pack("<QQ", FREE_NULL, 16)
        seg_addr = alloc(len(seg))
        buf[seg_addr:seg_addr + len(seg)] = seg
        heap = alloc(32)
        buf[heap:heap + 32] = b"HEAP" + bytes([0, 0, 0, 0]) + struct.pack("<QQQ", len(seg), free_off, seg_addr)
        # one symbol-table node with every member, sorted by name
        snod_size = 8 + 2 * LEAF_K * 40
        snod = alloc(snod_size)
        ent = bytearray(b"SNOD" + bytes([1, 0]) + struct.pack("<H", len(names)))
        for nm in names:
            hdr, bt, hp = child[nm]
            if bt is None:
                ent += struct.pack("<QQII16x", offs[nm], hdr, 0, 0)
            else:
                ent += struct.pack("<QQIIQQ", offs[nm], hdr, 1, 0, bt, hp)
        buf[snod:snod + len(ent)] = ent
        # version-1 B-tree, one leaf pointing at the node
        bt_size = 24 + (2 * INTERNAL_K + 1) * 8 + 2 * INTERNAL_K * 8
        bt = alloc(bt_size)
        last = offs[names[-1]] if names else 0
        node = b"TREE" + bytes([0, 0]) + struct.pack("<HQQ", 1 if names else 0, UNDEF, UNDEF)
        node += struct.pack("<QQQ", 0, snod, last) if names else b""
        buf[bt:bt + len(node)] = node
        msgs = [_message(0x0011, struct.pack("<QQ", bt, heap))]
        msgs += [_message(0x000C, _attr_msg(k, v)) for k, v in g.attrs.items()]
        oh = _object_header(msgs)
        hdr = alloc(len(oh))
        buf[hdr:hdr + len(oh)] = oh
        return hdr, bt, heap

    def _write_dataset(self, d: Dataset, buf, alloc):
        raw = d.data.tobytes()
        addr = alloc(len(raw))
        buf[addr:addr + len(raw)] = raw
        msgs = [_message(0x0001, _dspace_msg(d.data.shape)),
                _message(0x0003, _dtype_msg(_kind_of(d.data), d.data.itemsize), flags=1),
                _message(0x0008, struct.pack("<BBQQ", 3, 1, addr, len(raw)))]
        msgs += [_message(0x000C, _attr_msg(k, v)) for k, v in d.attrs.items()]
        oh = _object_header(msgs)
        hdr = alloc(len(oh))
        buf[hdr:hdr + len(oh)] = oh
        return hdr


# ---- reader of the same subset ----------------------------------------------

def _parse_dtype(b: bytes):
    cls, ver = b[0] & 0x0F, b[0] >> 4
    bits = b[1:4]
    size = struct.unpack_from("<I", b, 4)[0]
    if ver != 1:
        raise ValueError("datatype version")
    if cls == 1:
        off, prec, eloc, esz, mloc, msz, bias = struct.unpack_from("<HHBBBBI", b, 8)
        if (size, prec, eloc, esz, msz, bias, bits[1]) != (8, 64, 52, 11, 52, 1023, 63) or bits[0] & 1:
            raise ValueError("only little-endian IEEE double")
        return np.dtype("<f8"), 20
    if cls == 0:
        signed = bool(bits[0] & 0x08)
        return np.dtype(("<i" if signed else "<u") + str(size)), 12
    if cls == 3:
        return ("str", size), 8
    raise ValueError(f"datatype class {cls}")


def _parse_dspace(b: bytes):
    ver, rank = b[0], b[1]
    if ver != 1:
        raise ValueError("dataspace version")
    return tuple(struct.unpack_from("<Q", b, 8 + 8 * i)[0] for i in range(rank))


def read_h5(path) -> dict:
    """Tree of a file written by H5File: {"attrs": {...}, "children": {...}}
    for groups, {"attrs": {...}, "data": ndarray} for datasets."""
    data = open(path, "rb").read()
    if data[:8] != SIG or data[8] != 0:
        raise ValueError("not a version-0 HDF5 superblock")
    if data[13] != 8 or data[14] != 8:
        raise ValueError("offset/length sizes")
    leaf_k, internal_k = struct.unpack_from("<HH", data, 16)
    eof = struct.unpack_from("<Q", data, 40)[0]
    if eof != len(data):
        raise ValueError("end-of-file address")
    _, root_hdr, cache, _ = struct.unpack_from("<QQII", data, 56)
    return _read_object(data, root_hdr, leaf_k)


def _read_object(data, addr, leaf_k):
    ver, _, nmsg, _, size = struct.unpack_from("<BBHII", data, addr)
    if ver != 1:
        raise ValueError("object header version")
    p, end = addr + 16, addr + 16 + size
    attrs, node, shape, dtype, layout = {}, {}, None, None, None
    stab = None
    for _ in range(nmsg):
        mtype, msize, _flags = struct.unpack_from("<HHB", data, p)
        body = data[p + 8:p + 8 + msize]
        p += 8 + msize
        if mtype == 0x0011:
            stab = struct.unpack_from("<QQ", body, 0)
        elif mtype == 0x0001:
            shape = _parse_dspace(body)
        elif mtype == 0x0003:
            dtype, _ = _parse_dtype(body)
        elif mtype == 0x0008:
            v, cls, a, n = struct.unpack_from("<BBQQ", body, 0)
            if v != 3 or cls != 1:
                raise ValueError("only contiguous layout")
            layout = (a, n)
        elif mtype == 0x000C:
            v, _, nlen, dlen, slen = struct.unpack_from("<BBHHH", body, 0)
            q = 8
            name = body[q:q + nlen - 1].decode()
            q += nlen + (-nlen % 8)
            adt, _ = _parse_dtype(body[q:q + dlen])
            q += dlen + (-dlen % 8)
            ashape = _parse_dspace(body[q:q + slen])
            q += slen + (-slen % 8)
            if isinstance(adt, tuple):
                attrs[name] = body[q:q + adt[1]].rstrip(b"\0").decode("ascii")
            else:
                cnt = int(np.prod(ashape)) if ashape else 1
                attrs[name] = np.frombuffer(body, dtype=adt, count=cnt, offset=q).reshape(ashape)
    if p != end:
        raise ValueError("object header size")
    if stab is not None:
        node["children"] = _read_group(data, stab[0], stab[1], leaf_k)
    else:
        a, n = layout
        cnt = int(np.prod(shape))
        if n != cnt * dtype.itemsize:
            raise ValueError("layout size")
        node["data"] = np.frombuffer(data, dtype=dtype, count=cnt, offset=a).reshape(shape)
    node["attrs"] = attrs
    return node


def _read_group(data, btree, heap, leaf_k):
    if data[heap:heap + 4] != b"HEAP":
        raise ValueError("local heap signature")
    seg_size, free_head, seg_addr = struct.unpack_from("<QQQ", data, heap + 8)
    seg = data[seg_addr:seg_addr + seg_size]
    if free_head != FREE_NULL and free_head >= seg_size:
        raise ValueError("local heap free list")

    def name_at(off):
        return seg[off:seg.index(b"\0", off)].decode()

    if data[btree:btree + 4] != b"TREE" or data[btree + 4] != 0 or data[btree + 5] != 0:
        raise ValueError("group B-tree leaf")
    used = struct.unpack_from("<H", data, btree + 6)[0]
    out = {}
    p = btree + 24
    prev_key = struct.unpack_from("<Q", data, p)[0]
    for _ in range(used):
        child, key = struct.unpack_from("<QQ", data, p + 8)
        p += 16
        if data[child:child + 4] != b"SNOD":
            raise ValueError("symbol-table node signature")
        nsym = struct.unpack_from("<H", data, child + 6)[0]
        if nsym > 2 * leaf_k:
            raise ValueError("symbol-table node overflow")
        names = []
        for i in range(nsym):
            off, hdr = struct.unpack_from("<QQ", data, child + 8 + 40 * i)
            nm = name_at(off)
            names.append(nm)
            out[nm] = _read_object(data, hdr, leaf_k)
        if names != sorted(names) or (names and name_at(key) != names[-1]) or name_at(prev_key) > (names[0] if names else ""):
            raise ValueError("symbol-table order / B-tree keys")
        prev_key = key
    return out


# ---- the reference's frames ---------------------------------------------------

def image_data_tree(nx, ny, nz, rho, ux, uy, uz, order, director, populations=None) -> H5File:
    """The ImageData frame of SimIO::ExportVTKHDF (fields in cell order
    (z*ny + y)*nx + x; director [V][3]). With populations (15 blocks of V,
    f_i of every cell) the frame also holds the datasets f0..f14 that the
    reference writes when Params::kDebugLogging is set (src/sim_io.cc:185-196)."""
    f = H5File()
    vt = f.root.group("VTKHDF", {
        "Type": "ImageData",
        "Version": np.array([2, 0], dtype=np.int64),
        "WholeExtent": np.array([0, nx - 1, 0, ny - 1, 0, nz - 1], dtype=np.int64),
        "Origin": np.zeros(3),
        "Spacing": np.ones(3),
        "Direction": np.eye(3).reshape(9),
    })
    pd = vt.group("PointData")
    shape = (nz, ny, nx)
    pd.dataset("rho", np.asarray(rho, dtype=np.float64).reshape(shape))
    if populations is not None:
        pops = np.asarray(populations, dtype=np.float64).reshape(15, -1)
        for i in range(15):
            pd.dataset(f"f{i}", pops[i].reshape(shape))
    for name, arr in (("order", order), ("ux", ux), ("uy", uy), ("uz", uz)):
        pd.dataset(name, np.asarray(arr, dtype=np.float64).reshape(shape))
    pd.dataset("director", np.asarray(director, dtype=np.float64).reshape(shape + (3,)))
    return f


def step_width(num_steps: int) -> int:
    """Digits of the step number in file names (kStepWidth, src/sim_io.cc:170-174)."""
    w, n = 1, num_steps - 1
    while n >= 10:
        n //= 10
        w += 1
    return w


def export_frame(sim, directory: str, frame: int, num_steps: int, debug: bool = False) -> str:
    """Write ``<directory>/lbm_<frame>.vtkhdf`` from a device solver's state;
    ``debug`` adds the populations f0..f14 (the post-streaming f of the last
    step, the host populations the reference's CPU path writes)."""
    cfg = sim.cfg
    st = sim.state()
    order, director = sim.order_director()
    pops = sim.populations() if debug else None
    f = image_data_tree(cfg.nx, cfg.ny, cfg.nz, st["rho"], st["u"][0], st["u"][1], st["u"][2], order, director,
                        pops)
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, f"lbm_{frame:0{step_width(num_steps)}d}.vtkhdf")
    f.save(path)
    return path
