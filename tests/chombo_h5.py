"""Writer of small HDF5 files in the layout of Chombo's checkpoint / plot
files, for the reader's tests (no HDF5 library in this image).

It writes what the HDF5 library writes by default for Chombo: a version 0
superblock, version 1 object headers, symbol-table groups (one B-tree leaf,
one symbol node, one local heap per group), contiguous datasets and
version 1 attribute messages; integer, IEEE double, fixed-length string and
compound datatypes.  A user block can be put in front, as some writers do.
"""
import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF


def pad8(b):
    return b + b"\0" * (-len(b) % 8)


# ------------------------------------------------------------- datatypes --
def t_int(size=4, signed=True):
    bits = 0x08 if signed else 0
    return struct.pack("<B3sI", 0x10, bits.to_bytes(3, "little"), size) + \
        struct.pack("<HH", 0, 8 * size)


def t_double():
    bits = 0x20 | (63 << 8)  # little endian, implied msb, sign at bit 63
    return struct.pack("<B3sI", 0x11, bits.to_bytes(3, "little"), 8) + \
        struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)


def t_string(n):
    return struct.pack("<B3sI", 0x13, b"\0\0\0", n)


def t_compound(members):
    """members: [(name, offset, member type bytes)], total size from the
    last member (4-byte ints only here)."""
    size = max(off for _, off, _ in members) + 4
    out = struct.pack("<B3sI", 0x16, len(members).to_bytes(3, "little"), size)
    for name, off, mt in members:
        out += pad8(name.encode() + b"\0")
        out += struct.pack("<IB3sI4s16s", off, 0, b"", 0, b"", b"")
        out += mt
    return out


def box_type(dim=3):
    names = ["lo_i", "lo_j", "lo_k"][:dim] + ["hi_i", "hi_j", "hi_k"][:dim]
    return t_compound([(n, 4 * i, t_int()) for i, n in enumerate(names)])


def intvect_type(dim=3):
    return t_compound([(n, 4 * i, t_int()) for i, n in enumerate(["intvecti", "intvectj",
                                                                   "intvectk"][:dim])])


def space(dims):
    return struct.pack("<BBB5s", 1, len(dims), 0, b"") + b"".join(struct.pack("<Q", d)
                                                                   for d in dims)


# ------------------------------------------------------------ attributes --
def attr(name, dtype, data, dims=()):
    nm = name.encode() + b"\0"
    sp = space(dims)
    return struct.pack("<BBHHH", 1, 0, len(nm), len(dtype), len(sp)) + \
        pad8(nm) + pad8(dtype) + pad8(sp) + data


def a_int(name, v):
    return attr(name, t_int(), struct.pack("<i", v))


def a_double(name, v):
    return attr(name, t_double(), struct.pack("<d", v))


def a_string(name, s):
    b = s.encode()
    return attr(name, t_string(len(b) + 1), b + b"\0")


def a_box(name, lo, hi):
    return attr(name, box_type(), struct.pack("<6i", *lo, *hi))


def a_intvect(name, v):
    return attr(name, intvect_type(), struct.pack("<3i", *v))


class Group:
    def __init__(self):
        self.attrs = []
        self.items = {}   # name -> Group or Dataset

    def group(self, name):
        g = Group()
        self.items[name] = g
        return g

    def dataset(self, name, dtype_bytes, dims, data):
        self.items[name] = Dataset(dtype_bytes, dims, data)


class Dataset:
    def __init__(self, dtype_bytes, dims, data):
        self.dtype, self.dims, self.data = dtype_bytes, dims, data
        self.attrs = []


class Writer:
    def __init__(self, userblock=0):
        self.ub = userblock
        self.buf = bytearray(b"\0" * (userblock + 96))

    def alloc(self, b):
        a = len(self.buf) - self.ub
        self.buf += pad8(bytes(b))
        return a

    def header(self, msgs):
        body = b""
        for t, d in msgs:
            d = pad8(d)
            body += struct.pack("<HHB3s", t, len(d), 0, b"") + d
        return self.alloc(struct.pack("<BBHII", 1, 0, len(msgs), 1, len(body)) + b"\0" * 4 + body)

    def write_dataset(self, ds):
        addr = self.alloc(ds.data) if ds.data else UNDEF
        layout = struct.pack("<BBQQ", 3, 1, addr, len(ds.data))
        msgs = [(0x01, space(ds.dims)), (0x03, ds.dtype), (0x08, layout)]
        msgs += [(0x0C, a) for a in ds.attrs]
        return self.header(msgs)

    def write_group(self, g):
        children = sorted((n, self.write_item(it)) for n, it in g.items.items())
        heap = b"\0" * 8
        offs = []
        for n, _ in children:
            offs.append(len(heap))
            heap += pad8(n.encode() + b"\0")
        heap_data = self.alloc(heap)
        heap_hdr = self.alloc(b"HEAP" + struct.pack("<B3sQQQ", 0, b"", len(heap), UNDEF,
                                                    heap_data))
        snod = b"SNOD" + struct.pack("<BBH", 1, 0, len(children))
        for (n, a), off in zip(children, offs):
            snod += struct.pack("<QQII16s", off, a, 0, 0, b"")
        snod_addr = self.alloc(snod)
        tree = b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF)
        tree += struct.pack("<QQQ", 0, snod_addr, offs[-1] if offs else 0)
        tree_addr = self.alloc(tree)
        msgs = [(0x11, struct.pack("<QQ", tree_addr, heap_hdr))]
        msgs += [(0x0C, a) for a in g.attrs]
        return self.header(msgs), tree_addr, heap_hdr

    def write_item(self, it):
        if isinstance(it, Group):
            return self.write_group(it)[0]
        return self.write_dataset(it)

    def finish(self, root, path):
        root_addr, tree, heap = self.write_group(root)
        sb = bytearray(b"\x89HDF\r\n\x1a\n")
        sb += struct.pack("<BBBBBBBBHHI", 0, 0, 0, 0, 0, 8, 8, 0, 4, 16, 0)
        eof = len(self.buf) - self.ub
        sb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
        sb += struct.pack("<QQII", 0, root_addr, 1, 0) + struct.pack("<QQ", tree, heap)
        assert len(sb) == 96
        self.buf[self.ub:self.ub + 96] = sb
        with open(path, "wb") as f:
            f.write(bytes(self.buf))


def write_chombo(path, levels, names, time=0.0, ghost=(0, 0, 0), userblock=0):
    """levels: [{"dx", "dt", "ref_ratio", "domain": (lo, hi), "boxes": [(lo, hi)],
    "fabs": [array (nc, nz, ny, nx) of box grown by ghost]}]"""
    root = Group()
    root.attrs += [a_int("num_levels", len(levels)), a_int("num_components", len(names)),
                   a_double("time", time), a_int("iteration", 7),
                   a_int("max_level", len(levels) - 1)]
    root.attrs += [a_string("component_%d" % c, n) for c, n in enumerate(names)]
    g = root.group("Chombo_global")
    g.attrs += [a_int("SpaceDim", 3), a_double("testReal", 0.0)]
    for l, lev in enumerate(levels):
        lg = root.group("level_%d" % l)
        lg.attrs += [a_int("ref_ratio", lev["ref_ratio"]), a_int("tag_buffer_size", 3),
                     a_double("dx", lev["dx"]), a_double("dt", lev["dt"]),
                     a_box("prob_domain", *lev["domain"])]
        lg.attrs += [a_int("is_periodic_%d" % d, 0) for d in range(3)]
        bx = b"".join(struct.pack("<6i", *lo, *hi) for lo, hi in lev["boxes"])
        lg.dataset("boxes", box_type(), (len(lev["boxes"]),), bx)
        offs, data = [0], []
        for a in lev["fabs"]:
            data.append(np.ascontiguousarray(a, dtype="<f8").ravel())
            offs.append(offs[-1] + data[-1].size)
        lg.dataset("data:offsets=0", t_int(8), (len(offs),), struct.pack("<%dq" % len(offs), *offs))
        flat = np.concatenate(data) if data else np.zeros(0)
        lg.dataset("data:datatype=0", t_double(), (flat.size,), flat.tobytes())
        da = lg.group("data_attributes")
        da.attrs += [a_int("comps", len(names)), a_intvect("ghost", ghost),
                     a_intvect("outputGhost", ghost), a_string("objectType", "FArrayBox")]
    w = Writer(userblock)
    w.finish(root, path)
