"""Independent reader of the HDF5 files the reference's checkpoints use.

TEST INFRASTRUCTURE ONLY: the checker that the library's HDF5 reader
(gale_amd/csrc/gs_checkpoint.cpp) is compared against.  It restates the
published HDF5 file format (superblock version 0/1, version-1 object headers
with continuation blocks, symbol-table groups: version-1 B-tree of group
nodes, SNOD symbol nodes and the local heap; dataspace, datatype, layout
versions 1-3 and attribute messages) in numpy, written separately from the
C++ reader so that a misread offset in either shows up as a disagreement.
Pinned by the closed forms the reference's files hold (mesh vertices of a
regular box, element connectivity, node index columns, the particle count
attribute against the dataset length) in tests/test_checkpoint.py.
"""
import struct

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"


class H5File:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.b = fh.read()
        if self.b[:8] != _SIG:
            raise ValueError("not an HDF5 file")
        ver = self.b[8]
        if ver not in (0, 1) or self.b[13] != 8 or self.b[14] != 8:
            raise ValueError("unsupported superblock")
        entry = (24 if ver == 0 else 28) + 32
        self.root = self._q(entry + 8)

    def _q(self, off):
        return struct.unpack_from("<Q", self.b, off)[0]

    def _messages(self, addr):
        if self.b[addr] != 1:
            raise ValueError("object header version")
        size = struct.unpack_from("<I", self.b, addr + 8)[0]
        blocks, out = [(addr + 16, size)], []
        while blocks:
            start, length = blocks.pop(0)
            o, end = start, start + length
            while o + 8 <= end:
                mtype, msize = struct.unpack_from("<HH", self.b, o)
                if mtype == 0x10:
                    blocks.append((self._q(o + 8), self._q(o + 16)))
                elif mtype:
                    out.append((mtype, o + 8, msize))
                o += 8 + msize
        return out

    def names(self):
        """{name: object header address} of the root group."""
        stab = [m for m in self._messages(self.root) if m[0] == 0x11][0]
        tree, heap = self._q(stab[1]), self._q(stab[1] + 8)
        assert self.b[heap:heap + 4] == b"HEAP"
        seg = self._q(heap + 24)
        out = {}

        def walk(node):
            sig = self.b[node:node + 4]
            if sig == b"TREE":
                n = struct.unpack_from("<H", self.b, node + 6)[0]
                for i in range(n):
                    walk(self._q(node + 32 + 16 * i))
            elif sig == b"SNOD":
                n = struct.unpack_from("<H", self.b, node + 6)[0]
                for j in range(n):
                    e = node + 8 + 40 * j
                    s = seg + self._q(e)
                    out[self.b[s:self.b.index(b"\0", s)].decode()] = self._q(e + 8)
            else:
                raise ValueError("bad group node")

        walk(tree)
        return out

    @staticmethod
    def _dtype(b, o):
        cls, bits, size = b[o] & 15, b[o + 1], struct.unpack_from("<I", b, o + 4)[0]
        order = ">" if bits & 1 else "<"
        if cls == 0 and bits & 8:
            return np.dtype(order + {4: "i4", 8: "i8"}[size])
        if cls == 1:
            return np.dtype(order + {4: "f4", 8: "f8"}[size])
        raise ValueError("unsupported datatype")

    def _space(self, o):
        ver, rank = self.b[o], self.b[o + 1]
        d0 = o + (8 if ver == 1 else 4)
        return tuple(self._q(d0 + 8 * i) for i in range(rank))

    def dataset(self, name):
        msgs = self._messages(self.names()[name.lstrip("/")])
        shape = dtype = addr = None
        for t, o, n in msgs:
            if t == 1:
                shape = self._space(o)
            elif t == 3:
                dtype = self._dtype(self.b, o)
            elif t == 8:
                ver = self.b[o]
                if ver in (1, 2):
                    nd, cls = self.b[o + 1], self.b[o + 2]
                    addr = self._q(o + 8) if cls == 1 else o + 8 + 4 * nd + 4
                else:
                    cls = self.b[o + 1]
                    addr = self._q(o + 2) if cls == 1 else o + 4
            elif t == 0x0B:
                raise ValueError("filtered dataset")
        count = int(np.prod(shape)) if shape else 1
        return np.frombuffer(self.b, dtype=dtype, count=count, offset=addr).reshape(shape)

    def attributes(self):
        """{name: values} of the root group's attributes."""
        out = {}
        for t, o, n in self._messages(self.root):
            if t != 0x0C:
                continue
            ver = self.b[o]
            nl, tl, sl = struct.unpack_from("<HHH", self.b, o + 2)
            pad = (lambda x: (x + 7) // 8 * 8) if ver == 1 else (lambda x: x)
            p = o + 8 + (1 if ver == 3 else 0)
            name = self.b[p:p + nl - 1].decode()
            p += pad(nl)
            dt = self._dtype(self.b, p)
            p += pad(tl)
            shape = self._space(p)
            p += pad(sl)
            cnt = int(np.prod(shape)) if shape else 1
            out[name] = np.frombuffer(self.b, dtype=dt, count=cnt, offset=p).copy()
        return out
