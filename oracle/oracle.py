"""Python binding of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

Type specs are the nested tuples described in tempi_amd/mpi.py; they are
turned into the oracle's constructor-node array here.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

OT_NAMED, OT_CONTIGUOUS, OT_VECTOR, OT_HVECTOR, OT_SUBARRAY, OT_RESIZED, OT_HINDEXED = range(7)
OT_MAXD = 8

NAMED_SIZES = {
    "MPI_BYTE": 1, "MPI_CHAR": 1, "MPI_UNSIGNED_CHAR": 1, "MPI_INT8_T": 1,
    "MPI_UINT8_T": 1, "MPI_SHORT": 2, "MPI_INT16_T": 2, "MPI_INT": 4,
    "MPI_UNSIGNED": 4, "MPI_INT32_T": 4, "MPI_FLOAT": 4, "MPI_LONG": 8,
    "MPI_LONG_LONG": 8, "MPI_INT64_T": 8, "MPI_UINT64_T": 8, "MPI_DOUBLE": 8,
}


class OtNode(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("child", ctypes.c_int32),
        ("count", ctypes.c_int64),
        ("blocklength", ctypes.c_int64),
        ("stride", ctypes.c_int64),
        ("size", ctypes.c_int64),
        ("ndims", ctypes.c_int32),
        ("order", ctypes.c_int32),
        ("sizes", ctypes.c_int64 * OT_MAXD),
        ("subsizes", ctypes.c_int64 * OT_MAXD),
        ("starts", ctypes.c_int64 * OT_MAXD),
        ("lb", ctypes.c_int64),
        ("extent", ctypes.c_int64),
        ("blens", ctypes.POINTER(ctypes.c_int64)),
        ("disps", ctypes.POINTER(ctypes.c_int64)),
    ]


class SbDesc(ctypes.Structure):
    _fields_ = [
        ("start", ctypes.c_int64),
        ("extent", ctypes.c_int64),
        ("ndims", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("counts", ctypes.c_int64 * 8),
        ("strides", ctypes.c_int64 * 8),
    ]


_LIB = None


def build():
    """Compile liboracle.so (gcc, via the Makefile next to this file)."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        N = ctypes.POINTER(OtNode)
        I32, I64 = ctypes.c_int32, ctypes.c_int64
        P = ctypes.c_void_p
        for f in ("ot_lb", "ot_extent", "ot_size"):
            getattr(L, f).argtypes = [N, I32]
            getattr(L, f).restype = I64
        L.ot_blocks.argtypes = [N, I32, I64, P, P, I64]
        L.ot_blocks.restype = I64
        L.ot_pack.argtypes = [N, I32, P, I64, P]
        L.ot_pack.restype = I64
        L.ot_unpack.argtypes = [N, I32, P, I64, P]
        L.ot_unpack.restype = I64
        L.sb_pack.argtypes = [ctypes.POINTER(SbDesc), P, I64, P]
        L.sb_pack.restype = I64
        L.sb_unpack.argtypes = [ctypes.POINTER(SbDesc), P, I64, P]
        L.sb_unpack.restype = I64
        L.oracle_fill_splitmix.argtypes = [P, I64, ctypes.c_uint64]
        L.oracle_fill_splitmix.restype = None
        _LIB = L
    return _LIB


class OracleType:
    """A type spec lowered to oracle nodes (root is the last node)."""

    def __init__(self, spec):
        self.spec = spec
        self._nodes = []
        self._keep = []
        self.root = self._add(spec)
        arr = (OtNode * len(self._nodes))(*self._nodes)
        self.nodes = arr

    def _add(self, spec) -> int:
        kind = spec[0]
        n = OtNode()
        n.child = -1
        if kind == "named":
            n.kind = OT_NAMED
            n.size = NAMED_SIZES[spec[1]]
        elif kind == "contiguous":
            n.kind = OT_CONTIGUOUS
            n.count = spec[1]
            n.child = self._add(spec[2])
        elif kind == "vector":
            n.kind = OT_VECTOR
            n.count, n.blocklength, n.stride = spec[1], spec[2], spec[3]
            n.child = self._add(spec[4])
        elif kind == "hvector":
            n.kind = OT_HVECTOR
            n.count, n.blocklength, n.stride = spec[1], spec[2], spec[3]
            n.child = self._add(spec[4])
        elif kind == "subarray":
            n.kind = OT_SUBARRAY
            sizes, subsizes, starts, order = spec[1], spec[2], spec[3], spec[4]
            n.ndims = len(sizes)
            n.order = 0 if order == "C" else 1
            for i in range(len(sizes)):
                n.sizes[i], n.subsizes[i], n.starts[i] = sizes[i], subsizes[i], starts[i]
            n.child = self._add(spec[5])
        elif kind == "resized":
            n.kind = OT_RESIZED
            n.lb, n.extent = spec[1], spec[2]
            n.child = self._add(spec[3])
        elif kind in ("hindexed", "indexed", "indexed_block", "hindexed_block"):
            child_spec = spec[3]
            c = self._add(child_spec)
            if kind in ("hindexed", "indexed"):
                blens, disps = list(spec[1]), list(spec[2])
            else:
                blens, disps = [spec[1]] * len(spec[2]), list(spec[2])
            if kind in ("indexed", "indexed_block"):
                ext = self.extent_of(c)
                disps = [d * ext for d in disps]
            b = (ctypes.c_int64 * max(len(blens), 1))(*blens)
            d = (ctypes.c_int64 * max(len(disps), 1))(*disps)
            self._keep += [b, d]
            n.kind = OT_HINDEXED
            n.count = len(blens)
            n.blens = ctypes.cast(b, ctypes.POINTER(ctypes.c_int64))
            n.disps = ctypes.cast(d, ctypes.POINTER(ctypes.c_int64))
            n.child = c
        elif kind == "dup":
            return self._add(spec[1])
        elif kind == "struct" and len(spec[1]) == 1:
            # one-block struct: bl copies of the child at disp
            c = self._add(spec[3][0])
            b = (ctypes.c_int64 * 1)(spec[1][0])
            d = (ctypes.c_int64 * 1)(spec[2][0])
            self._keep += [b, d]
            n.kind = OT_HINDEXED
            n.count = 1
            n.blens = ctypes.cast(b, ctypes.POINTER(ctypes.c_int64))
            n.disps = ctypes.cast(d, ctypes.POINTER(ctypes.c_int64))
            n.child = c
        else:
            raise ValueError(f"oracle cannot model spec kind {kind!r}")
        self._nodes.append(n)
        return len(self._nodes) - 1

    def extent_of(self, idx: int) -> int:
        arr = (OtNode * len(self._nodes))(*self._nodes)
        return lib().ot_extent(arr, idx)

    def lb(self) -> int:
        return lib().ot_lb(self.nodes, self.root)

    def extent(self) -> int:
        return lib().ot_extent(self.nodes, self.root)

    def size(self) -> int:
        return lib().ot_size(self.nodes, self.root)

    def blocks(self, incount: int = 1, cap: int = 1 << 20):
        disp = np.empty(cap, dtype=np.int64)
        ln = np.empty(cap, dtype=np.int64)
        n = lib().ot_blocks(self.nodes, self.root, incount, disp.ctypes.data,
                            ln.ctypes.data, cap)
        if n > cap:
            return self.blocks(incount, int(n))
        return list(zip(disp[:n].tolist(), ln[:n].tolist()))

    def span(self, incount: int = 1):
        """(lowest, highest+1) byte displacement touched by incount objects."""
        lo, hi = None, None
        for d, l in self.blocks(incount):
            lo = d if lo is None else min(lo, d)
            hi = d + l if hi is None else max(hi, d + l)
        return (lo or 0, hi or 0)

    def pack(self, buf: np.ndarray, origin: int, incount: int) -> np.ndarray:
        """Pack from `buf` whose byte `origin` is the type origin."""
        out = np.zeros(incount * self.size(), dtype=np.uint8)
        n = lib().ot_pack(self.nodes, self.root, buf.ctypes.data + origin, incount,
                          out.ctypes.data)
        assert n == out.size
        return out

    def unpack(self, packed: np.ndarray, buf: np.ndarray, origin: int, outcount: int) -> None:
        """Unpack into `buf` (modified in place) at byte `origin`."""
        n = lib().ot_unpack(self.nodes, self.root, packed.ctypes.data, outcount,
                            buf.ctypes.data + origin)
        assert n == outcount * self.size()


def sb_desc(start, extent, counts, strides) -> SbDesc:
    d = SbDesc()
    d.start, d.extent, d.ndims = start, extent, len(counts)
    for i, (c, s) in enumerate(zip(counts, strides)):
        d.counts[i], d.strides[i] = c, s
    return d


def sb_pack(desc: SbDesc, buf: np.ndarray, origin: int, incount: int) -> np.ndarray:
    size = 1
    for i in range(desc.ndims):
        size *= desc.counts[i]
    out = np.zeros(incount * size, dtype=np.uint8)
    n = lib().sb_pack(ctypes.byref(desc), buf.ctypes.data + origin, incount, out.ctypes.data)
    assert n == out.size
    return out


def sb_unpack(desc: SbDesc, packed: np.ndarray, buf: np.ndarray, origin: int, outcount: int):
    lib().sb_unpack(ctypes.byref(desc), packed.ctypes.data, outcount, buf.ctypes.data + origin)


def splitmix_bytes(n: int, seed: int) -> np.ndarray:
    out = np.empty(int(n), dtype=np.uint8)
    if n:
        lib().oracle_fill_splitmix(out.ctypes.data, int(n), seed & 0xFFFFFFFFFFFFFFFF)
    return out
