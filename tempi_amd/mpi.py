"""MPI access for Python hosts: build derived datatypes from nested tuple
specs and call MPI_Pack / MPI_Unpack through the interposer or directly in
the underlying MPI library.

A type spec is a nested tuple mirroring the MPI constructors::

    ("named", "MPI_FLOAT")
    ("contiguous", count, child)
    ("vector", count, blocklength, stride, child)        # stride in child extents
    ("hvector", count, blocklength, stride_bytes, child)
    ("subarray", sizes, subsizes, starts, "C" | "F", child)
    ("resized", lb, extent, child)
    ("indexed", blocklengths, displacements, child)      # displacements in extents
    ("hindexed", blocklengths, displacements, child)     # bytes
    ("indexed_block", blocklength, displacements, child)
    ("hindexed_block", blocklength, displacements, child)
    ("struct", blocklengths, displacements, children)
    ("dup", child)

The same specs feed the CPU oracle (oracle/oracle.py), so one description
drives TEMPI, the underlying library and the oracle.
"""
from __future__ import annotations

import ctypes

from . import lib as _tempi_lib

NAMED_SIZES = {
    "MPI_BYTE": 1, "MPI_CHAR": 1, "MPI_UNSIGNED_CHAR": 1, "MPI_INT8_T": 1,
    "MPI_UINT8_T": 1, "MPI_SHORT": 2, "MPI_INT16_T": 2, "MPI_INT": 4,
    "MPI_UNSIGNED": 4, "MPI_INT32_T": 4, "MPI_FLOAT": 4, "MPI_LONG": 8,
    "MPI_LONG_LONG": 8, "MPI_INT64_T": 8, "MPI_UINT64_T": 8, "MPI_DOUBLE": 8,
}


class MpiError(RuntimeError):
    pass


class Mpi:
    """ctypes view of the MPI library as seen from libtempi_amd.so.

    ``interposed`` calls (MPI_Pack, MPI_Unpack, MPI_Type_commit, MPI_Type_free,
    MPI_Init, MPI_Finalize, and the point-to-point calls MPI_Send, MPI_Recv,
    MPI_Isend, MPI_Irecv, MPI_Wait, MPI_Waitall, MPI_Test) resolve to
    libtempi_amd.so's definitions; the
    ``library_*`` calls go straight to the underlying libmpi (the reference
    test's oracle, test/pack_unpack.cpp:64-66).
    """

    def __init__(self):
        self.tempi = _tempi_lib()
        ok = ctypes.c_int()
        hsize = self.tempi.tempi_mpi_handle_size()
        self.H = ctypes.c_int32 if hsize == 4 else ctypes.c_int64

        def const(name):
            v = self.tempi.tempi_mpi_constant(name.encode(), ctypes.byref(ok))
            if not ok.value:
                raise MpiError(f"unknown MPI constant {name}")
            return v

        self.const = const
        self.COMM_WORLD = const("MPI_COMM_WORLD")
        self.ORDER_C = const("MPI_ORDER_C")
        self.ORDER_FORTRAN = const("MPI_ORDER_FORTRAN")
        self.ERRORS_RETURN = const("MPI_ERRORS_RETURN")
        self.ERR_TRUNCATE = const("MPI_ERR_TRUNCATE")
        # the underlying library, found through the object that defines a
        # function libtempi_amd.so does not override
        info = _dl_info(self.tempi.MPI_Type_get_envelope)
        self.libmpi = ctypes.CDLL(info, mode=ctypes.RTLD_GLOBAL)
        self._proto()

    def _proto(self):
        H, I, P = self.H, ctypes.c_int, ctypes.c_void_p
        Hp, Ip = ctypes.POINTER(self.H), ctypes.POINTER(ctypes.c_int)
        Aint = ctypes.c_ssize_t
        Count = ctypes.c_longlong
        pack_args = [P, I, H, P, I, Ip, H]
        unpack_args = [P, I, Ip, P, I, H, H]
        for L in (self.tempi, self.libmpi):
            L.MPI_Pack.argtypes = pack_args
            L.MPI_Unpack.argtypes = unpack_args
            L.MPI_Type_commit.argtypes = [Hp]
            L.MPI_Type_free.argtypes = [Hp]
        t = self.tempi
        t.MPI_Init.argtypes = [P, P]
        t.MPI_Finalize.argtypes = []
        t.MPI_Type_contiguous.argtypes = [I, H, Hp]
        t.MPI_Type_vector.argtypes = [I, I, I, H, Hp]
        t.MPI_Type_create_hvector.argtypes = [I, I, Aint, H, Hp]
        t.MPI_Type_create_subarray.argtypes = [I, Ip, Ip, Ip, I, H, Hp]
        t.MPI_Type_create_resized.argtypes = [H, Aint, Aint, Hp]
        t.MPI_Type_indexed.argtypes = [I, Ip, Ip, H, Hp]
        t.MPI_Type_create_hindexed.argtypes = [I, Ip, ctypes.POINTER(Aint), H, Hp]
        t.MPI_Type_create_indexed_block.argtypes = [I, I, Ip, H, Hp]
        t.MPI_Type_create_hindexed_block.argtypes = [I, I, ctypes.POINTER(Aint), H, Hp]
        t.MPI_Type_create_struct.argtypes = [I, Ip, ctypes.POINTER(Aint), Hp, Hp]
        t.MPI_Type_dup.argtypes = [H, Hp]
        t.MPI_Type_get_extent_x.argtypes = [H, ctypes.POINTER(Count), ctypes.POINTER(Count)]
        t.MPI_Type_size_x.argtypes = [H, ctypes.POINTER(Count)]
        t.MPI_Pack_size.argtypes = [I, H, H, Ip]
        t.MPI_Comm_set_errhandler.argtypes = [H, H]
        t.MPI_Initialized.argtypes = [Ip]
        t.MPI_Finalized.argtypes = [Ip]
        # point to point (interposed) and what completes it
        Rp = ctypes.POINTER(self.H)
        t.MPI_Send.argtypes = [P, I, H, I, I, H]
        t.MPI_Recv.argtypes = [P, I, H, I, I, H, P]
        t.MPI_Isend.argtypes = [P, I, H, I, I, H, Rp]
        t.MPI_Irecv.argtypes = [P, I, H, I, I, H, Rp]
        t.MPI_Wait.argtypes = [Rp, P]
        t.MPI_Waitall.argtypes = [I, Rp, P]
        t.MPI_Test.argtypes = [Rp, Ip, P]
        t.MPI_Get_count.argtypes = [P, H, Ip]
        t.MPI_Dist_graph_create_adjacent.argtypes = [H, I, Ip, P, I, Ip, P, H, I, Hp]
        Ap = ctypes.POINTER(ctypes.c_ssize_t)
        t.MPI_Neighbor_alltoallw.argtypes = [P, Ip, Ap, Hp, P, Ip, Ap, Hp, H]
        t.MPI_Comm_free.argtypes = [Hp]
        t.MPI_Error_class.argtypes = [I, Ip]
        self.status_size = self.const("sizeof(MPI_Status)")
        self.REQUEST_NULL = self.const("MPI_REQUEST_NULL")

    # -- lifetime ------------------------------------------------------------
    def init(self):
        flag = ctypes.c_int()
        self.tempi.MPI_Initialized(ctypes.byref(flag))
        if not flag.value:
            _ok(self.tempi.MPI_Init(None, None), "MPI_Init")

    def finalize(self):
        flag = ctypes.c_int()
        self.tempi.MPI_Finalized(ctypes.byref(flag))
        if not flag.value:
            _ok(self.tempi.MPI_Finalize(), "MPI_Finalize")

    def errors_return(self):
        _ok(self.tempi.MPI_Comm_set_errhandler(self.COMM_WORLD, self.ERRORS_RETURN),
            "MPI_Comm_set_errhandler")

    # -- types ---------------------------------------------------------------
    def build(self, spec, temps=None):
        """Create (uncommitted) the datatype for `spec`; intermediate handles
        are appended to `temps` so the caller can free them."""
        t = self.tempi
        out = self.H()
        kind = spec[0]
        own = temps is None
        temps = [] if temps is None else temps

        def child(s):
            h = self.build(s, temps)
            if s[0] != "named":
                temps.append(h)
            return h

        def iarr(xs):
            return (ctypes.c_int * max(len(xs), 1))(*[int(x) for x in xs])

        def aarr(xs):
            return (ctypes.c_ssize_t * max(len(xs), 1))(*[int(x) for x in xs])

        if kind == "named":
            return self.H(self.const(spec[1]))
        elif kind == "contiguous":
            _ok(t.MPI_Type_contiguous(spec[1], child(spec[2]), ctypes.byref(out)), kind)
        elif kind == "vector":
            _ok(t.MPI_Type_vector(spec[1], spec[2], spec[3], child(spec[4]),
                                  ctypes.byref(out)), kind)
        elif kind == "hvector":
            _ok(t.MPI_Type_create_hvector(spec[1], spec[2], spec[3], child(spec[4]),
                                          ctypes.byref(out)), kind)
        elif kind == "subarray":
            sizes, subsizes, starts, order = spec[1], spec[2], spec[3], spec[4]
            o = self.ORDER_C if order == "C" else self.ORDER_FORTRAN
            _ok(t.MPI_Type_create_subarray(len(sizes), iarr(sizes), iarr(subsizes),
                                           iarr(starts), o, child(spec[5]),
                                           ctypes.byref(out)), kind)
        elif kind == "resized":
            _ok(t.MPI_Type_create_resized(child(spec[3]), spec[1], spec[2],
                                          ctypes.byref(out)), kind)
        elif kind == "indexed":
            _ok(t.MPI_Type_indexed(len(spec[1]), iarr(spec[1]), iarr(spec[2]),
                                   child(spec[3]), ctypes.byref(out)), kind)
        elif kind == "hindexed":
            _ok(t.MPI_Type_create_hindexed(len(spec[1]), iarr(spec[1]), aarr(spec[2]),
                                           child(spec[3]), ctypes.byref(out)), kind)
        elif kind == "indexed_block":
            _ok(t.MPI_Type_create_indexed_block(len(spec[2]), spec[1], iarr(spec[2]),
                                                child(spec[3]), ctypes.byref(out)), kind)
        elif kind == "hindexed_block":
            _ok(t.MPI_Type_create_hindexed_block(len(spec[2]), spec[1], aarr(spec[2]),
                                                 child(spec[3]), ctypes.byref(out)), kind)
        elif kind == "struct":
            kids = [child(s) for s in spec[3]]
            harr = (self.H * max(len(kids), 1))(*[k.value for k in kids])
            _ok(t.MPI_Type_create_struct(len(spec[1]), iarr(spec[1]), aarr(spec[2]), harr,
                                         ctypes.byref(out)), kind)
        elif kind == "dup":
            _ok(t.MPI_Type_dup(child(spec[1]), ctypes.byref(out)), kind)
        else:
            raise MpiError(f"unknown spec kind {kind!r}")
        if own:
            for h in temps:
                hh = self.H(h.value)
                t.MPI_Type_free(ctypes.byref(hh))
        return out

    def commit(self, h, interposed=True):
        L = self.tempi if interposed else self.libmpi
        _ok(L.MPI_Type_commit(ctypes.byref(h)), "MPI_Type_commit")
        return h

    def type_from_spec(self, spec, interposed=True):
        """Build and commit; returns the handle."""
        temps = []
        h = self.build(spec, temps)
        self.commit(h, interposed)
        for x in temps:
            hh = self.H(x.value)
            self.tempi.MPI_Type_free(ctypes.byref(hh))
        return h

    def free(self, h):
        hh = self.H(h.value)
        _ok(self.tempi.MPI_Type_free(ctypes.byref(hh)), "MPI_Type_free")

    def extent(self, h):
        lb, ext = ctypes.c_longlong(), ctypes.c_longlong()
        _ok(self.tempi.MPI_Type_get_extent_x(h, ctypes.byref(lb), ctypes.byref(ext)), "extent")
        return lb.value, ext.value

    def size(self, h):
        s = ctypes.c_longlong()
        _ok(self.tempi.MPI_Type_size_x(h, ctypes.byref(s)), "size")
        return s.value

    def pack_size(self, count, h):
        s = ctypes.c_int()
        _ok(self.tempi.MPI_Pack_size(count, h, self.COMM_WORLD, ctypes.byref(s)), "pack_size")
        return s.value

    # -- pack / unpack -------------------------------------------------------
    def pack(self, inbuf, incount, h, outbuf, outsize, position=0, library=False):
        """MPI_Pack; returns (error code, new position)."""
        L = self.libmpi if library else self.tempi
        pos = ctypes.c_int(position)
        err = L.MPI_Pack(inbuf, incount, h, outbuf, outsize, ctypes.byref(pos),
                         self.COMM_WORLD)
        return err, pos.value

    def unpack(self, inbuf, insize, outbuf, outcount, h, position=0, library=False):
        L = self.libmpi if library else self.tempi
        pos = ctypes.c_int(position)
        err = L.MPI_Unpack(inbuf, insize, ctypes.byref(pos), outbuf, outcount, h,
                           self.COMM_WORLD)
        return err, pos.value

    # -- point to point -------------------------------------------------------
    def _status_dict(self, raw, h=None):
        def field(name):
            off = self.const(f"offsetof(MPI_Status,{name})")
            return int.from_bytes(bytes(raw[off:off + 4]), "little", signed=True)

        out = {"source": field("MPI_SOURCE"), "tag": field("MPI_TAG"),
               "error": field("MPI_ERROR")}
        if h is not None:
            n = ctypes.c_int()
            self.tempi.MPI_Get_count(raw, h, ctypes.byref(n))
            out["count"] = n.value
        return out

    def send(self, buf, count, h, dest, tag):
        return self.tempi.MPI_Send(buf, count, h, dest, tag, self.COMM_WORLD)

    def recv(self, buf, count, h, source, tag):
        """MPI_Recv; returns (error code, status dict)."""
        raw = (ctypes.c_char * self.status_size)()
        err = self.tempi.MPI_Recv(buf, count, h, source, tag, self.COMM_WORLD, raw)
        return err, self._status_dict(raw, h)

    def isend(self, buf, count, h, dest, tag):
        """MPI_Isend; returns (error code, request handle)."""
        r = self.H()
        err = self.tempi.MPI_Isend(buf, count, h, dest, tag, self.COMM_WORLD, ctypes.byref(r))
        return err, r.value

    def irecv(self, buf, count, h, source, tag):
        r = self.H()
        err = self.tempi.MPI_Irecv(buf, count, h, source, tag, self.COMM_WORLD, ctypes.byref(r))
        return err, r.value

    def wait(self, request, h=None):
        """MPI_Wait; returns (error code, status dict, request after)."""
        r = self.H(request)
        raw = (ctypes.c_char * self.status_size)()
        err = self.tempi.MPI_Wait(ctypes.byref(r), raw)
        return err, self._status_dict(raw, h), r.value

    def waitall(self, requests, types=None):
        """MPI_Waitall; returns (error code, [status dict], [request after])."""
        n = len(requests)
        arr = (self.H * max(n, 1))(*requests)
        raw = (ctypes.c_char * (self.status_size * max(n, 1)))()
        err = self.tempi.MPI_Waitall(n, arr, raw)
        sts = []
        for i in range(n):
            one = (ctypes.c_char * self.status_size).from_buffer(raw, i * self.status_size)
            sts.append(self._status_dict(one, types[i] if types else None))
        return err, sts, list(arr[:n])

    def test(self, request, h=None):
        """MPI_Test; returns (error code, flag, status dict, request after)."""
        r = self.H(request)
        flag = ctypes.c_int()
        raw = (ctypes.c_char * self.status_size)()
        err = self.tempi.MPI_Test(ctypes.byref(r), ctypes.byref(flag), raw)
        return err, flag.value, self._status_dict(raw, h), r.value

    def error_class(self, code):
        c = ctypes.c_int()
        _ok(self.tempi.MPI_Error_class(code, ctypes.byref(c)), "MPI_Error_class")
        return c.value

    # -- neighbourhood collectives -------------------------------------------
    def dist_graph(self, sources, destinations):
        """MPI_Dist_graph_create_adjacent over MPI_COMM_WORLD (unweighted)."""
        out = self.H()
        src = (ctypes.c_int * max(len(sources), 1))(*sources)
        dst = (ctypes.c_int * max(len(destinations), 1))(*destinations)
        unweighted = ctypes.c_void_p(self.const("MPI_UNWEIGHTED"))
        _ok(self.tempi.MPI_Dist_graph_create_adjacent(
            self.COMM_WORLD, len(sources), src, unweighted, len(destinations), dst,
            unweighted, self.const("MPI_INFO_NULL"), 0, ctypes.byref(out)),
            "MPI_Dist_graph_create_adjacent")
        return out

    def comm_free(self, comm):
        c = self.H(comm.value)
        _ok(self.tempi.MPI_Comm_free(ctypes.byref(c)), "MPI_Comm_free")

    def neighbor_alltoallw(self, sendbuf, sendcounts, sdispls, sendtypes, recvbuf,
                           recvcounts, rdispls, recvtypes, comm):
        def ints(xs):
            return (ctypes.c_int * max(len(xs), 1))(*xs)

        def aints(xs):
            return (ctypes.c_ssize_t * max(len(xs), 1))(*xs)

        def hs(xs):
            return (self.H * max(len(xs), 1))(*[x.value for x in xs])

        return self.tempi.MPI_Neighbor_alltoallw(
            sendbuf, ints(sendcounts), aints(sdispls), hs(sendtypes), recvbuf,
            ints(recvcounts), aints(rdispls), hs(recvtypes), comm)

    def strided_block(self, h):
        """(StridedBlock, size) of a type committed through the interposer, or
        None when it has no strided form."""
        from . import StridedBlock

        sb = StridedBlock()
        size = ctypes.c_int64()
        err = self.tempi.tempi_type_strided_block(h, ctypes.byref(sb), ctypes.byref(size))
        if err != 0:
            return None
        return sb, size.value


def _ok(err, what):
    if err != 0:
        raise MpiError(f"{what} failed with MPI error {err}")
    return err


class _DlInfo(ctypes.Structure):
    _fields_ = [("dli_fname", ctypes.c_char_p), ("dli_fbase", ctypes.c_void_p),
                ("dli_sname", ctypes.c_char_p), ("dli_saddr", ctypes.c_void_p)]


def _dl_info(fn) -> str:
    libdl = ctypes.CDLL(None)
    libdl.dladdr.argtypes = [ctypes.c_void_p, ctypes.POINTER(_DlInfo)]
    info = _DlInfo()
    addr = ctypes.cast(fn, ctypes.c_void_p).value
    if not libdl.dladdr(addr, ctypes.byref(info)) or not info.dli_fname:
        raise MpiError("cannot locate the MPI library")
    return info.dli_fname.decode()


_MPI = None


def get() -> Mpi:
    """Process-wide Mpi (MPI_Init'ed on first use, as a singleton run)."""
    global _MPI
    if _MPI is None:
        m = Mpi()
        m.init()
        _MPI = m
    return _MPI
