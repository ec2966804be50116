"""tempi_amd -- MI355X (gfx950) derived-datatype MPI_Pack / MPI_Unpack.

The product is the C-ABI shared library ``libtempi_amd.so`` (declared in
``include/tempi_amd.h``): it interposes MPI_Type_commit / MPI_Pack /
MPI_Unpack in front of the system MPI and runs hand-written HIP kernels for
device-resident buffers.  This package only loads it and gives Python hosts
(tests, the benchmark) typed access:

* :func:`lib` -- the loaded library, with ctypes prototypes;
* :mod:`tempi_amd.mpi` -- build MPI datatypes from nested tuple specs and call
  MPI_Pack / MPI_Unpack either through the interposer or straight into the
  underlying MPI library;
* :class:`DeviceBuffer` -- device or mapped host memory from the library's
  own allocators.

There is no fallback: if the library is missing, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtempi_amd.so")

TEMPI_SUCCESS = 0
TEMPI_ERR_ARG = -1
TEMPI_ERR_UNSUPPORTED = -2
TEMPI_ERR_TRUNCATE = -3
TEMPI_ERR_HIP = -4
TEMPI_ERR_NODEVICE = -5
TEMPI_ERR_NOMEM = -6
TEMPI_ERR_IO = -7

DIR_PACK = 0
DIR_UNPACK = 1

SELECT_HEURISTIC = 0
SELECT_MODEL = 1
SELECT_AUTOTUNE = 2
SELECT_DEFAULT = 3

MAX_DIMS = 8


class StridedBlock(ctypes.Structure):
    """tempi_strided_block (include/tempi_amd.h)."""

    _fields_ = [
        ("start", ctypes.c_int64),
        ("extent", ctypes.c_int64),
        ("ndims", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("counts", ctypes.c_int64 * MAX_DIMS),
        ("strides", ctypes.c_int64 * MAX_DIMS),
    ]

    @classmethod
    def make(cls, start, extent, counts, strides=None):
        sb = cls()
        sb.start = int(start)
        sb.extent = int(extent)
        sb.ndims = len(counts)
        strides = list(strides) if strides is not None else [1] + [0] * (len(counts) - 1)
        for i, c in enumerate(counts):
            sb.counts[i] = int(c)
            sb.strides[i] = int(strides[i])
        return sb

    def as_tuple(self):
        n = self.ndims
        return (self.start, self.extent, tuple(self.counts[:n]), tuple(self.strides[:n]))

    def size(self) -> int:
        s = 1
        for i in range(self.ndims):
            s *= self.counts[i]
        return s

    def __repr__(self):
        st, ex, c, s = self.as_tuple()
        return f"StridedBlock(start={st}, extent={ex}, counts={list(c)}, strides={list(s)})"


class VariantInfo(ctypes.Structure):
    _fields_ = [
        ("lanes_per_row", ctypes.c_int32),
        ("word_bytes", ctypes.c_int32),
        ("unroll", ctypes.c_int32),
        ("nontemporal_stores", ctypes.c_int32),
        ("nontemporal_loads", ctypes.c_int32),
    ]


class Counters(ctypes.Structure):
    _fields_ = [
        (n, ctypes.c_int64)
        for n in (
            "pack_kernel_launches",
            "unpack_kernel_launches",
            "batch_launches",
            "graph_launches",
            "library_pack_calls",
            "library_unpack_calls",
            "commits_canonical",
            "commits_unsupported",
            "model_inferences",
            "autotune_runs",
            "verified_choices",
            "p2p_device",
            "p2p_oneshot",
            "p2p_staged",
            "p2p_library",
            "trial_launches",
            "method_model_choices",
            "host_staged_calls",
            "signalled_waits",
        )
    ]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class CopyEmulation(ctypes.Structure):
    """tempi_copy_emulation (include/tempi_amd.h)."""

    _fields_ = [
        ("launches", ctypes.c_int64),
        ("max_units", ctypes.c_int64),
        ("emulated", ctypes.c_int32),
        ("wrapped", ctypes.c_int32),
        ("read_lo", ctypes.c_int64),
        ("read_hi", ctypes.c_int64),
        ("write_lo", ctypes.c_int64),
        ("write_hi", ctypes.c_int64),
        ("bytes_read", ctypes.c_int64),
        ("bytes_written", ctypes.c_int64),
        ("exact", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


class BatchRequest(ctypes.Structure):
    """tempi_batch_request (include/tempi_amd.h)."""

    _fields_ = [
        ("sb", StridedBlock),
        ("strided", ctypes.c_void_p),
        ("packed", ctypes.c_void_p),
        ("count", ctypes.c_int64),
        ("direction", ctypes.c_int32),
        ("variant", ctypes.c_int32),
    ]


class TempiError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: error {code}: {last_error()}")
        self.code = code


_LIB = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64


def _prototypes(L):
    sb_p = ctypes.POINTER(StridedBlock)
    i64_p = ctypes.POINTER(ctypes.c_int64)
    protos = {
        "tempi_version": (_I, []),
        "tempi_last_error": (ctypes.c_char_p, []),
        "tempi_device_count": (_I, []),
        "tempi_mpi_constant": (_I64, [ctypes.c_char_p, ctypes.POINTER(_I)]),
        "tempi_mpi_handle_size": (_I, []),
        "tempi_num_variants": (_I, []),
        "tempi_variant_info_get": (_I, [_I, ctypes.POINTER(VariantInfo)]),
        "tempi_pack_sb": (_I, [sb_p, _P, _I64, _P, _I64, i64_p, _I, _P]),
        "tempi_unpack_sb": (_I, [sb_p, _P, _I64, i64_p, _P, _I64, _I, _P]),
        "tempi_select_variant": (_I, [sb_p, _I64, _P, _P, _I, _I, ctypes.POINTER(_I)]),
        "tempi_stream": (_P, []),
        "tempi_stream_synchronize": (_I, [_P]),
        "tempi_alloc_device": (_P, [_I64]),
        "tempi_free_device": (_I, [_P]),
        "tempi_alloc_host_mapped": (_P, [_I64]),
        "tempi_free_host_mapped": (_I, [_P]),
        "tempi_memcpy": (_I, [_P, _P, _I64]),
        "tempi_memset": (_I, [_P, _I, _I64]),
        "tempi_is_device_accessible": (_I, [_P]),
        "tempi_event_create": (_I, [ctypes.POINTER(_P)]),
        "tempi_event_destroy": (_I, [_P]),
        "tempi_event_record": (_I, [_P, _P]),
        "tempi_event_synchronize": (_I, [_P]),
        "tempi_event_elapsed_ms": (_I, [_P, _P, ctypes.POINTER(ctypes.c_float)]),
        "tempi_counters_get": (_I, [ctypes.POINTER(Counters)]),
        "tempi_counters_reset": (_I, []),
        "tempi_fastdiv_check": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32]),
        "tempi_num_grid_caps": (_I, []),
        "tempi_time_variant": (_I, [sb_p, _I64, _P, _P, _I, _I, _I, _I,
                                    ctypes.POINTER(ctypes.c_float)]),
        "tempi_autotune": (_I, [sb_p, _I64, _P, _P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                ctypes.POINTER(ctypes.c_float)]),
        "tempi_copy_emulate": (_I, [sb_p, _I64, ctypes.c_uint64, ctypes.c_uint64, _I, _I, _I, _I,
                                    _I64, ctypes.POINTER(CopyEmulation)]),
        "tempi_flush_caches": (_I, [_I64]),
        "tempi_launch_latency_us": (_I, [_I, ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(ctypes.c_float)]),
        "tempi_batch_create": (_I, [ctypes.POINTER(BatchRequest), _I64, _I, ctypes.POINTER(_P)]),
        "tempi_batch_launch": (_I, [_P, _P]),
        "tempi_batch_graph_launch": (_I, [_P, _P]),
        "tempi_batch_info": (_I, [_P, i64_p, i64_p, i64_p]),
        "tempi_batch_variant": (_I, [_P, _I64, ctypes.POINTER(_I)]),
        "tempi_batch_destroy": (_I, [_P]),
        "tempi_batch_run_noalloc": (_I, [ctypes.POINTER(BatchRequest), _I64, _P]),
        "tempi_graph_create": (_I, [ctypes.POINTER(_P), _I64, ctypes.POINTER(_P)]),
        "tempi_graph_launch": (_I, [_P, _P]),
        "tempi_graph_destroy": (_I, [_P]),
        "tempi_mlp_num_features": (_I, []),
        "tempi_mlp_num_outputs": (_I, []),
        "tempi_mlp_num_params": (_I, []),
        "tempi_mlp_features": (_I, [sb_p, _I64, _P, _P, _I, _P]),
        "tempi_mlp_grad": (_I, [_P, _P, _P, _P, _I, _P, _P]),
        "tempi_mlp_train": (_I, [_P, _P, _P, _I64, _I, ctypes.c_float, ctypes.c_uint64, _P, _P]),
        "tempi_mlp_predict": (_I, [_P, _P, _P, _I64, _P, _P]),
        "tempi_mlp_save": (_I, [ctypes.c_char_p, _P, _P]),
        "tempi_mlp_load": (_I, [ctypes.c_char_p]),
        "tempi_mlp_loaded": (_I, []),
        "tempi_interp_time": (ctypes.c_double, [_P, _I, _I64]),
        "tempi_interp_2d": (ctypes.c_double, [_P, _I, _I, _I64, _I64]),
        "tempi_perf_clear": (_I, []),
        "tempi_perf_get_curve": (_I, [_I, _P, _I, ctypes.POINTER(_I)]),
        "tempi_perf_set_curve": (_I, [_I, _P, _I]),
        "tempi_perf_get_table": (_I, [_I, _P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
        "tempi_perf_set_table": (_I, [_I, _P, _I, _I]),
        "tempi_perf_kernel_launch": (ctypes.c_double, []),
        "tempi_perf_set_kernel_launch": (_I, [ctypes.c_double]),
        "tempi_perf_load": (_I, [ctypes.c_char_p]),
        "tempi_perf_save": (_I, [ctypes.c_char_p]),
        "tempi_perf_measure": (_I, [_I, ctypes.c_double]),
        "tempi_iid_statistics": (_I, [_P, _I64, _P]),
        "tempi_iid_test": (_I, [_P, _I64, ctypes.c_uint64, _P]),
        "tempi_perf_model": (_I, [_I, _I, _I64, _I64, ctypes.POINTER(ctypes.c_double)]),
        "tempi_choose_method": (_I, [_I, _I64, _I64]),
        "tempi_p2p_method": (_I, [_I, _I64, _I64, _I64]),
        "tempi_set_datatype_method": (_I, [_I]),
        "tempi_method_features": (_I, [_I64, _I64, _I64, _I, ctypes.POINTER(ctypes.c_float)]),
        "tempi_method_model_load": (_I, [ctypes.c_char_p]),
        "tempi_method_model_unload": (_I, []),
        "tempi_method_model_loaded": (_I, []),
    }
    for name, (res, args) in protos.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    # the two functions whose first argument is an MPI handle
    h = ctypes.c_int32 if L.tempi_mpi_handle_size() == 4 else ctypes.c_int64
    for name in ("tempi_type_strided_block", "tempi_type_canonicalize"):
        fn = getattr(L, name)
        fn.restype = _I
        fn.argtypes = [h, sb_p, i64_p]


def lib():
    """The loaded libtempi_amd.so (loads on first call; raises if absent)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m tempi_amd.build` "
                "(there is no non-HIP fallback)"
            )
        L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        _prototypes(L)
        _LIB = L
    return _LIB


def last_error() -> str:
    if _LIB is None:
        return ""
    s = _LIB.tempi_last_error()
    return s.decode() if s else ""


def check(code, what="tempi call"):
    if code != TEMPI_SUCCESS:
        raise TempiError(code, what)
    return code


def device_count() -> int:
    return lib().tempi_device_count()


def counters() -> dict:
    c = Counters()
    check(lib().tempi_counters_get(ctypes.byref(c)))
    return c.as_dict()


def reset_counters():
    check(lib().tempi_counters_reset())


def iid_statistics(samples) -> list:
    """The eight SP 800-90B section 5.1 statistics of `samples` (see
    tempi_iid_statistics)."""
    import numpy as np

    a = np.ascontiguousarray(samples, np.float64)
    out = (ctypes.c_double * 8)()
    check(lib().tempi_iid_statistics(a.ctypes.data, a.size, out), "iid_statistics")
    return list(out)


def iid_test(samples, seed: int = 0):
    """(independent?, counts) of the SP 800-90B section 5.1 permutation test
    on `samples`: counts[2 k], counts[2 k + 1] = C0, C1 of statistic k (see
    tempi_iid_test)."""
    import numpy as np

    a = np.ascontiguousarray(samples, np.float64)
    c = (ctypes.c_int64 * 16)()
    r = lib().tempi_iid_test(a.ctypes.data, a.size, ctypes.c_uint64(seed), c)
    if r < 0:
        check(r, "iid_test")
    return r == 1, list(c)


def num_variants() -> int:
    return lib().tempi_num_variants()


def variant_info(v: int) -> dict:
    info = VariantInfo()
    check(lib().tempi_variant_info_get(v, ctypes.byref(info)), "variant_info")
    return {n: getattr(info, n) for n, _ in VariantInfo._fields_}


class DeviceBuffer:
    """Device memory (or pinned mapped host memory with ``mapped=True``)
    allocated by the library's allocator."""

    def __init__(self, nbytes: int, mapped: bool = False):
        L = lib()
        self.nbytes = int(nbytes)
        self.mapped = mapped
        n = max(self.nbytes, 1)
        self.ptr = L.tempi_alloc_host_mapped(n) if mapped else L.tempi_alloc_device(n)
        if not self.ptr:
            raise TempiError(TEMPI_ERR_NOMEM, f"allocate {n} bytes")

    def free(self):
        if self.ptr:
            L = lib()
            (L.tempi_free_host_mapped if self.mapped else L.tempi_free_device)(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def addr(self, offset: int = 0) -> int:
        return int(self.ptr) + int(offset)

    def upload(self, data, offset: int = 0):
        import numpy as np

        a = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
        assert offset + a.size <= self.nbytes
        check(lib().tempi_memcpy(self.addr(offset), a.ctypes.data, a.size), "upload")

    def download(self, nbytes: int | None = None, offset: int = 0):
        import numpy as np

        n = self.nbytes - offset if nbytes is None else int(nbytes)
        out = np.empty(n, dtype=np.uint8)
        check(lib().tempi_memcpy(out.ctypes.data, self.addr(offset), n), "download")
        return out

    def fill(self, value: int):
        check(lib().tempi_memset(self.addr(), int(value), self.nbytes), "memset")


class Event:
    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().tempi_event_create(ctypes.byref(h)), "event_create")
        self.h = h

    def record(self, stream=None):
        check(lib().tempi_event_record(self.h, stream), "event_record")

    def synchronize(self):
        check(lib().tempi_event_synchronize(self.h), "event_synchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        check(lib().tempi_event_elapsed_ms(self.h, end.h, ctypes.byref(ms)), "elapsed")
        return float(ms.value)

    def __del__(self):
        try:
            lib().tempi_event_destroy(self.h)
        except Exception:
            pass


def stream():
    s = lib().tempi_stream()
    if not s:
        raise TempiError(TEMPI_ERR_NODEVICE, "tempi_stream")
    return s


def synchronize(stream_=None):
    check(lib().tempi_stream_synchronize(stream_), "stream_synchronize")


def pack_sb(sb: StridedBlock, inbuf: int, incount: int, outbuf: int, outsize: int,
            position: int = 0, variant: int = -1, stream_=None) -> int:
    """Asynchronous pack on the library stream; returns the new position."""
    pos = ctypes.c_int64(position)
    check(
        lib().tempi_pack_sb(ctypes.byref(sb), inbuf, incount, outbuf, outsize,
                            ctypes.byref(pos), variant, stream_),
        "tempi_pack_sb",
    )
    return pos.value


def unpack_sb(sb: StridedBlock, inbuf: int, insize: int, outbuf: int, outcount: int,
              position: int = 0, variant: int = -1, stream_=None) -> int:
    pos = ctypes.c_int64(position)
    check(
        lib().tempi_unpack_sb(ctypes.byref(sb), inbuf, insize, ctypes.byref(pos), outbuf,
                              outcount, variant, stream_),
        "tempi_unpack_sb",
    )
    return pos.value


def time_variant(sb: StridedBlock, incount: int, strided: int, packed: int, direction: int,
                 variant: int, grid_cap: int = 0, reps: int = 10) -> float:
    us = ctypes.c_float()
    check(lib().tempi_time_variant(ctypes.byref(sb), incount, strided, packed, direction,
                                   variant, grid_cap, reps, ctypes.byref(us)), "time_variant")
    return float(us.value)


def autotune(sb: StridedBlock, incount: int, strided: int, packed: int, direction: int):
    """(best variant, best grid cap index, best time per variant in us)."""
    v, c = ctypes.c_int(), ctypes.c_int()
    per = (ctypes.c_float * num_variants())()
    check(lib().tempi_autotune(ctypes.byref(sb), incount, strided, packed, direction,
                               ctypes.byref(v), ctypes.byref(c), per), "autotune")
    return v.value, c.value, list(per)


def copy_emulate(sb: StridedBlock, incount: int, strided_addr: int, packed_addr: int,
                 direction: int, variant: int, max_blocks: int = 0, cu_count: int = 256,
                 unit_limit: int = 1 << 22) -> dict:
    """Host-side emulation of the copy's launches and memory accesses (no GPU,
    nothing dereferenced); see tempi_copy_emulate."""
    out = CopyEmulation()
    check(lib().tempi_copy_emulate(ctypes.byref(sb), incount, strided_addr, packed_addr,
                                   direction, variant, max_blocks, cu_count, unit_limit,
                                   ctypes.byref(out)), "copy_emulate")
    return out.as_dict()


def flush_caches(nbytes: int = 0):
    """Evict L2 / Infinity Cache with a read-only sweep (for cold timings)."""
    check(lib().tempi_flush_caches(nbytes), "flush_caches")


def launch_latency_us(reps: int = 1000):
    h, d = ctypes.c_float(), ctypes.c_float()
    check(lib().tempi_launch_latency_us(reps, ctypes.byref(h), ctypes.byref(d)), "latency")
    return float(h.value), float(d.value)


def batch_requests(requests):
    """A ctypes array of tempi_batch_request from (StridedBlock, strided_addr,
    packed_addr, count, direction[, variant]) tuples."""
    reqs = list(requests)
    arr = (BatchRequest * max(len(reqs), 1))()
    for i, r in enumerate(reqs):
        sb, strided, packed, count, direction = r[:5]
        arr[i].sb = sb
        arr[i].strided = strided
        arr[i].packed = packed
        arr[i].count = count
        arr[i].direction = direction
        arr[i].variant = r[5] if len(r) > 5 else -1
    return arr, len(reqs)


def batch_run_noalloc(requests, stream_=None):
    """tempi_batch_run_noalloc: resolve and launch the requests as one batch
    whose tables sit in the library's static device array (no allocation);
    asynchronous on the stream."""
    arr, n = batch_requests(requests)
    check(lib().tempi_batch_run_noalloc(arr, n, stream_), "batch_run_noalloc")


class Batch:
    """A fixed set of pack/unpack requests run by one persistent kernel.

    requests: iterable of (StridedBlock, strided_addr, packed_addr, count,
    direction[, variant])."""

    def __init__(self, requests, method: int = SELECT_DEFAULT):
        arr, n = batch_requests(requests)
        h = ctypes.c_void_p()
        check(lib().tempi_batch_create(arr, n, method, ctypes.byref(h)), "batch_create")
        self.h = h
        self.n = n

    def launch(self, stream_=None):
        check(lib().tempi_batch_launch(self.h, stream_), "batch_launch")

    def graph_launch(self, stream_=None):
        check(lib().tempi_batch_graph_launch(self.h, stream_), "batch_graph_launch")

    def info(self):
        r, t, b = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(lib().tempi_batch_info(self.h, ctypes.byref(r), ctypes.byref(t), ctypes.byref(b)))
        return {"requests": r.value, "tiles": t.value, "packed_bytes": b.value}

    def variant(self, i: int) -> int:
        v = ctypes.c_int()
        check(lib().tempi_batch_variant(self.h, i, ctypes.byref(v)))
        return v.value

    def destroy(self):
        if self.h:
            lib().tempi_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Graph:
    """One executable graph replaying the launches of several batches in
    order (the batches must outlive it)."""

    def __init__(self, batches):
        self.batches = list(batches)  # kept alive with the graph
        arr = (ctypes.c_void_p * max(len(self.batches), 1))(
            *[b.h.value for b in self.batches])
        h = ctypes.c_void_p()
        check(lib().tempi_graph_create(arr, len(self.batches), ctypes.byref(h)),
              "graph_create")
        self.h = h

    def launch(self, stream_=None):
        check(lib().tempi_graph_launch(self.h, stream_), "graph_launch")

    def destroy(self):
        if self.h:
            lib().tempi_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def select_variant(sb: StridedBlock, incount: int, strided: int, packed: int,
                   direction: int = DIR_PACK, method: int = SELECT_HEURISTIC) -> int:
    v = ctypes.c_int()
    check(
        lib().tempi_select_variant(ctypes.byref(sb), incount, strided, packed, direction,
                                   method, ctypes.byref(v)),
        "select_variant",
    )
    return v.value
