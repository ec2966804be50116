"""Synchronous MPI_Pack / MPI_Unpack of copies of a few workgroups return when
the copy's own completion signal reaches the host, before the runtime's
end-of-kernel processing.  The data must be there at that point: each test
reads the result right after the call returns with a plain hipMemcpy on the
null stream, which does not wait for the library's (non-blocking) stream,
and compares it with the oracle -- 300 rounds of fresh seeded bytes each."""
import ctypes

import numpy as np
import pytest

import tempi_amd as ta
import typecatalog as tc
from conftest import signalled
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

D2H = 2  # hipMemcpyDeviceToHost


@pytest.fixture(scope="module")
def hip():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipMemcpy.restype = ctypes.c_int
    return lib


def fetch(hip, buf, n):
    out = np.empty(n, np.uint8)
    assert hip.hipMemcpy(out.ctypes.data, buf.addr(), n, D2H) == 0
    return out


@pytest.mark.parametrize("spec", [
    ("vector", 64, 4, 12, tc.FLOAT),           # 64 rows of 16 B: one workgroup
    ("vector", 1024, 256, 512, tc.BYTE),       # the 256 KiB 2D benchmark type
    ("subarray", [40, 30, 20], [3, 30, 20], [0, 0, 0], "C", tc.DOUBLE),
])
def test_results_are_there_when_the_call_returns(mpi, hip, spec):
    h = mpi.type_from_spec(spec)
    sb, size = mpi.strided_block(h)
    o = orc.OracleType(spec)
    lo, hi = o.span(1)
    assert lo == 0
    src = ta.DeviceBuffer(hi)
    dst = ta.DeviceBuffer(hi)
    packed = ta.DeviceBuffer(size)
    n_sig = (signalled(sb, src.addr(), packed.addr(), ta.DIR_PACK)
             + signalled(sb, dst.addr(), packed.addr(), ta.DIR_UNPACK))
    assert n_sig >= 1, spec  # each case runs at least one direction signalled
    before = ta.counters()["signalled_waits"]
    for it in range(300):
        host = orc.splitmix_bytes(hi, 7000 + it)
        src.upload(host)
        err, pos = mpi.pack(src.addr(), 1, h, packed.addr(), size)
        assert err == 0 and pos == size
        exp = o.pack(host, 0, 1)
        assert (fetch(hip, packed, size) == exp).all(), (spec, it)
        dst.fill(0xAB)
        err, pos = mpi.unpack(packed.addr(), size, dst.addr(), 1, h)
        assert err == 0 and pos == size
        ref = np.full(hi, 0xAB, np.uint8)
        o.unpack(exp, ref, 0, 1)
        assert (fetch(hip, dst, hi) == ref).all(), (spec, it)
    assert ta.counters()["signalled_waits"] - before == 300 * n_sig
    mpi.free(h)


def test_packs_into_pageable_memory(mpi):
    """The kernel writes a pinned staging buffer; the library copies it out
    once the signal arrives."""
    spec = ("vector", 200, 3, 7, tc.FLOAT)
    h = mpi.type_from_spec(spec)
    o = orc.OracleType(spec)
    lo, hi = o.span(1)
    size = o.size()
    src = ta.DeviceBuffer(hi)
    before = ta.counters()["signalled_waits"]
    for it in range(100):
        host = orc.splitmix_bytes(hi, 9000 + it)
        src.upload(host)
        out = np.zeros(size, np.uint8)
        err, pos = mpi.pack(src.addr(), 1, h, out.ctypes.data, size)
        assert err == 0 and pos == size
        assert (out == o.pack(host, 0, 1)).all(), it
    assert ta.counters()["signalled_waits"] - before == 100
    mpi.free(h)


def test_large_copies_wait_on_the_stream(mpi, hip):
    """A copy of more than 128 workgroups waits with the stream (its result is
    checked the same way)."""
    spec = ("vector", 1 << 16, 64, 96, tc.BYTE)  # 4 MiB in 64-B rows
    h = mpi.type_from_spec(spec)
    sb, size = mpi.strided_block(h)
    o = orc.OracleType(spec)
    lo, hi = o.span(1)
    host = orc.splitmix_bytes(hi, 31)
    src = ta.DeviceBuffer(hi)
    src.upload(host)
    packed = ta.DeviceBuffer(size)
    assert not signalled(sb, src.addr(), packed.addr(), ta.DIR_PACK)
    before = ta.counters()["signalled_waits"]
    err, pos = mpi.pack(src.addr(), 1, h, packed.addr(), size)
    assert err == 0 and pos == size
    assert (fetch(hip, packed, size) == o.pack(host, 0, 1)).all()
    assert ta.counters()["signalled_waits"] == before
    mpi.free(h)
