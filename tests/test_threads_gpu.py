"""MPI_Pack / MPI_Unpack of device data from several threads at once (what a
caller with MPI_THREAD_MULTIPLE does): the type cache, the variant choice of
shapes seen for the first time, the staging pools and the library stream are
shared; every thread's bytes must come out exactly as the oracle packs them.
ctypes releases the interpreter lock during each call, so the threads'
calls overlap inside the library."""
import threading

import numpy as np
import pytest

import tempi_amd as ta
import typecatalog as tc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SPECS = [
    tc.grid_face_subarray(40, 3, "x"),
    tc.grid_face_subarray(40, 3, "y"),
    tc.grid_face_subarray(40, 3, "z"),
    ("vector", 3000, 12, 100, tc.BYTE),
    ("vector", 700, 256, 1100, tc.BYTE),
    ("hvector", 50, 1, 4000, ("vector", 20, 3, 7, tc.FLOAT)),
    ("subarray", [64, 48, 40], [30, 20, 10], [5, 7, 9], "C", tc.DOUBLE),
    ("vector", 5, 2, -3, tc.FLOAT),
]


def test_concurrent_packs_and_unpacks(mpi):
    assert ta.device_count() > 0
    handles = [mpi.type_from_spec(s) for s in SPECS]
    jobs = []
    for i, (spec, h) in enumerate(zip(SPECS, handles)):
        o = orc.OracleType(spec)
        incount = 2
        lo, hi = o.span(incount)
        origin = -min(lo, 0)
        length = origin + hi + 16
        host = orc.splitmix_bytes(length, 70 + i)
        expected = o.pack(host, origin, incount)
        image = np.full(length, 0xAB, np.uint8)
        o.unpack(expected, image, origin, incount)
        src = ta.DeviceBuffer(length)
        src.upload(host)
        jobs.append(dict(h=h, incount=incount, origin=origin, length=length,
                         expected=expected, image=image, src=src,
                         packed=ta.DeviceBuffer(expected.size),
                         dst=ta.DeviceBuffer(length)))
    errors = []
    start = threading.Barrier(len(jobs))

    def worker(j):
        try:
            start.wait()
            size = j["expected"].size
            for it in range(20):
                err, pos = mpi.pack(j["src"].addr(j["origin"]), j["incount"], j["h"],
                                    j["packed"].addr(), size)
                assert err == 0 and pos == size
                assert (j["packed"].download() == j["expected"]).all(), "pack"
                j["dst"].fill(0xAB)
                err, pos = mpi.unpack(j["packed"].addr(), size, j["dst"].addr(j["origin"]),
                                      j["incount"], j["h"])
                assert err == 0 and pos == size
                assert (j["dst"].download() == j["image"]).all(), "unpack"
        except BaseException as e:  # reported by the main thread
            errors.append((j["h"], repr(e)))

    before = ta.counters()
    threads = [threading.Thread(target=worker, args=(j,)) for j in jobs]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a worker did not finish"
    assert not errors, errors
    after = ta.counters()
    n = 20 * len(jobs)
    assert after["pack_kernel_launches"] - before["pack_kernel_launches"] >= n
    assert after["library_pack_calls"] == before["library_pack_calls"]
    for h in handles:
        mpi.free(h)
