"""Point-to-point transfers of non-contiguous device data on the GPU, this
process sending to itself: every staging method, blocking and non-blocking
calls, completion by MPI_Wait / MPI_Waitall / MPI_Test, statuses, short
receives.  Received images are checked byte for byte against the oracle
(the type map unpacked into a 0xAB-filled buffer)."""
import numpy as np
import pytest

import tempi_amd as ta
import typecatalog as tc
from conftest import signalled
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

DEVICE, ONESHOT, STAGED, AUTO = 1, 2, 3, 0
COUNTER = {ONESHOT: "p2p_oneshot", STAGED: "p2p_staged"}


@pytest.fixture(autouse=True)
def method_default():
    assert ta.device_count() > 0
    yield
    ta.lib().tempi_set_datatype_method(-1)


def setup(mpi, spec, count, seed):
    h = mpi.type_from_spec(spec)
    o = orc.OracleType(spec)
    n = o.extent() * count
    host = orc.splitmix_bytes(n, seed)
    src = ta.DeviceBuffer(n)
    src.upload(host)
    dst = ta.DeviceBuffer(n)
    dst.fill(0xAB)
    expected = np.full(n, 0xAB, np.uint8)
    o.unpack(o.pack(host, 0, count), expected, 0, count)
    return h, src, dst, expected


FACE = tc.grid_face_subarray(32, 3, "x")  # 12-B rows of a 38^3 float grid


@pytest.mark.parametrize("method", [ONESHOT, STAGED, AUTO])
def test_isend_irecv_self(mpi, method):
    ta.lib().tempi_set_datatype_method(method)
    h, src, dst, expected = setup(mpi, FACE, 2, 3)
    c0 = ta.counters()
    err, rr = mpi.irecv(dst.addr(), 2, h, 0, 11)
    assert err == 0
    err, sr = mpi.isend(src.addr(), 2, h, 0, 11)
    assert err == 0
    err, sts, reqs = mpi.waitall([rr, sr], [h, h])
    assert err == 0 and reqs == [mpi.REQUEST_NULL] * 2
    assert sts[0]["count"] == 2 and sts[0]["source"] == 0 and sts[0]["tag"] == 11
    assert (dst.download() == expected).all()
    c1 = ta.counters()
    if method in COUNTER:
        assert c1[COUNTER[method]] == c0[COUNTER[method]] + 2
    assert c1["p2p_library"] == c0["p2p_library"]
    mpi.free(h)


def test_small_isend_irecv_complete_on_the_signal(mpi):
    """ONESHOT Isend/Irecv of a small type: a pack or unpack of a few
    workgroups completes its request as soon as its own completion signal
    arrives.  The received bytes are read with a plain hipMemcpy on the null
    stream right after MPI_Waitall returns, 50 rounds of fresh bytes."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    ta.lib().tempi_set_datatype_method(ONESHOT)
    spec = ("vector", 300, 3, 7, tc.FLOAT)
    h = mpi.type_from_spec(spec)
    sb, size = mpi.strided_block(h)
    o = orc.OracleType(spec)
    lo, hi = o.span(1)
    src, dst = ta.DeviceBuffer(hi), ta.DeviceBuffer(hi)
    stage = ta.DeviceBuffer(size)  # stands in for the staging buffer's alignment
    n_sig = (signalled(sb, src.addr(), stage.addr(), ta.DIR_PACK)
             + signalled(sb, dst.addr(), stage.addr(), ta.DIR_UNPACK))
    assert n_sig == 2
    c0 = ta.counters()["signalled_waits"]
    for it in range(50):
        host = orc.splitmix_bytes(hi, 500 + it)
        src.upload(host)
        dst.fill(0xAB)
        err, rr = mpi.irecv(dst.addr(), 1, h, 0, 31)
        assert err == 0
        err, sr = mpi.isend(src.addr(), 1, h, 0, 31)
        assert err == 0
        err, sts, reqs = mpi.waitall([rr, sr], [h, h])
        assert err == 0 and sts[0]["count"] == 1
        got = np.empty(hi, np.uint8)
        assert hip.hipMemcpy(got.ctypes.data, dst.addr(), hi, 2) == 0
        ref = np.full(hi, 0xAB, np.uint8)
        o.unpack(o.pack(host, 0, 1), ref, 0, 1)
        assert (got == ref).all(), it
    # each request completes on its signal unless its event was seen first
    done = ta.counters()["signalled_waits"] - c0
    assert 0.9 * 50 * n_sig <= done <= 50 * n_sig, done
    mpi.free(h)


@pytest.mark.parametrize("method", [ONESHOT, STAGED])
def test_wire_format_is_the_libraries_packed_form(mpi, method):
    """A message travels as MPI_PACKED bytes in type-map order, the MPI
    library's own packed form: a device-buffer side (packed or unpacked here)
    talks to a host-buffer side whose derived type the MPI library handles
    itself, in both directions (`sender.cpp:109`,
    `async_operation.cpp:127,261`)."""
    ta.lib().tempi_set_datatype_method(method)
    h, src, dst, expected = setup(mpi, FACE, 2, 41)
    host_src = src.download()
    c0 = ta.counters()
    # device sender -> host receiver: the library unpacks the packed bytes
    host_dst = np.full(expected.size, 0xAB, np.uint8)
    err, rr = mpi.irecv(host_dst.ctypes.data, 2, h, 0, 51)
    assert err == 0
    err, sr = mpi.isend(src.addr(), 2, h, 0, 51)
    assert err == 0
    err, sts, reqs = mpi.waitall([rr, sr], [h, h])
    assert err == 0 and reqs == [mpi.REQUEST_NULL] * 2
    assert sts[0]["count"] == 2
    assert (host_dst == expected).all()
    # host sender -> device receiver: the library packs, the kernel unpacks
    err, rr = mpi.irecv(dst.addr(), 2, h, 0, 52)
    assert err == 0
    err, sr = mpi.isend(host_src.ctypes.data, 2, h, 0, 52)
    assert err == 0
    err, sts, reqs = mpi.waitall([rr, sr], [h, h])
    assert err == 0 and reqs == [mpi.REQUEST_NULL] * 2
    assert sts[0]["count"] == 2
    assert (dst.download() == expected).all()
    c1 = ta.counters()
    assert c1[COUNTER[method]] == c0[COUNTER[method]] + 2
    assert c1["p2p_library"] == c0["p2p_library"] + 2
    mpi.free(h)


@pytest.mark.parametrize("method", [ONESHOT, STAGED])
def test_blocking_send_and_recv_self(mpi, method):
    """Blocking sends and receives; with ONESHOT a pack or unpack of a few
    workgroups returns on its own completion signal (STAGED copies to the
    host after the kernel and waits on the stream)."""
    ta.lib().tempi_set_datatype_method(method)
    spec = ("vector", 4096, 48, 200, tc.BYTE)
    h, src, dst, expected = setup(mpi, spec, 1, 5)
    sig0 = ta.counters()["signalled_waits"]
    # blocking send matched by a receive posted first
    err, rr = mpi.irecv(dst.addr(), 1, h, 0, 21)
    assert err == 0
    assert mpi.send(src.addr(), 1, h, 0, 21) == 0
    err, st, rr = mpi.wait(rr, h)
    assert err == 0 and st["count"] == 1 and rr == mpi.REQUEST_NULL
    assert (dst.download() == expected).all()
    # blocking receive of a send issued first (still packing when it starts)
    dst.fill(0xAB)
    err, sr = mpi.isend(src.addr(), 1, h, 0, 22)
    assert err == 0
    err, st = mpi.recv(dst.addr(), 1, h, 0, 22)
    assert err == 0 and st["count"] == 1 and st["tag"] == 22
    err, _, sr = mpi.wait(sr)
    assert err == 0
    assert (dst.download() == expected).all()
    # the blocking send's pack and the blocking receive's unpack (the
    # staging buffer's device address decides only alignment, 256-B here)
    sb, _ = mpi.strided_block(h)
    stage = ta.DeviceBuffer(sb.size())
    want = 0
    if method == ONESHOT:
        want = (signalled(sb, src.addr(), stage.addr(), ta.DIR_PACK)
                + signalled(sb, dst.addr(), stage.addr(), ta.DIR_UNPACK))
    assert ta.counters()["signalled_waits"] - sig0 == want
    mpi.free(h)


def test_completion_by_test(mpi):
    ta.lib().tempi_set_datatype_method(ONESHOT)
    h, src, dst, expected = setup(mpi, FACE, 1, 7)
    err, rr = mpi.irecv(dst.addr(), 1, h, 0, 31)
    err, sr = mpi.isend(src.addr(), 1, h, 0, 31)
    done = {}
    for _ in range(200000):
        for name, r in (("recv", rr), ("send", sr)):
            if name in done:
                continue
            err, flag, st, after = mpi.test(r, h)
            assert err == 0
            if flag:
                done[name] = st
                assert after == mpi.REQUEST_NULL
        if len(done) == 2:
            break
    assert len(done) == 2, "MPI_Test never completed"
    assert done["recv"]["count"] == 1
    assert (dst.download() == expected).all()
    mpi.free(h)


def test_short_receive_unpacks_what_arrived(mpi):
    """A receive posted for 3 objects that gets 1: the status says 1 and only
    the first object's type map is written."""
    ta.lib().tempi_set_datatype_method(STAGED)
    spec = ("vector", 100, 8, 24, tc.BYTE)
    h = mpi.type_from_spec(spec)
    o = orc.OracleType(spec)
    host = orc.splitmix_bytes(o.extent() * 3, 9)
    src = ta.DeviceBuffer(host.size)
    src.upload(host)
    dst = ta.DeviceBuffer(host.size)
    dst.fill(0xAB)
    err, rr = mpi.irecv(dst.addr(), 3, h, 0, 41)
    err, sr = mpi.isend(src.addr(), 1, h, 0, 41)
    err, sts, _ = mpi.waitall([sr, rr], [h, h])
    assert err == 0 and sts[1]["count"] == 1
    expected = np.full(host.size, 0xAB, np.uint8)
    o.unpack(o.pack(host, 0, 1), expected, 0, 1)
    assert (dst.download() == expected).all()
    mpi.free(h)


def test_device_method_without_gpu_aware_library_stages(mpi):
    ta.lib().tempi_set_datatype_method(DEVICE)
    h, src, dst, expected = setup(mpi, FACE, 1, 13)
    c0 = ta.counters()["p2p_staged"]
    err, rr = mpi.irecv(dst.addr(), 1, h, 0, 51)
    err, sr = mpi.isend(src.addr(), 1, h, 0, 51)
    err, _, _ = mpi.waitall([rr, sr])
    assert err == 0
    assert ta.counters()["p2p_staged"] == c0 + 2
    assert (dst.download() == expected).all()
    mpi.free(h)


def test_grid_faces_exchange_self(mpi):
    """The six faces of a 134^3 float64 subdomain (radius 3) sent to itself
    with every request outstanding at once, as a halo exchange posts them."""
    halos = [(d, tc.halo_type(3, (128, 128, 128), d, 8))
             for d in [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]]
    n = 134 ** 3 * 8
    host = orc.splitmix_bytes(n, 17)
    src = ta.DeviceBuffer(n)
    src.upload(host)
    dst = ta.DeviceBuffer(n)
    dst.fill(0xAB)
    expected = np.full(n, 0xAB, np.uint8)
    handles, reqs = [], []
    for i, (d, spec) in enumerate(halos):
        h = mpi.type_from_spec(spec)
        handles.append(h)
        o = orc.OracleType(spec)
        o.unpack(o.pack(host, 0, 1), expected, 0, 1)
        err, r = mpi.irecv(dst.addr(), 1, h, 0, 100 + i)
        assert err == 0
        reqs.append(r)
    for i, h in enumerate(handles):
        err, r = mpi.isend(src.addr(), 1, h, 0, 100 + i)
        assert err == 0
        reqs.append(r)
    err, sts, after = mpi.waitall(reqs)
    assert err == 0 and after == [mpi.REQUEST_NULL] * len(reqs)
    assert (dst.download() == expected).all()
    for h in handles:
        mpi.free(h)


@pytest.mark.parametrize("method", [ONESHOT, STAGED])
def test_neighbor_alltoallw_halo_exchange_self(mpi, method):
    """A periodic halo exchange of one 134^3 float64 subdomain (radius 3)
    whose 26 neighbours are all this process: the interior halo sent in
    direction d lands in the exterior halo on the opposite side.  Called
    twice with fresh data (the second call reuses the kept batches)."""
    ta.lib().tempi_set_datatype_method(method)
    dirs = tc.halo_directions()
    lcr = (128, 128, 128)
    stypes = [mpi.type_from_spec(tc.halo_type(3, lcr, d, 8)) for d in dirs]
    rspecs = [tc.halo_type(3, lcr, tuple(-x for x in d), 8, exterior=True) for d in dirs]
    rtypes = [mpi.type_from_spec(sp) for sp in rspecs]
    comm = mpi.dist_graph([0] * len(dirs), [0] * len(dirs))
    n = 134 ** 3 * 8
    src = ta.DeviceBuffer(n)
    dst = ta.DeviceBuffer(n)
    c0 = ta.counters()[COUNTER[method]]
    for rep in range(2):
        host = orc.splitmix_bytes(n, 100 + rep)
        src.upload(host)
        dst.fill(0xAB)
        expected = np.full(n, 0xAB, np.uint8)
        for d, rs in zip(dirs, rspecs):
            packed = orc.OracleType(tc.halo_type(3, lcr, d, 8)).pack(host, 0, 1)
            orc.OracleType(rs).unpack(packed, expected, 0, 1)
        k = len(dirs)
        err = mpi.neighbor_alltoallw(src.addr(), [1] * k, [0] * k, stypes,
                                     dst.addr(), [1] * k, [0] * k, rtypes, comm)
        assert err == 0, rep
        assert (dst.download() == expected).all(), rep
    # one argument set built, then reused
    assert ta.counters()[COUNTER[method]] == c0 + 1
    mpi.comm_free(comm)
    for h in stypes + rtypes:
        mpi.free(h)


def test_neighbor_alltoallw_with_mixed_buffers_is_an_error(mpi):
    """One buffer in device memory and the other in host memory: refused
    with MPI_ERR_BUFFER instead of handing the MPI library (not GPU-aware
    here) a device pointer; nothing is written."""
    mpi.errors_return()
    h = mpi.type_from_spec(FACE)
    comm = mpi.dist_graph([0], [0])
    o = orc.OracleType(FACE)
    src = ta.DeviceBuffer(o.extent())
    src.fill(1)
    host_dst = np.full(o.extent(), 0xAB, np.uint8)
    err = mpi.neighbor_alltoallw(src.addr(), [1], [0], [h], host_dst.ctypes.data, [1], [0],
                                 [h], comm)
    assert err == mpi.const("MPI_ERR_BUFFER")
    assert (host_dst == 0xAB).all()
    mpi.comm_free(comm)
    mpi.free(h)


def test_wildcards_proc_null_and_truncation(mpi):
    """MPI_ANY_SOURCE / MPI_ANY_TAG receives report the real source and tag;
    MPI_PROC_NULL goes to the library; a message longer than the receive is
    a truncation error and leaves the receive buffer untouched."""
    mpi.errors_return()
    ta.lib().tempi_set_datatype_method(ONESHOT)
    h, src, dst, expected = setup(mpi, FACE, 2, 19)
    any_src, any_tag = mpi.const("MPI_ANY_SOURCE"), mpi.const("MPI_ANY_TAG")
    proc_null = mpi.const("MPI_PROC_NULL")
    one = np.full(dst.nbytes, 0xAB, np.uint8)
    o = orc.OracleType(FACE)
    host = orc.splitmix_bytes(dst.nbytes, 19)
    o.unpack(o.pack(host, 0, 1), one, 0, 1)
    err, rr = mpi.irecv(dst.addr(), 1, h, any_src, any_tag)
    err, sr = mpi.isend(src.addr(), 1, h, 0, 77)
    err, sts, _ = mpi.waitall([rr, sr], [h, h])
    assert err == 0 and sts[0]["source"] == 0 and sts[0]["tag"] == 77
    assert (dst.download() == one).all()
    lib0 = ta.counters()["p2p_library"]
    err, nr = mpi.isend(src.addr(), 1, h, proc_null, 78)
    assert err == 0
    err, _, nr = mpi.wait(nr)
    assert err == 0 and ta.counters()["p2p_library"] == lib0 + 1
    dst.fill(0xAB)
    err, rr = mpi.irecv(dst.addr(), 1, h, 0, 79)
    err, sr = mpi.isend(src.addr(), 2, h, 0, 79)
    err, _, rr = mpi.wait(rr, h)
    assert err != 0 and mpi.error_class(err) == mpi.ERR_TRUNCATE
    err, _, sr = mpi.wait(sr)
    assert err == 0
    assert (dst.download() == 0xAB).all()
    mpi.free(h)


RANK_SCRIPT = r'''
import ctypes, sys
sys.path.insert(0, ROOT); sys.path.insert(0, ROOT + "/tests")
import numpy as np
import tempi_amd as ta, typecatalog as tc
from tempi_amd import mpi as tmpi
from oracle import oracle as orc

m = tmpi.get()
rank = ctypes.c_int()
m.tempi.MPI_Comm_rank.argtypes = [m.H, ctypes.POINTER(ctypes.c_int)]
m.tempi.MPI_Comm_rank(m.COMM_WORLD, ctypes.byref(rank))
me = rank.value
peer = 1 - me
print(f"rank {me} up", flush=True)


def grid(seed, n):
    host = orc.splitmix_bytes(n, seed)
    d = ta.DeviceBuffer(n)
    d.upload(host)
    return host, d


# each rank sends one type and receives the other's: the x face of a 38^3
# float grid (12-B rows) one way, a byte vector the other way
specs = [tc.grid_face_subarray(32, 3, "x"), ("vector", 300, 40, 77, tc.BYTE)]
for method in (2, 3):  # ONESHOT, STAGED
    ta.lib().tempi_set_datatype_method(method)
    mine, theirs = specs[me], specs[peer]
    ho = orc.OracleType(mine)
    hs = m.type_from_spec(mine)
    hr = m.type_from_spec(theirs)
    tr = orc.OracleType(theirs)
    src_host, src = grid(100 * method + me, ho.extent() * 2)
    dst = ta.DeviceBuffer(tr.extent() * 2)
    dst.fill(0xAB)
    c0 = ta.counters()
    err, rr = m.irecv(dst.addr(), 2, hr, peer, 7)
    assert err == 0
    err, sr = m.isend(src.addr(), 2, hs, peer, 7)
    assert err == 0
    err, sts, after = m.waitall([rr, sr], [hr, hs])
    assert err == 0 and sts[0]["count"] == 2 and sts[0]["source"] == peer, sts
    peer_host = orc.splitmix_bytes(tr.extent() * 2, 100 * method + peer)
    exp = np.full(tr.extent() * 2, 0xAB, np.uint8)
    tr.unpack(tr.pack(peer_host, 0, 2), exp, 0, 2)
    assert (dst.download() == exp).all(), ("p2p", method)
    key = {2: "p2p_oneshot", 3: "p2p_staged"}[method]
    assert ta.counters()[key] == c0[key] + 2, (key, ta.counters())
    m.free(hs)
    m.free(hr)
    print(f"rank {me} p2p method {method} ok", flush=True)

# neighbourhood exchange over a 2-rank graph: the +x interior halo of a 20^3
# float64 subdomain (radius 3) goes to the peer, whose -x exterior halo takes it
lcr = (14, 14, 14)
n = 20 ** 3 * 8
comm = m.dist_graph([peer], [peer])
st = m.type_from_spec(tc.halo_type(3, lcr, (1, 0, 0), 8))
rspec = tc.halo_type(3, lcr, (-1, 0, 0), 8, exterior=True)
rt = m.type_from_spec(rspec)
for method in (2, 3):
    ta.lib().tempi_set_datatype_method(method)
    src_host, src = grid(500 + 10 * method + me, n)
    dst = ta.DeviceBuffer(n)
    dst.fill(0xAB)
    err = m.neighbor_alltoallw(src.addr(), [1], [0], [st], dst.addr(), [1], [0], [rt], comm)
    assert err == 0, err
    peer_host = orc.splitmix_bytes(n, 500 + 10 * method + peer)
    exp = np.full(n, 0xAB, np.uint8)
    packed = orc.OracleType(tc.halo_type(3, lcr, (1, 0, 0), 8)).pack(peer_host, 0, 1)
    orc.OracleType(rspec).unpack(packed, exp, 0, 1)
    assert (dst.download() == exp).all(), ("neighbor", method)
    print(f"rank {me} neighbor method {method} ok", flush=True)
m.comm_free(comm)
m.free(st)
m.free(rt)
ta.lib().tempi_set_datatype_method(-1)
m.finalize()
print(f"rank {me} done", flush=True)
'''


def test_two_ranks_device_buffers(tmp_path):
    """Two processes on the one GPU under mpiexec: device-resident derived
    types exchanged both ways through the interposed MPI_Isend / MPI_Irecv
    (ONESHOT and STAGED) and MPI_Neighbor_alltoallw on a 2-rank graph; each
    receiver checks the bytes against the oracle, gaps untouched."""
    import os
    import shutil
    import signal
    import subprocess
    import sys

    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if not os.path.exists(mpiexec):
        pytest.skip("no mpiexec")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = tmp_path / "ranks.py"
    script.write_text(f"ROOT = {root!r}\n" + RANK_SCRIPT)
    print("two ranks: p2p ONESHOT/STAGED, neighbor_alltoallw", flush=True)
    p = subprocess.Popen([mpiexec, "-n", "2", "-hosts", "127.0.0.1", sys.executable, str(script)],
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         start_new_session=True)
    try:
        out, _ = p.communicate(timeout=120)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        out, _ = p.communicate()
        raise AssertionError("two-rank exchange timed out:\n" + out)
    print(out, flush=True)
    assert p.returncode == 0, out
    for r in (0, 1):
        assert f"rank {r} done" in out, out


def test_method_choice_follows_replaced_tables(mpi):
    """With no method forced and no method model, a send of device data takes
    the method the measured tables price cheapest; replacing the tables
    re-prices the next send of the same size (cached choices belong to the
    tables they were made on)."""
    from tempi_amd import system as tsys
    ta.lib().tempi_set_datatype_method(0)  # by model
    ta.lib().tempi_method_model_unload()  # the tables decide
    h, src, dst, expected = setup(mpi, FACE, 1, 23)

    def tables(pack_host_s, pack_device_s):
        tsys.clear()
        for c in (tsys.D2H, tsys.H2D, tsys.INTRA_CPU_CPU):
            tsys.set_curve(c, np.full(tsys.CURVE_LEN, 1e-6))
        tsys.set_table(tsys.PACK_HOST, np.full((9, 9), pack_host_s))
        tsys.set_table(tsys.UNPACK_HOST, np.full((9, 9), pack_host_s))
        tsys.set_table(tsys.PACK_DEVICE, np.full((9, 9), pack_device_s))
        tsys.set_table(tsys.UNPACK_DEVICE, np.full((9, 9), pack_device_s))

    try:
        for cheap, key in (("host", "p2p_oneshot"), ("device", "p2p_staged")):
            tables(1e-6 if cheap == "host" else 1.0, 1e-6 if cheap == "device" else 1.0)
            dst.fill(0xAB)
            c0 = ta.counters()[key]
            err, rr = mpi.irecv(dst.addr(), 1, h, 0, 61)
            err, sr = mpi.isend(src.addr(), 1, h, 0, 61)
            err, _, _ = mpi.waitall([rr, sr], [h, h])
            assert err == 0
            assert ta.counters()[key] == c0 + 2, (cheap, ta.counters())
            assert (dst.download() == expected).all(), cheap
    finally:
        tsys.clear()
        ta.lib().tempi_method_model_load(None)  # the installed model, if any
        mpi.free(h)


def test_method_model_drives_the_choice(mpi, tmp_path):
    """With no method forced, a send to a peer on this node takes the method
    the method model prices cheapest: two crafted weight files (all weights
    zero, output biases favouring ONE_SHOT in one and STAGED in the other)
    each decide the method of the next exchange, whose bytes arrive intact;
    loading a model re-prices shapes chosen before it."""
    L = ta.lib()
    n, out = L.tempi_mlp_num_params(), L.tempi_mlp_num_outputs()
    norm = np.concatenate([np.zeros(16, np.float32), np.ones(16, np.float32)])
    ta.lib().tempi_set_datatype_method(0)
    h, src, dst, expected = setup(mpi, FACE, 1, 29)
    try:
        for fast, key in ((ONESHOT, "p2p_oneshot"), (STAGED, "p2p_staged")):
            params = np.zeros(n, np.float32)
            params[n - out:] = 1.0
            params[n - out + fast - 1] = 0.0  # output k: method k + 1
            path = str(tmp_path / f"method{fast}.txt")
            assert L.tempi_mlp_save(path.encode(), params.ctypes.data, norm.ctypes.data) == 0
            assert L.tempi_method_model_load(path.encode()) == 0
            c0 = ta.counters()
            assert L.tempi_p2p_method(1, 4096, 64, 192) == fast
            dst.fill(0xAB)
            err, rr = mpi.irecv(dst.addr(), 1, h, 0, 71)
            err, sr = mpi.isend(src.addr(), 1, h, 0, 71)
            err, _, _ = mpi.waitall([rr, sr], [h, h])
            assert err == 0
            c1 = ta.counters()
            assert c1[key] == c0[key] + 2, (fast, c1)
            assert c1["method_model_choices"] > c0["method_model_choices"]
            assert (dst.download() == expected).all(), fast
        # a peer on another node is priced by the tables, not the model
        c0 = ta.counters()["method_model_choices"]
        L.tempi_p2p_method(0, 4096, 64, 192)
        assert ta.counters()["method_model_choices"] == c0
    finally:
        L.tempi_method_model_load(None)  # the installed model, or none
        mpi.free(h)


def test_installed_method_model_is_consulted():
    """The shipped method model (tempi_amd/method_mlp.txt) loads, and the
    library asks it for a same-node message whose method is not forced."""
    L = ta.lib()
    assert L.tempi_method_model_load(None) == 0, "tempi_amd/method_mlp.txt is missing"
    ta.lib().tempi_set_datatype_method(0)
    c0 = ta.counters()["method_model_choices"]
    # a shape no other test prices, so the choice is not cached
    m = L.tempi_p2p_method(1, 12345 * 1024, 1024, 3000)
    assert m in (ONESHOT, STAGED)
    assert ta.counters()["method_model_choices"] == c0 + 1


STRUCT2 = ("struct", [3, 2], [0, 20], [("named", "MPI_FLOAT"), ("named", "MPI_DOUBLE")])


def test_device_data_without_strided_form_sent_through_host_images(mpi):
    """Isend/Irecv and blocking Send of a two-block struct on device buffers
    (no strided form; the library cannot read device memory): the library
    moves pinned host images of the buffers' spans, the receiver's bytes
    between the blocks keep 0xAB, and the result equals MPICH's unpack of
    MPICH's pack of the host copy."""
    h = mpi.type_from_spec(STRUCT2)
    lb, ext = mpi.extent(h)
    count = 4
    n = ext * count
    host = orc.splitmix_bytes(n, 41)
    size = mpi.pack_size(count, h)
    packed = np.zeros(size, np.uint8)
    assert mpi.pack(host.ctypes.data, count, h, packed.ctypes.data, size, library=True)[0] == 0
    expected = np.full(n, 0xAB, np.uint8)
    assert mpi.unpack(packed.ctypes.data, size, expected.ctypes.data, count, h,
                      library=True)[0] == 0
    src = ta.DeviceBuffer(n)
    src.upload(host)
    dst = ta.DeviceBuffer(n)
    dst.fill(0xAB)
    c0 = ta.counters()
    err, rr = mpi.irecv(dst.addr(), count, h, 0, 31)
    assert err == 0
    err, sr = mpi.isend(src.addr(), count, h, 0, 31)
    assert err == 0
    err, sts, reqs = mpi.waitall([rr, sr], [h, h])
    assert err == 0 and reqs == [mpi.REQUEST_NULL] * 2 and sts[0]["count"] == count
    assert (dst.download() == expected).all()
    assert ta.counters()["host_staged_calls"] == c0["host_staged_calls"] + 2
    dst.fill(0xAB)
    err, rr = mpi.irecv(dst.addr(), count, h, 0, 32)
    assert err == 0
    assert mpi.send(src.addr(), count, h, 0, 32) == 0
    flag = 0
    while not flag:
        err, flag, st, rr = mpi.test(rr, h)
        assert err == 0
    assert (dst.download() == expected).all()
    mpi.free(h)


def test_pending_receive_keeps_device_writes_between_the_blocks(mpi):
    """A receive of a two-block struct into device memory (host image, the
    library cannot read device memory) is posted; while it is pending every
    byte between the type map's blocks is rewritten on the device (a kernel
    updating the rest of the buffer, as in a halo exchange overlapped with
    interior work).  Completion writes the received type map and nothing
    else: the rewritten bytes keep their new values."""
    h = mpi.type_from_spec(STRUCT2)
    lb, ext = mpi.extent(h)
    count = 4
    n = ext * count
    size = mpi.pack_size(count, h)
    # the type map's bytes: MPICH unpacks all-ones packed bytes into zeros
    ones = np.full(size, 0xFF, np.uint8)
    tm = np.zeros(n, np.uint8)
    assert mpi.unpack(ones.ctypes.data, size, tm.ctypes.data, count, h, library=True)[0] == 0
    in_map = tm != 0
    assert in_map.sum() == size and (~in_map).sum() > 0
    host = orc.splitmix_bytes(n, 47)
    packed = np.zeros(size, np.uint8)
    assert mpi.pack(host.ctypes.data, count, h, packed.ctypes.data, size, library=True)[0] == 0
    received = np.zeros(n, np.uint8)
    assert mpi.unpack(packed.ctypes.data, size, received.ctypes.data, count, h,
                      library=True)[0] == 0
    expected = np.where(in_map, received, 0xCD).astype(np.uint8)
    src = ta.DeviceBuffer(n)
    src.upload(host)
    dst = ta.DeviceBuffer(n)
    dst.fill(0xAB)
    err, rr = mpi.irecv(dst.addr(), count, h, 0, 35)
    assert err == 0
    dst.upload(np.where(in_map, 0xAB, 0xCD).astype(np.uint8))
    err, sr = mpi.isend(src.addr(), count, h, 0, 35)
    assert err == 0
    err, sts, _ = mpi.waitall([rr, sr], [h, h])
    assert err == 0 and sts[0]["count"] == count
    got = dst.download()
    assert (got[~in_map] == 0xCD).all()
    assert (got == expected).all()
    mpi.free(h)


def test_contiguous_device_data_is_staged_without_gpu_aware_library(mpi):
    """MPI_FLOAT device buffers (contiguous) are staged through the host
    instead of reaching a library that cannot read device memory."""
    h = mpi.build(("named", "MPI_FLOAT"))
    host = orc.splitmix_bytes(4096, 43)
    src = ta.DeviceBuffer(4096)
    src.upload(host)
    dst = ta.DeviceBuffer(4096)
    dst.fill(0)
    c0 = ta.counters()
    err, rr = mpi.irecv(dst.addr(), 1024, h, 0, 33)
    assert err == 0
    err, sr = mpi.isend(src.addr(), 1024, h, 0, 33)
    assert err == 0
    err, _, _ = mpi.waitall([rr, sr], [h, h])
    assert err == 0
    assert (dst.download() == host).all()
    assert ta.counters()["p2p_library"] == c0["p2p_library"]
