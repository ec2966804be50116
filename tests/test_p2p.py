"""Point-to-point calls on the CPU: host buffers go to the MPI library as
they are (reference: src/internal/send.cpp:24-36, isend.cpp:20-26), with
this process sending to itself; and the argument checks of the method
override."""
import ctypes

import numpy as np

import tempi_amd as ta
import typecatalog as tc
from oracle import oracle as orc


def test_host_buffers_take_the_library_path(mpi):
    spec = ("vector", 64, 3, 7, tc.BYTE)
    h = mpi.type_from_spec(spec)
    o = orc.OracleType(spec)
    src = orc.splitmix_bytes(o.extent(), 11)
    dst = np.full(o.extent(), 0xAB, np.uint8)
    before = ta.counters()["p2p_library"]
    err, rr = mpi.irecv(dst.ctypes.data, 1, h, 0, 5)
    assert err == 0
    err, sr = mpi.isend(src.ctypes.data, 1, h, 0, 5)
    assert err == 0
    err, sts, reqs = mpi.waitall([sr, rr], [h, h])
    assert err == 0 and all(r == mpi.REQUEST_NULL for r in reqs)
    assert sts[1]["count"] == 1 and sts[1]["source"] == 0 and sts[1]["tag"] == 5
    expected = np.full(o.extent(), 0xAB, np.uint8)
    o.unpack(o.pack(src, 0, 1), expected, 0, 1)
    assert (dst == expected).all()
    assert ta.counters()["p2p_library"] == before + 2
    # blocking send into a receive posted first, then MPI_Test to completion
    dst[:] = 0xAB
    err, rr = mpi.irecv(dst.ctypes.data, 1, h, 0, 6)
    assert mpi.send(src.ctypes.data, 1, h, 0, 6) == 0
    flag = 0
    for _ in range(10000):
        err, flag, st, rr = mpi.test(rr, h)
        if flag:
            break
    assert flag and st["count"] == 1 and rr == mpi.REQUEST_NULL
    assert (dst == expected).all()
    mpi.free(h)


def test_method_override_arguments():
    L = ta.lib()
    assert L.tempi_set_datatype_method(99) == ta.TEMPI_ERR_ARG
    prev = L.tempi_set_datatype_method(2)  # ONESHOT
    assert L.tempi_set_datatype_method(prev) == 2


def test_neighbor_alltoallw_host_buffers_go_to_the_library(mpi):
    """Host buffers: the library's own MPI_Neighbor_alltoallw, on a graph
    whose two edges lead from this process to itself."""
    spec = ("vector", 16, 4, 9, tc.BYTE)
    h = mpi.type_from_spec(spec)
    o = orc.OracleType(spec)
    comm = mpi.dist_graph([0, 0], [0, 0])
    src = orc.splitmix_bytes(2 * o.extent(), 21)
    dst = np.full(src.size, 0xAB, np.uint8)
    before = ta.counters()["p2p_library"]
    err = mpi.neighbor_alltoallw(src.ctypes.data, [1, 1], [0, o.extent()], [h, h],
                                 dst.ctypes.data, [1, 1], [o.extent(), 0], [h, h], comm)
    assert err == 0
    assert ta.counters()["p2p_library"] == before + 1
    # block i sent to the first neighbour lands at receive block i
    expected = np.full(src.size, 0xAB, np.uint8)
    o.unpack(o.pack(src, 0, 1), expected, o.extent(), 1)
    o.unpack(o.pack(src, o.extent(), 1), expected, 0, 1)
    assert (dst == expected).all()
    mpi.comm_free(comm)
    mpi.free(h)


def test_two_ranks_host_buffers(tmp_path):
    """Two processes under mpiexec (CPU only): MPI_Init through the library
    builds its node map with one MPI_Allgather, and host-buffer sends of a
    derived type between the ranks go to the MPI library and arrive intact."""
    import os
    import shutil
    import subprocess
    import sys

    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if not os.path.exists(mpiexec):
        import pytest
        pytest.skip("no mpiexec")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = tmp_path / "two.py"
    script.write_text(
        "import sys, ctypes\n"
        f"sys.path.insert(0, {root!r}); sys.path.insert(0, {os.path.join(root, 'tests')!r})\n"
        "import numpy as np\n"
        "import tempi_amd as ta, typecatalog as tc\n"
        "from tempi_amd import mpi as tmpi\n"
        "from oracle import oracle as orc\n"
        "m = tmpi.get()\n"
        "rank = ctypes.c_int()\n"
        "m.tempi.MPI_Comm_rank.argtypes = [m.H, ctypes.POINTER(ctypes.c_int)]\n"
        "m.tempi.MPI_Comm_rank(m.COMM_WORLD, ctypes.byref(rank))\n"
        "spec = ('vector', 32, 5, 11, tc.BYTE)\n"
        "h = m.type_from_spec(spec)\n"
        "o = orc.OracleType(spec)\n"
        "src = orc.splitmix_bytes(o.extent(), 31)\n"
        "if rank.value == 0:\n"
        "    assert m.send(src.ctypes.data, 1, h, 1, 9) == 0\n"
        "else:\n"
        "    dst = np.full(o.extent(), 0xAB, np.uint8)\n"
        "    err, st = m.recv(dst.ctypes.data, 1, h, 0, 9)\n"
        "    exp = np.full(o.extent(), 0xAB, np.uint8)\n"
        "    o.unpack(o.pack(src, 0, 1), exp, 0, 1)\n"
        "    assert err == 0 and st['count'] == 1 and st['source'] == 0\n"
        "    assert (dst == exp).all()\n"
        "    print('rank 1 ok', flush=True)\n"
        "m.free(h)\n"
        "m.finalize()\n")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="", ROCR_VISIBLE_DEVICES="")
    r = subprocess.run([mpiexec, "-n", "2", "-hosts", "127.0.0.1", sys.executable, str(script)],
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "rank 1 ok" in r.stdout


def test_two_ranks_measure_the_pingpong_curve(tmp_path):
    """Two processes under mpiexec (CPU only) measure the CPU-CPU ping-pong
    curve together (TEMPI_MEASURE_PINGPONG): both end up with the same 24
    points (rank 0's measurement, broadcast), finite and positive, filed as
    the intra-node curve since the ranks share this node, and 8 MiB takes
    longer than 1 B."""
    import os
    import shutil
    import subprocess
    import sys

    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if not os.path.exists(mpiexec):
        import pytest
        pytest.skip("no mpiexec")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = tmp_path / "pingpong.py"
    script.write_text(
        "import sys, ctypes\n"
        f"sys.path.insert(0, {root!r})\n"
        "import numpy as np\n"
        "import tempi_amd as ta\n"
        "from tempi_amd import mpi as tmpi, system\n"
        "m = tmpi.get()\n"
        "rank = ctypes.c_int()\n"
        "m.tempi.MPI_Comm_rank.argtypes = [m.H, ctypes.POINTER(ctypes.c_int)]\n"
        "m.tempi.MPI_Comm_rank(m.COMM_WORLD, ctypes.byref(rank))\n"
        "system.clear()\n"
        "system.measure_performance(system.MEASURE_PINGPONG, 0.01)\n"
        "c = system.get_curve(system.INTRA_CPU_CPU)\n"
        "assert system.get_curve(system.INTER_CPU_CPU).size == 0\n"
        "assert c.shape == (24,) and np.isfinite(c).all() and (c > 0).all(), c\n"
        "assert c[23] > c[0], c\n"
        "print(f'rank {rank.value} curve ' + ' '.join(f'{x:.9e}' for x in c), flush=True)\n"
        "m.finalize()\n")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="", ROCR_VISIBLE_DEVICES="")
    r = subprocess.run([mpiexec, "-n", "2", "-hosts", "127.0.0.1", sys.executable, str(script)],
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    curves = {}
    for line in r.stdout.splitlines():
        if " curve " in line:
            who, vals = line.split(" curve ")
            curves[who] = vals
    assert set(curves) == {"rank 0", "rank 1"}, r.stdout
    assert curves["rank 0"] == curves["rank 1"]
