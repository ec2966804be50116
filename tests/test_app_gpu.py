"""The drop-in use on the GPU: a native MPI program (tests/c/pack_unpack_app.cpp,
after the reference's test/pack_unpack.cpp) run with libtempi_amd.so
preloaded in front of the MPI library."""
import os
import re
import subprocess

import pytest

import tempi_amd as ta

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "tests", "c", "pack_unpack_app")


def test_native_mpi_program_with_preload():
    assert os.path.exists(APP), "build it with __graft_entry__.build()"
    env = dict(os.environ, LD_PRELOAD=ta.LIB_PATH, TEMPI_LOG_LEVEL="1")
    r = subprocess.run([APP], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 14 and "FAIL" not in r.stdout, r.stdout
    m = re.search(r"pack kernels (\d+) unpack kernels (\d+) library packs (\d+)", r.stdout)
    assert m, r.stdout
    # every device pack / unpack ran a kernel; the 14 host-copy packs (the
    # expected bytes) went to the library
    assert int(m.group(1)) == 14 and int(m.group(2)) == 14 and int(m.group(3)) == 14
