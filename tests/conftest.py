import json
import os
import sys

import pytest

TESTS = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(TESTS)
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and runs the HIP path")


@pytest.fixture(scope="session")
def mpi():
    """The process's MPI (singleton MPI_Init through libtempi_amd.so)."""
    from tempi_amd import mpi as tmpi

    m = tmpi.get()
    yield m
    m.finalize()


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(TESTS, "golden", "pack_golden.json")) as f:
        return json.load(f)["cases"]


def spec_from_json(s):
    """JSON turns tuples into lists: restore the tuple structure of a spec."""
    kind = s[0]
    if kind == "named":
        return ("named", s[1])
    if kind == "struct":
        return ("struct", list(s[1]), list(s[2]), [spec_from_json(c) for c in s[3]])
    if kind == "subarray":
        return ("subarray", list(s[1]), list(s[2]), list(s[3]), s[4], spec_from_json(s[5]))
    if kind in ("hindexed", "indexed"):
        return (kind, list(s[1]), list(s[2]), spec_from_json(s[3]))
    if kind in ("indexed_block", "hindexed_block"):
        return (kind, s[1], list(s[2]), spec_from_json(s[3]))
    if kind == "dup":
        return ("dup", spec_from_json(s[1]))
    # contiguous / vector / hvector / resized: scalars then the child
    return tuple(s[:-1]) + (spec_from_json(s[-1]),)


def signalled(sb, strided, packed, direction):
    """Whether the library's copy of one object runs as one launch of at most
    128 workgroups -- the copies that signal their own completion -- from the
    variant it picks and the grid its launch code forms (256-thread
    workgroups of 256 / lanes-per-row groups, each moving 1 unit, or 4 / U
    units when U < 4, at most 16 workgroups per CU of MI355X's 256)."""
    import tempi_amd as ta

    v = ta.select_variant(sb, 1, strided, packed, direction, ta.SELECT_DEFAULT)
    info = ta.variant_info(v)
    e = ta.copy_emulate(sb, 1, strided, packed, direction, v)
    k = 1 if info["unroll"] >= 4 else 4 // info["unroll"]
    groups = (256 // info["lanes_per_row"]) * k
    blocks = min(-(-e["max_units"] // groups), 16 * 256)
    return e["launches"] == 1 and blocks <= 128
