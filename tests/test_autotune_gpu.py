"""The autotuner on the GPU: its pick is the fastest of its own timings, the
per-shape cache makes a second request free, the MPI path under the
autotune policy packs the same bytes, and the timing helpers answer."""
import numpy as np
import pytest

import tempi_amd as ta
import typecatalog as tc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    assert ta.device_count() > 0


def test_autotune_picks_its_fastest_timing():
    sb = ta.StridedBlock.make(0, 4096 * 200, [48, 4096], [1, 200])
    src = ta.DeviceBuffer(sb.extent + 64)
    out = ta.DeviceBuffer(48 * 4096)
    v, cap, per = ta.autotune(sb, 1, src.addr(), out.addr(), ta.DIR_PACK)
    assert 0 <= v < ta.num_variants() and 0 <= cap < ta.lib().tempi_num_grid_caps()
    per = np.array(per)
    assert per.shape == (ta.num_variants(),) and np.isfinite(per).all() and (per > 0).all()
    # the chosen variant is the one with the best time over all grid caps
    assert per[v] == per.min()


def test_autotune_policy_caches_per_shape_and_packs_exactly():
    spec = ("vector", 3000, 20, 77, tc.BYTE)
    o = orc.OracleType(spec)
    host = orc.splitmix_bytes(o.extent() + 8, 4)
    sb = ta.StridedBlock.make(0, o.extent(), [20, 3000], [1, 77])
    src = ta.DeviceBuffer(host.size)
    src.upload(host)
    out = ta.DeviceBuffer(o.size())
    runs0 = ta.counters()["autotune_runs"]
    v1 = ta.select_variant(sb, 1, src.addr(8), out.addr(), ta.DIR_PACK, ta.SELECT_AUTOTUNE)
    assert ta.counters()["autotune_runs"] == runs0 + 1
    v2 = ta.select_variant(sb, 1, src.addr(8), out.addr(), ta.DIR_PACK, ta.SELECT_AUTOTUNE)
    assert v2 == v1 and ta.counters()["autotune_runs"] == runs0 + 1
    # a different alignment of the strided side is a different shape
    ta.select_variant(sb, 1, src.addr(3), out.addr(), ta.DIR_PACK, ta.SELECT_AUTOTUNE)
    assert ta.counters()["autotune_runs"] == runs0 + 2
    ta.pack_sb(sb, src.addr(8), 1, out.addr(), o.size(), 0, v1)
    ta.synchronize()
    assert (out.download() == o.pack(host, 8, 1)).all()


def test_timing_helpers():
    host_us, dev_us = ta.launch_latency_us(200)
    assert 0 < host_us < 1000 and 0 < dev_us < 1000
    ta.flush_caches(64 << 20)
    sb = ta.StridedBlock.make(0, 1 << 20, [1 << 20], [1])
    a, b = ta.DeviceBuffer(1 << 20), ta.DeviceBuffer(1 << 20)
    us = ta.time_variant(sb, 1, a.addr(), b.addr(), ta.DIR_PACK, 0, 0, 5)
    assert 0 < us < 10000


def test_first_cold_timing_after_settling_is_not_inflated():
    """Settling a 256 MiB copy's variant times candidates cold and then frees
    the cache-sweep buffer, so the next cold timing allocates and fills it
    again.  That fill is ordered before the sweep and the timed copy: the
    first timing agrees with the ones after it (with the fill on another
    stream it ran under the first timed copy, 40-70 % slower)."""
    n = 256 << 20
    bl = 64
    rows = n // bl
    stride = 2 * bl + 64
    sb = ta.StridedBlock.make(0, rows * stride, [bl, rows], [1, stride])
    src = ta.DeviceBuffer(rows * stride + 64)
    dst = ta.DeviceBuffer(n)
    s = ta.stream()
    ta.unpack_sb(sb, dst.addr(), n, src.addr(), 1, 0, -1, s)
    ta.synchronize(s)
    v = ta.select_variant(sb, 1, src.addr(), dst.addr(), ta.DIR_UNPACK, ta.SELECT_DEFAULT)
    t = [ta.time_variant(sb, 1, src.addr(), dst.addr(), ta.DIR_UNPACK, v, 0, 1)
         for _ in range(6)]
    rest = float(np.median(t[1:]))
    print("first", t[0], "then", t[1:], flush=True)
    assert t[0] <= 1.25 * rest, t
    src.free()
    dst.free()
