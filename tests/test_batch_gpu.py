"""GPU parity of the batched queue kernel and its hipGraph replay."""
import random

import numpy as np
import pytest

import tempi_amd as ta
import typecatalog as tc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    assert ta.device_count() > 0, "no GPU visible"


def halo_requests(mpi, nsub, lcr, radius, quant, grids, packed, direction, exterior=False):
    """26 halo subarrays of each of `nsub` subdomains; returns (requests,
    [(sub, dir, sb, packed offset, size)])."""
    reqs, meta = [], []
    off = 0
    for s in range(nsub):
        for d in tc.halo_directions():
            h = mpi.type_from_spec(tc.halo_type(radius, lcr, d, quant, exterior=exterior))
            sb, size = mpi.strided_block(h)
            mpi.free(h)
            reqs.append((sb, grids[s], packed.addr(off), 1, direction))
            meta.append((s, d, sb, off, size))
            off += size
    return reqs, meta, off


def test_halo_batch_pack_and_unpack(mpi):
    nsub, lcr, radius, quant = 4, (16, 12, 10), 3, 8
    ext = [n + 2 * radius for n in lcr]
    sub_bytes = ext[0] * ext[1] * ext[2] * quant
    hosts = [orc.splitmix_bytes(sub_bytes, 100 + s) for s in range(nsub)]
    dev = [ta.DeviceBuffer(sub_bytes) for _ in range(nsub)]
    for d, h in zip(dev, hosts):
        d.upload(h)
    total = 0
    for d in tc.halo_directions():
        o = orc.OracleType(tc.halo_type(radius, lcr, d, quant))
        total += o.size()
    packed = ta.DeviceBuffer(total * nsub)
    packed.fill(0x5A)
    reqs, meta, size = halo_requests(mpi, nsub, lcr, radius, quant,
                                     [d.addr() for d in dev], packed, ta.DIR_PACK)
    assert size == total * nsub
    b = ta.Batch(reqs)
    assert b.info()["requests"] == 26 * nsub
    before = ta.counters()
    b.launch()
    ta.synchronize()
    assert ta.counters()["batch_launches"] == before["batch_launches"] + 1
    assert ta.counters()["pack_kernel_launches"] == before["pack_kernel_launches"]
    got = packed.download()
    for s, d, sb, off, sz in meta:
        exp = orc.sb_pack(orc.sb_desc(*sb.as_tuple()), hosts[s], 0, 1)
        assert (got[off:off + sz] == exp).all(), (s, d)
    # unpack the packed halos into the exterior ghost cells of fresh grids
    outs = [ta.DeviceBuffer(sub_bytes) for _ in range(nsub)]
    for o in outs:
        o.fill(0xAB)
    ureqs, umeta, usize = halo_requests(mpi, nsub, lcr, radius, quant,
                                        [o.addr() for o in outs], packed, ta.DIR_UNPACK,
                                        exterior=True)
    assert usize == size
    ub = ta.Batch(ureqs)
    ub.launch()
    ta.synchronize()
    for s in range(nsub):
        ref = np.full(sub_bytes, 0xAB, np.uint8)
        for s2, d, sb, off, sz in umeta:
            if s2 == s:
                orc.sb_unpack(orc.sb_desc(*sb.as_tuple()), got[off:off + sz], ref, 0, 1)
        assert (outs[s].download() == ref).all(), s


def random_sb(rng):
    """A random 1D/2D/3D strided block (start, extent, counts, strides) and
    its span in bytes."""
    nd = rng.choice([1, 2, 2, 3, 3])
    bl = rng.choice([1, 3, 4, 8, 12, 16, 24, 64, 100, 256, 1000, 2048, 5000])
    counts, strides = [bl], [1]
    inner = bl
    for _ in range(nd - 1):
        c = rng.randint(1, 40)
        st = inner + rng.randint(0, 300)
        counts.append(c)
        strides.append(st)
        inner = st * (c - 1) + inner
    start = rng.randint(0, 31)
    extent = start + inner + rng.randint(0, 64)
    return ta.StridedBlock.make(start, extent, counts, strides), extent


def test_mixed_random_requests():
    rng = random.Random(77)
    n = 300
    sbs = []
    span_total = 0
    for _ in range(n):
        sb, extent = random_sb(rng)
        count = rng.randint(1, 3)
        base = rng.randint(0, 15)
        sbs.append((sb, count, base, span_total))
        span_total += extent * count + 64
    host = orc.splitmix_bytes(span_total, 5)
    src = ta.DeviceBuffer(span_total)
    src.upload(host)
    sizes = [sb.size() * c for sb, c, _, _ in sbs]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(int)
    packed = ta.DeviceBuffer(int(sum(sizes)) + 16)
    reqs = [(sb, src.addr(o + base), packed.addr(int(po) + 3), c, ta.DIR_PACK)
            for (sb, c, base, o), po in zip(sbs, offs)]
    b = ta.Batch(reqs)
    b.launch()
    ta.synchronize()
    got = packed.download()
    for (sb, c, base, o), po, sz in zip(sbs, offs, sizes):
        exp = orc.sb_pack(orc.sb_desc(*sb.as_tuple()), host, o + base, c)
        assert (got[po + 3:po + 3 + sz] == exp).all(), (sb, c, base)


def test_graph_replays_track_fresh_inputs():
    """Each replay (graph or plain launch, interleaved with eager uploads)
    must pack the source as it is at that moment: a queue counter left
    behind by an earlier launch would skip tiles and fail this."""
    sbs = [ta.StridedBlock.make(8, 64 * 2072, [12, 64], [1, 2072]),
           ta.StridedBlock.make(0, 40000, [2048, 3], [1, 4096]),
           ta.StridedBlock.make(3, 3000, [7, 20, 5], [1, 11, 600])]
    span = 200000
    src = ta.DeviceBuffer(span)
    sizes = [sb.size() for sb in sbs]
    packed = ta.DeviceBuffer(sum(sizes))
    offs = [0, sizes[0], sizes[0] + sizes[1]]
    reqs = [(sb, src.addr(1000 * i), packed.addr(o), 1, ta.DIR_PACK)
            for i, (sb, o) in enumerate(zip(sbs, offs))]
    b = ta.Batch(reqs)
    for rep in range(6):
        host = orc.splitmix_bytes(span, 1000 + rep)
        src.upload(host)
        packed.fill(0)
        if rep % 3 == 2:
            b.launch()
        else:
            b.graph_launch()
        ta.synchronize()
        got = packed.download()
        for i, (sb, o, sz) in enumerate(zip(sbs, offs, sizes)):
            exp = orc.sb_pack(orc.sb_desc(*sb.as_tuple()), host, 1000 * i, 1)
            assert (got[o:o + sz] == exp).all(), (rep, i)
    assert ta.counters()["graph_launches"] >= 4


def test_batch_matches_per_call():
    """The same requests through the batch and through one tempi_pack_sb each."""
    rng = random.Random(3)
    sbs = [random_sb(rng) for _ in range(40)]
    span = sum(e + 64 for _, e in sbs)
    host = orc.splitmix_bytes(span, 8)
    src = ta.DeviceBuffer(span)
    src.upload(host)
    sizes = [sb.size() for sb, _ in sbs]
    total = sum(sizes)
    a = ta.DeviceBuffer(total)
    c = ta.DeviceBuffer(total)
    reqs, pos, o = [], 0, 0
    for (sb, e), sz in zip(sbs, sizes):
        reqs.append((sb, src.addr(o), a.addr(pos), 1, ta.DIR_PACK))
        ta.pack_sb(sb, src.addr(o), 1, c.addr(), total, pos)
        pos += sz
        o += e + 64
    ta.Batch(reqs).launch()
    ta.synchronize()
    assert (a.download() == c.download()).all()


def test_empty_and_zero_count_requests():
    sb = ta.StridedBlock.make(0, 64, [16, 2], [1, 32])
    src = ta.DeviceBuffer(64)
    out = ta.DeviceBuffer(64)
    b = ta.Batch([(sb, src.addr(), out.addr(), 0, ta.DIR_PACK)])
    assert b.info()["tiles"] == 0
    b.launch()
    b.graph_launch()
    ta.synchronize()
    b2 = ta.Batch([])
    b2.launch()
    ta.synchronize()


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8, 9])
def test_small_batches(k):
    """Batches of up to 8 requests launch with the whole batch in the kernel
    arguments; 9 takes the table-driven launch.  One request is large enough
    that its tiles outnumber the persistent grid (several passes per
    workgroup), the others are random shapes of both directions."""
    rng = random.Random(500 + k)
    big = ta.StridedBlock.make(4, 4160 * 16384, [4096, 16384], [1, 4160])
    shapes = [(big, big.extent)] + [random_sb(rng) for _ in range(k - 1)]
    span = sum(e + 64 for _, e in shapes)
    host = orc.splitmix_bytes(span, 40 + k)
    sizes = [sb.size() for sb, _ in shapes]
    phost = orc.splitmix_bytes(sum(sizes), 80 + k)
    src = ta.DeviceBuffer(span)
    src.upload(host)
    packed = ta.DeviceBuffer(sum(sizes))
    packed.upload(phost)
    reqs, meta, o, po = [], [], 0, 0
    for i, ((sb, e), sz) in enumerate(zip(shapes, sizes)):
        d = ta.DIR_UNPACK if i % 2 else ta.DIR_PACK
        reqs.append((sb, src.addr(o), packed.addr(po), 1, d))
        meta.append((sb, d, o, po, sz))
        o += e + 64
        po += sz
    b = ta.Batch(reqs)
    # more tiles than 8 workgroups on each of MI355X's 256 CUs
    assert b.info()["tiles"] > 8 * 256
    b.launch()
    ta.synchronize()
    got_s, got_p = src.download(), packed.download()
    exp_s = host.copy()
    for sb, d, so, pk, sz in meta:
        desc = orc.sb_desc(*sb.as_tuple())
        if d == ta.DIR_PACK:
            assert (got_p[pk:pk + sz] == orc.sb_pack(desc, host, so, 1)).all(), (k, sb)
        else:
            orc.sb_unpack(desc, phost[pk:pk + sz], exp_s, so, 1)
    assert (got_s == exp_s).all(), k


def test_batch_creation_leaves_buffers_alone():
    """Creating a batch settles the variants of its large requests by timing
    trial copies; they write scratch, not the requests' buffers.  An unpack
    request's strided buffer and a pack request's packed buffer keep their
    contents until the batch runs, and then hold exactly the copy."""
    assert ta.lib().tempi_mlp_load(None) == 0
    c0 = ta.counters()
    # 64 MiB of 1-KiB rows at a 1088-B stride (a shape no other test uses)
    rows = 65536
    sb = ta.StridedBlock.make(16, 16 + (rows - 1) * 1088 + 1024, [1024, rows], [1, 1088])
    desc = orc.sb_desc(*sb.as_tuple())
    n = sb.extent + 64
    host = orc.splitmix_bytes(n, 61)
    phost = orc.splitmix_bytes(sb.size(), 62)
    src = ta.DeviceBuffer(n)
    src.upload(host)
    out = ta.DeviceBuffer(sb.size())
    out.fill(0x5A)
    pk = ta.DeviceBuffer(sb.size())
    pk.upload(phost)
    dst = ta.DeviceBuffer(n)
    dst.fill(0xAB)
    b = ta.Batch([(sb, src.addr(), out.addr(), 1, ta.DIR_PACK),
                  (sb, dst.addr(), pk.addr(), 1, ta.DIR_UNPACK)])
    ta.synchronize()
    c1 = ta.counters()
    assert c1["verified_choices"] >= c0["verified_choices"] + 2
    assert c1["trial_launches"] > c0["trial_launches"]
    assert c1["pack_kernel_launches"] == c0["pack_kernel_launches"]
    assert (out.download() == 0x5A).all(), "pack trial wrote the packed buffer"
    assert (dst.download() == 0xAB).all(), "unpack trial wrote the strided buffer"
    b.launch()
    ta.synchronize()
    assert (out.download() == orc.sb_pack(desc, host, 0, 1)).all()
    ref = np.full(n, 0xAB, np.uint8)
    orc.sb_unpack(desc, phost, ref, 0, 1)
    assert (dst.download() == ref).all()


def test_one_graph_of_a_pack_and_an_unpack_batch(mpi):
    """tempi_graph_create of a pack batch and the unpack batch that reads its
    output: each replay packs the grid's faces as they are at that moment
    and unpacks them into the second grid (eager uploads and plain launches
    between the replays)."""
    n, g = 40, 3
    spec_faces = [tc.grid_face_subarray(n, g, f) for f in "xyz"]
    blocks = []
    for spec in spec_faces:
        h = mpi.type_from_spec(spec)
        blocks.append(mpi.strided_block(h))
        mpi.free(h)
    nbytes = (n + 2 * g) ** 3 * 4
    grid, grid2 = ta.DeviceBuffer(nbytes), ta.DeviceBuffer(nbytes)
    packed = [ta.DeviceBuffer(size) for _, size in blocks]
    pb = ta.Batch([(sb, grid.addr(), p.addr(), 1, ta.DIR_PACK)
                   for (sb, _), p in zip(blocks, packed)])
    ub = ta.Batch([(sb, grid2.addr(), p.addr(), 1, ta.DIR_UNPACK)
                   for (sb, _), p in zip(blocks, packed)])
    graph = ta.Graph([pb, ub])
    for rnd in range(4):
        host = orc.splitmix_bytes(nbytes, 900 + rnd)
        grid.upload(host)
        grid2.fill(0xAB)
        c0 = ta.counters()["graph_launches"]
        graph.launch()
        ta.synchronize()
        assert ta.counters()["graph_launches"] == c0 + 1
        ref = np.full(nbytes, 0xAB, np.uint8)
        for (sb, size), p in zip(blocks, packed):
            desc = orc.sb_desc(*sb.as_tuple())
            exp = orc.sb_pack(desc, host, 0, 1)
            assert (p.download() == exp).all(), rnd
            orc.sb_unpack(desc, exp, ref, 0, 1)
        assert (grid2.download() == ref).all(), rnd
        # a plain launch of the pack batch between replays
        pb.launch()
        ta.synchronize()
    graph.destroy()
