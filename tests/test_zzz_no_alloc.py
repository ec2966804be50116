"""tempi_batch_run_noalloc on the GPU: the batch entry point whose tables
sit in the library's static device array.  Its output must equal
tempi_batch_create + tempi_batch_launch on the same requests (and the
oracle), on two calls in a row, the second with fewer requests; a request
set whose tables exceed the array is refused before any device work."""
import numpy as np
import pytest

import tempi_amd as ta
import typecatalog as tc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def halo_blocks(mpi, lcr, exterior=False):
    out = []
    for d in tc.halo_directions():
        h = mpi.type_from_spec(tc.halo_type(3, lcr, d, 8, exterior=exterior))
        out.append(mpi.strided_block(h))
        mpi.free(h)
    return out


def halo_case(mpi, nsub, lcr, seed):
    """Pack requests of the 26 halos of `nsub` subdomains (float64, radius 3)
    of seeded bytes, into one packed buffer: (requests, hosts, metadata,
    packed bytes, buffers)."""
    radius, quant = 3, 8
    sub_bytes = (lcr[0] + 6) * (lcr[1] + 6) * (lcr[2] + 6) * quant
    hosts = [orc.splitmix_bytes(sub_bytes, seed + s) for s in range(nsub)]
    subs = [ta.DeviceBuffer(sub_bytes) for _ in range(nsub)]
    for d, h in zip(subs, hosts):
        d.upload(h)
    blocks = halo_blocks(mpi, lcr)
    total = nsub * sum(size for _, size in blocks)
    reqs, meta, off = [], [], 0
    for s in range(nsub):
        for sb, size in blocks:
            meta.append((s, sb, off, size))
            reqs.append([sb, subs[s].addr(), off, 1, ta.DIR_PACK])
            off += size
    return reqs, hosts, meta, total, subs


def run_both(reqs, total, keep):
    """The packed bytes of the requests from tempi_batch_create + launch and
    from tempi_batch_run_noalloc (packed offsets relative to each output)."""
    outs = []
    for kind in ("batch", "noalloc"):
        packed = ta.DeviceBuffer(total + 64)
        packed.fill(0x5A)
        keep.append(packed)
        rq = [(r[0], r[1], packed.addr(r[2]), r[3], r[4]) for r in reqs]
        if kind == "batch":
            b = ta.Batch(rq)
            b.launch()
            ta.synchronize()
            b.destroy()
        else:
            ta.batch_run_noalloc(rq)
            ta.synchronize()
        outs.append(packed.download())
    return outs


def test_noalloc_matches_the_allocating_batch(mpi):
    keep = []
    # first call: 8 subdomains of 24 x 20 x 16 (208 requests, k_batch)
    reqs, hosts, meta, total, subs = halo_case(mpi, 8, (24, 20, 16), 600)
    ref, got = run_both(reqs, total, keep)
    assert (got == ref).all()
    assert (got[total:] == 0x5A).all()
    for s, sb, off, size in meta:
        exp = orc.sb_pack(orc.sb_desc(*sb.as_tuple()), hosts[s], 0, 1)
        assert (got[off:off + size] == exp).all(), (s, sb)
    # second call, smaller: 3 faces of one subdomain (k_batch_small), on
    # fresh bytes -- nothing of the first call's tables may leak into it
    reqs2, hosts2, meta2, total2, subs2 = halo_case(mpi, 1, (16, 12, 10), 700)
    reqs2, meta2 = reqs2[:3], meta2[:3]
    total2 = sum(m[3] for m in meta2)
    ref2, got2 = run_both(reqs2, total2, keep)
    assert (got2 == ref2).all()
    for s, sb, off, size in meta2:
        exp = orc.sb_pack(orc.sb_desc(*sb.as_tuple()), hosts2[s], 0, 1)
        assert (got2[off:off + size] == exp).all(), (s, sb)
    assert (got2[total2:] == 0x5A).all()


def test_noalloc_unpack_and_back_to_back_calls(mpi):
    """Unpack requests, and two calls enqueued back to back on one stream
    without a synchronization between them (the second upload waits for
    the first launch in stream order)."""
    keep = []
    reqs, hosts, meta, total, subs = halo_case(mpi, 2, (16, 12, 10), 800)
    packed = ta.DeviceBuffer(total)
    keep.append(packed)
    ta.batch_run_noalloc([(r[0], r[1], packed.addr(r[2]), r[3], r[4]) for r in reqs])
    outs = [ta.DeviceBuffer(h.size) for h in hosts]
    for o in outs:
        o.fill(0xAB)
    # unpack what the first call packed into the exterior ghost cells of
    # fresh subdomains (disjoint regions)
    ext = halo_blocks(mpi, (16, 12, 10), exterior=True)
    ureqs = [(ext[i % 26][0], outs[m[0]].addr(), packed.addr(r[2]), r[3], ta.DIR_UNPACK)
             for i, (r, m) in enumerate(zip(reqs, meta))]
    ta.batch_run_noalloc(ureqs)
    ta.synchronize()
    got = packed.download()
    for s, h in enumerate(hosts):
        ref = np.full(h.size, 0xAB, np.uint8)
        for i, (s2, sb, off, size) in enumerate(meta):
            if s2 == s:
                esb = ext[i % 26][0]
                orc.sb_unpack(orc.sb_desc(*esb.as_tuple()), got[off:off + size], ref, 0, 1)
        assert (outs[s].download() == ref).all(), s


def test_noalloc_refuses_tables_beyond_the_array():
    src = ta.DeviceBuffer(1 << 16)
    out = ta.DeviceBuffer(1 << 16)
    sb = ta.StridedBlock.make(0, 64, [16, 2], [1, 32])
    reqs = [(sb, src.addr(), out.addr(), 1, ta.DIR_PACK)] * 20000
    with pytest.raises(ta.TempiError) as e:
        ta.batch_run_noalloc(reqs)
    assert e.value.code == ta.TEMPI_ERR_UNSUPPORTED
    assert "static scratch" in str(e.value)
