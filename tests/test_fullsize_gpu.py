"""Byte checks of the batched and single-copy paths at the benchmark
configurations' full sizes (run on an MI355X: pytest -m gpu).

* the halo batch: 26 faces, edges and corners (radius 3, float64) of 64
  subdomains of 128^3 cells -- 1 664 requests, 158 183 424 packed bytes --
  in one k_batch launch and in graph replays on fresh inputs, then unpacked
  into the exterior ghost cells of 64 fresh subdomains prefilled with 0xAB;
* the 4 096 random 1D/2D/3D requests (packed sizes log-uniform in
  64 B - 16 MiB, 5.7 GB) in one launch;
* the alignment/shape sweep: 2D/3D byte vectors of 1 B - 1 KiB blocks up to
  256 MiB packed through the interposed MPI_Pack / MPI_Unpack, and 1/3/12-B
  rows at 2^20 rows at every base offset 0-15.

Every packed byte is compared with the C oracle (oracle/oracle.c, sb_pack /
ot_pack) on the host copy of the same seeded bytes; unpacks are compared
with the oracle's unpack into 0xAB, or, for strided spans over 1 GiB, by
packing the unpacked image again (every type-map byte) and counting the
bytes that are not 0xAB (no byte outside the type map was written).
"""
import random

import numpy as np
import pytest

import tempi_amd as ta
import typecatalog as tc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

NSUB, LCR, RADIUS, QUANT = 64, (128, 128, 128), 3, 8
SUB_BYTES = 134 ** 3 * QUANT


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    assert ta.device_count() > 0, "no GPU visible"


def halo_blocks(mpi, exterior):
    """(direction, strided block, packed size) of the 26 halos of one
    subdomain, committed through the interposer."""
    out = []
    for d in tc.halo_directions():
        h = mpi.type_from_spec(tc.halo_type(RADIUS, LCR, d, QUANT, exterior=exterior))
        sb, size = mpi.strided_block(h)
        mpi.free(h)
        out.append((d, sb, size))
    return out


def check_packed(got, hosts, meta):
    for s, d, sb, off, size in meta:
        exp = orc.sb_pack(orc.sb_desc(*sb.as_tuple()), hosts[s], 0, 1)
        assert (got[off:off + size] == exp).all(), ("subdomain", s, "direction", d)


def test_halo_batch_64_subdomains_of_128_cubed(mpi):
    halos = halo_blocks(mpi, exterior=False)
    per_sub = sum(size for _, _, size in halos)
    total = NSUB * per_sub
    assert total == 158183424
    hosts = [orc.splitmix_bytes(SUB_BYTES, 4000 + s) for s in range(NSUB)]
    subs = [ta.DeviceBuffer(SUB_BYTES) for _ in range(NSUB)]
    for d, h in zip(subs, hosts):
        d.upload(h)
    packed = ta.DeviceBuffer(total)
    packed.fill(0x5A)
    reqs, meta, off = [], [], 0
    for s in range(NSUB):
        for d, sb, size in halos:
            reqs.append((sb, subs[s].addr(), packed.addr(off), 1, ta.DIR_PACK))
            meta.append((s, d, sb, off, size))
            off += size
    b = ta.Batch(reqs)
    info = b.info()
    assert info["requests"] == 1664 and info["packed_bytes"] == total
    before = ta.counters()
    b.launch()
    ta.synchronize()
    after = ta.counters()
    assert after["batch_launches"] == before["batch_launches"] + 1
    assert after["pack_kernel_launches"] == before["pack_kernel_launches"]
    print("halo batch: plain launch", flush=True)
    check_packed(packed.download(), hosts, meta)
    # graph replays on fresh inputs, with eager uploads and a plain launch
    # between them
    for rnd, changed in enumerate(((0, 63), (1, 17, 40))):
        for s in changed:
            hosts[s] = orc.splitmix_bytes(SUB_BYTES, 5000 + 100 * rnd + s)
            subs[s].upload(hosts[s])
        packed.fill(0x5A)
        b.graph_launch()
        ta.synchronize()
        print(f"halo batch: graph replay {rnd}", flush=True)
        check_packed(packed.download(), hosts, meta)
        b.launch()
    ta.synchronize()
    final = packed.download()
    check_packed(final, hosts, meta)
    b.destroy()
    # unpack into the exterior ghost cells of 64 fresh subdomains
    ext_halos = halo_blocks(mpi, exterior=True)
    outs = [ta.DeviceBuffer(SUB_BYTES) for _ in range(NSUB)]
    ureqs, umeta, off = [], [], 0
    for s in range(NSUB):
        for (d, sb, size), (d0, _, size0) in zip(ext_halos, halos):
            assert d == d0 and size == size0
            ureqs.append((sb, outs[s].addr(), packed.addr(off), 1, ta.DIR_UNPACK))
            umeta.append((s, sb, off, size))
            off += size
    ub = ta.Batch(ureqs)
    for replay in (False, True):
        for o in outs:
            o.fill(0xAB)
        (ub.graph_launch if replay else ub.launch)()
        ta.synchronize()
        print(f"halo unpack batch: {'graph replay' if replay else 'plain launch'}", flush=True)
        for s in range(NSUB):
            ref = np.full(SUB_BYTES, 0xAB, np.uint8)
            for s2, sb, o_, size in umeta[26 * s:26 * (s + 1)]:
                assert s2 == s
                orc.sb_unpack(orc.sb_desc(*sb.as_tuple()), final[o_:o_ + size], ref, 0, 1)
            assert (outs[s].download() == ref).all(), ("subdomain", s, "replay", replay)
    ub.destroy()


def test_random_requests_4096():
    geo, span_max = tc.random_requests(seed=4, nreq=4096)
    sbs = [ta.StridedBlock.make(*g) for g in geo]
    sizes = [sb.size() for sb in sbs]
    total = sum(sizes)
    host = orc.splitmix_bytes(span_max, 46)
    src = ta.DeviceBuffer(span_max)
    src.upload(host)
    packed = ta.DeviceBuffer(total)
    packed.fill(0x5A)
    reqs, offs, off = [], [], 0
    for sb, size in zip(sbs, sizes):
        reqs.append((sb, src.addr(), packed.addr(off), 1, ta.DIR_PACK))
        offs.append(off)
        off += size
    b = ta.Batch(reqs)
    assert b.info()["requests"] == 4096
    b.launch()
    ta.synchronize()
    print(f"random batch: {total} B packed", flush=True)
    got = packed.download()
    for i, (sb, o, size) in enumerate(zip(sbs, offs, sizes)):
        exp = orc.sb_pack(orc.sb_desc(*sb.as_tuple()), host, 0, 1)
        assert (got[o:o + size] == exp).all(), (i, sb)
    b.destroy()


CHUNK = 256 << 20


def count_not(buf, nbytes, value):
    """Bytes of a device buffer that differ from `value`, read in chunks."""
    n = 0
    for o in range(0, nbytes, CHUNK):
        part = buf.download(min(CHUNK, nbytes - o), o)
        n += int(np.count_nonzero(part != value))
    return n


def sweep_roundtrip(mpi, spec, base, src, host, span_cap=1 << 30):
    """MPI_Pack of `spec` from the shared source at `base`, checked against
    the oracle; MPI_Unpack of the packed bytes into a 0xAB-filled buffer,
    checked against the oracle's unpack (spans up to `span_cap`) or by
    packing the image again and counting the bytes that are not 0xAB."""
    o = orc.OracleType(spec)
    lo, hi = o.span(1)
    assert lo >= 0 and hi + base <= host.size
    h = mpi.type_from_spec(spec)
    sb, size = mpi.strided_block(h)
    expected = orc.sb_pack(orc.sb_desc(*sb.as_tuple()), host, base, 1)
    assert expected.size == size
    packed = ta.DeviceBuffer(size + 64)
    packed.fill(0x5A)
    err, pos = mpi.pack(src.addr(base), 1, h, packed.addr(), size)
    assert err == 0 and pos == size
    got = packed.download(size + 64)
    assert (got[:size] == expected).all(), (spec, base)
    assert (got[size:] == 0x5A).all()
    span = hi + base
    dst = ta.DeviceBuffer(span + 64)
    dst.fill(0xAB)
    err, pos = mpi.unpack(packed.addr(), size, dst.addr(base), 1, h)
    assert err == 0 and pos == size
    if span <= span_cap:
        ref = np.full(span + 64, 0xAB, np.uint8)
        orc.sb_unpack(orc.sb_desc(*sb.as_tuple()), expected, ref, base, 1)
        assert (dst.download() == ref).all(), (spec, base)
    else:
        again = ta.DeviceBuffer(size)
        err, pos = mpi.pack(dst.addr(base), 1, h, again.addr(), size)
        assert err == 0 and (again.download() == expected).all(), (spec, base)
        again.free()
        assert count_not(dst, span + 64, 0xAB) == int(np.count_nonzero(expected != 0xAB)), \
            (spec, base)
    dst.free()
    packed.free()
    mpi.free(h)
    return size


def test_sweep_up_to_256mib(mpi):
    """48 cases of the sweep generator (seeds 1000 + k, the benchmark's
    cases) whose strided span fits 2 GiB, through the interposed calls with
    the library's default variant choice."""
    span_max = 2 << 30
    host = orc.splitmix_bytes(span_max + 64, 48)
    src = ta.DeviceBuffer(host.size)
    src.upload(host)
    done, k, biggest = 0, 0, 0
    while done < 48:
        spec, base, packed_bytes = tc.sweep_type(random.Random(1000 + k))
        k += 1
        o = orc.OracleType(spec)
        if o.span(1)[1] + base > span_max:
            continue
        print(f"sweep case {k - 1}: {spec} base {base}", flush=True)
        biggest = max(biggest, sweep_roundtrip(mpi, spec, base, src, host))
        done += 1
    assert biggest >= 128 << 20


def test_sweep_case_of_256mib(mpi):
    """A 256 MiB 3D case: 8 planes of 2^17 rows of 256 B."""
    spec = ("hvector", 8, 1, (1 << 17) * 300 + 4000, ("vector", 1 << 17, 256, 300, tc.BYTE))
    o = orc.OracleType(spec)
    span = o.span(1)[1] + 16
    host = orc.splitmix_bytes(span, 49)
    src = ta.DeviceBuffer(span)
    src.upload(host)
    assert sweep_roundtrip(mpi, spec, 7, src, host) == 256 << 20


@pytest.mark.parametrize("bl", [1, 3, 12])
def test_short_rows_at_two_to_the_20_every_base_offset(mpi, bl):
    """Rows of 1, 3 and 12 B, 2^20 of them, stride = bl + U{0..4096} capped
    so that the span stays under 2 GiB, at every base offset 0-15 (pack at
    each; unpack at four of them)."""
    rng = random.Random(50 + bl)
    stride = bl + rng.randint(0, 2000)
    spec = ("vector", 1 << 20, bl, stride, tc.BYTE)
    o = orc.OracleType(spec)
    span = o.span(1)[1] + 16
    host = orc.splitmix_bytes(span, 51 + bl)
    src = ta.DeviceBuffer(span)
    src.upload(host)
    h = mpi.type_from_spec(spec)
    sb, size = mpi.strided_block(h)
    packed = ta.DeviceBuffer(size)
    for base in range(16):
        expected = orc.sb_pack(orc.sb_desc(*sb.as_tuple()), host, base, 1)
        packed.fill(0x5A)
        err, pos = mpi.pack(src.addr(base), 1, h, packed.addr(), size)
        assert err == 0 and pos == size
        assert (packed.download() == expected).all(), (bl, stride, base)
    mpi.free(h)
    for base in (0, 5, 11, 15):
        sweep_roundtrip(mpi, spec, base, src, host)
