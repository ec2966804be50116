"""Brute force on the GPU: every tiny copy of tests/typecatalog.py -- 2D:
rows of 1-16 B, 1-4 rows at strides up to 16 B, one and two objects; 3D:
1-3 rows by 1-3 planes, negative plane strides among them; the strided side
at each byte offset 0-15 -- packed and unpacked by every kernel variant,
through the single-copy kernel and through the batch kernel, against the
oracle's gather and scatter.  Packed slices sit at every alignment with
guard bytes between them; each unpack writes its own 0xAB-filled region, so
a byte written outside a type map shows.  The host emulation of the same
copies (tests/test_copy_emulation.py) shows every access in bounds."""
import numpy as np
import pytest

import tempi_amd as ta
import typecatalog as tc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

@pytest.fixture(scope="module", params=["2d", "3d"])
def tiny(request):
    if request.param == "2d":
        copies, span = tc.tiny_copies(), tc.TINY_SPAN
    else:
        copies, span = tc.tiny_copies_3d(), tc.TINY3_SPAN
    region = span  # destination bytes per unpack
    host = orc.splitmix_bytes(span, 31)
    sbs, exp, offs = [], [], []
    pos = 0
    for k, (start, ext, counts, strides, off, inc) in enumerate(copies):
        sb = ta.StridedBlock.make(start, ext, counts, strides)
        e = orc.sb_pack(orc.sb_desc(*sb.as_tuple()), host, off, inc)
        sbs.append(sb)
        exp.append(e)
        offs.append(pos)
        pos += e.size + k % 3  # 0-2 guard bytes: slices at every alignment
    packed = np.full(pos + 64, 0x5A, np.uint8)
    for e, o in zip(exp, offs):
        packed[o:o + e.size] = e
    image = np.full(len(copies) * region, 0xAB, np.uint8)
    for k, ((_, _, _, _, off, inc), sb, e) in enumerate(zip(copies, sbs, exp)):
        orc.sb_unpack(orc.sb_desc(*sb.as_tuple()), e, image[k * region:(k + 1) * region],
                      off, inc)
    src = ta.DeviceBuffer(span)
    src.upload(host)
    psrc = ta.DeviceBuffer(packed.size)
    psrc.upload(packed)
    print(f"{len(copies)} tiny {request.param} copies, {pos} packed bytes", flush=True)
    return dict(copies=copies, sbs=sbs, exp=exp, offs=offs, packed=packed, image=image,
                src=src, psrc=psrc, region=region)


def check(t, pk, dst, what):
    got = pk.download()
    bad = np.nonzero(got != t["packed"])[0]
    assert bad.size == 0, (what, "pack", int(bad[0]), len(bad))
    got = dst.download()
    bad = np.nonzero(got != t["image"])[0]
    assert bad.size == 0, (what, "unpack", t["copies"][int(bad[0]) // t["region"]], len(bad))


@pytest.mark.parametrize("variant", range(19))
def test_every_tiny_copy_single_copy_kernel(tiny, variant):
    t, rg = tiny, tiny["region"]
    pk = ta.DeviceBuffer(t["packed"].size)
    pk.fill(0x5A)
    dst = ta.DeviceBuffer(len(t["copies"]) * rg)
    dst.fill(0xAB)
    for k, ((_, _, _, _, off, inc), sb, e, o) in enumerate(
            zip(t["copies"], t["sbs"], t["exp"], t["offs"])):
        assert ta.pack_sb(sb, t["src"].addr(off), inc, pk.addr(o), e.size, 0, variant) == e.size
        assert ta.unpack_sb(sb, t["psrc"].addr(o), e.size, dst.addr(k * rg + off), inc, 0,
                            variant) == e.size
    ta.synchronize()
    check(t, pk, dst, variant)


@pytest.mark.parametrize("variant", range(19))
def test_every_tiny_copy_batch_kernel(tiny, variant):
    t, rg = tiny, tiny["region"]
    pk = ta.DeviceBuffer(t["packed"].size)
    pk.fill(0x5A)
    dst = ta.DeviceBuffer(len(t["copies"]) * rg)
    dst.fill(0xAB)
    packs, unpacks = [], []
    for k, ((_, _, _, _, off, inc), sb, o) in enumerate(zip(t["copies"], t["sbs"], t["offs"])):
        packs.append((sb, t["src"].addr(off), pk.addr(o), inc, ta.DIR_PACK, variant))
        unpacks.append((sb, dst.addr(k * rg + off), t["psrc"].addr(o), inc,
                        ta.DIR_UNPACK, variant))
    bp, bu = ta.Batch(packs), ta.Batch(unpacks)
    bp.launch()
    bu.launch()
    ta.synchronize()
    check(t, pk, dst, variant)
    # every 37th request again, in batches of at most 8 (the launch that
    # carries its tile index in the kernel arguments)
    pk.fill(0x5A)
    dst.fill(0xAB)
    sample = list(range(variant, len(packs), 37))
    small = []
    for i in range(0, len(sample), 8):
        grp = sample[i:i + 8]
        small += [ta.Batch([packs[j] for j in grp]), ta.Batch([unpacks[j] for j in grp])]
        small[-2].launch()
        small[-1].launch()
    ta.synchronize()
    want_p = np.full_like(t["packed"], 0x5A)
    want_u = np.full_like(t["image"], 0xAB)
    for j in sample:
        o, e = t["offs"][j], t["exp"][j]
        want_p[o:o + e.size] = e
        want_u[j * rg:(j + 1) * rg] = t["image"][j * rg:(j + 1) * rg]
    assert (pk.download() == want_p).all(), (variant, "small pack batches")
    assert (dst.download() == want_u).all(), (variant, "small unpack batches")
