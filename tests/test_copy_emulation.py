"""The kernels' memory accesses, emulated on the host (no GPU).

tempi_copy_emulate runs the launch plan, grid, 32-bit grid-stride index
sequence and per-unit address arithmetic of a copy with the kernels' own
address code, recording every byte range loaded and stored.  These tests pin,
for every variant, that a copy stores exactly the bytes it must (each once,
nothing else) and loads only inside the type map's span -- at misaligned
strided and packed addresses, including the 20-48 B rows the lane-per-row
variants handle with head and tail stores -- and that copies whose unit count
would not fit the kernels' 32-bit unit index are split into launches of fewer
than 2^31 units, so the index never wraps."""
import itertools
import random

import pytest

import tempi_amd as ta
import typecatalog as tc

SB = ta.StridedBlock.make
BASE_S = 1 << 40  # addresses only: nothing is dereferenced
BASE_P = 1 << 41


def typemap_span(sb, count):
    """[lo, hi) of the type map's bytes from the strided origin."""
    lo = hi = sb.start
    hi += sb.counts[0]
    for d in range(1, sb.ndims):
        reach = (sb.counts[d] - 1) * sb.strides[d]
        if reach < 0:
            lo += reach
        else:
            hi += reach
    reach = (count - 1) * sb.extent
    if reach < 0:
        lo += reach
    else:
        hi += reach
    return lo, hi


def check_copy(sb, count, s_off, p_off, direction, variant, max_blocks=0):
    e = ta.copy_emulate(sb, count, BASE_S + s_off, BASE_P + p_off, direction, variant,
                        max_blocks=max_blocks)
    packed = sb.size() * count
    lo, hi = typemap_span(sb, count)
    where = (sb, count, s_off, p_off, direction, variant, max_blocks, e)
    assert e["emulated"] == 1 and e["wrapped"] == 0, where
    assert e["exact"] == 1, where
    assert e["bytes_read"] == packed and e["bytes_written"] == packed, where
    strided_rng = (e["read_lo"], e["read_hi"]) if direction == ta.DIR_PACK else \
        (e["write_lo"], e["write_hi"])
    packed_rng = (e["write_lo"], e["write_hi"]) if direction == ta.DIR_PACK else \
        (e["read_lo"], e["read_hi"])
    assert lo <= strided_rng[0] and strided_rng[1] <= hi, where
    assert packed_rng == (0, packed), where
    return e


def test_every_variant_misaligned_short_rows():
    """Rows of 1-70 B (the 20-48 B rows of the lane-per-row variants among
    them), both sides at every offset modulo 16, 1-3 dimensions, incount
    1-2, pack and unpack: stores exactly the type map / packed stream, loads
    inside the span."""
    nv = ta.num_variants()
    rng = random.Random(19)
    rows = [1, 3, 4, 7, 12, 13, 20, 24, 28, 33, 36, 44, 48, 63, 64, 70]
    n = 0
    for L in rows:
        for variant in range(nv):
            offs = rng.sample(range(16), 4) + [0]
            for s_off in offs:
                p_off = rng.randrange(16)
                stride = L + rng.randrange(0, 40)
                cnt = rng.randrange(1, 9)
                if rng.random() < 0.5:
                    sb = SB(rng.randrange(0, 8), (cnt - 1) * stride + L + rng.randrange(0, 24),
                            [L, cnt], [1, stride])
                else:
                    c2 = rng.randrange(1, 4)
                    s2 = cnt * stride + rng.randrange(0, 30)
                    sb = SB(rng.randrange(0, 8), (c2 - 1) * s2 + (cnt - 1) * stride + L + 5,
                            [L, cnt, c2], [1, stride, s2])
                count = rng.choice([1, 2])
                for direction in (ta.DIR_PACK, ta.DIR_UNPACK):
                    check_copy(sb, count, s_off, p_off, direction, variant)
                    n += 1
    assert n > 2000


def test_lane_per_row_variants_at_the_probe_shapes():
    """Variants 6, 7, 8, 9, 18 on rows of 20-48 B whose strided side sits
    1-15 B past a 16-B boundary, under every grid cap: each copy stays inside
    its type map and packed stream."""
    for variant, L, s_off, p_off, cap in itertools.product(
            (6, 7, 8, 9, 18), (20, 24, 28, 36, 40, 44, 48), (1, 4, 7, 13, 15), (0, 3),
            (0, 512, 1024, 2048, 8192)):
        stride = 2 * L + 64
        sb = SB(0, 257 * stride, [L, 257], [1, stride])
        for direction in (ta.DIR_PACK, ta.DIR_UNPACK):
            check_copy(sb, 1, s_off, p_off, direction, variant, max_blocks=cap)


def test_long_rows_and_negative_strides():
    rng = random.Random(5)
    for variant in range(ta.num_variants()):
        for L in (128, 1000, 4096, 5000):
            s_off = rng.randrange(16)
            sb = SB(3 * 6000, 6000, [L, 3], [1, -6000])
            for direction in (ta.DIR_PACK, ta.DIR_UNPACK):
                check_copy(sb, 2, s_off, rng.randrange(16), direction, variant)


def test_negative_stride_catalog():
    """The strided blocks the negative-stride vectors and hvectors lower to
    (blocks below the origin, reverse address order): every variant, both
    directions, incount 1 and 3, misaligned origins."""
    for spec, (start, extent, counts, strides) in tc.NEGATIVE_STRIDES:
        sb = SB(start, extent, list(counts), list(strides))
        for variant in range(ta.num_variants()):
            for count in (1, 3):
                for direction in (ta.DIR_PACK, ta.DIR_UNPACK):
                    check_copy(sb, count, 5, 3, direction, variant)


def test_units_past_two_to_the_31_are_split():
    """vector(2^31 - 1, 1, 2, MPI_BYTE) at incount 2 is 2^32 - 2 one-byte
    rows: one launch of that many units would step the kernels' 32-bit unit
    index past 2^32.  The copy runs as launches of fewer than 2^31 units."""
    rows = (1 << 31) - 1
    sb = SB(0, 2 * (rows - 1) + 1, [1, rows], [1, 2])
    for variant in (11, 9, 0):
        e = ta.copy_emulate(sb, 2, BASE_S, BASE_P, ta.DIR_PACK, variant, unit_limit=0)
        assert e["emulated"] == 0
        assert e["launches"] == 2 and e["max_units"] == rows, e
    # one dimension of 2^32 + 5 rows: cut into runs of the dimension
    big = (1 << 32) + 5
    sb = SB(0, 3 * big, [1, big], [1, 3])
    e = ta.copy_emulate(sb, 1, BASE_S, BASE_P, ta.DIR_UNPACK, 11, unit_limit=0)
    assert e["launches"] == 3 and e["max_units"] < (1 << 31), e
    # a dimension whose single index holds 2^31 units or more is cut below it
    sb = SB(0, 1 << 40, [1, 1 << 31, 3], [1, 2, 1 << 33])
    e = ta.copy_emulate(sb, 1, BASE_S, BASE_P, ta.DIR_PACK, 11, unit_limit=0)
    assert e["launches"] == 6 and e["max_units"] < (1 << 31), e


def test_split_copies_stay_exact():
    """The split itself, emulated in full: a copy forced into several
    launches stores the same bytes as one launch would."""
    # 3 x 5 x 7 rows with the unit budget crossed at every level is not
    # reachable at test sizes, so check the pieces' arithmetic through a
    # copy of rows > 1 GiB (cut into 1 GiB pieces) on the plan only, and
    # a multi-dimension copy in full
    sb = SB(0, 3 << 31, [(1 << 30) + 7, 2], [1, 3 << 30])
    e = ta.copy_emulate(sb, 1, BASE_S, BASE_P, ta.DIR_PACK, 0, unit_limit=0)
    assert e["launches"] == 4
    sb = SB(11, 10000, [9, 5, 7], [1, 13, 100])
    for variant in range(ta.num_variants()):
        check_copy(sb, 3, 5, 9, ta.DIR_UNPACK, variant)


@pytest.mark.parametrize("kind", ["2d", "3d"])
def test_every_tiny_copy_every_variant(kind):
    """Every tiny copy of tests/typecatalog.py -- 2D: rows of 1-16 B, 1-4
    rows, strides up to 16 B, one and two objects; 3D: 1-3 rows by 1-3
    planes, negative plane strides among them -- with the strided side at
    each offset 0-15 and the packed side at every offset modulo 16, under
    every variant in both directions: stores exactly the type map and the
    packed stream, loads inside them -- the host-side check of the copies
    that tests/test_tiny_exhaustive_gpu.py runs on the GPU."""
    copies = tc.tiny_copies() if kind == "2d" else tc.tiny_copies_3d()
    n = 0
    for k, (start, ext, counts, strides, off, inc) in enumerate(copies):
        sb = SB(start, ext, counts, strides)
        for variant in range(ta.num_variants()):
            for direction in (ta.DIR_PACK, ta.DIR_UNPACK):
                check_copy(sb, inc, off, (k + variant) % 16, direction, variant)
                n += 1
    assert n == len(copies) * ta.num_variants() * 2
