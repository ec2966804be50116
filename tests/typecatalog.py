"""Type specs used by the tests, the golden-vector generator and the bench.

The factories restate the reference's support/type.cpp layouts as nested
tuple specs (see tempi_amd/mpi.py for the spec grammar); the random
generators produce seeded nested types of the kinds the reference lowers.
"""
from __future__ import annotations

import random

BYTE = ("named", "MPI_BYTE")
FLOAT = ("named", "MPI_FLOAT")
DOUBLE = ("named", "MPI_DOUBLE")
INT = ("named", "MPI_INT")

# ---------------------------------------------------------------------------
# support/type.cpp factories (copyExt / allocExt are (x, y, z) in bytes)
# ---------------------------------------------------------------------------


def make_byte_vn_hv_hv(copy, alloc):  # support/type.cpp:3-32
    row = ("vector", copy[0], 1, 1, BYTE)
    plane = ("hvector", copy[1], 1, alloc[0], row)
    return ("hvector", copy[2], 1, alloc[0] * alloc[1], plane)


def make_byte_v1_hv_hv(copy, alloc):  # support/type.cpp:34-65
    row = ("vector", 1, copy[0], alloc[0], BYTE)
    plane = ("hvector", copy[1], 1, alloc[0], row)
    return ("hvector", copy[2], 1, alloc[0] * alloc[1], plane)


def make_byte_v_hv(copy, alloc):  # support/type.cpp:67-88
    plane = ("vector", copy[1], copy[0], alloc[0], BYTE)
    return ("hvector", copy[2], 1, alloc[0] * alloc[1], plane)


def make_float_v_hv(copy, alloc):  # support/type.cpp:90-111
    plane = ("vector", copy[1], copy[0] // 4, alloc[0] // 4, FLOAT)
    return ("hvector", copy[2], 1, alloc[0] * alloc[1], plane)


def make_hi(copy, alloc):  # support/type.cpp:113-136
    disps = [z * alloc[1] * alloc[0] + y * alloc[0]
             for z in range(copy[2]) for y in range(copy[1])]
    return ("hindexed", [copy[0]] * len(disps), disps, BYTE)


def make_hib(copy, alloc):  # support/type.cpp:138-156
    disps = [z * alloc[1] * alloc[0] + y * alloc[0]
             for z in range(copy[2]) for y in range(copy[1])]
    return ("hindexed_block", copy[0], disps, BYTE)


def make_subarray(copy, alloc):  # support/type.cpp:158-170
    return ("subarray", [alloc[2], alloc[1], alloc[0]], [copy[2], copy[1], copy[0]],
            [0, 0, 0], "C", BYTE)


def make_subarray_v(copy, alloc):  # support/type.cpp:172-197
    plane = ("subarray", [alloc[1], alloc[0]], [copy[1], copy[0]], [0, 0], "C", BYTE)
    return ("vector", copy[2], 1, 1, plane)


def make_off_subarray(copy, alloc, off):  # support/type.cpp:199-211
    return ("subarray", [alloc[0], alloc[1], alloc[2]], [copy[0], copy[1], copy[2]],
            [off[0], off[1], off[2]], "C", BYTE)


def make_2d_byte_vector(nb, bl, st):  # support/type.cpp:216-222
    return ("vector", nb, bl, st, BYTE)


def make_2d_byte_hvector(nb, bl, st):  # support/type.cpp:224-230
    return ("hvector", nb, bl, st, BYTE)


def make_2d_byte_subarray(nb, bl, st):  # support/type.cpp:232-242
    return ("subarray", [nb, st], [nb, bl], [0, 0], "C", BYTE)


def make_2d_hv_by_rows(block, c1, s1, c2, s2):  # support/type.cpp:245-258
    b = ("contiguous", block, BYTE)
    row = ("hvector", c1, 1, s1, b)
    return ("hvector", c2, 1, s2, row)


def make_2d_hv_by_cols(block, c1, s1, c2, s2):  # support/type.cpp:261-274
    b = ("contiguous", block, BYTE)
    col = ("hvector", c2, 1, s2, b)
    return ("hvector", c1, 1, s1, col)


def make_contiguous_byte_v1(n):  # support/type.cpp:279-286
    return ("vector", n, 1, 1, BYTE)


def make_contiguous_byte_vn(n):  # support/type.cpp:287-294
    return ("vector", 1, n, n, BYTE)


def make_contiguous_subarray(n):  # support/type.cpp:295-304
    return ("subarray", [n], [n], [0], "C", BYTE)


def make_contiguous_contiguous(n):  # support/type.cpp:305-310
    return ("contiguous", n, BYTE)


# ---------------------------------------------------------------------------
# halo subarrays (bin/bench_halo_exchange.cpp:87-169): the region that
# direction `dirn` sends, on a domain of `lcr` interior cells with a ghost
# ring of `radius`, elements of `quant` bytes, rows `pitch` bytes long.
# ---------------------------------------------------------------------------


def halo_type(radius, lcr, dirn, quant, pitch=None, ysize=None, exterior=False):
    px = []
    ext = []
    for d, n in zip(dirn, lcr):
        if d == -1:
            px.append(0 if exterior else radius)
        elif d == 1:
            px.append(n + (radius if exterior else 0))
        else:
            px.append(radius)
        ext.append(n if d == 0 else radius)
    pitch = pitch if pitch is not None else (lcr[0] + 2 * radius) * quant
    ysize = ysize if ysize is not None else lcr[1] + 2 * radius
    zsize = lcr[2] + 2 * radius
    sizes = [zsize, ysize, pitch]
    subsizes = [ext[2], ext[1], ext[0] * quant]
    starts = [px[2], px[1], px[0] * quant]
    return ("subarray", sizes, subsizes, starts, "C", BYTE)


def halo_directions():
    return [(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)
            if (x, y, z) != (0, 0, 0)]


def grid_face_subarray(n, ghost, face, elem=FLOAT):
    """Subarray of a (n+2*ghost)^3 grid: `face` in {'x','y','z'} selects the
    ghost-thick slab normal to that axis, (z, y, x) C order."""
    full = n + 2 * ghost
    sub = {"x": [n, n, ghost], "y": [n, ghost, n], "z": [ghost, n, n]}[face]
    return ("subarray", [full] * 3, sub, [ghost] * 3, "C", elem)


# ---------------------------------------------------------------------------
# seeded random nested types
# ---------------------------------------------------------------------------

# Vectors and hvectors with negative strides (blocks in reverse address
# order; MPI allows them): the type map starts at the origin and reaches
# below it.  With the strided block each lowers to, as (start, extent,
# counts, strides).
NEGATIVE_STRIDES = [
    (("vector", 5, 2, -3, FLOAT), (0, 56, (8, 5), (1, -12))),
    (("hvector", 4, 3, -40, ("vector", 2, 1, 3, INT)), (0, 168, (4, 2, 3, 4), (1, 12, 16, -40))),
    (("vector", 3, 1, -2, ("hvector", 4, 2, -12, BYTE)), (0, 190, (2, 4, 3), (1, -12, -76))),
    (("hvector", 6, 1, -1000, ("contiguous", 100, BYTE)), (0, 5100, (100, 6), (1, -1000))),
    (("vector", 64, 3, -5, DOUBLE), (0, 2544, (24, 64), (1, -40))),
    (("hvector", 3, 2, -7, ("subarray", [4, 6], [2, 3], [1, 2], "C", FLOAT)),
     (32, 206, (12, 2, 2, 3), (1, 24, 96, -7))),
]

_NAMED = [BYTE, BYTE, BYTE, FLOAT, DOUBLE, INT, ("named", "MPI_SHORT")]


def random_type(rng: random.Random, depth: int = 0, max_bytes: int = 1 << 16):
    """A random nested contiguous/vector/hvector/subarray/resized/indexed type
    whose packed size stays below max_bytes."""
    if depth >= 3 or rng.random() < 0.25:
        base = rng.choice(_NAMED)
    else:
        base = random_type(rng, depth + 1, max_bytes // 4)
    kind = rng.choice(["contiguous", "vector", "vector", "hvector", "hvector",
                       "subarray", "resized", "indexed_block", "hindexed"])
    if kind == "contiguous":
        return ("contiguous", rng.randint(1, 6), base)
    if kind == "vector":
        bl = rng.randint(1, 4)
        return ("vector", rng.randint(1, 6), bl, bl + rng.randint(-1, 5), base)
    if kind == "hvector":
        bl = rng.randint(1, 3)
        return ("hvector", rng.randint(1, 6), bl, rng.randint(0, 200), base)
    if kind == "subarray":
        nd = rng.randint(1, 3)
        sizes = [rng.randint(1, 6) for _ in range(nd)]
        subsizes = [rng.randint(1, s) for s in sizes]
        starts = [rng.randint(0, s - ss) for s, ss in zip(sizes, subsizes)]
        return ("subarray", sizes, subsizes, starts, rng.choice("CF"), base)
    if kind == "resized":
        return ("resized", 0, rng.randint(1, 64) * 4, base)
    if kind == "indexed_block":
        n = rng.randint(1, 4)
        step = rng.randint(1, 5)
        regular = rng.random() < 0.6
        disps = [i * step if regular else rng.randint(0, 20) for i in range(n)]
        if not regular:
            disps = sorted(set(disps))
        return ("indexed_block", rng.randint(1, 3), disps, base)
    # hindexed, regular or not
    n = rng.randint(1, 4)
    step = rng.randint(1, 64)
    regular = rng.random() < 0.5
    bl = rng.randint(1, 3)
    disps = [i * step for i in range(n)] if regular else sorted(
        set(rng.randint(0, 256) for _ in range(n)))
    blens = [bl] * len(disps) if (regular or rng.random() < 0.5) else [
        rng.randint(1, 3) for _ in disps]
    return ("hindexed", blens, disps, base)


def sweep_type(rng: random.Random, max_packed: int = 256 << 20, min_packed: int = 1024):
    """One case of the alignment/shape sweep: a 2D vector of bytes or a 3D
    hvector of that vector, block length from the sweep set, stride =
    blocklength + U{0..4096}, base offset U{0..15}.  Returns
    (spec, base_offset, packed_bytes)."""
    bls = [1, 3, 4, 8, 12, 16, 32, 64, 128, 256, 512, 1024]
    while True:
        bl = rng.choice(bls)
        rows = 1 << rng.randint(4, 20)
        stride = bl + rng.randint(0, 4096)
        if rng.random() < 0.5:
            spec = ("vector", rows, bl, stride, BYTE)
            packed = rows * bl
        else:
            planes = rng.randint(2, 8)
            inner = ("vector", rows, bl, stride, BYTE)
            s2 = rows * stride + rng.randint(0, 4096)
            spec = ("hvector", planes, 1, s2, inner)
            packed = rows * bl * planes
        if min_packed <= packed <= max_packed:
            return spec, rng.randint(0, 15), packed


def random_requests(seed: int = 4, nreq: int = 4096):
    """The batched random-request workload: `nreq` 1D/2D/3D strided blocks
    with packed sizes log-uniform in [64 B, 16 MiB], starts 0-12 B, all
    reading one source buffer.  Returns ([(start, extent, counts, strides)],
    bytes the source buffer needs)."""
    import math

    rng = random.Random(seed)
    out = []
    span_max = 0
    for _ in range(nreq):
        target = int(math.exp(rng.uniform(math.log(64), math.log(16 << 20))))
        nd = rng.choice([1, 2, 3])
        if nd == 1:
            counts, strides = [target], [1]
        else:
            bl = rng.choice([4, 8, 12, 16, 32, 64, 128, 256, 512, 1024, 2048])
            bl = min(bl, target)
            rows = max(1, target // bl)
            if nd == 2:
                counts, strides = [bl, rows], [1, bl + rng.randint(0, 512)]
            else:
                r1 = max(1, int(math.sqrt(rows)))
                r2 = max(1, rows // r1)
                s1 = bl + rng.randint(0, 512)
                counts, strides = [bl, r1, r2], [1, s1, s1 * r1 + rng.randint(0, 4096)]
        span = 1 + sum((c - 1) * s for c, s in zip(counts, strides)) + counts[0]
        out.append((rng.randint(0, 3) * 4, span + 16, counts, strides))
        span_max = max(span_max, span + 64)
    return out, span_max


# Every tiny 2D byte copy: rows of 1-16 B, 1-4 rows at strides from the row
# length up to 16 B, the strided side at each of the 16 byte offsets below a
# 16-B boundary, one object; and each shape again as two objects (extent =
# rows x stride + 4) at offset 5.  The largest strided span is
# 15 + 3 * 16 + 16 = 79 bytes from the buffer start (two objects: 5 + 68 + 64
# = 137).
TINY_SPAN = 160


def tiny_copies():
    """[(start, extent, counts, strides, offset, incount)]."""
    out = []
    for bl in range(1, 17):
        for rows in range(1, 5):
            for stride in range(bl, 17):
                extent = rows * stride
                for off in range(16):
                    out.append((0, extent, [bl, rows], [1, stride], off, 1))
                out.append((0, extent + 4, [bl, rows], [1, stride], 5, 2))
    return out


# Every tiny 3D byte copy of a coarser grid: rows of 1, 3, 4, 8, 12 or 16 B;
# 1-3 rows at strides of L, L + 1 and L + 5 bytes; 1-3 planes at strides of
# rows x stride, that plus 3, and minus (that plus 2) -- a negative plane
# stride starts the type map (planes - 1) x |stride| bytes in; the strided
# side at each offset 0-15.  Reach: at most 15 + 132 + 42 + 16 = 205 bytes.
TINY3_SPAN = 256


def tiny_copies_3d():
    """[(start, extent, counts, strides, offset, incount)]."""
    out = []
    for bl in (1, 3, 4, 8, 12, 16):
        for rows in range(1, 4):
            for s1 in (bl, bl + 1, bl + 5):
                for planes in range(1, 4):
                    for s2 in (rows * s1, rows * s1 + 3, -(rows * s1 + 2)):
                        start = (planes - 1) * -s2 if s2 < 0 else 0
                        extent = start + (planes - 1) * max(s2, 0) + rows * s1
                        for off in range(16):
                            out.append((start, extent, [bl, rows, planes], [1, s1, s2],
                                        off, 1))
    return out
