"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

* golden vectors recorded from MPICH's MPI_Pack / MPI_Unpack (no MPI needed);
* known answers probed from MPICH 3.3.2 by hand (vector(3,2,4,BYTE), a
  column-major hvector, a padded subarray, a count-1 subarray dimension);
* closed forms: packed size, vector / subarray extents, subarray start;
* the two oracle halves (typemap walker, strided-block gather) against each
  other, and against live MPICH on seeded random types when it is present;
* brute force over every tiny vector type;
* a planted mistake in the oracle's input is caught by the fixtures.
"""
import hashlib
import itertools
import random

import numpy as np
import pytest

import typecatalog as tc
from conftest import spec_from_json
from oracle import oracle as orc


def digest_ok(expected, actual: np.ndarray) -> bool:
    if "hex" in expected:
        return actual.tobytes().hex() == expected["hex"]
    return hashlib.sha256(actual.tobytes()).hexdigest() == expected["sha256"]


def run_oracle(case):
    spec = spec_from_json(case["spec"])
    o = orc.OracleType(spec)
    buf = orc.splitmix_bytes(case["buffer_bytes"], case["seed"])
    packed = o.pack(buf, case["origin"], case["incount"])
    dst = np.full(case["buffer_bytes"], 0xAB, dtype=np.uint8)
    o.unpack(packed, dst, case["origin"], case["incount"])
    return o, packed, dst


def test_golden_vectors(golden):
    assert len(golden) > 200
    for case in golden:
        o, packed, dst = run_oracle(case)
        assert o.size() == case["size"], case["name"]
        assert o.extent() == case["extent"], case["name"]
        assert o.lb() == case["lb"], case["name"]
        assert digest_ok(case["packed"], packed), case["name"]
        assert digest_ok(case["unpacked_into_ab"], dst), case["name"]


def test_planted_mistake_is_caught(golden):
    """An oracle with a stride off by one must fail the fixtures."""
    case = next(c for c in golden if c["name"] == "kat_hv_by_cols(13,3,16,5,53)")
    spec = spec_from_json(case["spec"])
    wrong = spec[:3] + (spec[3] + 1,) + spec[4:]
    o = orc.OracleType(wrong)
    buf = orc.splitmix_bytes(case["buffer_bytes"] + 64, case["seed"])
    assert not digest_ok(case["packed"], o.pack(buf, case["origin"], case["incount"]))


def ramp(n):
    return (np.arange(n) & 0xFF).astype(np.uint8)


def test_known_answers():
    # MPICH probes: vector(3,2,4,BYTE) -> 0 1 4 5 8 9
    o = orc.OracleType(("vector", 3, 2, 4, tc.BYTE))
    assert o.pack(ramp(16), 0, 1).tolist() == [0, 1, 4, 5, 8, 9]
    # hv_by_cols(13,3,16,5,53) packs blocks at 0, 53, 106, 159, 212, 16, 69, ...
    o = orc.OracleType(tc.make_2d_hv_by_cols(13, 3, 16, 5, 53))
    starts = [d for d, _ in orc.OracleType(tc.make_2d_hv_by_cols(13, 3, 16, 5, 53)).blocks()]
    assert starts[:7] == [0, 53, 106, 159, 212, 16, 69]
    # padded 1-D subarray at incount 2 -> 0 1 2 4 5 6
    o = orc.OracleType(("subarray", [1, 4], [1, 3], [0, 0], "C", tc.BYTE))
    assert o.pack(ramp(8), 0, 2).tolist() == [0, 1, 2, 4, 5, 6]
    # subarray with a subsize-1 dim at start {2,1}: bytes 17-21
    o = orc.OracleType(("subarray", [4, 8], [1, 5], [2, 1], "C", tc.BYTE))
    assert o.pack(ramp(32), 0, 1).tolist() == [17, 18, 19, 20, 21]


def test_closed_forms():
    rng = random.Random(7)
    for _ in range(200):
        count, bl, stride = rng.randint(1, 9), rng.randint(1, 5), rng.randint(1, 12)
        old = rng.choice([tc.BYTE, tc.FLOAT, tc.DOUBLE])
        e = orc.NAMED_SIZES[old[1]]
        o = orc.OracleType(("vector", count, bl, stride, old))
        assert o.size() == count * bl * e
        if stride >= bl:
            assert o.extent() == ((count - 1) * stride + bl) * e
        nd = rng.randint(1, 4)
        sizes = [rng.randint(1, 7) for _ in range(nd)]
        sub = [rng.randint(1, s) for s in sizes]
        st = [rng.randint(0, s - u) for s, u in zip(sizes, sub)]
        o = orc.OracleType(("subarray", sizes, sub, st, "C", old))
        assert o.extent() == int(np.prod(sizes)) * e
        assert o.size() == int(np.prod(sub)) * e
        first = sum(st[i] * e * int(np.prod(sizes[i + 1:])) for i in range(nd))
        assert o.blocks()[0][0] == first
        # incount objects pack incount times the size, object k = pack at k*extent
        buf = orc.splitmix_bytes(o.extent() * 3, 99)
        p3 = o.pack(buf, 0, 3)
        assert p3.size == 3 * o.size()
        for k in range(3):
            assert (p3[k * o.size():(k + 1) * o.size()] == o.pack(buf, k * o.extent(), 1)).all()


def test_walker_and_strided_gather_agree():
    """Hand-written strided blocks for the reference factories, gathered by
    sb_pack, equal the typemap walker."""
    cases = [
        # (spec, start, counts, strides)
        (tc.make_byte_v_hv((2, 3, 4), (10, 10, 10)), 0, [2, 3, 4], [1, 10, 100]),
        (tc.make_off_subarray((2, 3, 4), (16, 16, 16), (1, 1, 4)), 16 * 16 + 16 + 4,
         [4, 3, 2], [1, 16, 256]),
        (("vector", 1024, 256, 512, tc.BYTE), 0, [256, 1024], [1, 512]),
        (tc.grid_face_subarray(8, 3, "x"), (3 * 14 * 14 + 3 * 14 + 3) * 4, [12, 8, 8],
         [1, 14 * 4, 14 * 14 * 4]),
    ]
    for spec, start, counts, strides in cases:
        o = orc.OracleType(spec)
        d = orc.sb_desc(start, o.extent(), counts, strides)
        buf = orc.splitmix_bytes(o.extent() * 2 + 64, 5)
        for inc in (1, 2):
            assert (orc.sb_pack(d, buf, 0, inc) == o.pack(buf, 0, inc)).all()
            a = np.full(buf.size, 0xAB, np.uint8)
            b = a.copy()
            p = o.pack(buf, 0, inc)
            orc.sb_unpack(d, p, a, 0, inc)
            o.unpack(p, b, 0, inc)
            assert (a == b).all()


def test_unpack_leaves_gaps_and_restores_typemap():
    rng = random.Random(11)
    for _ in range(50):
        spec = tc.random_type(rng)
        o = orc.OracleType(spec)
        lo, hi = o.span(2)
        origin = -min(lo, 0)
        buf = orc.splitmix_bytes(hi + origin, 3)
        p = o.pack(buf, origin, 2)
        dst = np.full(buf.size, 0xAB, np.uint8)
        o.unpack(p, dst, origin, 2)
        mask = np.zeros(buf.size, bool)
        for d, ln in o.blocks(2):
            mask[origin + d: origin + d + ln] = True
        assert (dst[mask] == buf[mask]).all()
        assert (dst[~mask] == 0xAB).all()


@pytest.mark.parametrize("seed", range(4))
def test_oracle_matches_live_mpich_random(mpi, seed):
    rng = random.Random(seed * 7919)
    for _ in range(60):
        spec = tc.random_type(rng)
        inc = rng.randint(1, 3)
        o = orc.OracleType(spec)
        h = mpi.type_from_spec(spec)
        assert mpi.extent(h) == (o.lb(), o.extent())
        assert mpi.size(h) == o.size()
        lo, hi = o.span(inc)
        origin = -min(lo, 0)
        buf = orc.splitmix_bytes(hi + origin, seed)
        exp = np.zeros(max(inc * o.size(), 1), np.uint8)
        err, pos = mpi.pack(buf.ctypes.data + origin, inc, h, exp.ctypes.data, exp.size,
                            library=True)
        assert err == 0 and pos == inc * o.size()
        assert (o.pack(buf, origin, inc) == exp[:pos]).all(), spec
        mpi.free(h)


def test_brute_force_tiny_vectors(mpi):
    """Every vector(count<=3, bl<=3, stride in -4..6) of bytes and shorts at
    incount 1..2 against MPICH."""
    for count, bl, stride, old, inc in itertools.product(
            range(1, 4), range(1, 4), range(-4, 7), [tc.BYTE, ("named", "MPI_SHORT")], (1, 2)):
        spec = ("vector", count, bl, stride, old)
        o = orc.OracleType(spec)
        h = mpi.type_from_spec(spec)
        lo, hi = o.span(inc)
        origin = -min(lo, 0)
        buf = ramp(hi + origin + 1)
        exp = np.zeros(inc * o.size(), np.uint8)
        err, pos = mpi.pack(buf.ctypes.data + origin, inc, h, exp.ctypes.data, exp.size,
                            library=True)
        assert err == 0
        assert (o.pack(buf, origin, inc) == exp).all(), spec
        mpi.free(h)


def test_splitmix_stream():
    # SplitMix64 seeded with 0: first outputs of the reference generator
    # (Steele et al.; the same constants as java.util.SplittableRandom)
    b = orc.splitmix_bytes(16, 0)
    w = np.frombuffer(b.tobytes(), dtype="<u8")
    assert int(w[0]) == 0xE220A8397B1DCDAF
    assert int(w[1]) == 0x6E789E6AA1B965F4
    # prefix-stable and seed-sensitive
    assert (orc.splitmix_bytes(1000, 5)[:333] == orc.splitmix_bytes(333, 5)).all()
    assert not (orc.splitmix_bytes(64, 5) == orc.splitmix_bytes(64, 6)).all()
