"""The product canonicalizer (MPI_Type_commit -> tempi_strided_block) on the
CPU: its strided block must gather exactly the bytes MPI packs.

The strided block each type lowers to is checked by gathering with the
oracle's sb_pack and comparing with the oracle's typemap walk (itself pinned
to MPICH in test_oracle.py) and with MPICH's own MPI_Pack.
"""
import random

import numpy as np
import pytest

import typecatalog as tc
from conftest import spec_from_json
from oracle import oracle as orc


def check_type(mpi, spec, incounts=(1, 2, 3), expect_canonical=True):
    h = mpi.type_from_spec(spec)
    got = mpi.strided_block(h)
    if not expect_canonical:
        assert got is None, spec
        mpi.free(h)
        return None
    assert got is not None, (spec, "no strided form")
    o = orc.OracleType(spec)
    sb, size = got
    assert size == mpi.size(h)
    assert sb.extent == mpi.extent(h)[1]
    d = orc.sb_desc(*sb.as_tuple())
    for inc in incounts:
        lo, hi = o.span(inc)
        origin = -min(lo, 0)
        buf = orc.splitmix_bytes(hi + origin, 17)
        exp = o.pack(buf, origin, inc)
        assert (orc.sb_pack(d, buf, origin, inc) == exp).all(), (spec, sb)
        lib = np.zeros(max(exp.size, 1), np.uint8)
        err, pos = mpi.pack(buf.ctypes.data + origin, inc, h, lib.ctypes.data, lib.size,
                            library=True)
        assert err == 0 and (lib[:pos] == exp).all()
    mpi.free(h)
    return sb


def test_headline_forms(mpi):
    """Strided blocks of the benchmark types."""
    sb = check_type(mpi, ("vector", 1024, 256, 512, tc.BYTE))
    assert sb.as_tuple() == (0, 524032, (256, 1024), (1, 512))
    f = tc.grid_face_subarray(512, 3, "x")
    h = mpi.type_from_spec(f)
    sb, size = mpi.strided_block(h)
    mpi.free(h)
    assert sb.as_tuple() == (3226116, 555967328, (12, 512, 512), (1, 2072, 1073296))
    assert size == 3145728
    h = mpi.type_from_spec(tc.grid_face_subarray(512, 3, "y"))
    assert mpi.strided_block(h)[0].as_tuple() == (3226116, 555967328, (2048, 3, 512),
                                                  (1, 2072, 1073296))
    mpi.free(h)
    h = mpi.type_from_spec(tc.grid_face_subarray(512, 3, "z"))
    assert mpi.strided_block(h)[0].as_tuple() == (3226116, 555967328, (2048, 512, 3),
                                                  (1, 2072, 1073296))
    mpi.free(h)


def test_reference_parity_bugs_not_reproduced(mpi):
    # A1: hv_by_cols keeps type-map order (no stride swap)
    sb = check_type(mpi, tc.make_2d_hv_by_cols(13, 3, 16, 5, 53))
    assert list(sb.strides[: sb.ndims]) == [1, 53, 16]
    # A2: count-1 subarray dimension keeps its start offset
    sb = check_type(mpi, ("subarray", [4, 8], [1, 5], [2, 1], "C", tc.BYTE))
    assert sb.start == 17
    # A3: padded 1-D subarray at incount 2 (extent != size)
    check_type(mpi, ("subarray", [1, 4], [1, 3], [0, 0], "C", tc.BYTE), incounts=(2, 5))
    # A5: hvector strides beyond 2^31 stay 64-bit
    sb = check_type(mpi, ("hvector", 2, 1, 3 << 31, ("contiguous", 8, tc.BYTE)), incounts=())
    assert sb.strides[1] == 3 << 31


def test_reference_factories(mpi):
    for cp, al in (((2, 3, 4), (10, 10, 10)), ((12, 5, 3), (16, 8, 4)), ((16, 16, 16), (16, 16, 16)),
                   ((1, 7, 1), (9, 9, 9)), ((100, 100, 100), (200, 200, 200))):
        for f in (tc.make_byte_vn_hv_hv, tc.make_byte_v1_hv_hv, tc.make_byte_v_hv, tc.make_hi,
                  tc.make_hib, tc.make_subarray, tc.make_subarray_v):
            check_type(mpi, f(cp, al), incounts=(1, 2) if cp[0] < 100 else (1,))
    check_type(mpi, tc.make_off_subarray((2, 3, 4), (16, 16, 16), (1, 1, 4)))
    check_type(mpi, tc.make_float_v_hv((8, 3, 2), (16, 4, 4)))
    for f in (tc.make_2d_byte_vector, tc.make_2d_byte_hvector, tc.make_2d_byte_subarray):
        for nb, bl, st in ((1024, 1, 512), (4, 8, 16), (2, 256, 256), (7, 5, 9)):
            check_type(mpi, f(nb, bl, st))
    for f in (tc.make_contiguous_byte_v1, tc.make_contiguous_byte_vn, tc.make_contiguous_subarray,
              tc.make_contiguous_contiguous):
        sb = check_type(mpi, f(64))
        assert sb.as_tuple() == (0, 64, (64,), (1,))
    check_type(mpi, tc.make_2d_hv_by_rows(13, 3, 16, 5, 53))


def test_halo_types(mpi):
    for d in tc.halo_directions():
        for ext in (False, True):
            check_type(mpi, tc.halo_type(3, (10, 9, 8), d, 8, exterior=ext))


def test_golden_types_lower_exactly(mpi, golden):
    lowered = 0
    for case in golden:
        spec = spec_from_json(case["spec"])
        h = mpi.type_from_spec(spec)
        got = mpi.strided_block(h)
        mpi.free(h)
        if got is None:
            continue
        lowered += 1
        sb, _ = got
        o = orc.OracleType(spec)
        buf = orc.splitmix_bytes(case["buffer_bytes"], case["seed"])
        d = orc.sb_desc(*sb.as_tuple())
        assert (orc.sb_pack(d, buf, case["origin"], case["incount"])
                == o.pack(buf, case["origin"], case["incount"])).all(), case["name"]
    assert lowered >= 150


@pytest.mark.parametrize("seed", range(3))
def test_random_types(mpi, seed):
    rng = random.Random(4242 + seed)
    n_ok = 0
    for _ in range(150):
        spec = tc.random_type(rng)
        h = mpi.type_from_spec(spec)
        got = mpi.strided_block(h)
        mpi.free(h)
        if got is not None:
            check_type(mpi, spec, incounts=(1, 2))
            n_ok += 1
    assert n_ok > 40


def test_unsupported_types_fall_back(mpi):
    # irregular hindexed, multi-block struct, named pair type
    check_type(mpi, ("hindexed", [1, 2], [0, 7], tc.BYTE), expect_canonical=False)
    check_type(mpi, ("hindexed", [2, 2, 2], [0, 7, 9], tc.BYTE), expect_canonical=False)
    check_type(mpi, ("struct", [1, 1], [0, 8], [tc.INT, tc.DOUBLE]), expect_canonical=False)
    check_type(mpi, ("contiguous", 2, ("named", "MPI_DOUBLE_INT")), expect_canonical=False)


def test_supported_extras(mpi):
    # dup, resized, one-block struct, regular indexed
    check_type(mpi, ("dup", ("vector", 3, 2, 5, tc.FLOAT)))
    check_type(mpi, ("resized", 0, 100, ("vector", 3, 2, 5, tc.FLOAT)))
    check_type(mpi, ("struct", [3], [16], [tc.DOUBLE]))
    check_type(mpi, ("indexed_block", 2, [0, 5, 10, 15], tc.INT))
    check_type(mpi, ("hindexed", [4, 4, 4, 4], [0, 64, 1000, 1064], tc.BYTE))
    check_type(mpi, ("hvector", 4, 2, -24, ("contiguous", 3, tc.BYTE)))


def test_negative_strides(mpi):
    """Vectors and hvectors with negative strides lower to strided blocks
    with negative strides (no reordering: packed order is type-map order)
    that gather MPICH's bytes at incounts 1-3."""
    for spec, form in tc.NEGATIVE_STRIDES:
        sb = check_type(mpi, spec)
        assert sb.as_tuple() == form, spec


def test_equivalent_constructions_share_one_canonical_form(mpi):
    """The reference's type_equivalence test (test/type_equivalence.cpp:51-61)
    only compares sizes; here every construction of the same 3D layout must
    lower to the same (start, counts, strides) and pack the same bytes."""
    factories = (tc.make_byte_vn_hv_hv, tc.make_byte_v1_hv_hv, tc.make_byte_v_hv, tc.make_hi,
                 tc.make_hib, tc.make_subarray, tc.make_subarray_v)
    for cp, al in (((2, 3, 4), (10, 10, 10)), ((12, 5, 3), (16, 8, 4)), ((7, 1, 5), (9, 9, 9)),
                   ((16, 16, 16), (16, 16, 16)), ((1, 4, 4), (3, 5, 7))):
        forms = set()
        packs = set()
        for f in factories:
            spec = f(cp, al)
            h = mpi.type_from_spec(spec)
            got = mpi.strided_block(h)
            assert got is not None, (f.__name__, cp, al)
            start, _, counts, strides = got[0].as_tuple()
            forms.add((start, counts, strides))
            o = orc.OracleType(spec)
            buf = orc.splitmix_bytes(al[0] * al[1] * al[2] + 64, 3)
            packs.add(o.pack(buf, 0, 1).tobytes())
            mpi.free(h)
        assert len(forms) == 1, (cp, al, forms)
        assert len(packs) == 1


def test_fortran_order_subarray(mpi):
    # Fortran order: the first dimension is the fastest
    sb = check_type(mpi, ("subarray", [6, 5, 4], [2, 3, 2], [1, 2, 1], "F", tc.INT))
    assert list(sb.counts[: sb.ndims]) == [8, 3, 2]
    assert list(sb.strides[: sb.ndims]) == [1, 24, 120]
    assert sb.start == (1 + 2 * 6 + 1 * 30) * 4
