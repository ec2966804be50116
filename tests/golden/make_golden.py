"""Generate tests/golden/pack_golden.json from the underlying MPI library.

Every case packs `incount` objects of a type spec from a buffer of seeded
bytes (SplitMix64 stream, oracle/oracle.c) with the MPI library's own
MPI_Pack, and unpacks the packed bytes with its MPI_Unpack into a buffer
pre-filled with 0xAB.  The fixture stores the spec, the buffer geometry and
the expected bytes (hex when short, SHA-256 otherwise), so the oracle and the
GPU path can be checked without MPI.  Run in a container with MPICH:

    python tests/golden/make_golden.py

The expected values come from MPICH alone: the buffer span is derived from
MPI_Type_get_true_extent, and libtempi_amd.so is used only as the route to the
library's type constructors (TEMPI_DISABLE=1 makes it a pass-through).
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import random
import sys

os.environ["TEMPI_DISABLE"] = "1"
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import typecatalog as tc  # noqa: E402
from oracle.oracle import splitmix_bytes  # noqa: E402  (input bytes only)
from tempi_amd import mpi as tmpi  # noqa: E402

HEX_LIMIT = 512
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pack_golden.json")


def true_span(m, h, incount):
    lb, ext = m.extent(h)
    tlb, text = ctypes.c_longlong(), ctypes.c_longlong()
    m.tempi.MPI_Type_get_true_extent_x.argtypes = [m.H, ctypes.POINTER(ctypes.c_longlong),
                                                   ctypes.POINTER(ctypes.c_longlong)]
    assert m.tempi.MPI_Type_get_true_extent_x(h, ctypes.byref(tlb), ctypes.byref(text)) == 0
    lo = tlb.value + min(0, (incount - 1) * ext)
    hi = tlb.value + text.value + max(0, (incount - 1) * ext)
    return lo, hi


def digest(a: np.ndarray):
    if a.size <= HEX_LIMIT:
        return {"hex": a.tobytes().hex()}
    return {"sha256": hashlib.sha256(a.tobytes()).hexdigest()}


def case(m, name, spec, incount, seed):
    h = m.type_from_spec(spec)
    lo, hi = true_span(m, h, incount)
    origin = -min(lo, 0)
    length = max(hi, 0) + origin
    buf = splitmix_bytes(length, seed)
    size = m.size(h)
    packed = np.zeros(max(incount * size, 1), dtype=np.uint8)
    err, pos = m.pack(buf.ctypes.data + origin, incount, h, packed.ctypes.data,
                      packed.size, 0, library=True)
    assert err == 0 and pos == incount * size, (name, err, pos)
    dst = np.full(length, 0xAB, dtype=np.uint8)
    err, pos2 = m.unpack(packed.ctypes.data, packed.size, dst.ctypes.data + origin, incount,
                         h, 0, library=True)
    assert err == 0 and pos2 == pos
    lb, ext = m.extent(h)
    m.free(h)
    return {
        "name": name,
        "spec": spec,
        "incount": incount,
        "seed": seed,
        "buffer_bytes": int(length),
        "origin": int(origin),
        "size": int(size),
        "lb": int(lb),
        "extent": int(ext),
        "packed": digest(packed[: incount * size]),
        "unpacked_into_ab": digest(dst),
    }


def catalog():
    """(name, spec, incount) triples."""
    cases = []
    # test/pack_unpack.cpp:160-302
    cases.append(("contiguous_contiguous(10)", tc.make_contiguous_contiguous(10), 30))
    for c in (1, 2):
        cases.append((f"2d_byte_vector(2,3,4)x{c}", tc.make_2d_byte_vector(2, 3, 4), c))
        cases.append((f"2d_byte_subarray(2,3,4)x{c}", tc.make_2d_byte_subarray(2, 3, 4), c))
        cases.append((f"off_subarray((2,3,4),(16,16,16),(1,1,4))x{c}",
                      tc.make_off_subarray((2, 3, 4), (16, 16, 16), (1, 1, 4)), c))
        cases.append((f"byte_v_hv((2,3,4),(10,10,10))x{c}",
                      tc.make_byte_v_hv((2, 3, 4), (10, 10, 10)), c))
    cases.append(("byte_v_hv((10,10,1),(10,10,10))x1", tc.make_byte_v_hv((10, 10, 1), (10, 10, 10)), 1))
    cases.append(("byte_v_hv((4,3,4),(200,200,200))x1", tc.make_byte_v_hv((4, 3, 4), (200, 200, 200)), 1))
    for c in (1, 3):
        cases.append((f"byte_v_hv((100,100,100),(200,200,200))x{c}",
                      tc.make_byte_v_hv((100, 100, 100), (200, 200, 200)), c))
    # probes recorded with MPICH (known answers)
    cases.append(("kat_vector(3,2,4)", ("vector", 3, 2, 4, tc.BYTE), 1))
    cases.append(("kat_hv_by_cols(13,3,16,5,53)", tc.make_2d_hv_by_cols(13, 3, 16, 5, 53), 1))
    cases.append(("kat_padded_subarray({1,4},{1,3})x2", ("subarray", [1, 4], [1, 3], [0, 0], "C", tc.BYTE), 2))
    cases.append(("kat_subarray_count1_start({4,8},{1,5},{2,1})",
                  ("subarray", [4, 8], [1, 5], [2, 1], "C", tc.BYTE), 1))
    # the remaining support/type.cpp factories
    for cp, al in (((2, 3, 4), (10, 10, 10)), ((12, 5, 3), (16, 8, 4)), ((1, 1, 1), (4, 4, 4))):
        for f in (tc.make_byte_vn_hv_hv, tc.make_byte_v1_hv_hv, tc.make_byte_v_hv, tc.make_hi,
                  tc.make_hib, tc.make_subarray, tc.make_subarray_v):
            for c in (1, 2):
                cases.append((f"{f.__name__}({cp},{al})x{c}", f(cp, al), c))
    cases.append(("float_v_hv((8,3,2),(16,4,4))x2", tc.make_float_v_hv((8, 3, 2), (16, 4, 4)), 2))
    for f in (tc.make_2d_byte_vector, tc.make_2d_byte_hvector, tc.make_2d_byte_subarray):
        for nb, bl, st in ((4, 1, 16), (4, 8, 16), (3, 128, 256), (2, 256, 256), (7, 5, 9)):
            cases.append((f"{f.__name__}({nb},{bl},{st})x2", f(nb, bl, st), 2))
    cases.append(("hv_by_rows(13,3,16,5,53)", tc.make_2d_hv_by_rows(13, 3, 16, 5, 53), 2))
    for f in (tc.make_contiguous_byte_v1, tc.make_contiguous_byte_vn, tc.make_contiguous_subarray,
              tc.make_contiguous_contiguous):
        cases.append((f"{f.__name__}(37)x3", f(37), 3))
    # halo faces/edges/corners of a small float64 domain, radius 2
    for d in tc.halo_directions():
        cases.append((f"halo{d}", tc.halo_type(2, (6, 5, 4), d, 8), 1))
    # the 2D headline case (vector(1024,256,512) of bytes)
    cases.append(("c1_vector(1024,256,512)", ("vector", 1024, 256, 512, tc.BYTE), 1))
    # float grid faces at a reduced size (the full 518^3 grid is checked
    # against the oracle on the GPU box)
    for face in "xyz":
        cases.append((f"grid_face_{face}(32,3)", tc.grid_face_subarray(32, 3, face), 1))
    # negative strides, resized and struct
    cases.append(("neg_hvector", ("hvector", 4, 2, -24, ("contiguous", 3, tc.BYTE)), 2))
    cases.append(("resized_vector", ("resized", 0, 40, ("vector", 3, 2, 4, tc.INT)), 3))
    cases.append(("struct_one_block", ("struct", [3], [16], [tc.DOUBLE]), 2))
    # seeded random nested types
    rng = random.Random(1234)
    for k in range(120):
        cases.append((f"random{k}", tc.random_type(rng), rng.randint(1, 3)))
    return cases


def main():
    m = tmpi.get()
    out = []
    for i, (name, spec, incount) in enumerate(catalog()):
        out.append(case(m, name, spec, incount, seed=1000 + i))
    with open(OUT, "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py",
                   "library": "MPICH 3.3.2 (MPI_Pack/MPI_Unpack, homogeneous)",
                   "cases": out}, f, indent=0, separators=(",", ":"))
    print(f"wrote {len(out)} cases to {OUT}")
    m.finalize()


if __name__ == "__main__":
    main()
