"""CPU oracle of the measured system-performance model.

TEST INFRASTRUCTURE ONLY: the product interpolates in C++
(tempi_amd/csrc/perf.hip); this restatement is what it is checked against.
It follows the reference's definitions, found here by scanning the measured
points instead of by bit arithmetic:

* interp_time (src/internal/measure_system.cpp:182-214): entry i is the time
  of 2^i bytes; linear between the two entries around `bytes`; past the last
  entry the last time is scaled by bytes / 2^(n-1); empty -> +inf.
* interp_2d (measure_system.cpp:216-297): row i is 2^(2i+6) bytes, column j a
  block length of 2^j; linear in the block length between the two columns
  around it (clamped to the last column), then linear in bytes between the
  row at or below and the next row; past the last row the last row's value
  is scaled by bytes / (its bytes).  Sizes of 32-63 B use rows 0 and 1 with
  a negative weight (what the reference's integer arithmetic gives); below
  32 B the reference indexes out of its table, and the value at 32 B is
  used.
* the send models (src/internal/sender.cpp:128-137, 180-191, 238-248) and
  the DEVICE / ONE_SHOT choice (sender.cpp:251-273): pack + send + unpack,
  the smaller wins, ties and unknowns go to DEVICE.

Pinned by the reference's own known answers (test/measure_system.cpp:13-92)
and by closed forms (tables linear in bytes and in block length
interpolate exactly) in tests/test_perf.py.
"""
from __future__ import annotations

import math

INF = math.inf


def _bracket_pow2(v):
    """(lo, hi) exponents with 2^lo <= v <= 2^hi, hi - lo <= 1 (v >= 1)."""
    lo = 0
    while (1 << (lo + 1)) <= v:
        lo += 1
    hi = lo if (1 << lo) == v else lo + 1
    return lo, hi


def interp_time(curve, nbytes):
    n = len(curve)
    if n == 0 or nbytes < 0:
        return INF
    if nbytes == 0:
        return float(curve[0])
    lo, hi = _bracket_pow2(nbytes)
    if hi >= n:
        return float(curve[-1]) * nbytes / 2 ** (n - 1)
    if lo == hi:
        return float(curve[lo])
    w = (nbytes - 2 ** lo) / (2 ** hi - 2 ** lo)
    return float(curve[lo]) * (1 - w) + float(curve[hi]) * w


def _row_bytes(i):
    return 64 * 4 ** i


def interp_2d(table, nbytes, block):
    rows = len(table)
    if rows == 0 or nbytes < 0 or block < 0:
        return INF
    cols = len(table[0])
    nbytes = max(nbytes, 32)
    # row at or below (row 0 for 32..63 B), and the next one unless exact
    r1 = 0
    while r1 + 1 < 64 and _row_bytes(r1 + 1) <= nbytes:
        r1 += 1
    r2 = r1 if _row_bytes(r1) == nbytes else r1 + 1
    # columns around the block length, clamped to the table
    if block == 0:
        c1 = c2 = 0
        wx = 0.0
    else:
        c1, c2 = _bracket_pow2(block)
        x1, x2 = 2 ** c1, 2 ** c2
        if c2 >= cols:
            c1 = c2 = cols - 1
        wx = 0.0 if c1 == c2 else (block - x1) / (x2 - x1)

    def along_x(r):
        return (1 - wx) * table[r][c1] + wx * table[r][c2]

    if r2 >= rows:
        return along_x(rows - 1) / _row_bytes(rows - 1) * nbytes
    wy = 0.0 if r1 == r2 else (nbytes - _row_bytes(r1)) / (_row_bytes(r2) - _row_bytes(r1))
    return (1 - wy) * along_x(r1) + wy * along_x(r2)


def model(perf, method, colocated, nbytes, block):
    """perf: dict with the reference's perf.json keys (times only)."""
    cpu = perf["intraNodeCpuCpuPingpong" if colocated else "interNodeCpuCpuPingpong"]
    gpu = perf["intraNodeGpuGpuPingpong" if colocated else "interNodeGpuGpuPingpong"]
    if method == "device":
        parts = [interp_2d(perf["packDevice"], nbytes, block), interp_time(gpu, nbytes),
                 interp_2d(perf["unpackDevice"], nbytes, block)]
    elif method == "oneshot":
        parts = [interp_2d(perf["packHost"], nbytes, block), interp_time(cpu, nbytes),
                 interp_2d(perf["unpackHost"], nbytes, block)]
    elif method == "staged":
        parts = [interp_2d(perf["packDevice"], nbytes, block), interp_time(perf["d2h"], nbytes),
                 interp_time(cpu, nbytes), interp_time(perf["h2d"], nbytes),
                 interp_2d(perf["unpackDevice"], nbytes, block)]
    else:
        raise ValueError(method)
    return sum(parts)


def choose(perf, colocated, nbytes, block):
    o = model(perf, "oneshot", colocated, nbytes, block)
    d = model(perf, "device", colocated, nbytes, block)
    return "oneshot" if o < d else "device"
