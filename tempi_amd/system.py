"""The measured system-performance model, mirroring the reference's
``tempi::system`` namespace (include/measure_system.hpp:19-132).

Everything here calls libtempi_amd.so; the tables live in the library
(process-wide, loaded from ``<cache dir>/perf.json`` at MPI_Init).

    interp_time(curve, bytes)          measure_system.cpp:182-214
    interp_2d(table, bytes, block)     measure_system.cpp:216-297
    measure_performance(...)           measure_system.cu:381-609
    import_performance / export_performance   measure_system.cpp:140-180
    model_device / model_oneshot / model_staged   sender.cpp:128-137,180-191,238-248
    choose_method(colocated, bytes, block)        sender.cpp:251-273
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import check, lib

D2H, H2D, INTRA_CPU_CPU, INTRA_GPU_GPU, INTER_CPU_CPU, INTER_GPU_GPU = range(6)
CURVE_KEYS = ("d2h", "h2d", "intraNodeCpuCpuPingpong", "intraNodeGpuGpuPingpong",
              "interNodeCpuCpuPingpong", "interNodeGpuGpuPingpong")
PACK_DEVICE, UNPACK_DEVICE, PACK_HOST, UNPACK_HOST = range(4)
TABLE_KEYS = ("packDevice", "unpackDevice", "packHost", "unpackHost")
TABLE_DIM = 9
CURVE_LEN = 24
METHOD_UNKNOWN, METHOD_DEVICE, METHOD_ONESHOT, METHOD_STAGED = range(4)

MEASURE_LAUNCH, MEASURE_COPIES, MEASURE_DEVICE_TABLES, MEASURE_HOST_TABLES = 1, 2, 4, 8
MEASURE_ALL = 15
# collective over MPI_COMM_WORLD, at least two ranks; needs no GPU
MEASURE_PINGPONG = 16


def _dptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def interp_time(curve, nbytes: int) -> float:
    """Time of `nbytes` from a curve whose entry i is the time of 2^i bytes;
    +inf for an empty curve."""
    a = np.ascontiguousarray(curve, dtype=np.float64)
    return float(lib().tempi_interp_time(_dptr(a), int(a.size), int(nbytes)))


def interp_2d(table, nbytes: int, block_length: int) -> float:
    """Time from a table whose row i is 2^(2i+6) bytes and column j is a
    block length of 2^j bytes; +inf for an empty table."""
    a = np.ascontiguousarray(table, dtype=np.float64)
    if a.size == 0:
        return math.inf
    rows, cols = a.shape
    return float(lib().tempi_interp_2d(_dptr(a), rows, cols, int(nbytes), int(block_length)))


def clear():
    check(lib().tempi_perf_clear(), "perf_clear")


def get_curve(curve: int) -> np.ndarray:
    n = ctypes.c_int()
    check(lib().tempi_perf_get_curve(curve, None, 0, ctypes.byref(n)), "get_curve")
    out = np.zeros(n.value, np.float64)
    check(lib().tempi_perf_get_curve(curve, _dptr(out), out.size, ctypes.byref(n)), "get_curve")
    return out


def set_curve(curve: int, times):
    a = np.ascontiguousarray(times, dtype=np.float64)
    check(lib().tempi_perf_set_curve(curve, _dptr(a) if a.size else None, int(a.size)),
          "set_curve")


def get_table(table: int) -> np.ndarray:
    r, c = ctypes.c_int(), ctypes.c_int()
    check(lib().tempi_perf_get_table(table, None, 0, ctypes.byref(r), ctypes.byref(c)),
          "get_table")
    out = np.zeros((r.value, c.value), np.float64)
    check(lib().tempi_perf_get_table(table, _dptr(out), out.size, ctypes.byref(r),
                                     ctypes.byref(c)), "get_table")
    return out


def set_table(table: int, times):
    a = np.ascontiguousarray(times, dtype=np.float64)
    rows, cols = (a.shape if a.size else (0, 0))
    check(lib().tempi_perf_set_table(table, _dptr(a) if a.size else None, rows, cols),
          "set_table")


def kernel_launch() -> float:
    return float(lib().tempi_perf_kernel_launch())


def set_kernel_launch(seconds: float):
    check(lib().tempi_perf_set_kernel_launch(float(seconds)), "set_kernel_launch")


def import_performance(path: str | None = None):
    """Load perf.json (default: the cache directory's)."""
    check(lib().tempi_perf_load(path.encode() if path else None), "perf_load")


def export_performance(path: str | None = None):
    check(lib().tempi_perf_save(path.encode() if path else None), "perf_save")


def measure_performance(what: int = MEASURE_ALL, seconds_per_point: float = 0.05):
    """Fill the empty entries this GPU can measure (needs a GPU); with
    MEASURE_PINGPONG, collectively with every rank of MPI_COMM_WORLD, the
    CPU-CPU ping-pong curve (two ranks or more, no GPU needed)."""
    check(lib().tempi_perf_measure(what, seconds_per_point), "perf_measure")


def model(method: int, colocated: bool, nbytes: int, block_length: int):
    """Modelled seconds, or None when an input table is missing."""
    s = ctypes.c_double()
    err = lib().tempi_perf_model(method, int(bool(colocated)), int(nbytes), int(block_length),
                                 ctypes.byref(s))
    return None if err else float(s.value)


def model_device(colocated, nbytes, block_length):
    return model(METHOD_DEVICE, colocated, nbytes, block_length)


def model_oneshot(colocated, nbytes, block_length):
    return model(METHOD_ONESHOT, colocated, nbytes, block_length)


def model_staged(colocated, nbytes, block_length):
    return model(METHOD_STAGED, colocated, nbytes, block_length)


def choose_method(colocated: bool, nbytes: int, block_length: int) -> int:
    return int(lib().tempi_choose_method(int(bool(colocated)), int(nbytes), int(block_length)))
