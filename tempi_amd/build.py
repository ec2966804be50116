"""Build libtempi_amd.so in-tree with hipcc for gfx950.

    python -m tempi_amd.build            # build (always recompiles)

The library links the underlying MPI (MPICH by default, found through
$MPI_HOME or /opt/conda) so that it can decode datatypes; it is meant to be
linked before libmpi or LD_PRELOADed in front of it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtempi_amd.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = [
    "common.cpp",
    "canonicalize.cpp",
    "runtime.cpp",
    "interpose.cpp",
    "kernels.hip",
    "autotune.hip",
    "batch.hip",
    "mlp.hip",
    "perf.hip",
    "iid.cpp",
    "pool.cpp",
    "p2p.cpp",
    "neighbor.cpp",
]

ARCH = os.environ.get("TEMPI_ARCH", "gfx950")
# ROCm libraries beside the HIP runtime (the ROCTx range API)
ROCM_LIB = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")


def mpi_home() -> str:
    for cand in (os.environ.get("MPI_HOME"), "/opt/conda"):
        if cand and os.path.exists(os.path.join(cand, "include", "mpi.h")):
            return cand
    raise RuntimeError("mpi.h not found: set MPI_HOME")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def system_libdir() -> str | None:
    """Directory the loader resolves the system libstdc++ from.  It goes
    ahead of the MPI directory in RUNPATH, so that an older libstdc++ shipped
    next to libmpi (conda) never shadows the one the HIP runtime needs."""
    try:
        out = subprocess.run(["ldconfig", "-p"], capture_output=True, text=True).stdout
    except OSError:
        return None
    for line in out.splitlines():
        if line.strip().startswith("libstdc++.so.6 ") and "x86-64" in line:
            return os.path.dirname(line.split("=>")[-1].strip())
    return None


def common_flags() -> list[str]:
    mpi = mpi_home()
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-fvisibility=hidden",
        # packed fp32 instructions off for all device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-command-line-argument",
        f"-I{os.path.join(ROOT, 'include')}",
        f"-I{os.path.join(mpi, 'include')}",
    ]


def compile_one(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [hipcc(), *common_flags(), "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd.insert(1, "-x")
        cmd.insert(2, "hip")
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    mpi = mpi_home()
    tmp = OUT + ".tmp"
    rpath = [d for d in (system_libdir(), os.path.join(mpi, "lib")) if d]
    cmd = [
        hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs,
        f"-L{os.path.join(mpi, 'lib')}", "-lmpi",
        f"-L{ROCM_LIB}", "-lrocprofiler-sdk-roctx",
        f"-Wl,-rpath,{':'.join(rpath + [ROCM_LIB])}", "-Wl,--enable-new-dtags",
        "-ldl", "-o", tmp,
    ]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


TEST_APP_SRC = os.path.join(ROOT, "tests", "c", "pack_unpack_app.cpp")
TEST_APP = os.path.join(ROOT, "tests", "c", "pack_unpack_app")
SWEEP_APP_SRC = os.path.join(ROOT, "tests", "c", "mpi_pack_sweep.cpp")
SWEEP_APP = os.path.join(ROOT, "tests", "c", "mpi_pack_sweep")


def build_test_app(verbose: bool = False) -> str:
    """The native MPI programs (tests/c/pack_unpack_app.cpp, the parity
    test; tests/c/mpi_pack_sweep.cpp, the reference's pack and commit
    benchmarks), linked against the MPI library and the HIP runtime only:
    they run with libtempi_amd.so preloaded, the drop-in way."""
    mpi = mpi_home()
    rpath = [d for d in (system_libdir(), os.path.join(mpi, "lib")) if d]
    for src, out in ((TEST_APP_SRC, TEST_APP), (SWEEP_APP_SRC, SWEEP_APP)):
        # libmpi by path: a -L of the MPI directory would also resolve
        # libstdc++ from there (conda ships an older one)
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-O2", "-std=c++17", src,
               f"-I{os.path.join(mpi, 'include')}", f"-I{os.path.join(ROOT, 'include')}",
               f"-Wl,{os.path.join(mpi, 'lib', 'libmpi.so')}",
               f"-Wl,-rpath,{':'.join(rpath)}", "-Wl,--enable-new-dtags", "-o", out]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"native program build failed: {' '.join(cmd)}\n{res.stderr}")
        if verbose:
            print(f"built {out}")
    return TEST_APP


if __name__ == "__main__":
    build(verbose=True)
    build_test_app(verbose=True)
    sys.exit(0)
