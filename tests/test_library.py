"""libtempi_amd.so as a library, on the CPU: it loads, exports every entry
point include/tempi_amd.h declares, forwards host buffers and disabled paths
to the underlying MPI, and reports errors.  No kernel is launched here."""
import ctypes
import os
import random
import re
import subprocess
import sys

import numpy as np
import pytest

import tempi_amd as ta
import typecatalog as tc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tempi_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, re.M):
        name = m.group(1)
        if name.startswith(("tempi_", "MPI_")):
            names.add(name)
    return sorted(names)


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", ta.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_exports_every_declared_function():
    decl = declared_functions()
    assert "MPI_Pack" in decl and "tempi_pack_sb" in decl and len(decl) > 30
    missing = [n for n in decl if n not in exported_symbols()]
    assert not missing, missing


def test_library_loads_and_reports():
    L = ta.lib()
    assert L.tempi_version() >= 100
    assert ta.num_variants() >= 12
    for v in range(ta.num_variants()):
        info = ta.variant_info(v)
        assert info["word_bytes"] in (1, 2, 4, 8, 12, 16)
        assert 64 % info["lanes_per_row"] == 0
    with pytest.raises(ta.TempiError):
        ta.variant_info(ta.num_variants())


def test_fastdiv_matches_integer_division():
    L = ta.lib()
    rng = random.Random(3)
    divisors = [1, 2, 3, 5, 7, 12, 100, 511, 512, 513, 2072, 65535, 65536, 65537,
                (1 << 31) - 1, 1 << 31, (1 << 31) + 1, (1 << 32) - 1]
    divisors += [rng.randint(1, (1 << 32) - 1) for _ in range(300)]
    numerators = [0, 1, 2, (1 << 31) - 1, 1 << 31, (1 << 32) - 2, (1 << 32) - 1]
    for d in divisors:
        ns = numerators + [rng.randint(0, (1 << 32) - 1) for _ in range(50)] + [d - 1, d, d + 1]
        for n in ns:
            n &= 0xFFFFFFFF
            assert L.tempi_fastdiv_check(n, d) == n // d, (n, d)


def test_host_buffers_go_to_the_library(mpi):
    ta.reset_counters()
    spec = ("vector", 8, 3, 7, tc.BYTE)
    h = mpi.type_from_spec(spec)
    src = (np.arange(64) & 0xFF).astype(np.uint8)
    out = np.zeros(24, np.uint8)
    err, pos = mpi.pack(src.ctypes.data, 1, h, out.ctypes.data, out.size)
    assert err == 0 and pos == 24
    assert out.tolist() == [b for r in range(8) for b in range(7 * r, 7 * r + 3)]
    back = np.full(64, 0xAB, np.uint8)
    err, pos = mpi.unpack(out.ctypes.data, out.size, back.ctypes.data, 1, h)
    assert err == 0 and pos == 24
    c = ta.counters()
    assert c["library_pack_calls"] == 1 and c["library_unpack_calls"] == 1
    assert c["pack_kernel_launches"] == 0
    assert c["commits_canonical"] >= 1
    mpi.free(h)


@pytest.mark.skipif(ta.device_count() > 0, reason="checks the no-GPU error path")
def test_no_device_errors():
    sb = ta.StridedBlock.make(0, 8, [4, 2], [1, 8])
    with pytest.raises(ta.TempiError) as e:
        ta.pack_sb(sb, 0, 1, 0, 8)
    assert e.value.code == ta.TEMPI_ERR_NODEVICE
    assert not ta.lib().tempi_alloc_device(16)


def test_argument_errors():
    L = ta.lib()
    pos = ctypes.c_int64(0)
    assert L.tempi_pack_sb(None, None, 1, None, 8, ctypes.byref(pos), -1, None) == ta.TEMPI_ERR_ARG
    sb = ta.StridedBlock.make(0, 8, [4, 2], [1, 8])
    # truncation is detected before any device work
    assert L.tempi_pack_sb(ctypes.byref(sb), None, 2, None, 15, ctypes.byref(pos), -1,
                           None) == ta.TEMPI_ERR_TRUNCATE
    # descriptors are validated before their size is taken: a dimension count
    # beyond the arrays, and counts whose product overflows 64 bits
    wide = ta.StridedBlock.make(0, 8, [4, 2], [1, 8])
    wide.ndims = 99
    huge = ta.StridedBlock.make(0, 1 << 40, [1 << 40, 1 << 40], [1, 1 << 40])
    for d in (wide, huge):
        assert L.tempi_pack_sb(ctypes.byref(d), None, 1, None, 8, ctypes.byref(pos), -1,
                               None) == ta.TEMPI_ERR_ARG
        assert L.tempi_unpack_sb(ctypes.byref(d), None, 8, ctypes.byref(pos), None, 1, -1,
                                 None) == ta.TEMPI_ERR_ARG
    assert L.tempi_pack_sb(ctypes.byref(sb), None, 1 << 62, None, 8, ctypes.byref(pos), -1,
                           None) == ta.TEMPI_ERR_ARG
    assert pos.value == 0
    # a direction other than pack (0) or unpack (1) is refused, not guessed
    v = ctypes.c_int()
    assert L.tempi_select_variant(ctypes.byref(sb), 1, None, None, 7, 0,
                                  ctypes.byref(v)) == ta.TEMPI_ERR_ARG
    feats = (ctypes.c_float * L.tempi_mlp_num_features())()
    assert L.tempi_mlp_features(ctypes.byref(sb), 1, None, None, -1, feats) == ta.TEMPI_ERR_ARG
    for direction, count, code in ((7, 1, ta.TEMPI_ERR_ARG), (ta.DIR_PACK, -1, ta.TEMPI_ERR_ARG)):
        with pytest.raises(ta.TempiError) as e:
            ta.Batch([(sb, 0, 0, count, direction)])
        assert e.value.code == code
    bad = ta.StridedBlock.make(0, 8, [4, 2], [2, 8])
    v = ctypes.c_int()
    assert L.tempi_select_variant(ctypes.byref(bad), 1, None, None, 0, 0,
                                  ctypes.byref(v)) == ta.TEMPI_ERR_ARG


def test_heuristic_selection_is_host_logic():
    # strided / packed addresses only enter through their alignment
    sb = ta.StridedBlock.make(3226116, 555967328, [12, 512, 512], [1, 2072, 1073296])
    v = ta.select_variant(sb, 1, 1 << 20, 1 << 20)
    assert ta.variant_info(v)["word_bytes"] == 12  # 3-float rows, 4-B aligned
    assert ta.variant_info(v)["lanes_per_row"] == 1
    sb = ta.StridedBlock.make(3226116 + 1, 555967328, [12, 512, 512], [1, 2072, 1073296])
    v = ta.select_variant(sb, 1, 1 << 20, 1 << 20)
    assert ta.variant_info(v)["word_bytes"] == 1  # byte-aligned start
    sb = ta.StridedBlock.make(0, 524032, [256, 1024], [1, 512])
    v = ta.select_variant(sb, 1, 1 << 20, 1 << 20)
    assert ta.variant_info(v)["word_bytes"] == 16


def run_py(code, env_extra):
    env = dict(os.environ, **env_extra)
    return subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True,
                          text=True, timeout=120)


def test_disable_env_is_pass_through():
    code = (
        "import sys; sys.path.insert(0, 'tests');"
        "import tempi_amd as ta; from tempi_amd import mpi as tm; import typecatalog as tc\n"
        "m = tm.get(); h = m.type_from_spec(('vector', 4, 2, 3, tc.BYTE))\n"
        "print(ta.counters()['commits_canonical'], m.strided_block(h) is None); m.finalize()\n"
    )
    r = run_py(code, {"TEMPI_DISABLE": "1"})
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["0", "True"]
    r = run_py(code, {"TEMPI_NO_TYPE_COMMIT": "1"})
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["0", "True"]
    r = run_py(code, {})
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["1", "False"]


APP = r"""
#include <mpi.h>
#include <stdio.h>
int main(int argc, char **argv) {
  MPI_Init(&argc, &argv);
  MPI_Datatype t;
  MPI_Type_vector(3, 2, 4, MPI_BYTE, &t);
  MPI_Type_commit(&t);
  unsigned char src[16], dst[6];
  for (int i = 0; i < 16; ++i) src[i] = (unsigned char)i;
  int pos = 0;
  MPI_Pack(src, 1, t, dst, 6, &pos, MPI_COMM_WORLD);
  for (int i = 0; i < 6; ++i) printf("%d ", dst[i]);
  printf("%d\n", pos);
  MPI_Type_free(&t);
  MPI_Finalize();
  return 0;
}
"""


def test_ld_preload_in_front_of_mpi(tmp_path):
    """The drop-in use: an unmodified MPI program with the library preloaded."""
    src = tmp_path / "app.c"
    src.write_text(APP)
    exe = tmp_path / "app"
    mpi_home = "/opt/conda"
    subprocess.run(["gcc", str(src), "-o", str(exe), f"-I{mpi_home}/include", f"-L{mpi_home}/lib",
                    "-lmpi", f"-Wl,-rpath,{mpi_home}/lib"], check=True)
    env = dict(os.environ, LD_PRELOAD=ta.LIB_PATH, TEMPI_LOG_LEVEL="4")
    r = subprocess.run([str(exe)], env=env, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["0", "1", "4", "5", "8", "9", "6"]
    # at debug level MPI_Finalize reports every counter (the reference's
    # counter dump, src/internal/counters.cpp:30-120): here one commit with
    # a strided form and one host pack forwarded to the MPI library
    assert "rank 0 counters::commits_canonical: 1" in r.stderr, r.stderr
    assert "rank 0 counters::library_pack_calls: 1" in r.stderr, r.stderr
    assert r.stderr.count("rank 0 counters::") == 19


STRUCTS = {  # C ABI struct -> (ctypes mirror, its last field)
    "tempi_strided_block": ("StridedBlock", "strides"),
    "tempi_variant_info": ("VariantInfo", None),
    "tempi_copy_emulation": ("CopyEmulation", "reserved"),
    "tempi_batch_request": ("BatchRequest", "variant"),
    "tempi_counters": ("Counters", "signalled_waits"),
}


def test_ctypes_mirrors_match_the_header(tmp_path):
    """Every struct of include/tempi_amd.h has the size (and last-field
    offset) its Python ctypes mirror assumes -- a caller's buffer shorter
    than the struct the library fills is written past its end."""
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "tempi_amd.h"',
             "int main(void) {"]
    for c, (_, last) in STRUCTS.items():
        off = f"offsetof({c}, {last})" if last else "0"
        lines.append(f'  printf("{c} %zu %zu\\n", sizeof({c}), (size_t){off});')
    lines += ["  return 0;", "}"]
    src = tmp_path / "sizes.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "sizes"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run(["gcc", str(src), "-o", str(exe), f"-I{root}/include",
                    "-I/opt/conda/include"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    for line in out.splitlines():
        c, size, off = line.split()
        py, last = STRUCTS[c]
        cls = getattr(ta, py)
        assert ctypes.sizeof(cls) == int(size), (c, ctypes.sizeof(cls), size)
        if last:
            assert getattr(cls, last).offset == int(off), (c, last)
