"""The measured system-performance model on the CPU: interpolation against the
reference's own known answers (test/measure_system.cpp:13-92) and closed
forms, the library against the oracle, perf.json in the reference's schema,
the send models and the DEVICE / ONE_SHOT choice."""
import json
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import tempi_amd as ta
from oracle import perf_ref
from tempi_amd import system as tsys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "perf_reference_schema.json")

# test/measure_system.cpp:15-40 and :42-91
KAT_CURVE = [2, 4, 8, 16]
KAT_TIME = [(1, 2), (2, 4), (3, 6), (5, 10), (7, 14)]
KAT_TABLE = [[14, 18, 22], [16, 20, 24]]  # rows 64 B, 256 B; blocks 1, 2, 4
KAT_2D = [((64, 1), 14), ((160, 1), 15), ((256, 1), 16), ((64, 2), 18), ((256, 2), 20),
          ((64, 3), 20), ((64, 4), 22), ((160, 3), 21), ((512, 1), 32), ((512, 3), 44)]


@pytest.fixture(autouse=True)
def _clean_tables():
    tsys.clear()
    yield
    tsys.clear()


def test_reference_known_answers_library():
    for b, want in KAT_TIME:
        assert tsys.interp_time(KAT_CURVE, b) == want
    for (b, x), want in KAT_2D:
        assert tsys.interp_2d(KAT_TABLE, b, x) == want, (b, x)


def test_reference_known_answers_oracle():
    for b, want in KAT_TIME:
        assert perf_ref.interp_time(KAT_CURVE, b) == want
    for (b, x), want in KAT_2D:
        assert perf_ref.interp_2d(KAT_TABLE, b, x) == want, (b, x)


def test_planted_axis_swap_is_caught():
    """An interpolation that reads the table transposed (bytes along the
    columns) must fail the known answers."""
    t = np.array(KAT_TABLE, dtype=float)
    wrong = [perf_ref.interp_2d(t.T.tolist(), b, x) for (b, x), _ in KAT_2D]
    assert wrong != [w for _, w in KAT_2D]


def test_empty_inputs_are_infinite():
    assert tsys.interp_time([], 100) == math.inf
    assert tsys.interp_2d(np.zeros((0, 0)), 100, 4) == math.inf
    assert perf_ref.interp_time([], 100) == math.inf


def test_closed_forms():
    """A curve linear in bytes and a table bilinear in (bytes, block) are
    reproduced exactly wherever the interpolation is linear, and beyond the
    last row (scaling) for the bytes-linear part."""
    c = 3e-9
    curve = [c * 2 ** i for i in range(24)]
    rng = np.random.default_rng(7)
    for b in list(rng.integers(1, 1 << 30, 200)) + [1, 2, 3, 1 << 23, (1 << 23) + 1]:
        assert tsys.interp_time(curve, int(b)) == pytest.approx(c * int(b), rel=1e-12)
    alpha, beta = 2e-10, 5e-11
    table = [[64 * 4 ** i * (alpha + beta * 2 ** j) for j in range(9)] for i in range(9)]
    for _ in range(300):
        b = int(rng.integers(64, 1 << 26))
        x = int(rng.integers(1, 257))
        want = b * (alpha + beta * x)
        assert tsys.interp_2d(table, b, x) == pytest.approx(want, rel=1e-12), (b, x)
    # past the last column the block length clamps to it
    for b in (64, 5000, 1 << 22):
        assert tsys.interp_2d(table, b, 400) == pytest.approx(b * (alpha + beta * 256), rel=1e-12)


def test_library_matches_oracle_everywhere():
    rng = np.random.default_rng(11)
    for trial in range(40):
        n = int(rng.integers(1, 25))
        curve = rng.uniform(1e-6, 1e-3, n)
        rows, cols = int(rng.integers(1, 10)), int(rng.integers(1, 10))
        table = rng.uniform(1e-6, 1e-3, (rows, cols))
        for _ in range(50):
            b = int(rng.choice([0, 1, 17, 31, 32, 48, 63, 64, 65, 255, 256, 257]) if
                    rng.random() < 0.3 else rng.integers(0, 1 << 34))
            x = int(rng.integers(0, 600))
            assert tsys.interp_time(curve, b) == pytest.approx(
                perf_ref.interp_time(list(curve), b), rel=1e-12)
            assert tsys.interp_2d(table, b, x) == pytest.approx(
                perf_ref.interp_2d(table.tolist(), b, x), rel=1e-12, abs=1e-18), (rows, cols, b, x)


def _fixture_times():
    with open(FIXTURE) as fh:
        raw = json.load(fh)
    out = {}
    for k, v in raw.items():
        if k == "cudaKernelLaunch":
            out[k] = v
        elif v and isinstance(v[0], list):
            out[k] = [[e["time"] for e in row] for row in v]
        else:
            out[k] = [e["time"] for e in v]
    return out


def test_reference_schema_file_loads():
    ref = _fixture_times()
    tsys.import_performance(FIXTURE)
    assert tsys.kernel_launch() == ref["cudaKernelLaunch"]
    for c, key in enumerate(tsys.CURVE_KEYS):
        assert tsys.get_curve(c).tolist() == ref[key], key
    for t, key in enumerate(tsys.TABLE_KEYS):
        assert tsys.get_table(t).tolist() == ref[key], key


def test_save_load_round_trip(tmp_path):
    tsys.import_performance(FIXTURE)
    before = [tsys.get_table(t) for t in range(4)] + [tsys.get_curve(c) for c in range(6)]
    path = str(tmp_path / "perf.json")
    tsys.export_performance(path)
    text = json.load(open(path))  # valid JSON with every reference key
    assert set(text) == set(tsys.CURVE_KEYS) | set(tsys.TABLE_KEYS) | {"cudaKernelLaunch"}
    assert all(set(e) == {"iid", "time"} for e in text["packDevice"][3])
    assert [e["iid"] for e in text["d2h"][:4]] == [True, False, False, True]
    tsys.clear()
    assert tsys.get_table(0).size == 0
    tsys.import_performance(path)
    after = [tsys.get_table(t) for t in range(4)] + [tsys.get_curve(c) for c in range(6)]
    for a, b in zip(before, after):
        assert a.shape == b.shape and (a == b).all()


def test_malformed_files_are_rejected_and_keep_the_tables(tmp_path):
    tsys.import_performance(FIXTURE)
    keep = tsys.get_table(tsys.PACK_HOST)
    raw = json.load(open(FIXTURE))
    cases = {}
    missing = dict(raw)
    del missing["packHost"]
    cases["missing"] = json.dumps(missing)
    ragged = dict(raw)
    ragged["packDevice"] = [row[: 9 - i] for i, row in enumerate(raw["packDevice"])]
    cases["ragged"] = json.dumps(ragged)
    badiid = dict(raw)
    badiid["d2h"] = [{"time": 1.0}]
    cases["no_iid"] = json.dumps(badiid)
    cases["truncated"] = open(FIXTURE).read()[:5000]
    cases["not_object"] = "[1, 2]"
    for name, text in cases.items():
        p = tmp_path / f"{name}.json"
        p.write_text(text)
        assert ta.lib().tempi_perf_load(str(p).encode()) == ta.TEMPI_ERR_IO, name
        assert (tsys.get_table(tsys.PACK_HOST) == keep).all(), name
    assert ta.lib().tempi_perf_load(b"/nonexistent/perf.json") == ta.TEMPI_ERR_IO


def test_mutated_files_load_or_fail_cleanly(tmp_path):
    """Seeded mutations of a valid perf.json (characters replaced by JSON
    punctuation, digits and letters; spans deleted or repeated; truncation;
    deep nesting): each either loads whole or is refused with TEMPI_ERR_IO
    and leaves the tables as they were."""
    text = open(FIXTURE).read()
    rng = np.random.default_rng(11)
    alphabet = list('{}[],:"-+.eE0123456789 tfnul\\x')
    cases = ["[" * 100000, "{" + '"a":' * 5000 + "1" + "}" * 5000]
    for _ in range(300):
        t = list(text)
        for _ in range(int(rng.integers(1, 4))):
            i = int(rng.integers(0, len(t)))
            op = int(rng.integers(0, 4))
            if op == 0:
                t[i] = alphabet[int(rng.integers(0, len(alphabet)))]
            elif op == 1:
                del t[i:i + int(rng.integers(1, 40))]
            elif op == 2:
                t[i:i] = t[i:i + int(rng.integers(1, 40))]
            else:
                del t[i:]
        cases.append("".join(t))
    tsys.import_performance(FIXTURE)
    keep = [tsys.get_table(k) for k in range(4)]
    loaded = 0
    path = tmp_path / "perf.json"
    for k, case in enumerate(cases):
        path.write_text(case)
        err = ta.lib().tempi_perf_load(str(path).encode())
        assert err in (ta.TEMPI_SUCCESS, ta.TEMPI_ERR_IO), k
        if err == ta.TEMPI_SUCCESS:
            loaded += 1
            assert all(tsys.get_table(t).shape == (9, 9) for t in range(4)), k
            tsys.import_performance(FIXTURE)
        else:
            assert all((tsys.get_table(t) == keep[t]).all() for t in range(4)), k
    assert 0 < loaded < len(cases)


def test_models_and_choice_match_the_oracle():
    tsys.import_performance(FIXTURE)
    ref = _fixture_times()
    rng = np.random.default_rng(5)
    seen = set()
    for _ in range(400):
        b = int(rng.integers(1, 1 << 24))
        x = int(2 ** rng.uniform(0, 9))
        for name, meth in (("device", tsys.METHOD_DEVICE), ("oneshot", tsys.METHOD_ONESHOT),
                           ("staged", tsys.METHOD_STAGED)):
            got = tsys.model(meth, True, b, x)
            assert got == pytest.approx(perf_ref.model(ref, name, True, b, x), rel=1e-12)
        want = perf_ref.choose(ref, True, b, x)
        got = tsys.choose_method(True, b, x)
        assert got == (tsys.METHOD_ONESHOT if want == "oneshot" else tsys.METHOD_DEVICE)
        seen.add(want)
    # the fixture's host link is slower than its device pack for big messages
    # and cheaper in latency for small ones: both answers occur
    assert seen == {"device", "oneshot"}
    # inter-node curves are empty in the fixture: no model, DEVICE by default
    assert tsys.model(tsys.METHOD_DEVICE, False, 4096, 8) is None
    assert tsys.choose_method(False, 4096, 8) == tsys.METHOD_DEVICE


def test_choice_cache_follows_table_changes():
    tsys.import_performance(FIXTURE)
    first = tsys.choose_method(True, 1 << 20, 256)
    assert first == tsys.METHOD_DEVICE
    # make the host path free: the cached answer must not survive the change
    tsys.set_table(tsys.PACK_HOST, np.zeros((9, 9)))
    tsys.set_table(tsys.UNPACK_HOST, np.zeros((9, 9)))
    tsys.set_curve(tsys.INTRA_CPU_CPU, np.zeros(24))
    assert tsys.choose_method(True, 1 << 20, 256) == tsys.METHOD_ONESHOT


def test_argument_errors():
    L = ta.lib()
    assert L.tempi_perf_set_curve(6, None, 0) == ta.TEMPI_ERR_ARG
    assert L.tempi_perf_set_table(-1, None, 0, 0) == ta.TEMPI_ERR_ARG
    assert L.tempi_perf_set_table(0, None, 17, 9) == ta.TEMPI_ERR_ARG
    assert L.tempi_perf_measure(0x20, 0.0) == ta.TEMPI_ERR_ARG
    # the ping-pong needs a second rank
    assert L.tempi_perf_measure(0x10, 0.0) == ta.TEMPI_ERR_UNSUPPORTED
    assert tsys.model(7, True, 1, 1) is None


def test_mpi_init_loads_the_cache_dir(tmp_path):
    """MPI_Init reads <TEMPI_CACHE_DIR>/perf.json, as the reference's
    system::init does (src/internal/measure_system.cu:32, src/init.cpp:64)."""
    (tmp_path / "perf.json").write_text(open(FIXTURE).read())
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from tempi_amd import mpi as tmpi, system as tsys\n"
        "tmpi.get()\n"
        "print(tsys.get_table(tsys.PACK_DEVICE).shape, tsys.kernel_launch())\n" % ROOT
    )
    env = dict(os.environ, TEMPI_CACHE_DIR=str(tmp_path))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split("\n")[0] == "(9, 9) 2.5e-06"
