"""The NIST SP 800-90B section 5.1 permutation test that decides the `iid`
flag of every measured table entry (reference: sp_800_90B,
src/internal/iid.cpp:184-245; its test, test/iid.cpp:22-57), on the CPU.

Pins: the standard's worked example for each statistic; closed forms
(constant and monotone sequences); the reference test's two cases; the
oracle's counts, equal shuffle for shuffle; uniformity of the shuffle."""
import math

import numpy as np
import pytest

import tempi_amd as ta
from oracle import iid_ref


def stats(s):
    a = np.ascontiguousarray(s, np.float64)
    out = np.zeros(8)
    assert ta.lib().tempi_iid_statistics(a.ctypes.data, len(a), out.ctypes.data) == 0, \
        ta.last_error()
    return out.tolist()


def perm_test(s, seed):
    a = np.ascontiguousarray(s, np.float64)
    c = np.zeros(16, np.int64)
    r = ta.lib().tempi_iid_test(a.ctypes.data, len(a), seed, c.ctypes.data)
    assert r in (0, 1), ta.last_error()
    return r == 1, c.tolist()


# (sequence, statistic index, value): the worked examples of SP 800-90B
# 5.1.1-5.1.8
NIST_EXAMPLES = [
    ([2, 15, 4, 10, 9], 0, 6.0),                      # excursion
    ([2, 2, 2, 5, 7, 7, 9, 3, 1, 4, 4], 1, 3.0),      # directional runs
    ([2, 2, 2, 5, 7, 7, 9, 3, 1, 4, 4], 2, 6.0),      # longest directional run
    ([2, 2, 2, 5, 7, 7, 9, 3, 1, 4, 4], 3, 8.0),      # increases / decreases
    ([5, 15, 12, 1, 13, 9, 4], 4, 5.0),               # runs about the median
    ([5, 15, 12, 1, 13, 9, 4], 5, 2.0),               # longest run about the median
    ([2, 1, 1, 2, 0, 1, 0, 1, 1, 2], 6, 3.0),         # average collision
    ([2, 1, 1, 2, 0, 1, 0, 1, 1, 2], 7, 4.0),         # maximum collision
]


@pytest.mark.parametrize("seq,k,want", NIST_EXAMPLES)
def test_statistics_match_the_standards_examples(seq, k, want):
    assert stats(seq)[k] == want
    assert iid_ref.statistics([float(v) for v in seq])[k] == want


def test_statistics_of_closed_forms():
    # constant: no excursion, one run each way, a collision every 2 samples
    assert stats([3.5] * 9) == [0.0, 1.0, 8.0, 8.0, 1.0, 9.0, 2.0, 2.0]
    # strictly increasing: excursion at the middle, one directional run,
    # below/above the median in two runs, no collisions
    n = 10
    s = list(range(n))
    exc = max(abs(sum(s[:i + 1]) - (i + 1) * (n - 1) / 2) for i in range(n))
    assert stats(s) == [exc, 1.0, 9.0, 9.0, 2.0, 5.0, 0.0, 0.0]
    # an even-length median is the mean of the middle pair (2.5 here): the 2
    # falls below it, + - + -, four runs (the lower middle value as the
    # median would make it + - + +, three)
    assert stats([4, 1, 3, 2])[4] == 4.0


def test_counts_match_the_oracle_shuffle_for_shuffle():
    rng = np.random.default_rng(11)
    cases = [
        rng.integers(0, 6, 12).astype(float),            # many ties
        rng.normal(size=16),
        np.round(rng.exponential(size=20), 1),
        np.arange(7.0) - 1,                              # the reference's vector
        rng.normal(size=65),                             # several worker threads
        np.array([1e300, -1e300, 1.0, 2.0, 1e-300]),     # rounding in the excursion
    ]
    for s in cases:
        seed = int(rng.integers(0, 2 ** 63))
        ok, counts = perm_test(s, seed)
        want_ok, want = iid_ref.permutation_test(list(s), seed)
        assert counts == want, len(s)
        assert ok == want_ok
        assert sum(counts[1::2]) + sum(counts[0::2]) <= 8 * 10000


def test_constant_samples_are_exchangeable():
    ok, counts = perm_test([2.0] * 30, 1)
    assert ok and counts[0::2] == [0] * 8 and counts[1::2] == [10000] * 8


def test_monotone_vectors_fail():
    """The reference's fixed vector (test/iid.cpp:22-28), read through a
    closed form.  A monotone order of n distinct values has n - 1 increases;
    only the 2 monotone shuffles tie that, so C0 = 0 and
    C1 ~ Binomial(10 000, 2 / n!), and the test rejects whenever C1 <= 5.
    For n = 7 that is P(Poisson(3.97) <= 5) = 0.79 of seeds (other
    statistics can only add rejections); for n = 20, every seed."""
    s = np.arange(7.0) - 1
    lam = 10000 * 2 / math.factorial(7)
    c1, rejected = [], 0
    for seed in range(120):
        ok, counts = perm_test(s, seed)
        assert counts[6] == 0
        c1.append(counts[7])
        rejected += not ok
    assert abs(np.mean(c1) - lam) < 4 * math.sqrt(lam / len(c1)), np.mean(c1)
    p = sum(math.exp(-lam) * lam ** k / math.factorial(k) for k in range(6))
    assert rejected / 120 >= p - 4 * math.sqrt(p * (1 - p) / 120)
    for seed in range(10):
        assert not perm_test(np.arange(20.0), seed)[0]
        assert not perm_test(np.arange(20.0)[::-1], seed)[0]


def test_uniform_random_vectors_pass():
    """The reference's random loop (test/iid.cpp:30-55) waits for one pass;
    with eight statistics at a 0.001 two-sided level each, nearly all pass."""
    rng = np.random.default_rng(3)
    passed = sum(perm_test(rng.uniform(0, 1e4, 10), seed)[0] for seed in range(40))
    assert passed >= 36
    passed = sum(perm_test(rng.normal(size=400), seed)[0] for seed in range(20))
    assert passed >= 18


def test_dependent_sequences_fail():
    rng = np.random.default_rng(9)
    ar = np.zeros(300)
    for i in range(1, 300):
        ar[i] = 0.9 * ar[i - 1] + rng.normal()
    assert not perm_test(ar, 5)[0]
    drift = np.linspace(10, 11, 200) + rng.normal(scale=0.05, size=200)
    assert not perm_test(drift, 6)[0]
    alternating = np.tile([1.0, 5.0], 50) + rng.normal(scale=0.01, size=100)
    assert not perm_test(alternating, 7)[0]


def test_stream_is_splitmix64():
    """The shuffles' generator: SplitMix64's published first outputs from
    seed 0 (the constants of java.util.SplittableRandom)."""
    g = iid_ref.Splitmix64(0)
    assert g.next() == 0xE220A8397B1DCDAF
    assert g.next() == 0x6E789E6AA1B965F4
    # an index below n never comes from the biased low end of the range
    g = iid_ref.Splitmix64(7)
    assert all(0 <= g.below(n) < n for n in (1, 2, 3, 7, 1000, (1 << 63) + 1))


def test_shuffle_is_uniform():
    """Every order of 3 values appears about 1/6 of the time over 6 000 seeds
    (oracle stream; the library's counts equal the oracle's above)."""
    orders = {}
    for seed in range(6000):
        key = tuple(iid_ref.shuffled([0, 1, 2], seed * iid_ref.GOLDEN & iid_ref.MASK))
        orders[key] = orders.get(key, 0) + 1
    assert len(orders) == 6
    assert all(abs(v - 1000) < 4 * math.sqrt(1000 * 5 / 6) for v in orders.values())


def test_bad_input_is_an_error():
    L = ta.lib()
    one = np.array([1.0])
    assert L.tempi_iid_test(one.ctypes.data, 1, 0, None) == ta.TEMPI_ERR_ARG
    nan = np.array([1.0, float("nan"), 2.0])
    assert L.tempi_iid_test(nan.ctypes.data, 3, 0, None) == ta.TEMPI_ERR_ARG
    big = np.zeros((1 << 16) + 1)
    assert L.tempi_iid_test(big.ctypes.data, len(big), 0, None) == ta.TEMPI_ERR_ARG
    assert L.tempi_iid_test(None, 5, 0, None) == ta.TEMPI_ERR_ARG
    two = np.array([1.0, 2.0])
    assert L.tempi_iid_statistics(two.ctypes.data, 2, None) == ta.TEMPI_ERR_ARG


def test_trial_count_comes_from_the_environment(monkeypatch):
    """$TEMPI_MEASURE_TRIALS is checked when a measurement starts (mask 0
    measures nothing and needs no GPU)."""
    L = ta.lib()
    for bad in ("0", "-2", "x", "3x"):
        monkeypatch.setenv("TEMPI_MEASURE_TRIALS", bad)
        assert L.tempi_perf_measure(0, 0.0) == ta.TEMPI_ERR_ARG, bad
    monkeypatch.setenv("TEMPI_MEASURE_TRIALS", "3")
    assert L.tempi_perf_measure(0, 0.0) == 0
    monkeypatch.delenv("TEMPI_MEASURE_TRIALS")
    assert L.tempi_perf_measure(0, 0.0) == 0


def test_python_wrappers():
    assert ta.iid_statistics([2, 15, 4, 10, 9])[0] == 6.0
    ok, counts = ta.iid_test([2.0] * 30, seed=1)
    assert ok and counts[1::2] == [10000] * 8
    with pytest.raises(ta.TempiError):
        ta.iid_test([1.0])
