"""Oracle (test infrastructure only): the permutation test of NIST SP 800-90B
(January 2018), section 5.1, restated in plain Python from the standard's
definitions -- the eight statistics for non-binary data that the reference's
benchmark loop runs on every trial (src/internal/iid.cpp:184-245,
src/internal/benchmark.cpp:44-83) -- with the seeded shuffles the library
documents, so that its counts can be compared exactly.

Pinned by:
* the worked examples the standard gives for each statistic
  (5.1.1 excursion, 5.1.2-5.1.4 directional runs, 5.1.5-5.1.6 runs about the
  median, 5.1.7-5.1.8 collisions), which the tests hold as known answers;
* closed forms: a constant sequence (every shuffle ties every statistic), a
  monotone sequence of n distinct values (its increases statistic is tied
  only by the 2 monotone orders, so C1 ~ Binomial(10 000, 2 / n!));
* the reference's own test (test/iid.cpp:22-57): a monotone vector fails,
  uniform random vectors pass;
* uniformity of the shuffle over the 6 orders of 3 values.

Only tests/ may import this module.
"""
from __future__ import annotations

SHUFFLES = 10000
GOLDEN = 0x9E3779B97F4A7C15
MASK = (1 << 64) - 1


def median(s):
    x = sorted(s)
    n = len(x)
    return x[n // 2] if n % 2 else (x[n // 2 - 1] + x[n // 2]) / 2


def runs_of(signs):
    """(number of runs, longest run) of a list of symbols."""
    runs, longest, cur = 0, 0, 0
    for k, v in enumerate(signs):
        if k == 0 or v != signs[k - 1]:
            runs += 1
            cur = 1
        else:
            cur += 1
        longest = max(longest, cur)
    return runs, longest


def collisions(s):
    """C: for i = 1, 2, ..., the length of the shortest window s_i..s_j with
    two equal values; the next window starts at j + 1 (5.1.7 step 3)."""
    c = []
    i, n = 0, len(s)
    while i < n - 1:
        j = None
        for end in range(i + 1, n):
            window = s[i:end + 1]
            if len(set(window)) < len(window):
                j = end
                break
        if j is None:
            break
        c.append(j - i + 1)
        i = j + 1
    return c


def mean_of(s):
    """Left-to-right sum over n, as the library adds."""
    total = 0.0
    for v in s:
        total += v
    return total / len(s)


def statistics(s, mean=None):
    """The eight statistics, in the library's order.  `mean` is the samples'
    mean (a shuffle keeps the original's, as the standard's X-bar)."""
    n = len(s)
    if mean is None:
        mean = mean_of(s)
    run, exc = 0.0, 0.0
    for i in range(n):
        run += s[i]
        expect = float(i + 1) * mean
        exc = max(exc, abs(run - expect))
    d = ["+" if s[i] <= s[i + 1] else "-" for i in range(n - 1)]
    dr, dl = runs_of(d)
    ups = d.count("+")
    m = median(s)
    hm = ["+" if v >= m else "-" for v in s]
    mr, ml = runs_of(hm)
    c = collisions(s)
    avg = (sum(c) / len(c)) if c else 0.0
    mx = max(c) if c else 0
    return [exc, float(dr), float(dl), float(max(ups, (n - 1) - ups)),
            float(mr), float(ml), float(avg), float(mx)]


class Splitmix64:
    def __init__(self, seed):
        self.s = seed & MASK

    def next(self):
        self.s = (self.s + GOLDEN) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def below(self, n):
        reject = (1 << 64) % n
        while True:
            x = self.next()
            if x >= reject:
                return x % n


def shuffled(s, seed):
    """Fisher-Yates over a copy of s, from the last position down."""
    p = list(s)
    g = Splitmix64(seed)
    for i in range(len(p) - 1, 0, -1):
        j = g.below(i + 1)
        p[i], p[j] = p[j], p[i]
    return p


def permutation_test(s, seed, shuffles=SHUFFLES):
    """(independent?, counts): counts[2 k], counts[2 k + 1] = C0, C1 of
    statistic k over the shuffles; shuffle k uses seed + k * GOLDEN."""
    s = [float(v) for v in s]
    mean = mean_of(s)
    t = statistics(s, mean)
    counts = [0] * 16
    for k in range(shuffles):
        tp = statistics(shuffled(s, (seed + k * GOLDEN) & MASK), mean)
        for i in range(8):
            counts[2 * i] += tp[i] > t[i]
            counts[2 * i + 1] += tp[i] == t[i]
    ok = all(not (counts[2 * i] + counts[2 * i + 1] <= 5 or counts[2 * i] >= shuffles - 5)
             for i in range(8))
    return ok, counts
