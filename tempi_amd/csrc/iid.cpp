// iid.cpp -- the permutation test of NIST SP 800-90B (January 2018),
// section 5.1, that decides whether a measurement's samples look independent
// and identically distributed.
//
// reference: src/internal/iid.cpp:184-245 (sp_800_90B) runs it on every
// trial of a benchmark point and src/internal/benchmark.cpp:44-83 repeats
// trials until it passes (at most RunConfig::maxTrials = 10); the result is
// the `iid` flag of every perf.json entry.  The reference uses eight of the
// standard's statistics for non-binary data (5.1.1-5.1.8), and so does this
// file, written from the standard's definitions:
//
//   0 excursion                max_i |s_1 + ... + s_i - i * mean|
//   1 directional runs         runs in s'_i = (s_i <= s_{i+1}), i < L
//   2 longest directional run  longest run in that s'
//   3 increases / decreases    max(#(s_i <= s_{i+1}), #(s_i > s_{i+1}))
//   4 runs about the median    runs in s'_i = (s_i >= median)
//   5 longest median run       longest run in that s'
//   6 average collision        mean of C, below
//   7 maximum collision        max of C
//
// C lists, walking from i = 1, the length j - i + 1 of the shortest window
// s_i .. s_j holding two equal values, the next window starting at j + 1
// (an empty C gives 0 for both collision statistics).  The median of an
// even-length sequence is the mean of its two middle values.
//
// The test: compute the eight statistics T of the samples; shuffle the
// samples 10 000 times and compute them again (T'); per statistic count
// C0 = #(T' > T) and C1 = #(T' == T); reject independence when
// C0 + C1 <= 5 or C0 >= 9995 for any statistic.
//
// Where this differs from the reference's code (which the standard decides):
// the reference stores every statistic in an int64 (truncating the
// excursion and the average collision), takes the middle element pair of an
// odd-length sequence as its median, leaves the last run out of both
// longest-run statistics, and counts only collisions with the window's
// first sample.
//
// Shuffle k (of 10 000) permutes the original samples with a Fisher-Yates
// shuffle driven by its own splitmix64 stream (Steele, Lea and Flood, "Fast
// splittable pseudorandom number generators", OOPSLA 2014) seeded with
// seed + k * 0x9E3779B97F4A7C15, drawing an index below n by rejecting
// draws under 2^64 mod n.  The shuffles are independent, so they run on
// several threads and the counts do not depend on how many.
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace tempi {

namespace {

constexpr int kShuffles = 10000;
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

struct Sample {
  double v;    // the value
  int32_t r;   // dense rank of the value: equal values, equal ranks
  uint8_t hi;  // v >= median
};

struct Stream {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += kGolden);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [0, n): draws below 2^64 mod n are rejected
  uint64_t below(uint64_t n) {
    const uint64_t reject = (0 - n) % n;
    for (;;) {
      const uint64_t x = next();
      if (x >= reject) {
        return x % n;
      }
    }
  }
};

// The eight statistics of a[0, n).  `mean` is the samples' mean;
// `seen` (one entry per distinct value) and `stamp` track the collision
// windows.  Contraction off: the excursion sums must round as written.
void statistics(const Sample *a, int64_t n, double mean,
                std::vector<int64_t> &seen, int64_t &stamp, double *t) {
#pragma clang fp contract(off)
  double run = 0, exc = 0;
  for (int64_t i = 0; i < n; ++i) {
    run += a[i].v;
    const double expect = (double)(i + 1) * mean;
    const double d = std::fabs(run - expect);
    exc = std::max(exc, d);
  }
  t[0] = exc;

  int64_t runs = 1, longest = 1, cur = 1, up_n = 0;
  bool prev = false;
  for (int64_t i = 0; i + 1 < n; ++i) {
    const bool up = a[i].r <= a[i + 1].r;
    up_n += up ? 1 : 0;
    if (i > 0) {
      if (up == prev) {
        ++cur;
      } else {
        ++runs;
        cur = 1;
      }
    }
    longest = std::max(longest, cur);
    prev = up;
  }
  t[1] = (double)runs;
  t[2] = (double)longest;
  t[3] = (double)std::max(up_n, (n - 1) - up_n);

  runs = 1;
  longest = 1;
  cur = 1;
  for (int64_t i = 1; i < n; ++i) {
    if (a[i].hi == a[i - 1].hi) {
      ++cur;
    } else {
      ++runs;
      cur = 1;
    }
    longest = std::max(longest, cur);
  }
  t[4] = (double)runs;
  t[5] = (double)longest;

  int64_t windows = 0, total = 0, widest = 0;
  int64_t i = 0;
  while (i < n - 1) {
    ++stamp;
    int64_t j = -1;
    for (int64_t k = i; k < n; ++k) {
      int64_t &s = seen[(size_t)a[k].r];
      if (s == stamp) {
        j = k;
        break;
      }
      s = stamp;
    }
    if (j < 0) {
      break;
    }
    const int64_t len = j - i + 1;
    ++windows;
    total += len;
    widest = std::max(widest, len);
    i = j + 1;
  }
  t[6] = windows ? (double)total / (double)windows : 0.0;
  t[7] = (double)widest;
}

// Samples with ranks and median flags, and their mean; TEMPI_ERR_ARG for
// fewer than 2 samples, more than kIidMaxSamples or a non-finite one.
int prepare(const double *s, int64_t n, std::vector<Sample> *out,
            double *mean, int64_t *distinct) {
#pragma clang fp contract(off)
  if (!s || n < 2 || n > kIidMaxSamples) {
    return set_error(TEMPI_ERR_ARG, "IID test of %lld samples", (long long)n);
  }
  std::vector<double> sorted(s, s + n);
  double sum = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (!std::isfinite(s[i])) {
      return set_error(TEMPI_ERR_ARG, "sample %lld is not finite", (long long)i);
    }
    sum += s[i];
  }
  std::sort(sorted.begin(), sorted.end());
  const double median =
      (n % 2) ? sorted[(size_t)(n / 2)]
              : (sorted[(size_t)(n / 2 - 1)] + sorted[(size_t)(n / 2)]) / 2;
  sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
  out->resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    Sample &x = (*out)[(size_t)i];
    x.v = s[i];
    x.r = (int32_t)(std::lower_bound(sorted.begin(), sorted.end(), s[i]) -
                    sorted.begin());
    x.hi = s[i] >= median ? 1 : 0;
  }
  *mean = sum / (double)n;
  *distinct = (int64_t)sorted.size();
  return TEMPI_SUCCESS;
}

} // namespace

int iid_statistics(const double *samples, int64_t n, double *t) {
  std::vector<Sample> a;
  double mean = 0;
  int64_t distinct = 0;
  int err = prepare(samples, n, &a, &mean, &distinct);
  if (err) {
    return err;
  }
  std::vector<int64_t> seen((size_t)distinct, 0);
  int64_t stamp = 0;
  statistics(a.data(), n, mean, seen, stamp, t);
  return TEMPI_SUCCESS;
}

int iid_test(const double *samples, int64_t n, uint64_t seed,
             int64_t *counts) {
  std::vector<Sample> a;
  double mean = 0;
  int64_t distinct = 0;
  int err = prepare(samples, n, &a, &mean, &distinct);
  if (err) {
    return err;
  }
  double t[kIidStats];
  {
    std::vector<int64_t> seen((size_t)distinct, 0);
    int64_t stamp = 0;
    statistics(a.data(), n, mean, seen, stamp, t);
  }
  // about n * 10 000 * 10 operations: threads once that passes ~10 ms
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int workers =
      (int)std::min<int64_t>(std::min<unsigned>(hw, 8u), 1 + n / 64);
  std::vector<std::vector<int64_t>> c((size_t)workers,
                                      std::vector<int64_t>(2 * kIidStats, 0));
  auto work = [&](int w) {
    std::vector<Sample> p((size_t)n);
    std::vector<int64_t> seen((size_t)distinct, 0);
    int64_t stamp = 0;
    double tp[kIidStats];
    for (int k = w; k < kShuffles; k += workers) {
      std::copy(a.begin(), a.end(), p.begin());
      Stream g{seed + (uint64_t)k * kGolden};
      for (int64_t i = n - 1; i > 0; --i) {
        const int64_t j = (int64_t)g.below((uint64_t)i + 1);
        std::swap(p[(size_t)i], p[(size_t)j]);
      }
      statistics(p.data(), n, mean, seen, stamp, tp);
      for (int s = 0; s < kIidStats; ++s) {
        c[(size_t)w][(size_t)(2 * s)] += tp[s] > t[s] ? 1 : 0;
        c[(size_t)w][(size_t)(2 * s + 1)] += tp[s] == t[s] ? 1 : 0;
      }
    }
  };
  // worker w takes shuffles w, w + workers, ...; a worker whose thread
  // cannot be started runs on this one (the counts are the same either way)
  std::vector<std::thread> th;
  int started = 1;
  try {
    for (int w = 1; w < workers; ++w) {
      th.emplace_back(work, w);
      ++started;
    }
  } catch (...) {
  }
  work(0);
  for (int w = started; w < workers; ++w) {
    work(w);
  }
  for (std::thread &x : th) {
    x.join();
  }
  bool iid = true;
  for (int s = 0; s < kIidStats; ++s) {
    int64_t c0 = 0, c1 = 0;
    for (int w = 0; w < workers; ++w) {
      c0 += c[(size_t)w][(size_t)(2 * s)];
      c1 += c[(size_t)w][(size_t)(2 * s + 1)];
    }
    if (counts) {
      counts[2 * s] = c0;
      counts[2 * s + 1] = c1;
    }
    if (c0 + c1 <= 5 || c0 >= kShuffles - 5) {
      iid = false;
    }
  }
  return iid ? 1 : 0;
}

} // namespace tempi

using namespace tempi;

// (no exception leaves the C ABI: a failed allocation is TEMPI_ERR_NOMEM)
TEMPI_EXPORT int tempi_iid_statistics(const double *samples, int64_t n,
                                      double *stats) {
  if (!stats) {
    return set_error(TEMPI_ERR_ARG, "null statistics array");
  }
  try {
    return iid_statistics(samples, n, stats);
  } catch (...) {
    return set_error(TEMPI_ERR_NOMEM, "IID statistics of %lld samples",
                     (long long)n);
  }
}

TEMPI_EXPORT int tempi_iid_test(const double *samples, int64_t n,
                                uint64_t seed, int64_t *counts) {
  try {
    return iid_test(samples, n, seed, counts);
  } catch (...) {
    return set_error(TEMPI_ERR_NOMEM, "IID test of %lld samples", (long long)n);
  }
}
