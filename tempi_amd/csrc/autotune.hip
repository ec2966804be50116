// autotune.hip -- exhaustive timing of kernel variants and grid caps, the
// per-shape choice cache, and the variant selection policy.
//
// The reference picks its kernel with a fixed rule at commit time
// (include/pack_kernels.cuh:129-166, src/internal/packer_3d.cu:31-69) and
// only measures whole methods offline (src/internal/measure_system.cu:
// 514-609, the 9x9 pack tables).  Here every (variant, grid cap) pair is
// timed on the caller's own buffers with HIP events, the winner is cached
// per shape, and the timings are what the learned selector is trained on.
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <unordered_map>

namespace tempi {

// multiples of the CU count; 0 = the launch code's default (16 per CU)
const int kGridCaps[] = {0, 2, 4, 8, 32};

// Copies of at least kColdBytes packed bytes are timed one launch at a time,
// each after a read-only sweep of kColdFlushBytes (twice the 256 MB
// Infinity Cache) empties the caches: data that large cannot stay cached
// between real calls, and back-to-back repeats would rank the variants by
// how they cope with the previous repeat's dirty lines instead.
constexpr int64_t kColdBytes = 64ll << 20;
constexpr int64_t kColdFlushBytes = kColdFlushBytesAll;
const int kNumGridCaps = sizeof(kGridCaps) / sizeof(kGridCaps[0]);

// Reads one dword per 64 B of `buf` and writes one word per workgroup: the
// caches end up full of clean lines of `buf`, so a kernel timed after it
// starts from HBM without write-backs of someone else's dirty data.
__global__ void __launch_bounds__(256) k_touch(const uint32_t *__restrict__ buf,
                                               uint64_t words, uint32_t *sink) {
  uint32_t acc = 0;
  for (uint64_t i = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * 16; i < words;
       i += (uint64_t)gridDim.x * 256 * 16) {
    acc ^= buf[i];
  }
  __shared__ uint32_t s;
  if (threadIdx.x == 0) {
    s = 0;
  }
  __syncthreads();
  atomicXor(&s, acc);
  __syncthreads();
  if (threadIdx.x == 0) {
    sink[blockIdx.x] = s;
  }
}

struct FlushBuf {
  std::mutex mu;
  char *buf = nullptr;
  int64_t bytes = 0;
};
FlushBuf &flush_buf() {
  static FlushBuf f;
  return f;
}

// Release the sweep buffer (after the stream's work is done).
void flush_release() {
  FlushBuf &f = flush_buf();
  std::lock_guard<std::mutex> lk(f.mu);
  if (f.buf) {
    (void)hipDeviceSynchronize();
    (void)hipFree(f.buf);
    f.buf = nullptr;
    f.bytes = 0;
  }
}

// Sweep `bytes` (<= 0: 1 GiB) of a library-owned buffer with k_touch on
// `stream`; the buffer is allocated and filled on first use.
int flush_caches_on(int64_t bytes, hipStream_t s) {
  if (bytes <= 0) {
    bytes = 1ll << 30;
  }
  FlushBuf &f = flush_buf();
  std::lock_guard<std::mutex> lk(f.mu);
  const int blocks = 2048;
  if (f.bytes < bytes) {
    if (f.buf) {
      TEMPI_HIP_CHECK(hipStreamSynchronize(s));
      TEMPI_HIP_CHECK(hipFree(f.buf));
      f.buf = nullptr;
      f.bytes = 0;
    }
    TEMPI_HIP_CHECK(hipMalloc(&f.buf, (size_t)bytes + blocks * 4));
    // filled on the sweep's own stream: a fill on the null stream is not
    // ordered with this (non-blocking) stream and ran on under the first
    // sweep and the copy timed after it (that copy measured 40-70 % slow)
    TEMPI_HIP_CHECK(hipMemsetAsync(f.buf, 1, (size_t)bytes, s));
    f.bytes = bytes;
  }
  hipLaunchKernelGGL(k_touch, dim3(blocks), dim3(256), 0, s,
                     reinterpret_cast<const uint32_t *>(f.buf),
                     (uint64_t)bytes / 4,
                     reinterpret_cast<uint32_t *>(f.buf + f.bytes));
  TEMPI_HIP_CHECK(hipGetLastError());
  return TEMPI_SUCCESS;
}

namespace {
std::mutex g_tune_mu;
std::unordered_map<std::string, Choice> g_tune_cache;
std::unordered_map<std::string, Choice> g_verify_cache; // default policy, large copies
} // namespace

// The key is the raw bytes of the facts (direction, row length, row
// dimensions, the alignments modulo 16 of both sides), not text: it is
// built for every copy a caller makes and never stored, and formatting it
// with a stream took about ten times as long as the lookup it served.
std::string shape_key_of(const Geom &g, const char *strided,
                         const char *packed, int dir) {
  int64_t w[4 + 2 * kMaxRowDims];
  int n = 0;
  w[n++] = dir;
  w[n++] = g.L;
  w[n++] = g.nrd;
  for (int k = 0; k < g.nrd; ++k) {
    w[n++] = g.cnt[k];
    w[n++] = g.str[k];
  }
  w[n++] = (int64_t)(((uintptr_t)(strided + g.start) & 15) |
                     (((uintptr_t)packed & 15) << 8));
  return std::string(reinterpret_cast<const char *>(w),
                     (size_t)n * sizeof(int64_t));
}

bool times_cold(const Geom &g) { return g.L * g.rows >= kColdBytes; }
bool times_cold_bytes(int64_t bytes) { return bytes >= kColdBytes; }

int time_config(const Geom &g, const char *strided, char *packed, int dir,
                int variant, int max_blocks, int reps, hipStream_t stream,
                float *us) {
  TrialScope trial;
  const bool cold = times_cold(g);
  hipEvent_t a = nullptr, b = nullptr;
  TEMPI_HIP_CHECK(hipEventCreate(&a));
  TEMPI_HIP_CHECK(hipEventCreate(&b));
  int err = launch_copy(g, strided, packed, dir, variant, stream, max_blocks);
  float total_ms = 0;
  if (cold) {
    // one launch at a time, each after the caches are emptied
    for (int r = 0; r < std::max(reps, 1) && !err; ++r) {
      err = flush_caches_on(kColdFlushBytes, stream);
      if (!err) {
        (void)hipEventRecord(a, stream);
        err = launch_copy(g, strided, packed, dir, variant, stream, max_blocks);
        (void)hipEventRecord(b, stream);
      }
      float ms = 0;
      if (!err && (hipSuccess != hipEventSynchronize(b) ||
                   hipSuccess != hipEventElapsedTime(&ms, a, b))) {
        err = set_error(TEMPI_ERR_HIP, "event timing failed");
      }
      total_ms += ms;
    }
  } else if (!err) {
    (void)hipEventRecord(a, stream);
    for (int r = 0; r < reps && !err; ++r) {
      err = launch_copy(g, strided, packed, dir, variant, stream, max_blocks);
    }
    (void)hipEventRecord(b, stream);
    if (!err && (hipSuccess != hipEventSynchronize(b) ||
                 hipSuccess != hipEventElapsedTime(&total_ms, a, b))) {
      err = set_error(TEMPI_ERR_HIP, "event timing failed");
    }
  }
  if (!err) {
    *us = total_ms * 1e3f / (float)std::max(reps, 1);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return err;
}

int autotune(const Geom &g, const char *strided, char *packed, int dir,
             hipStream_t stream, Choice *best, float *per_variant_us) {
  ++counters().autotune_runs;
  // enough launches to see past timer resolution, fewer for big copies
  const int64_t bytes = g.L * g.rows;
  const int reps = bytes >= (64ll << 20) ? 3 : (bytes >= (4ll << 20) ? 5 : 10);
  float best_us = 1e30f;
  for (int v = 0; v < kNumVariants; ++v) {
    float vbest = 1e30f;
    for (int c = 0; c < kNumGridCaps; ++c) {
      const int cap = kGridCaps[c] * device_cu_count();
      float us = 0;
      int err = time_config(g, strided, packed, dir, v, cap, reps, stream, &us);
      if (err) {
        return err;
      }
      if (us < vbest) {
        vbest = us;
      }
      if (us < best_us) {
        best_us = us;
        best->variant = v;
        best->max_blocks = cap;
      }
    }
    if (per_variant_us) {
      per_variant_us[v] = vbest;
    }
  }
  return TEMPI_SUCCESS;
}

// Where trial copies of the library's own choices (verification and the
// autotune policy) write.  A copy that runs right after its choice (MPI_Pack,
// MPI_Unpack, tempi_pack_sb, tempi_unpack_sb) is timed on the caller's own
// buffers: each trial stores exactly the bytes the real copy stores next,
// from the same source, so nothing extra is allocated.  A batch resolves
// its variants when it is created, long before it runs -- an unpack trial
// there would overwrite the caller's strided buffer with whatever the
// packed buffer held then -- so its trials write a scratch image of the
// destination from the staging pool instead, at most kMaxTrialScratch
// bytes.  The image has the written side's length and its address modulo
// 256, so the trials see the same alignment (and the same row, line and
// page pattern) as the real copy; the read side is the caller's, untouched.
TrialBuffers &TrialBuffers::operator=(TrialBuffers &&o) noexcept {
  if (this != &o) {
    this->~TrialBuffers();
    mem = o.mem;
    stream = o.stream;
    strided = o.strided;
    packed = o.packed;
    o.mem = nullptr;
  }
  return *this;
}

TrialBuffers::~TrialBuffers() {
  if (mem) {
    (void)hipStreamSynchronize(stream);
    pool_release(false, mem);
    mem = nullptr;
  }
}

int64_t written_span(const Geom &g, int dir) {
  if (dir != TEMPI_DIR_UNPACK) {
    return g.rows * g.L;
  }
  int64_t lo = 0, hi = g.L;
  for (int k = 0; k < g.nrd; ++k) {
    const int64_t reach = (g.cnt[k] - 1) * g.str[k];
    (reach < 0 ? lo : hi) += reach;
  }
  return hi - lo;
}

// largest scratch image a deferred trial may take from the pool (a
// 256 MiB packed stream and its alignment slack)
constexpr size_t kMaxTrialScratch = (256ull << 20) + 256;

int trial_buffers(const Geom &g, const char *strided, char *packed, int dir,
                  hipStream_t stream, TrialBuffers *t) {
  // bytes the copy writes: the packed stream, or the strided rows' span
  int64_t lo = 0, hi = g.L;
  const char *first = packed;
  if (dir == TEMPI_DIR_UNPACK) {
    for (int k = 0; k < g.nrd; ++k) {
      const int64_t reach = (g.cnt[k] - 1) * g.str[k];
      (reach < 0 ? lo : hi) += reach;
    }
    first = strided + g.start + lo;
  } else {
    hi = g.rows * g.L;
  }
  const size_t bytes = (size_t)(hi - lo) + 256;
  t->stream = stream;
  if (bytes > kMaxTrialScratch) {
    // trials of a copy this large are not worth the memory; it keeps the
    // model's pick
    return set_error(TEMPI_ERR_NOMEM,
                     "trial copies would need %zu B of scratch", bytes);
  }
  t->mem = pool_acquire(false, bytes);
  if (!t->mem) {
    return set_error(TEMPI_ERR_NOMEM, "no scratch for trial copies (%zu B)",
                     bytes);
  }
  char *base = static_cast<char *>(t->mem);
  char *image =
      base + (((uintptr_t)first - (uintptr_t)base) & (uintptr_t)255);
  if (dir == TEMPI_DIR_UNPACK) {
    t->strided = image - g.start - lo;
    t->packed = packed;
  } else {
    t->strided = strided;
    t->packed = image;
  }
  return TEMPI_SUCCESS;
}

// The default policy for copies timed cold (see kColdBytes): the model's
// three best variants and the closed-form rule's pick, each timed once on
// the copy, the fastest kept for the shape.  About a millisecond once per
// shape, against 10-20 % that a wrong pick costs each call of a copy this
// size.  Smaller copies keep the model's pick: a warm trial times the copy
// repeated on its own, and which variant wins there is not which wins where
// the copy runs -- a variant with nontemporal stores that times best alone
// sends the packed bytes out of the caches just before the transfer or
// unpack that reads them.

static int verify_candidates(const Geom &g, const char *strided, char *packed,
                             int dir, int rule, hipStream_t stream,
                             bool in_place, Choice *out) {
  int cand[4];
  int n = model_ranked(g, strided, packed, dir, 3, cand);
  if (n == 0) {
    return TEMPI_ERR_UNSUPPORTED;
  }
  if (std::find(cand, cand + n, rule) == cand + n) {
    cand[n++] = rule;
  }
  TrialBuffers t;
  const char *ts = strided;
  char *tp = packed;
  if (!in_place) {
    int err = trial_buffers(g, strided, packed, dir, stream, &t);
    if (err) {
      return err;
    }
    ts = t.strided;
    tp = t.packed;
  }
  float best_us = 1e30f;
  int err = TEMPI_SUCCESS;
  for (int i = 0; i < n && !err; ++i) {
    float us = 0;
    err = time_config(g, ts, tp, dir, cand[i], 0, 1, stream, &us);
    if (!err && us < best_us) {
      best_us = us;
      out->variant = cand[i];
      out->max_blocks = 0;
    }
  }
  // the cold trials' cache sweep buffer is not kept for later calls
  flush_release();
  if (err) {
    return err;
  }
  ++counters().verified_choices;
  return TEMPI_SUCCESS;
}

Choice choose(const Geom &g, const char *strided, char *packed, int dir,
              int method, hipStream_t stream, bool in_place) {
  Choice c;
  c.variant = heuristic_variant(g, strided, packed, dir);
  if (method == TEMPI_SELECT_MODEL ||
      (method == TEMPI_SELECT_DEFAULT && model_available())) {
    const std::string key = shape_key_of(g, strided, packed, dir);
    if (method == TEMPI_SELECT_DEFAULT && times_cold(g)) {
      {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_verify_cache.find(key);
        if (it != g_verify_cache.end()) {
          return it->second;
        }
      }
      Choice v;
      if (stream && !verify_candidates(g, strided, packed, dir, c.variant,
                                       stream, in_place, &v)) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_verify_cache[key] = v;
        return v;
      }
    }
    const int v = model_variant(g, strided, packed, dir, key);
    if (v >= 0) {
      c.variant = v;
    }
    return c;
  }
  if (method != TEMPI_SELECT_AUTOTUNE || !stream) {
    return c;
  }
  const std::string key = shape_key_of(g, strided, packed, dir);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      return it->second;
    }
  }
  Choice best = c;
  TrialBuffers t;
  if ((!in_place && trial_buffers(g, strided, packed, dir, stream, &t)) ||
      autotune(g, in_place ? strided : t.strided,
               in_place ? packed : t.packed, dir, stream, &best, nullptr) ||
      hipSuccess != hipStreamSynchronize(stream)) {
    TEMPI_LOG(LOG_WARN, "autotune failed (%s); using the closed-form rule",
              tempi_last_error());
    return c;
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = best;
  return best;
}

__global__ void k_empty(int *p) {
  if (p) {
    *p = threadIdx.x;
  }
}

} // namespace tempi

using namespace tempi;

static int resolve_buffers(const tempi_strided_block *sb, int64_t incount,
                           const void *strided, const void *packed,
                           int direction, Geom *g, void **ds, void **dp) {
  if (!sb) {
    return set_error(TEMPI_ERR_ARG, "null strided block");
  }
  int err = check_direction(direction);
  if (err) {
    return err;
  }
  err = make_geom(*sb, incount, g);
  if (err) {
    return err;
  }
  err = ensure_runtime();
  if (err) {
    return err;
  }
  if (!device_address(strided, ds) || !device_address(packed, dp)) {
    return set_error(TEMPI_ERR_ARG, "buffers must be device-accessible");
  }
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_num_grid_caps(void) { return kNumGridCaps; }

TEMPI_EXPORT int tempi_time_variant(const tempi_strided_block *sb,
                                    int64_t incount, const void *strided,
                                    void *packed, int direction, int variant,
                                    int grid_cap, int reps, float *us) {
  Geom g;
  void *ds = nullptr, *dp = nullptr;
  int err = resolve_buffers(sb, incount, strided, packed, direction, &g, &ds,
                            &dp);
  if (err) {
    return err;
  }
  if (grid_cap < 0 || grid_cap >= kNumGridCaps) {
    return set_error(TEMPI_ERR_ARG, "grid cap index %d", grid_cap);
  }
  return time_config(g, static_cast<const char *>(ds), static_cast<char *>(dp),
                     direction, variant, kGridCaps[grid_cap] * device_cu_count(),
                     std::max(reps, 1), library_stream(), us);
}

TEMPI_EXPORT int tempi_autotune(const tempi_strided_block *sb, int64_t incount,
                                const void *strided, void *packed,
                                int direction, int *best_variant,
                                int *best_grid_cap, float *per_variant_us) {
  Geom g;
  void *ds = nullptr, *dp = nullptr;
  int err = resolve_buffers(sb, incount, strided, packed, direction, &g, &ds,
                            &dp);
  if (err) {
    return err;
  }
  Choice best;
  err = autotune(g, static_cast<const char *>(ds), static_cast<char *>(dp),
                 direction, library_stream(), &best, per_variant_us);
  if (err) {
    return err;
  }
  if (best_variant) {
    *best_variant = best.variant;
  }
  if (best_grid_cap) {
    *best_grid_cap = 0;
    for (int c = 0; c < kNumGridCaps; ++c) {
      if (kGridCaps[c] * device_cu_count() == best.max_blocks) {
        *best_grid_cap = c;
      }
    }
  }
  return TEMPI_SUCCESS;
}

// Evict L2 and the Infinity Cache by reading `bytes` (default 1 GiB) of a
// library-owned buffer; synchronous.
TEMPI_EXPORT int tempi_flush_caches(int64_t bytes) {
  int err = ensure_runtime();
  if (err) {
    return err;
  }
  hipStream_t s = library_stream();
  if ((err = flush_caches_on(bytes, s))) {
    return err;
  }
  TEMPI_HIP_CHECK(hipStreamSynchronize(s));
  return TEMPI_SUCCESS;
}

// reference: src/internal/measure_system.cu:36-64 (empty kernel <<<1,1>>>)
TEMPI_EXPORT int tempi_launch_latency_us(int reps, float *host_us,
                                         float *device_us) {
  int err = ensure_runtime();
  if (err) {
    return err;
  }
  reps = std::max(reps, 1);
  hipStream_t s = library_stream();
  hipEvent_t a, b;
  TEMPI_HIP_CHECK(hipEventCreate(&a));
  TEMPI_HIP_CHECK(hipEventCreate(&b));
  hipLaunchKernelGGL(k_empty, dim3(1), dim3(1), 0, s, nullptr);
  TEMPI_HIP_CHECK(hipStreamSynchronize(s));
  TEMPI_HIP_CHECK(hipEventRecord(a, s));
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < reps; ++i) {
    hipLaunchKernelGGL(k_empty, dim3(1), dim3(1), 0, s, nullptr);
  }
  auto t1 = std::chrono::steady_clock::now();
  TEMPI_HIP_CHECK(hipEventRecord(b, s));
  TEMPI_HIP_CHECK(hipEventSynchronize(b));
  float ms = 0;
  TEMPI_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
  if (host_us) {
    *host_us =
        (float)std::chrono::duration<double, std::micro>(t1 - t0).count() /
        reps;
  }
  if (device_us) {
    *device_us = ms * 1e3f / reps;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return TEMPI_SUCCESS;
}
