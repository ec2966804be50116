// kernels.hip -- gfx950 strided-block pack/unpack kernels and their host-side
// geometry, alignment and launch code.
//
// Every pack or unpack is a set of rows of L bytes (geometry in
// internal.hpp).  A work unit is one segment of one row; a group of P lanes
// of a 64-wide wavefront moves a segment as U words of W bytes per lane, the
// P lanes of a group touching consecutive words so that each load and store
// instruction of the group covers P*W contiguous bytes.  The stores are
// always W-aligned: when the store side of a row does not start W-aligned,
// the lanes copy the head and tail bytes one at a time and the body words are
// loaded from wherever they sit on the other side (gfx9 global loads accept
// any byte alignment).  When the host has shown that both sides of every row
// are W-aligned and L % W == 0, the ALIGNED instantiation skips that logic.
//
// reference: include/pack_kernels.cuh:19-120 (grid_x_memcpy_aligned,
// pack_2d/unpack_2d) and :350-433 (pack_3d/unpack_3d).  The index map is the
// same -- packed[r*L + x] = strided[start + sum_k idx_k(r)*stride_k + x],
// with the object index as the outermost dimension -- but offsets are 64-bit
// (the reference's `unsigned bi = y*stride1`, :82/:382, wraps past 4 GiB),
// the word width honours the runtime pointers and strides, not only the
// block length and start (:129-157), and the object loop is part of the
// parallel row space instead of a serial loop (:372).
#include "copy_core.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>

namespace tempi {

FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  uint32_t l = 0;
  while (l < 32 && (1ull << l) < d) {
    ++l;
  }
  const unsigned __int128 num =
      ((unsigned __int128)1 << 32) * (unsigned __int128)((1ull << l) - d);
  f.mul = (uint32_t)(num / d + 1);
  f.shift = l;
  return f;
}

// Where signalled copies count their finished workgroups and report to the
// host; set once by the host (signal_ready) before the first signalled
// launch.
struct SignalSlots {
  unsigned long long *counter;   // finished workgroups of the running launch
  unsigned long long *completed; // signalled launches completed so far
  unsigned long long *flag;      // `completed`, in pinned host memory
};
static __device__ SignalSlots g_signal_slots;

// The end of a launch with p.signal set, run by every thread of the
// workgroup: the workgroup's stores are released at system scope (the
// barrier's workgroup-scope fences order the other waves' stores before
// thread 0's release, scopes composing), then counted.  The launch's last
// workgroup resets the counter for the next signalled launch -- launches on
// the library stream never overlap -- and stores the number of completed
// signalled launches to the host's flag.
__device__ __forceinline__ void signal_done(const KParams &p) {
  if (!p.signal) {
    return;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    const SignalSlots s = g_signal_slots;
    const unsigned long long done =
        __hip_atomic_fetch_add(s.counter, 1ull, __ATOMIC_ACQ_REL,
                               __HIP_MEMORY_SCOPE_AGENT) +
        1ull;
    if (done == gridDim.x) {
      __hip_atomic_store(s.counter, 0ull, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      const unsigned long long c =
          __hip_atomic_fetch_add(s.completed, 1ull, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT) +
          1ull;
      __hip_atomic_store(s.flag, c, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

template <int P, int W, int U, bool ALIGNED, int DIR, bool NTS, bool NTL>
__global__ void __launch_bounds__(kBlock) k_strided_copy(KParams p) {
  constexpr uint32_t GROUPS = kBlock / P;
  constexpr uint32_t K = units_in_flight(U);
  const uint32_t lane = threadIdx.x % P;
  const uint32_t step = gridDim.x * GROUPS;
  for (uint32_t u = blockIdx.x * GROUPS + threadIdx.x / P; u < p.units;
       u += step * K) {
    copy_units<P, W, U, ALIGNED, DIR, NTS, NTL>(p, u, step, p.units, lane);
  }
  signal_done(p);
}

// ---------------------------------------------------------------------------
// variant table
// ---------------------------------------------------------------------------

#define TEMPI_VDESC(I, P, W, U, NTS, NTL) {P, W, U, NTS ? 1 : 0, NTL ? 1 : 0},
const VariantDesc kVariants[] = {TEMPI_VARIANTS(TEMPI_VDESC)};
#undef TEMPI_VDESC
const int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

typedef void (*KernelFn)(KParams);

#define TEMPI_KROW(I, P, W, U, NTS, NTL)                                         \
  {{k_strided_copy<P, W, U, false, 0, NTS, NTL>,                               \
    k_strided_copy<P, W, U, false, 1, NTS, NTL>},                              \
   {k_strided_copy<P, W, U, true, 0, NTS, NTL>,                                \
    k_strided_copy<P, W, U, true, 1, NTS, NTL>}},

// [variant][aligned][dir]
static const KernelFn kKernels[][2][2] = {TEMPI_VARIANTS(TEMPI_KROW)};
#undef TEMPI_KROW
static_assert(sizeof(kKernels) / sizeof(kKernels[0]) ==
                  sizeof(kVariants) / sizeof(kVariants[0]),
              "kernel and variant tables differ");

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------

int make_geom(const tempi_strided_block &sb, int64_t count, Geom *g) {
  *g = Geom();
  if (sb.ndims < 1 || sb.ndims > TEMPI_MAX_DIMS) {
    return set_error(TEMPI_ERR_ARG, "strided block with %d dimensions",
                     sb.ndims);
  }
  if (count < 0) {
    return set_error(TEMPI_ERR_ARG, "negative count");
  }
  if (sb.strides[0] != 1 && sb.counts[0] > 1) {
    return set_error(TEMPI_ERR_ARG, "strides[0] must be 1");
  }
  std::vector<Dim> dims;
  dims.push_back(Dim{sb.counts[0], 1});
  for (int d = 1; d < sb.ndims; ++d) {
    dims.push_back(Dim{sb.counts[d], sb.strides[d]});
  }
  dims.push_back(Dim{count, sb.extent});
  for (const Dim &x : dims) {
    if (x.count < 0) {
      return set_error(TEMPI_ERR_ARG, "negative dimension count");
    }
    if (x.count == 0) {
      g->start = sb.start;
      return TEMPI_SUCCESS; // nothing to move
    }
  }
  // the packed size bounds every product of counts formed below
  int64_t packed = 1;
  for (const Dim &x : dims) {
    if (__builtin_mul_overflow(packed, x.count, &packed)) {
      return set_error(TEMPI_ERR_ARG, "packed size overflows 64 bits");
    }
  }
  normalize_dims(dims);
  if ((int)dims.size() - 1 > kMaxRowDims) {
    return set_error(TEMPI_ERR_UNSUPPORTED, "%d row dimensions",
                     (int)dims.size() - 1);
  }
  g->start = sb.start;
  g->L = dims[0].count;
  g->nrd = (int)dims.size() - 1;
  g->rows = 1;
  for (int k = 0; k < g->nrd; ++k) {
    g->cnt[k] = dims[k + 1].count;
    g->str[k] = dims[k + 1].stride;
    g->rows *= g->cnt[k];
  }
  return TEMPI_SUCCESS;
}

bool geom_aligned(const Geom &g, const char *strided, const char *packed,
                  int W) {
  const int A = word_align(W);
  if (g.L % W) {
    return false;
  }
  if ((uintptr_t)packed % A || (uintptr_t)(strided + g.start) % A) {
    return false;
  }
  for (int k = 0; k < g.nrd; ++k) {
    if (g.str[k] % A) {
      return false;
    }
  }
  return true;
}

static int max_aligned_word(const Geom &g, const char *strided,
                            const char *packed) {
  for (int W = 16; W > 1; W /= 2) {
    if (geom_aligned(g, strided, packed, W)) {
      return W;
    }
  }
  return 1;
}

int heuristic_variant(const Geom &g, const char *strided, const char *packed,
                      int dir) {
  (void)dir;
  const int64_t L = g.L;
  const int64_t bytes = L * g.rows;
  // Copies this large are read from HBM: nontemporal loads let their line
  // fills go first from the caches, ahead of the lines being written.
  const bool big = bytes >= (8ll << 20);
  if (L >= 1024) {
    // one long run: nontemporal stores
    if (g.rows == 1) {
      return big ? 2 : 0;
    }
    return big ? 13 : 0;
  }
  if (L >= 512) {
    return big ? 14 : 3;
  }
  if (L >= 256) {
    return big ? 15 : 4;
  }
  if (L >= 128) {
    return big ? 16 : 5;
  }
  if (L >= 64) {
    return big ? 17 : 6;
  }
  const int Wa = max_aligned_word(g, strided, packed);
  if (Wa >= 16) {
    return 7;
  }
  if (Wa == 8) {
    return big ? 18 : 8;
  }
  if (Wa == 4) {
    // rows of 3 or 6 dwords move as 12-B words
    return (L % 12 == 0) ? 12 : 9;
  }
  return L >= 16 ? 7 : 11;
}

int fill_kparams(const Geom &g, const char *strided, char *packed, int dir,
                 int variant, KParams *pp, bool *aligned) {
  const VariantDesc &v = kVariants[variant];
  *aligned = geom_aligned(g, strided, packed, v.W);
  const int64_t words = g.L / v.W;
  const int64_t per_unit = (int64_t)v.P * v.U;
  const int64_t nseg = std::max<int64_t>(1, (words + per_unit - 1) / per_unit);
  const int64_t units = g.rows * nseg;
  if (units >= kMaxUnits || words >= (1ll << 32)) {
    return set_error(TEMPI_ERR_UNSUPPORTED,
                     "copy of %lld rows of %lld B exceeds 2^31 work units",
                     (long long)g.rows, (long long)g.L);
  }
  KParams &p = *pp;
  p = KParams{};
  const char *sbase = strided + g.start;
  if (dir == TEMPI_DIR_PACK) {
    p.src = sbase;
    p.dst = packed;
  } else {
    p.src = packed;
    p.dst = const_cast<char *>(sbase);
  }
  p.L = g.L;
  p.units = (uint32_t)units;
  p.nseg = (uint32_t)nseg;
  p.segdiv = make_fastdiv((uint32_t)nseg);
  p.words = (uint32_t)words;
  p.nrd = g.nrd;
  for (int k = 0; k < g.nrd; ++k) {
    if (g.cnt[k] >= (1ll << 32)) {
      return set_error(TEMPI_ERR_UNSUPPORTED, "dimension count >= 2^32");
    }
    p.cnt[k] = (uint32_t)g.cnt[k];
    p.div[k] = make_fastdiv((uint32_t)g.cnt[k]);
    p.str[k] = g.str[k];
  }
  return TEMPI_SUCCESS;
}

// workgroups of one launch of `units` units
static int64_t grid_of(int variant, uint32_t units, int max_blocks,
                       int cu_count) {
  const int64_t groups = (int64_t)(kBlock / kVariants[variant].P) *
                         units_in_flight(kVariants[variant].U);
  int64_t blocks = ((int64_t)units + groups - 1) / groups;
  const int64_t cap = max_blocks > 0 ? max_blocks : (int64_t)cu_count * 16;
  blocks = std::min<int64_t>(blocks, cap);
  return std::max<int64_t>(blocks, 1);
}

// segments per row of `variant` (the unit count is rows times this)
static int64_t segments_of(const Geom &g, int variant) {
  const VariantDesc &v = kVariants[variant];
  const int64_t per_unit = (int64_t)v.P * v.U;
  return std::max<int64_t>(1, (g.L / v.W + per_unit - 1) / per_unit);
}

// Call piece(geometry, strided, packed) for each launch of a copy: rows
// longer than 1 GiB go as 1 GiB pieces, and a copy of kMaxUnits units or
// more is cut along its outer row dimensions -- a run of whole indices of
// the outermost dimension per launch, or one launch per index when a single
// index already holds too many units.
template <class F>
static int split_units(const Geom &g, const char *strided, char *packed,
                       int64_t nseg, F &piece) {
  if (g.rows * nseg < kMaxUnits || g.nrd == 0) {
    return piece(g, strided, packed);
  }
  const int k = g.nrd - 1;
  const int64_t c = g.cnt[k];
  const int64_t inner = g.rows / c;
  Geom sub = g;
  if (inner * nseg < kMaxUnits) {
    const int64_t m = (kMaxUnits - 1) / (inner * nseg);
    for (int64_t o = 0; o < c; o += m) {
      const int64_t n = std::min(m, c - o);
      sub.cnt[k] = n;
      sub.rows = inner * n;
      sub.start = g.start + o * g.str[k];
      int err = piece(sub, strided, packed + o * inner * g.L);
      if (err) {
        return err;
      }
    }
    return TEMPI_SUCCESS;
  }
  sub.nrd = k;
  sub.rows = inner;
  for (int64_t o = 0; o < c; ++o) {
    sub.start = g.start + o * g.str[k];
    int err = split_units(sub, strided, packed + o * inner * g.L, nseg, piece);
    if (err) {
      return err;
    }
  }
  return TEMPI_SUCCESS;
}

template <class F>
static int for_each_launch(const Geom &g, const char *strided, char *packed,
                           int variant, F &&piece) {
  constexpr int64_t kPiece = 1ll << 30;
  if (g.L <= kPiece) {
    return split_units(g, strided, packed, segments_of(g, variant), piece);
  }
  Geom row = g;
  row.nrd = 0;
  row.rows = 1;
  for (int64_t r = 0; r < g.rows; ++r) {
    // strided offset of row r
    int64_t off = g.start, q = r;
    for (int k = 0; k < g.nrd; ++k) {
      off += (q % g.cnt[k]) * g.str[k];
      q /= g.cnt[k];
    }
    for (int64_t x = 0; x < g.L; x += kPiece) {
      row.L = std::min(kPiece, g.L - x);
      row.start = off + x;
      int err = piece(row, strided, packed + r * g.L + x);
      if (err) {
        return err;
      }
    }
  }
  return TEMPI_SUCCESS;
}

// One launch; `signal`: it signals its completion (the caller holds the
// signal's lock).
static int launch_one(const Geom &g, const char *strided, char *packed,
                      int dir, int variant, hipStream_t stream,
                      int max_blocks, bool signal = false) {
  KParams p;
  bool aligned = false;
  int err = fill_kparams(g, strided, packed, dir, variant, &p, &aligned);
  if (err) {
    return err;
  }
  const int64_t blocks = grid_of(variant, p.units, max_blocks, device_cu_count());
  KernelFn fn = kKernels[variant][aligned ? 1 : 0][dir];
  p.signal = signal ? 1u : 0u;
  void *args[] = {&p};
  TEMPI_HIP_CHECK(hipLaunchKernel(reinterpret_cast<const void *>(fn),
                                  dim3((unsigned)blocks), dim3(kBlock), args,
                                  0, stream));
  if (in_trial()) {
    ++counters().trial_launches;
  } else if (dir == TEMPI_DIR_PACK) {
    ++counters().pack_kernel_launches;
  } else {
    ++counters().unpack_kernel_launches;
  }
  return TEMPI_SUCCESS;
}

namespace {
thread_local int t_trial_depth = 0;
} // namespace

TrialScope::TrialScope() { ++t_trial_depth; }
TrialScope::~TrialScope() { --t_trial_depth; }
bool in_trial() { return t_trial_depth > 0; }

int launch_copy(const Geom &g, const char *strided, char *packed, int dir,
                int variant, hipStream_t stream, int max_blocks) {
  if (g.rows == 0 || g.L == 0) {
    return TEMPI_SUCCESS;
  }
  if (variant < 0 || variant >= kNumVariants) {
    return set_error(TEMPI_ERR_ARG, "variant %d out of range", variant);
  }
  return for_each_launch(g, strided, packed, variant,
                         [&](const Geom &pg, const char *ps, char *pp) {
                           return launch_one(pg, ps, pp, dir, variant, stream,
                                             max_blocks);
                         });
}

// ---------------------------------------------------------------------------
// waiting for a synchronous copy
// ---------------------------------------------------------------------------
//
// MPI_Pack and MPI_Unpack return when their copy is done.  On MI355X an
// empty kernel's own store reaches the host 5.7 us after its launch begins,
// but hipStreamSynchronize returns after 11.1 us: the rest is the runtime's
// end-of-kernel processing.  A copy of one launch of at most
// kSignalMaxBlocks workgroups on the library stream therefore signals its
// own completion (KParams::signal) and the host spins on the flag.  Larger
// grids wait on the stream: the workgroups' releases and counter updates
// queue behind each other, and past about 150 of them they cost more than
// the wait saves (MPI_Pack of 64-B rows: 12.5-13.1 us signalled against
// 13.3-13.7 us at 128 workgroups, 14.9 against 13.7 us at 192).

namespace {

constexpr int64_t kSignalMaxBlocks = 128;

struct SignalState {
  std::mutex mu; // held from a launch's sequence number to its launch
  bool tried = false;
  unsigned long long *slots = nullptr;    // device memory: counter, completed
  unsigned long long *flag = nullptr;     // pinned host memory
  unsigned long long issued = 0;          // signalled launches so far
  // set after a signalled wait fails: the counter may no longer match the
  // launches, so later copies wait on the stream
  std::atomic<bool> off{false};
};

SignalState &signal_state() {
  static SignalState s;
  return s;
}

// Counter, completed count and flag, allocated and published to the kernels
// once, on the library stream's device (the state's lock held); false when
// they cannot be.
bool signal_ready(SignalState &st) {
  if (st.tried) {
    return st.flag != nullptr && !st.off.load(std::memory_order_relaxed);
  }
  st.tried = true;
  // the zeroed counters and the published slots are in place before the
  // first signalled launch: both go through the library stream (which does
  // not wait for the null stream), and it is drained here
  hipStream_t s = library_stream();
  void *c = nullptr, *f = nullptr, *fd = nullptr;
  bool ok = hipSuccess == hipMalloc(&c, 256) &&
            hipSuccess == hipMemsetAsync(c, 0, 256, s) &&
            hipSuccess == hipHostMalloc(&f, 256, hipHostMallocMapped |
                                                     hipHostMallocPortable) &&
            hipSuccess == hipHostGetDevicePointer(&fd, f, 0);
  if (ok) {
    __atomic_store_n(static_cast<unsigned long long *>(f), 0ull, __ATOMIC_RELEASE);
    unsigned long long *cs = static_cast<unsigned long long *>(c);
    const SignalSlots slots{cs, cs + 1, static_cast<unsigned long long *>(fd)};
    ok = hipSuccess == hipMemcpyToSymbolAsync(HIP_SYMBOL(g_signal_slots), &slots,
                                              sizeof slots, 0,
                                              hipMemcpyHostToDevice, s) &&
         hipSuccess == hipStreamSynchronize(s);
  }
  if (!ok) {
    (void)hipGetLastError();
    if (c) {
      (void)hipFree(c);
    }
    if (f) {
      (void)hipHostFree(f);
    }
    TEMPI_LOG(LOG_WARN, "no completion signal: synchronous copies wait on the stream");
    return false;
  }
  st.slots = static_cast<unsigned long long *>(c);
  st.flag = static_cast<unsigned long long *>(f);
  return true;
}

// Spin until the flag counts `seq` completed signalled launches.  A launch
// that never runs to its end leaves the flag short, so after 0.2 ms the
// stream is checked every 0.1 ms: finished without the flag is an error,
// and so is a failed stream.
int wait_signal(const SignalState &st, unsigned long long seq,
                hipStream_t stream) {
  const auto t0 = std::chrono::steady_clock::now();
  double check_at = 2e-4;
  for (;;) {
    if (__atomic_load_n(st.flag, __ATOMIC_ACQUIRE) >= seq) {
      return TEMPI_SUCCESS;
    }
    __builtin_ia32_pause();
    const double el = std::chrono::duration<double>(
                          std::chrono::steady_clock::now() - t0)
                          .count();
    if (el < check_at) {
      continue;
    }
    const hipError_t e = hipStreamQuery(stream);
    if (e == hipSuccess) {
      if (__atomic_load_n(st.flag, __ATOMIC_ACQUIRE) >= seq) {
        return TEMPI_SUCCESS;
      }
      return set_error(TEMPI_ERR_HIP, "a copy ended without its signal");
    }
    if (e != hipErrorNotReady) {
      (void)hipGetLastError();
      return set_error(TEMPI_ERR_HIP, "copy failed: %s", hipGetErrorString(e));
    }
    check_at = el + 1e-4;
  }
}

// one launch of at most kSignalMaxBlocks workgroups, on the library
// stream's device
bool signal_fits(const Geom &g, int variant, int max_blocks) {
  const int64_t units = g.rows * segments_of(g, variant);
  int dev = -1;
  return g.L <= (1ll << 30) && units < kMaxUnits &&
         grid_of(variant, (uint32_t)units, max_blocks, device_cu_count()) <=
             kSignalMaxBlocks &&
         hipSuccess == hipGetDevice(&dev) && dev == library_device();
}

} // namespace

int launch_copy_signalled(const Geom &g, const char *strided, char *packed,
                          int dir, int variant, int max_blocks,
                          unsigned long long *seq) {
  *seq = 0;
  if (g.rows == 0 || g.L == 0) {
    return TEMPI_SUCCESS;
  }
  if (variant < 0 || variant >= kNumVariants) {
    return set_error(TEMPI_ERR_ARG, "variant %d out of range", variant);
  }
  hipStream_t s = library_stream();
  if (signal_fits(g, variant, max_blocks)) {
    SignalState &st = signal_state();
    std::lock_guard<std::mutex> lk(st.mu);
    if (signal_ready(st)) {
      const int err = launch_one(g, strided, packed, dir, variant, s,
                                 max_blocks, /*signal=*/true);
      if (!err) {
        *seq = ++st.issued;
      }
      return err;
    }
  }
  return launch_copy(g, strided, packed, dir, variant, s, max_blocks);
}

bool signal_reached(unsigned long long seq) {
  const SignalState &st = signal_state();
  return seq && st.flag && __atomic_load_n(st.flag, __ATOMIC_ACQUIRE) >= seq;
}

int wait_signal_or_event(unsigned long long seq, hipEvent_t ev) {
  const auto t0 = std::chrono::steady_clock::now();
  double check_at = seq ? 2e-4 : 0.0;
  for (;;) {
    if (signal_reached(seq)) {
      ++counters().signalled_waits;
      return TEMPI_SUCCESS;
    }
    __builtin_ia32_pause();
    const double el = std::chrono::duration<double>(
                          std::chrono::steady_clock::now() - t0)
                          .count();
    if (el < check_at) {
      continue;
    }
    const hipError_t q = seq ? hipEventQuery(ev) : hipEventSynchronize(ev);
    if (q == hipSuccess) {
      return TEMPI_SUCCESS;
    }
    if (q != hipErrorNotReady) {
      (void)hipGetLastError();
      return set_error(TEMPI_ERR_HIP, "copy failed: %s", hipGetErrorString(q));
    }
    check_at = el + 1e-4;
  }
}

int launch_copy_wait(const Geom &g, const char *strided, char *packed,
                     int dir, int variant, int max_blocks) {
  if (g.rows == 0 || g.L == 0) {
    return TEMPI_SUCCESS;
  }
  if (variant < 0 || variant >= kNumVariants) {
    return set_error(TEMPI_ERR_ARG, "variant %d out of range", variant);
  }
  hipStream_t s = library_stream();
  if (signal_fits(g, variant, max_blocks)) {
    SignalState &st = signal_state();
    std::unique_lock<std::mutex> lk(st.mu);
    if (signal_ready(st)) {
      int err = launch_one(g, strided, packed, dir, variant, s, max_blocks,
                           /*signal=*/true);
      const unsigned long long seq = err ? st.issued : ++st.issued;
      lk.unlock();
      if (!err) {
        err = wait_signal(st, seq, s);
        if (err) {
          st.off.store(true);
          TEMPI_LOG(LOG_WARN, "completion signal switched off: %s",
                    tempi_last_error());
        }
      }
      if (err) {
        (void)hipStreamSynchronize(s);
        return err;
      }
      ++counters().signalled_waits;
      return TEMPI_SUCCESS;
    }
  }
  int err = launch_copy(g, strided, packed, dir, variant, s, max_blocks);
  if (hipSuccess != hipStreamSynchronize(s) && !err) {
    err = set_error(TEMPI_ERR_HIP, "stream synchronize failed");
  }
  return err;
}

// ---------------------------------------------------------------------------
// host-side emulation of a copy's memory accesses
// ---------------------------------------------------------------------------
//
// Runs the launches, grid, grid-stride index sequence and unit addresses
// that launch_copy and the kernels would, with the same code for the
// address arithmetic (unit_addr), and records every byte range loaded and
// stored instead of touching memory.  The buffer addresses are numbers
// only; nothing is dereferenced.  It shows on the CPU which bytes a launch
// reads and writes for any variant, alignment and shape.

namespace {

struct Recorder {
  uintptr_t rbase = 0, wbase = 0; // origins of the read and written sides
  int64_t rlo = INT64_MAX, rhi = INT64_MIN, wlo = INT64_MAX, whi = INT64_MIN;
  int64_t nread = 0, nwritten = 0;
  // per-byte store counts over written-side offsets [map_lo, map_lo + size)
  std::vector<uint8_t> map;
  int64_t map_lo = 0;
  bool outside_map = false;
  void read(const char *p, int64_t n) {
    const int64_t o = (int64_t)((uintptr_t)p - rbase);
    rlo = std::min(rlo, o);
    rhi = std::max(rhi, o + n);
    nread += n;
  }
  void write(const char *p, int64_t n) {
    const int64_t o = (int64_t)((uintptr_t)p - wbase);
    wlo = std::min(wlo, o);
    whi = std::max(whi, o + n);
    nwritten += n;
    if (!map.empty()) {
      for (int64_t i = o; i < o + n; ++i) {
        const int64_t j = i - map_lo;
        if (j < 0 || j >= (int64_t)map.size()) {
          outside_map = true;
        } else if (map[(size_t)j] < 255) {
          ++map[(size_t)j];
        }
      }
    }
  }
};

thread_local Recorder *t_rec = nullptr;

struct HostMem {
  static void copy_short(char *d, const char *s, uint32_t n) {
    t_rec->read(s, n);
    t_rec->write(d, n);
  }
  static void copy_byte(char *d, const char *s) {
    t_rec->read(s, 1);
    t_rec->write(d, 1);
  }
};

// one work unit of one group, all P lanes: the kernels' address phase, then
// the body words of the load and store phases
template <int P, int W, int U, bool ALIGNED, int DIR>
void emulate_unit(const KParams &p, uint32_t u) {
  for (uint32_t lane = 0; lane < (uint32_t)P; ++lane) {
    const UnitAddr a = unit_addr<HostMem, P, W, U, ALIGNED, DIR>(p, u, lane);
    for (int j = 0; j < U; ++j) {
      const uint32_t w = a.w + (uint32_t)(j * P);
      if (w < a.n) {
        t_rec->read(a.s + (size_t)w * W, W);
        t_rec->write(a.d + (size_t)w * W, W);
      }
    }
  }
}

typedef void (*EmulateFn)(const KParams &, uint32_t);
#define TEMPI_EROW(I, P, W, U, NTS, NTL)                                         \
  {{emulate_unit<P, W, U, false, 0>, emulate_unit<P, W, U, false, 1>},          \
   {emulate_unit<P, W, U, true, 0>, emulate_unit<P, W, U, true, 1>}},
const EmulateFn kEmulate[][2][2] = {TEMPI_VARIANTS(TEMPI_EROW)};
#undef TEMPI_EROW

} // namespace

} // namespace tempi

using namespace tempi;

// Emulates launch_copy of the copy on the host (see above); `unit_limit`
// bounds the work units emulated in full (beyond it only the launch plan is
// reported).
TEMPI_EXPORT int tempi_copy_emulate(const tempi_strided_block *sb,
                                    int64_t incount, uint64_t strided_addr,
                                    uint64_t packed_addr, int direction,
                                    int variant, int max_blocks, int cu_count,
                                    int64_t unit_limit,
                                    tempi_copy_emulation *out) {
  if (!sb || !out || cu_count <= 0 || max_blocks < 0) {
    return set_error(TEMPI_ERR_ARG, "bad argument");
  }
  int err = check_direction(direction);
  if (err) {
    return err;
  }
  if (variant < 0 || variant >= kNumVariants) {
    return set_error(TEMPI_ERR_ARG, "variant %d out of range", variant);
  }
  Geom g;
  if ((err = make_geom(*sb, incount, &g))) {
    return err;
  }
  *out = tempi_copy_emulation{};
  const char *strided = reinterpret_cast<const char *>((uintptr_t)strided_addr);
  char *packed = reinterpret_cast<char *>((uintptr_t)packed_addr);
  if (g.rows == 0 || g.L == 0) {
    out->exact = 1;
    return TEMPI_SUCCESS;
  }
  // plan: launches and units
  int64_t units_total = 0;
  err = for_each_launch(g, strided, packed, variant,
                        [&](const Geom &pg, const char *, char *) {
                          const int64_t u = pg.rows * segments_of(pg, variant);
                          ++out->launches;
                          out->max_units = std::max(out->max_units, u);
                          units_total += u;
                          return TEMPI_SUCCESS;
                        });
  if (err) {
    return err;
  }
  if (units_total > unit_limit) {
    out->emulated = 0;
    return TEMPI_SUCCESS;
  }
  Recorder rec;
  const bool pack = direction == TEMPI_DIR_PACK;
  rec.rbase = pack ? (uintptr_t)strided : (uintptr_t)packed;
  rec.wbase = pack ? (uintptr_t)packed : (uintptr_t)strided;
  // the bytes the copy must write, each once: the packed stream, or the
  // strided rows (a map of the written side, up to 256 MiB)
  int64_t lo = 0, hi = g.L;
  for (int k = 0; k < g.nrd; ++k) {
    const int64_t reach = (g.cnt[k] - 1) * g.str[k];
    (reach < 0 ? lo : hi) += reach;
  }
  const int64_t map_lo = pack ? 0 : g.start + lo;
  const int64_t map_len = pack ? g.rows * g.L : hi - lo;
  std::vector<uint8_t> want;
  if (map_len <= (256ll << 20)) {
    rec.map.assign((size_t)map_len, 0);
    rec.map_lo = map_lo;
    want.assign((size_t)map_len, 0);
    for (int64_t r = 0; r < g.rows; ++r) {
      int64_t off = pack ? r * g.L : g.start, q = r;
      for (int k = 0; !pack && k < g.nrd; ++k) {
        off += (q % g.cnt[k]) * g.str[k];
        q /= g.cnt[k];
      }
      for (int64_t x = 0; x < g.L; ++x) {
        uint8_t &c = want[(size_t)(off + x - map_lo)];
        c = (uint8_t)std::min(255, c + 1);
      }
    }
  }
  t_rec = &rec;
  err = for_each_launch(
      g, strided, packed, variant,
      [&](const Geom &pg, const char *ps, char *pp) -> int {
        KParams p;
        bool aligned = false;
        int e = fill_kparams(pg, ps, pp, direction, variant, &p, &aligned);
        if (e) {
          return e;
        }
        const VariantDesc &v = kVariants[variant];
        const EmulateFn fn = kEmulate[variant][aligned ? 1 : 0][direction];
        const uint32_t GROUPS = (uint32_t)(kBlock / v.P);
        const uint32_t K = (uint32_t)units_in_flight(v.U);
        const uint32_t blocks =
            (uint32_t)grid_of(variant, p.units, max_blocks, cu_count);
        const uint32_t step = blocks * GROUPS;
        // the kernels' index sequence, in their 32-bit arithmetic
        for (uint32_t b = 0; b < blocks; ++b) {
          for (uint32_t grp = 0; grp < GROUPS; ++grp) {
            for (uint32_t u = b * GROUPS + grp; u < p.units;) {
              for (uint32_t k = 0; k < K; ++k) {
                const uint32_t uu = u + k * step;
                if ((uint64_t)u + (uint64_t)k * step != (uint64_t)uu) {
                  out->wrapped = 1;
                } else if (uu < p.units) {
                  fn(p, uu);
                }
              }
              const uint64_t next = (uint64_t)u + (uint64_t)step * K;
              if (next >= (1ull << 32)) {
                out->wrapped = 1;
                break;
              }
              u = (uint32_t)next;
            }
          }
        }
        return TEMPI_SUCCESS;
      });
  t_rec = nullptr;
  if (err) {
    return err;
  }
  out->emulated = 1;
  out->read_lo = rec.nread ? rec.rlo : 0;
  out->read_hi = rec.nread ? rec.rhi : 0;
  out->write_lo = rec.nwritten ? rec.wlo : 0;
  out->write_hi = rec.nwritten ? rec.whi : 0;
  out->bytes_read = rec.nread;
  out->bytes_written = rec.nwritten;
  out->exact = rec.map.empty() ? -1 : (!rec.outside_map && rec.map == want);
  return TEMPI_SUCCESS;
}

namespace tempi {

// exported for the host-side unit test of the divisor arithmetic
extern "C" __attribute__((visibility("default"))) uint32_t
tempi_fastdiv_check(uint32_t n, uint32_t d) {
  return make_fastdiv(d).div(n);
}

} // namespace tempi
