// internal.hpp -- declarations shared by the translation units of
// libtempi_amd.so.  Nothing here is part of the public ABI.
#pragma once

#include "../../include/tempi_amd.h"

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <charconv>
#include <cstdint>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#define TEMPI_EXPORT extern "C" __attribute__((visibility("default")))

namespace tempi {

// ---------------------------------------------------------------------------
// logging and errors
// ---------------------------------------------------------------------------
enum { LOG_ERROR = 0, LOG_WARN = 1, LOG_INFO = 2, LOG_DEBUG = 3, LOG_SPEW = 4 };
extern int g_log_level;
void log_message(int level, const char *fmt, ...)
    __attribute__((format(printf, 2, 3)));
#define TEMPI_LOG(lvl, ...)                                                    \
  do {                                                                         \
    if ((lvl) <= ::tempi::g_log_level)                                         \
      ::tempi::log_message((lvl), __VA_ARGS__);                                \
  } while (0)

// record a message for tempi_last_error() and return `code`
int set_error(int code, const char *fmt, ...)
    __attribute__((format(printf, 2, 3)));

// Numbers in the files this library reads and writes (perf.json, model
// weights) are converted with std::from_chars / std::to_chars, which ignore
// the process locale: an application that calls setlocale() with a decimal
// comma neither breaks the files nor has them misread.
template <class T> bool parse_number(const char *b, const char *e, T *out) {
  const std::from_chars_result r = std::from_chars(b, e, *out);
  return r.ec == std::errc() && r.ptr == e;
}
// shortest form that reads back to the same value; "" when it does not fit
template <class T> std::string format_number(T v) {
  char buf[64];
  const std::to_chars_result r = std::to_chars(buf, buf + sizeof buf, v);
  return r.ec == std::errc() ? std::string(buf, r.ptr) : std::string();
}

// TEMPI_DIR_PACK or TEMPI_DIR_UNPACK; any other value is an argument error
// (it would otherwise select a copy direction silently)
inline int check_direction(int dir) {
  return dir == TEMPI_DIR_PACK || dir == TEMPI_DIR_UNPACK
             ? TEMPI_SUCCESS
             : set_error(TEMPI_ERR_ARG, "direction %d", dir);
}

#define TEMPI_HIP_CHECK(expr)                                                  \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (hipSuccess != e_) {                                                    \
      return ::tempi::set_error(TEMPI_ERR_HIP, "%s: %s (%s:%d)", #expr,        \
                                hipGetErrorString(e_), __FILE__, __LINE__);    \
    }                                                                          \
  } while (0)

// ---------------------------------------------------------------------------
// environment (reference: src/internal/env.cpp:23-107, include/env.hpp)
// ---------------------------------------------------------------------------
struct Env {
  bool disable = false;        // TEMPI_DISABLE: pure pass-through
  bool no_pack = false;        // TEMPI_NO_PACK: MPI_Pack/Unpack -> library
  bool no_type_commit = false; // TEMPI_NO_TYPE_COMMIT: no canonicalization
  // TEMPI_DATATYPE_{ONESHOT,DEVICE,STAGED,AUTO}: how non-contiguous device
  // data is sent (TEMPI_METHOD_*; UNKNOWN = choose by the system model)
  int datatype_method = TEMPI_METHOD_UNKNOWN;
  // TEMPI_CONTIGUOUS_{STAGED,AUTO}: contiguous device data (default: library)
  int contiguous_method = TEMPI_METHOD_UNKNOWN;
  bool contiguous_auto = false;
  bool gpu_aware_mpi = false;  // TEMPI_MPI_GPU_AWARE: library takes device pointers
  int select = TEMPI_SELECT_DEFAULT; // TEMPI_SELECT=heuristic|model|autotune
  std::string cache_dir;       // TEMPI_CACHE_DIR (autotune cache, model)
  std::string model_path;      // TEMPI_MODEL (MLP weights file)
  std::string method_model_path; // TEMPI_METHOD_MODEL (method MLP weights)
};
Env &env();
void read_environment();
void ensure_env(); // read_environment() once per process

// ---------------------------------------------------------------------------
// counters (reference: include/counters.hpp:86-193)
// ---------------------------------------------------------------------------
struct Counters {
  std::atomic<int64_t> pack_kernel_launches{0};
  std::atomic<int64_t> unpack_kernel_launches{0};
  std::atomic<int64_t> batch_launches{0};
  std::atomic<int64_t> graph_launches{0};
  std::atomic<int64_t> library_pack_calls{0};
  std::atomic<int64_t> library_unpack_calls{0};
  std::atomic<int64_t> commits_canonical{0};
  std::atomic<int64_t> commits_unsupported{0};
  std::atomic<int64_t> model_inferences{0};
  std::atomic<int64_t> autotune_runs{0};
  std::atomic<int64_t> verified_choices{0};
  std::atomic<int64_t> p2p_device{0};  // sends/receives by method
  std::atomic<int64_t> p2p_oneshot{0};
  std::atomic<int64_t> p2p_staged{0};
  std::atomic<int64_t> p2p_library{0}; // forwarded to the MPI library
  std::atomic<int64_t> trial_launches{0}; // copies timed by the library
  std::atomic<int64_t> method_model_choices{0}; // staging methods the MLP chose
  std::atomic<int64_t> host_staged_calls{0}; // pack/unpack staged for pageable memory
  std::atomic<int64_t> signalled_waits{0}; // copies waited for by their own signal
};
Counters &counters();
// log every counter, in rank order, at LOG_DEBUG (runtime.cpp)
void report_counters();

// While a TrialScope lives on a thread, the copies it launches are the
// library's timing runs (verification, autotuning, system tables): they
// count as trial launches, not as the caller's pack or unpack kernels.
struct TrialScope {
  TrialScope();
  ~TrialScope();
};
bool in_trial();

// ---------------------------------------------------------------------------
// the underlying MPI library (reference: include/symbols.hpp:180-215,
// src/internal/symbols.cpp:22-90)
// ---------------------------------------------------------------------------
struct RealMpi {
  int (*Init)(int *, char ***) = nullptr;
  int (*Init_thread)(int *, char ***, int, int *) = nullptr;
  int (*Finalize)(void) = nullptr;
  int (*Type_commit)(MPI_Datatype *) = nullptr;
  int (*Type_free)(MPI_Datatype *) = nullptr;
  int (*Pack)(const void *, int, MPI_Datatype, void *, int, int *,
              MPI_Comm) = nullptr;
  int (*Unpack)(const void *, int, int *, void *, int, MPI_Datatype,
                MPI_Comm) = nullptr;
  int (*Send)(const void *, int, MPI_Datatype, int, int, MPI_Comm) = nullptr;
  int (*Recv)(void *, int, MPI_Datatype, int, int, MPI_Comm,
              MPI_Status *) = nullptr;
  int (*Isend)(const void *, int, MPI_Datatype, int, int, MPI_Comm,
               MPI_Request *) = nullptr;
  int (*Irecv)(void *, int, MPI_Datatype, int, int, MPI_Comm,
               MPI_Request *) = nullptr;
  int (*Wait)(MPI_Request *, MPI_Status *) = nullptr;
  int (*Waitall)(int, MPI_Request *, MPI_Status *) = nullptr;
  int (*Test)(MPI_Request *, int *, MPI_Status *) = nullptr;
  int (*Neighbor_alltoallw)(const void *, const int *, const MPI_Aint *,
                            const MPI_Datatype *, void *, const int *,
                            const MPI_Aint *, const MPI_Datatype *,
                            MPI_Comm) = nullptr;
  int (*Comm_free)(MPI_Comm *) = nullptr;
};
// resolves on first call; aborts with a message if a symbol is missing
const RealMpi &real_mpi();

// ---------------------------------------------------------------------------
// canonical form (the commit-time lowering)
// ---------------------------------------------------------------------------
struct Canon {
  bool ok = false;
  std::string reason; // why not, when !ok
  int64_t size = 0;   // packed bytes per object
  int64_t lb = 0;
  int64_t extent = 0;
  tempi_strided_block sb{};
};
// decode `dt` with MPI_Type_get_envelope/contents and lower it
Canon canonicalize(MPI_Datatype dt);

// one dimension of a layout: `count` repeats `stride` bytes apart
struct Dim {
  int64_t count;
  int64_t stride;
};
// Order-preserving normalization of a layout whose dims[0] is the byte
// run {bytes, 1}: drop count-1 dimensions and fold a dimension into its inner
// neighbour when it continues exactly where the inner one ends.
void normalize_dims(std::vector<Dim> &dims);

// ---------------------------------------------------------------------------
// device runtime (reference: src/internal/streams.cpp, allocators.cpp)
// ---------------------------------------------------------------------------
bool have_device();           // lazily probes hipGetDeviceCount
int ensure_runtime();         // stream etc.; TEMPI_SUCCESS or error
hipStream_t library_stream(); // valid after ensure_runtime()
int library_device();          // the device of library_stream()
int device_cu_count();
// returns 1 and the device-usable address when the kernels can read/write p
int device_address(const void *p, void **dev);
// true for device memory the host cannot load or store (not host-mapped,
// not managed)
bool device_only(const void *p);
void finalize_runtime();

// ---------------------------------------------------------------------------
// copy geometry and kernels
// ---------------------------------------------------------------------------
constexpr int kMaxRowDims = 8;

// A pack or unpack call as rows of L bytes: row r sits at
// strided_base + sum_k idx_k(r) * stride[k] on the strided side and at
// packed_base + r * L on the packed side; dimension 0 is the fastest.
struct Geom {
  int64_t L = 0;
  int64_t rows = 0;
  int nrd = 0;
  int64_t cnt[kMaxRowDims] = {};
  int64_t str[kMaxRowDims] = {};
  int64_t start = 0; // byte offset of row 0 from the strided buffer origin
};
int make_geom(const tempi_strided_block &sb, int64_t count, Geom *g);

struct VariantDesc {
  int P, W, U;
  int NTS; // nontemporal stores
  int NTL; // nontemporal loads
};
extern const VariantDesc kVariants[];
extern const int kNumVariants;

// true when every row starts W-aligned on both sides and L % W == 0
bool geom_aligned(const Geom &g, const char *strided, const char *packed,
                  int W);

// enqueue the copy; `strided` / `packed` are device addresses of the
// buffer origin (strided) and of the first packed byte.
int launch_copy(const Geom &g, const char *strided, char *packed, int dir,
                int variant, hipStream_t stream, int max_blocks = 0);
// launch_copy on the library stream, then wait until the copy is done (a
// copy of a few workgroups signals its own completion to the host)
int launch_copy_wait(const Geom &g, const char *strided, char *packed, int dir,
                     int variant, int max_blocks = 0);
// launch_copy on the library stream without waiting; a copy that signals
// its completion gets *seq > 0, to test with signal_reached (else 0)
int launch_copy_signalled(const Geom &g, const char *strided, char *packed,
                          int dir, int variant, int max_blocks,
                          unsigned long long *seq);
bool signal_reached(unsigned long long seq);
// until signal `seq` arrives (seq 0: none) or `ev`, recorded after the
// copy, completes; a failed event is an error
int wait_signal_or_event(unsigned long long seq, hipEvent_t ev);

// closed-form variant rule (the "plain heuristic")
int heuristic_variant(const Geom &g, const char *strided, const char *packed,
                      int dir);

// a kernel variant and its grid cap (0: the default cap)
struct Choice {
  int variant = 0;
  int max_blocks = 0;
};

// variant selection entry used by the MPI and tempi_* paths; AUTOTUNE (and
// the default policy for copies of 64 MiB or more) times candidates on
// `stream` (results cached by shape): on the caller's buffers when
// `in_place` (the copy itself runs next, writing the same bytes), else on a
// scratch image of the written side
Choice choose(const Geom &g, const char *strided, char *packed, int dir,
              int method, hipStream_t stream, bool in_place);

// grid caps the autotuner tries, as multiples of the CU count (0 = default)
extern const int kGridCaps[];
extern const int kNumGridCaps;

// average device time of one launch of (variant, max_blocks), over `reps`
// back-to-back launches between two events on `stream`
int time_config(const Geom &g, const char *strided, char *packed, int dir,
                int variant, int max_blocks, int reps, hipStream_t stream,
                float *us);

// the shape and alignment facts a copy's speed depends on (cache key)
std::string shape_key_of(const Geom &g, const char *strided,
                         const char *packed, int dir);

// learned selector (mlp.hip)
void copy_features(const Geom &g, const char *strided, const char *packed,
                   int dir, float *f);
bool model_available();
int model_variant(const Geom &g, const char *strided, const char *packed,
                  int dir, const std::string &key); // -1: no model
void model_precommit(const tempi_strided_block &sb);
// one copy of a set whose variants are about to be chosen
struct CopyShape {
  Geom g;
  const char *strided;
  const char *packed;
  int dir;
};
// evaluates the selector once for every shape of `copies` not yet in its
// cache (copies timed cold are left to their trials), so the per-copy
// choices that follow are cache hits
void model_prefetch(const std::vector<CopyShape> &copies);
// waits for commit-time inferences still in flight (at MPI_Finalize)
void model_finalize();
// inputs of the selector networks (variants and staging methods)
constexpr int kModelInputs = 16;
// the staging-method model (mlp.hip): features of a message of `bytes`
// packed bytes in rows of `block` bytes `stride` apart, and the candidate the
// model prices cheapest (-1 without a model)
void method_features(int64_t bytes, int64_t block, int64_t stride,
                     bool colocated, float *f);
int method_model_pick(const float *f, const int *candidates, int n);
// changes whenever the method model is loaded or unloaded (mlp.hip)
uint64_t method_model_generation();
// the model's k best variants for this copy, best first (not cached);
// returns how many were written (0 without a model)
int model_ranked(const Geom &g, const char *strided, const char *packed,
                 int dir, int k, int *out);
// copies of at least this many packed bytes are timed cold (autotune.hip)
bool times_cold(const Geom &g);
bool times_cold_bytes(int64_t bytes);
// sweep `bytes` of a library-owned buffer so the caches hold nothing of the
// next timed launch (autotune.hip)
int flush_caches_on(int64_t bytes, hipStream_t s);
constexpr int64_t kColdFlushBytesAll = 512ll << 20;
// free the sweep buffer flush_caches_on allocated (autotune.hip)
void flush_release();

// Bytes from the first to the last byte a copy writes: the packed stream
// for a pack, the span of the strided rows for an unpack.
int64_t written_span(const Geom &g, int dir);

// A scratch image of the side a copy writes, for timing runs that must not
// touch the caller's buffers: `strided` / `packed` are the copy's buffers
// with the written one replaced by the image, which has the written span's
// length and its address modulo 256.  The image is freed once `stream` has
// finished (the destructor synchronizes it); it is not kept in the staging
// pool, which would hold on to memory a one-time trial no longer needs.
struct TrialBuffers {
  void *mem = nullptr;
  hipStream_t stream = nullptr;
  const char *strided = nullptr;
  char *packed = nullptr;
  TrialBuffers() = default;
  TrialBuffers(const TrialBuffers &) = delete;
  TrialBuffers &operator=(const TrialBuffers &) = delete;
  TrialBuffers(TrialBuffers &&o) noexcept { *this = std::move(o); }
  TrialBuffers &operator=(TrialBuffers &&o) noexcept;
  ~TrialBuffers();
};
int trial_buffers(const Geom &g, const char *strided, char *packed, int dir,
                  hipStream_t stream, TrialBuffers *t);

// system performance tables (perf.hip): load <cache dir>/perf.json if present
void perf_init();
// changes whenever the measured tables do (perf.hip)
uint64_t perf_generation();
// modelled seconds of a method (perf.hip); +inf when a table is missing
double perf_model_seconds(int method, bool colocated, int64_t bytes,
                          int64_t block_length);

// NIST SP 800-90B section 5.1 permutation test of measured samples
// (iid.cpp): the eight statistics of `n` samples into t[kIidStats], and the
// test itself -- 1 independent, 0 not, a TEMPI_ERR_* code for bad input;
// `counts` (if not null) receives C0 and C1 of each statistic.
constexpr int kIidStats = 8;
constexpr int64_t kIidMaxSamples = 1 << 16;
int iid_statistics(const double *samples, int64_t n, double *t);
int iid_test(const double *samples, int64_t n, uint64_t seed, int64_t *counts);

// the per-type record cache (interpose.cpp); predefined types are lowered
// on their first lookup
bool type_lookup(MPI_Datatype t, Canon *out);
// [lo, hi): the bytes, from the buffer origin, that `count` objects of `dt`
// occupy (true lower bound to true upper bound over the objects)
int type_span(MPI_Datatype dt, int count, int64_t *lo, int64_t *hi);
// A pinned host image of the type-map span of `count` objects of `dt` in
// device memory, for handing typed device data to an MPI library that
// cannot read device memory (interpose.cpp).  take() copies the span in;
// for an image the library will write (a receive, an unpack) it also keeps
// a snapshot, and put_back() writes to the device only the bytes that
// differ from it -- the bytes the library changed.  Bytes between the type
// map's blocks are never written back, so device writes to them made while
// the image was out (a kernel updating a subdomain's interior while its
// ghost cells are received) are kept.
struct SpanImage {
  char *host = nullptr;   // pinned host memory, `len` bytes
  char *before = nullptr; // the span as taken (images the library writes)
  char *dev = nullptr;    // first byte of the span on the device
  int64_t lo = 0;         // span start from the buffer origin
  int64_t len = 0;
  SpanImage() = default;
  SpanImage(const SpanImage &) = delete;
  SpanImage &operator=(const SpanImage &) = delete;
  ~SpanImage();
  // `written`: the library writes into the image (put_back follows)
  int take(const void *dev_buf, MPI_Datatype dt, int count, bool written);
  int put_back();
  // the buffer origin as the MPI library sees it
  void *base() const { return host - lo; }
};

// Range of the calling thread's work for tracing tools (ROCTx), opened on
// construction and closed on destruction (runtime.cpp).
struct TraceRange {
  explicit TraceRange(const char *name);
  ~TraceRange();
  TraceRange(const TraceRange &) = delete;
  TraceRange &operator=(const TraceRange &) = delete;
};
// pack (dir 0) or unpack `count` objects on `stream` with the library's
// variant choice (variant < 0) or the given one; device addresses
int run_copy(const tempi_strided_block &sb, int64_t count, const char *strided,
             char *packed, int dir, int variant, hipStream_t stream);
// run_copy on the library stream with the library's variant, returning when
// the copy is done (launch_copy_wait)
int run_copy_wait(const tempi_strided_block &sb, int64_t count,
                  const char *strided, char *packed, int dir);
// the same without waiting; *seq > 0 when the copy signals its completion
int run_copy_signalled(const tempi_strided_block &sb, int64_t count,
                       const char *strided, char *packed, int dir,
                       unsigned long long *seq);

// cached staging buffers (pool.cpp): device memory, or pinned host memory
// mapped into the device address space
void *pool_acquire(bool host, size_t bytes);
void pool_release(bool host, void *p);
void pool_trim();

// point-to-point state (p2p.cpp)
void p2p_init();     // at MPI_Init: which ranks share a node
void p2p_finalize(); // at MPI_Finalize: completes or reports pending work
// do two ranks of MPI_COMM_WORLD share a node (by the map MPI_Init built)?
bool world_ranks_share_node(int a, int b);
// the staging method forced by tempi_set_datatype_method or TEMPI_DATATYPE_*
// (DEVICE becomes STAGED without a GPU-aware library); UNKNOWN: by model
int forced_datatype_method();
// neighbour exchanges (neighbor.cpp): kept argument sets
void neighbor_finalize();
void neighbor_forget_type(MPI_Datatype t);
void neighbor_forget_comm(MPI_Comm c);

// exhaustive search; fills best and, when given, the best time per variant
int autotune(const Geom &g, const char *strided, char *packed, int dir,
             hipStream_t stream, Choice *best, float *per_variant_us);

} // namespace tempi
