// runtime.cpp -- device probe, the library stream, pointer classification,
// allocators and events.
//
// reference: src/internal/streams.cpp:16-42 (kernStream, non-blocking),
// include/allocator_device.hpp / allocator_host.hpp (device memory and
// mapped pinned host memory), src/pack.cpp:43-55 (device-accessibility test
// with cudaPointerGetAttributes).  Everything is created lazily so that the
// library loads, and forwards to MPI, on hosts without a GPU.
#include "internal.hpp"

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <atomic>

namespace tempi {

// ROCTx ranges (reference: the NVTX ranges around MPI_Pack and MPI_Unpack,
// src/pack.cpp:37,66 and src/unpack.cpp:29,58).  Without a tracing tool
// attached a push or pop returns at once.
TraceRange::TraceRange(const char *name) { roctxRangePushA(name); }
TraceRange::~TraceRange() { roctxRangePop(); }

// The counters, one line each, in rank order at MPI_Finalize when the log
// level is debug or higher (reference: counters::finalize,
// src/internal/counters.cpp:30-120, which prints at LOG_DEBUG).
void report_counters() {
  if (g_log_level < LOG_DEBUG) {
    return;
  }
  tempi_counters c{};
  (void)tempi_counters_get(&c);
  int rank = 0, size = 1, init = 0;
  if (MPI_SUCCESS == MPI_Initialized(&init) && init) {
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
  }
  const std::pair<const char *, int64_t> rows[] = {
      {"pack_kernel_launches", c.pack_kernel_launches},
      {"unpack_kernel_launches", c.unpack_kernel_launches},
      {"batch_launches", c.batch_launches},
      {"graph_launches", c.graph_launches},
      {"library_pack_calls", c.library_pack_calls},
      {"library_unpack_calls", c.library_unpack_calls},
      {"commits_canonical", c.commits_canonical},
      {"commits_unsupported", c.commits_unsupported},
      {"model_inferences", c.model_inferences},
      {"autotune_runs", c.autotune_runs},
      {"verified_choices", c.verified_choices},
      {"p2p_device", c.p2p_device},
      {"p2p_oneshot", c.p2p_oneshot},
      {"p2p_staged", c.p2p_staged},
      {"p2p_library", c.p2p_library},
      {"trial_launches", c.trial_launches},
      {"method_model_choices", c.method_model_choices},
      {"host_staged_calls", c.host_staged_calls},
      {"signalled_waits", c.signalled_waits},
  };
  for (int r = 0; r < size; ++r) {
    if (r == rank) {
      for (const auto &row : rows) {
        TEMPI_LOG(LOG_DEBUG, "rank %d counters::%s: %lld", rank, row.first,
                  (long long)row.second);
      }
    }
    if (size > 1) {
      MPI_Barrier(MPI_COMM_WORLD);
    }
  }
}

namespace {
struct Runtime {
  std::once_flag probe_once;
  int device_count = 0;
  std::mutex mu;
  std::atomic<bool> ready{false};
  hipStream_t stream = nullptr;
  int device = 0;
  int cus = 0;
};
Runtime &rt() {
  static Runtime r;
  return r;
}
} // namespace

bool have_device() {
  Runtime &r = rt();
  std::call_once(r.probe_once, [&r] {
    int n = 0;
    if (hipSuccess != hipGetDeviceCount(&n)) {
      n = 0;
      (void)hipGetLastError();
    }
    r.device_count = n;
  });
  return r.device_count > 0;
}

int ensure_runtime() {
  Runtime &r = rt();
  if (r.ready.load(std::memory_order_acquire)) {
    return TEMPI_SUCCESS;
  }
  if (!have_device()) {
    return set_error(TEMPI_ERR_NODEVICE, "no GPU visible to the HIP runtime");
  }
  std::lock_guard<std::mutex> lk(r.mu);
  if (r.ready) {
    return TEMPI_SUCCESS;
  }
  int dev = 0;
  TEMPI_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  TEMPI_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  r.cus = prop.multiProcessorCount;
  r.device = dev;
  // default priority, non-blocking: no implicit ordering with the caller's
  // streams, as the reference's kernStream
  TEMPI_HIP_CHECK(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking));
  TEMPI_LOG(LOG_INFO, "device %d (%s), %d CUs", dev, prop.gcnArchName, r.cus);
  r.ready.store(true, std::memory_order_release);
  return TEMPI_SUCCESS;
}

hipStream_t library_stream() { return rt().stream; }

int library_device() { return rt().device; }

int device_cu_count() { return rt().cus > 0 ? rt().cus : 256; }

int device_address(const void *p, void **dev) {
  *dev = nullptr;
  if (!p || !have_device()) {
    return 0;
  }
  hipPointerAttribute_t a{};
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError(); // pageable host memory: not an error for us
    return 0;
  }
  switch (a.type) {
  case hipMemoryTypeDevice:
  case hipMemoryTypeManaged:
  case hipMemoryTypeUnified:
    *dev = const_cast<void *>(p);
    return 1;
  case hipMemoryTypeHost:
    if (!a.devicePointer) {
      return 0;
    }
    if (a.hostPointer && a.hostPointer != p) {
      // attributes describe the allocation base
      *dev = static_cast<char *>(a.devicePointer) +
             (static_cast<const char *>(p) -
              static_cast<const char *>(a.hostPointer));
    } else {
      *dev = a.devicePointer;
    }
    return 1;
  default:
    return 0;
  }
}

bool device_only(const void *p) {
  if (!p || !have_device()) {
    return false;
  }
  hipPointerAttribute_t a{};
  if (hipSuccess != hipPointerGetAttributes(&a, p)) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}

void finalize_runtime() {
  flush_release();
  Runtime &r = rt();
  std::lock_guard<std::mutex> lk(r.mu);
  if (r.ready.load(std::memory_order_acquire)) {
    (void)hipStreamSynchronize(r.stream);
    (void)hipStreamDestroy(r.stream);
    r.stream = nullptr;
    r.ready.store(false, std::memory_order_release);
  }
}

} // namespace tempi

using namespace tempi;

static hipStream_t as_stream(void *s) {
  return s ? static_cast<hipStream_t>(s) : library_stream();
}

TEMPI_EXPORT int tempi_device_count(void) {
  if (!have_device()) {
    return 0;
  }
  int n = 0;
  (void)hipGetDeviceCount(&n);
  return n;
}

TEMPI_EXPORT void *tempi_stream(void) {
  if (ensure_runtime() != TEMPI_SUCCESS) {
    return nullptr;
  }
  return library_stream();
}

TEMPI_EXPORT int tempi_stream_synchronize(void *stream) {
  int err = ensure_runtime();
  if (err) {
    return err;
  }
  TEMPI_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT void *tempi_alloc_device(int64_t bytes) {
  if (bytes <= 0 || ensure_runtime() != TEMPI_SUCCESS) {
    return nullptr;
  }
  void *p = nullptr;
  if (hipSuccess != hipMalloc(&p, (size_t)bytes)) {
    (void)hipGetLastError();
    set_error(TEMPI_ERR_NOMEM, "hipMalloc(%lld) failed", (long long)bytes);
    return nullptr;
  }
  return p;
}

TEMPI_EXPORT int tempi_free_device(void *p) {
  if (!p) {
    return TEMPI_SUCCESS;
  }
  TEMPI_HIP_CHECK(hipFree(p));
  return TEMPI_SUCCESS;
}

// pinned, device-mapped host memory allocated as such (the target of the
// ONE_SHOT method; the reference registers existing memory instead)
TEMPI_EXPORT void *tempi_alloc_host_mapped(int64_t bytes) {
  if (bytes <= 0 || ensure_runtime() != TEMPI_SUCCESS) {
    return nullptr;
  }
  void *p = nullptr;
  if (hipSuccess != hipHostMalloc(&p, (size_t)bytes,
                                  hipHostMallocMapped | hipHostMallocPortable)) {
    (void)hipGetLastError();
    set_error(TEMPI_ERR_NOMEM, "hipHostMalloc(%lld) failed", (long long)bytes);
    return nullptr;
  }
  return p;
}

TEMPI_EXPORT int tempi_free_host_mapped(void *p) {
  if (!p) {
    return TEMPI_SUCCESS;
  }
  TEMPI_HIP_CHECK(hipHostFree(p));
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_memcpy(void *dst, const void *src, int64_t bytes) {
  if (bytes < 0) {
    return set_error(TEMPI_ERR_ARG, "negative size");
  }
  if (bytes == 0) {
    return TEMPI_SUCCESS;
  }
  int err = ensure_runtime();
  if (err) {
    return err;
  }
  TEMPI_HIP_CHECK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDefault,
                                 library_stream()));
  TEMPI_HIP_CHECK(hipStreamSynchronize(library_stream()));
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_memset(void *dst, int value, int64_t bytes) {
  if (bytes < 0) {
    return set_error(TEMPI_ERR_ARG, "negative size");
  }
  if (bytes == 0) {
    return TEMPI_SUCCESS;
  }
  int err = ensure_runtime();
  if (err) {
    return err;
  }
  TEMPI_HIP_CHECK(hipMemsetAsync(dst, value, (size_t)bytes, library_stream()));
  TEMPI_HIP_CHECK(hipStreamSynchronize(library_stream()));
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_is_device_accessible(const void *p) {
  void *d = nullptr;
  return device_address(p, &d);
}

TEMPI_EXPORT int tempi_event_create(void **event) {
  int err = ensure_runtime();
  if (err) {
    return err;
  }
  hipEvent_t e = nullptr;
  TEMPI_HIP_CHECK(hipEventCreate(&e));
  *event = e;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_event_destroy(void *event) {
  if (event) {
    TEMPI_HIP_CHECK(hipEventDestroy(static_cast<hipEvent_t>(event)));
  }
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_event_record(void *event, void *stream) {
  int err = ensure_runtime();
  if (err) {
    return err;
  }
  TEMPI_HIP_CHECK(
      hipEventRecord(static_cast<hipEvent_t>(event), as_stream(stream)));
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_event_synchronize(void *event) {
  TEMPI_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(event)));
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_event_elapsed_ms(void *start, void *stop, float *ms) {
  TEMPI_HIP_CHECK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                      static_cast<hipEvent_t>(stop)));
  return TEMPI_SUCCESS;
}
