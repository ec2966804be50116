// interpose.cpp -- the MPI functions this library overrides, the per-type
// record cache, and the tempi_* pack/unpack entry points.
//
// reference: src/init.cpp:22-67, src/init_thread.cpp:22-65,
// src/finalize.cpp:20-46, src/type_commit.cpp:16-113, src/type_free.cpp:14-25,
// src/pack.cpp:28-67, src/unpack.cpp:20-59, include/type_cache.hpp:23-30.
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_map>

namespace tempi {

namespace {

struct TypeRecord {
  Canon canon;
};

struct TypeCache {
  std::mutex mu;
  std::unordered_map<int64_t, TypeRecord> map;
};

TypeCache &type_cache() {
  static TypeCache c;
  return c;
}

int64_t key_of(MPI_Datatype t) { return (int64_t)(intptr_t)t; }

} // namespace

bool type_lookup(MPI_Datatype t, Canon *out) {
  TypeCache &c = type_cache();
  {
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.map.find(key_of(t));
    if (it != c.map.end()) {
      *out = it->second.canon;
      return true;
    }
  }
  // Predefined types (MPI_FLOAT, MPI_BYTE, ...) are never committed: they
  // are lowered on first use, so device data of such a type takes the
  // device path instead of reaching the MPI library as a device pointer.
  int ni = 0, na = 0, nd = 0, comb = 0, init = 0, fin = 0;
  if (env().no_type_commit || t == MPI_DATATYPE_NULL ||
      MPI_SUCCESS != MPI_Initialized(&init) || !init ||
      MPI_SUCCESS != MPI_Finalized(&fin) || fin ||
      MPI_SUCCESS != MPI_Type_get_envelope(t, &ni, &na, &nd, &comb) ||
      comb != MPI_COMBINER_NAMED) {
    return false;
  }
  TypeRecord rec;
  rec.canon = canonicalize(t);
  std::lock_guard<std::mutex> lk(c.mu);
  c.map[key_of(t)] = rec;
  *out = rec.canon;
  return true;
}

int type_span(MPI_Datatype dt, int count, int64_t *lo, int64_t *hi) {
  MPI_Count tlb = 0, text = 0, lb = 0, ext = 0;
  if (MPI_SUCCESS != MPI_Type_get_true_extent_x(dt, &tlb, &text) ||
      MPI_SUCCESS != MPI_Type_get_extent_x(dt, &lb, &ext)) {
    return set_error(TEMPI_ERR_ARG, "extent query failed");
  }
  const int64_t reach = (int64_t)std::max(count - 1, 0) * (int64_t)ext;
  *lo = (int64_t)tlb + std::min<int64_t>(0, reach);
  *hi = (int64_t)tlb + (int64_t)text + std::max<int64_t>(0, reach);
  return TEMPI_SUCCESS;
}

namespace {

typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

// Writes to dst[0, n) the bytes of `after` that differ from `before` (both
// 16-B aligned, read through the device mapping of pinned host memory):
// 16 bytes compared per thread and step, byte stores where they differ.
__global__ void __launch_bounds__(256)
    k_write_changed(char *dst, const char *__restrict__ after,
                    const char *__restrict__ before, int64_t n) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x * 16;
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 16;
       i < n; i += step) {
    if (n - i >= 16) {
      const u32x4_t a = *reinterpret_cast<const u32x4_t *>(after + i);
      const u32x4_t b = *reinterpret_cast<const u32x4_t *>(before + i);
      if (a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w) {
        continue;
      }
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const uint32_t wa = a[k >> 2] >> (8 * (k & 3));
        const uint32_t wb = b[k >> 2] >> (8 * (k & 3));
        if ((uint8_t)wa != (uint8_t)wb) {
          dst[i + k] = (char)wa;
        }
      }
    } else {
      for (int64_t k = i; k < n; ++k) {
        if (after[k] != before[k]) {
          dst[k] = after[k];
        }
      }
    }
  }
}

} // namespace

SpanImage::~SpanImage() {
  pool_release(true, host);
  pool_release(true, before);
}

int SpanImage::take(const void *dev_buf, MPI_Datatype dt, int count,
                    bool written) {
  int64_t hi = 0;
  int err = type_span(dt, count, &lo, &hi);
  if (err) {
    return err;
  }
  len = std::max<int64_t>(hi - lo, 0);
  dev = static_cast<char *>(const_cast<void *>(dev_buf)) + lo;
  if ((err = ensure_runtime())) {
    return err;
  }
  const size_t bytes = (size_t)std::max<int64_t>(len, 1);
  host = static_cast<char *>(pool_acquire(true, bytes));
  if (written && host) {
    before = static_cast<char *>(pool_acquire(true, bytes));
  }
  if (!host || (written && !before)) {
    return set_error(TEMPI_ERR_NOMEM, "host image of %lld B", (long long)len);
  }
  if (len > 0) {
    hipStream_t s = library_stream();
    TEMPI_HIP_CHECK(hipMemcpyAsync(host, dev, (size_t)len,
                                   hipMemcpyDeviceToHost, s));
    if (written) {
      TEMPI_HIP_CHECK(hipMemcpyAsync(before, dev, (size_t)len,
                                     hipMemcpyDeviceToHost, s));
    }
    TEMPI_HIP_CHECK(hipStreamSynchronize(s));
  }
  return TEMPI_SUCCESS;
}

int SpanImage::put_back() {
  if (len <= 0) {
    return TEMPI_SUCCESS;
  }
  void *a = nullptr, *b = nullptr;
  if (!before || !device_address(host, &a) || !device_address(before, &b)) {
    return set_error(TEMPI_ERR_ARG, "host image was not taken for writing");
  }
  hipStream_t s = library_stream();
  const int64_t threads = (len + 15) / 16;
  const int64_t blocks = std::min<int64_t>((threads + 255) / 256,
                                           (int64_t)device_cu_count() * 8);
  hipLaunchKernelGGL(k_write_changed, dim3((unsigned)blocks), dim3(256), 0, s,
                     dev, static_cast<const char *>(a),
                     static_cast<const char *>(b), len);
  TEMPI_HIP_CHECK(hipGetLastError());
  TEMPI_HIP_CHECK(hipStreamSynchronize(s));
  return TEMPI_SUCCESS;
}

namespace {
bool lookup(MPI_Datatype t, Canon *out) { return type_lookup(t, out); }
} // namespace

// Shared body of pack and unpack on device-addressable buffers.
// `strided` / `packed` are device addresses of the strided buffer origin and
// of the first packed byte.
int run_copy(const tempi_strided_block &sb, int64_t count, const char *strided,
             char *packed, int dir, int variant, hipStream_t stream) {
  Geom g;
  int err = make_geom(sb, count, &g);
  if (err) {
    return err;
  }
  if (g.rows == 0) {
    return TEMPI_SUCCESS;
  }
  Choice c;
  if (variant < 0) {
    c = choose(g, strided, packed, dir, env().select, stream,
               /*in_place=*/true);
  } else {
    c.variant = variant;
  }
  return launch_copy(g, strided, packed, dir, c.variant, stream, c.max_blocks);
}

// run_copy on the library stream, returning when the copy is done
int run_copy_wait(const tempi_strided_block &sb, int64_t count,
                         const char *strided, char *packed, int dir) {
  Geom g;
  int err = make_geom(sb, count, &g);
  if (err) {
    return err;
  }
  if (g.rows == 0) {
    return TEMPI_SUCCESS;
  }
  const Choice c = choose(g, strided, packed, dir, env().select,
                          library_stream(), /*in_place=*/true);
  return launch_copy_wait(g, strided, packed, dir, c.variant, c.max_blocks);
}

// run_copy on the library stream without waiting; *seq > 0 when the copy
// signals its completion (signal_reached)
int run_copy_signalled(const tempi_strided_block &sb, int64_t count,
                       const char *strided, char *packed, int dir,
                       unsigned long long *seq) {
  *seq = 0;
  Geom g;
  int err = make_geom(sb, count, &g);
  if (err) {
    return err;
  }
  if (g.rows == 0) {
    return TEMPI_SUCCESS;
  }
  const Choice c = choose(g, strided, packed, dir, env().select,
                          library_stream(), /*in_place=*/true);
  return launch_copy_signalled(g, strided, packed, dir, c.variant,
                               c.max_blocks, seq);
}

// A pinned, device-mapped host staging buffer from the pool, returned on
// scope exit.
struct HostStage {
  void *p = nullptr;
  ~HostStage() { pool_release(true, p); }
};

// MPI_Pack / MPI_Unpack on device-accessible buffers.  `strided` is the
// device address of the strided buffer; `packed` is the first packed byte,
// either a device address or, with `host_packed`, pageable host memory the
// device cannot address: then the kernel reads or writes a pinned host
// staging buffer through its device mapping (as ONE_SHOT does) and the host
// copies between it and the caller's memory -- a kernel store across the
// link plus a host memcpy, where a copy from device memory into pageable
// memory would go through the runtime's own bounce buffer.
int pack_or_unpack(const Canon &cn, int count, const char *strided,
                   char *packed, bool host_packed, int dir) {
  const int64_t bytes = (int64_t)count * cn.size;
  HostStage stage;
  char *kernel_packed = packed;
  if (host_packed) {
    stage.p = pool_acquire(true, (size_t)std::max<int64_t>(bytes, 1));
    void *mapped = nullptr;
    if (!stage.p || !device_address(stage.p, &mapped)) {
      return set_error(TEMPI_ERR_NOMEM, "host staging buffer of %lld B",
                       (long long)bytes);
    }
    kernel_packed = static_cast<char *>(mapped);
    ++counters().host_staged_calls;
    if (dir == TEMPI_DIR_UNPACK) {
      memcpy(stage.p, packed, (size_t)bytes);
    }
  }
  // returns once the copy is done (also on failure: the staging buffer is
  // read, or goes back to the pool, only after that)
  const int err = run_copy_wait(cn.sb, count, strided, kernel_packed, dir);
  if (!err && host_packed && dir == TEMPI_DIR_PACK) {
    memcpy(packed, stage.p, (size_t)bytes);
  }
  return err;
}

// MPI_Pack / MPI_Unpack by the MPI library when one of the buffers is in
// device memory the host cannot address and the library cannot read device
// memory (it is not GPU-aware): the type has no strided form, the buffers
// are host data on one side and device memory on the other, or packing is
// switched off.  A device-resident typed buffer is replaced by a pinned
// host image of its type map's span ([true lb, true ub) over the objects);
// after an unpack only the bytes the library changed in the image are
// written back, so the bytes between the type map's blocks are never
// written.  Device-resident packed
// bytes cross through a pinned host buffer.  Without this the library would
// be handed a device pointer.  MPI error code.
int library_staged(bool pack, const void *typed, int count,
                   MPI_Datatype datatype, void *packed, int packed_size,
                   int *position, MPI_Comm comm) {
  if (ensure_runtime()) {
    return MPI_ERR_INTERN;
  }
  ++counters().host_staged_calls;
  SpanImage image;
  void *typed_at = const_cast<void *>(typed);
  if (device_only(typed)) {
    if (image.take(typed, datatype, count, /*written=*/!pack)) {
      TEMPI_LOG(LOG_ERROR, "staging for the MPI library: %s",
                tempi_last_error());
      return MPI_ERR_NO_MEM;
    }
    typed_at = image.base();
  }
  int err = MPI_SUCCESS;
  if (!device_only(packed)) {
    // packed bytes the library can address: it checks sizes and positions
    err = pack ? real_mpi().Pack(typed_at, count, datatype, packed,
                                 packed_size, position, comm)
               : real_mpi().Unpack(packed, packed_size, position, typed_at,
                                   count, datatype, comm);
  } else {
    // packed bytes in device memory cross through a pinned host buffer of
    // the bytes this call packs (or at most as many as remain to unpack)
    int bytes = 0;
    if (MPI_SUCCESS != MPI_Pack_size(count, datatype, comm, &bytes) ||
        *position < 0 || *position > packed_size ||
        (pack && (int64_t)*position + bytes > (int64_t)packed_size)) {
      MPI_Comm_call_errhandler(comm, MPI_ERR_TRUNCATE);
      return MPI_ERR_TRUNCATE;
    }
    if (!pack) {
      bytes = std::min(bytes, packed_size - *position);
    }
    HostStage stage;
    stage.p = pool_acquire(true, (size_t)std::max(bytes, 1));
    if (!stage.p) {
      return MPI_ERR_NO_MEM;
    }
    char *dev_at = static_cast<char *>(packed) + *position;
    hipStream_t s = library_stream();
    if (!pack && (hipSuccess != hipMemcpyAsync(stage.p, dev_at, (size_t)bytes,
                                               hipMemcpyDeviceToHost, s) ||
                  hipSuccess != hipStreamSynchronize(s))) {
      return MPI_ERR_INTERN;
    }
    int pos = 0;
    err = pack ? real_mpi().Pack(typed_at, count, datatype, stage.p, bytes,
                                 &pos, comm)
               : real_mpi().Unpack(stage.p, bytes, &pos, typed_at, count,
                                   datatype, comm);
    if (err != MPI_SUCCESS) {
      return err;
    }
    if (pack && (hipSuccess != hipMemcpyAsync(dev_at, stage.p, (size_t)pos,
                                              hipMemcpyHostToDevice, s) ||
                 hipSuccess != hipStreamSynchronize(s))) {
      return MPI_ERR_INTERN;
    }
    *position += pos;
  }
  if (err == MPI_SUCCESS && !pack && image.host && image.put_back()) {
    return MPI_ERR_INTERN;
  }
  return err;
}

// would the MPI library be handed device memory it cannot address?
bool library_needs_staging(const void *a, const void *b) {
  return !env().gpu_aware_mpi && (device_only(a) || device_only(b));
}

} // namespace tempi

using namespace tempi;

// ---------------------------------------------------------------------------
// MPI interposition
// ---------------------------------------------------------------------------

TEMPI_EXPORT int MPI_Init(int *argc, char ***argv) {
  ensure_env();
  int err = real_mpi().Init(argc, argv);
  if (env().disable || err != MPI_SUCCESS) {
    return err;
  }
  // bring the stream up now so the first MPI_Pack does not pay for it
  if (have_device()) {
    (void)ensure_runtime();
  }
  perf_init();
  p2p_init();
  return err;
}

TEMPI_EXPORT int MPI_Init_thread(int *argc, char ***argv, int required,
                                 int *provided) {
  ensure_env();
  int err = real_mpi().Init_thread(argc, argv, required, provided);
  if (env().disable || err != MPI_SUCCESS) {
    return err;
  }
  if (have_device()) {
    (void)ensure_runtime();
  }
  perf_init();
  p2p_init();
  return err;
}

TEMPI_EXPORT int MPI_Finalize(void) {
  ensure_env();
  if (!env().disable) {
    report_counters();
    {
      TypeCache &c = type_cache();
      std::lock_guard<std::mutex> lk(c.mu);
      c.map.clear();
    }
    neighbor_finalize();
    p2p_finalize();
    model_finalize();
    finalize_runtime();
  }
  return real_mpi().Finalize();
}

TEMPI_EXPORT int MPI_Type_commit(MPI_Datatype *datatype) {
  ensure_env();
  int result = real_mpi().Type_commit(datatype);
  if (env().disable || env().no_type_commit || result != MPI_SUCCESS) {
    return result;
  }
  TypeRecord rec;
  rec.canon = canonicalize(*datatype);
  if (rec.canon.ok) {
    ++counters().commits_canonical;
    if (env().select == TEMPI_SELECT_MODEL ||
        (env().select == TEMPI_SELECT_DEFAULT && model_available())) {
      // evaluate the learned selector for this type now, not in MPI_Pack
      model_precommit(rec.canon.sb);
    }
  } else {
    ++counters().commits_unsupported;
    TEMPI_LOG(LOG_DEBUG, "type %lld: library path (%s)",
              (long long)key_of(*datatype), rec.canon.reason.c_str());
  }
  TypeCache &c = type_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  c.map[key_of(*datatype)] = rec;
  return result;
}

TEMPI_EXPORT int MPI_Type_free(MPI_Datatype *datatype) {
  ensure_env();
  if (!env().disable && datatype) {
    // the library may change the handle, and may reuse its value for a new
    // type afterwards: drop the record first (reference: type_free.cpp:20)
    {
      TypeCache &c = type_cache();
      std::lock_guard<std::mutex> lk(c.mu);
      c.map.erase(key_of(*datatype));
    }
    neighbor_forget_type(*datatype);
  }
  return real_mpi().Type_free(datatype);
}

TEMPI_EXPORT int MPI_Pack(const void *inbuf, int incount, MPI_Datatype datatype,
                          void *outbuf, int outsize, int *position,
                          MPI_Comm comm) {
  ensure_env();
  const RealMpi &lib = real_mpi();
  if (env().disable) {
    return lib.Pack(inbuf, incount, datatype, outbuf, outsize, position, comm);
  }
  TraceRange trace("MPI_Pack");
  Canon cn;
  void *din = nullptr, *dout = nullptr;
  // device data packed into memory the device cannot address (pageable host
  // memory) is staged on the device; the reference hands that case to the
  // MPI library, which cannot read device memory
  if (env().no_pack || incount <= 0 || !position || !outbuf ||
      !lookup(datatype, &cn) || !cn.ok || !device_address(inbuf, &din)) {
    if (incount > 0 && position && inbuf && outbuf &&
        library_needs_staging(inbuf, outbuf)) {
      return library_staged(true, inbuf, incount, datatype, outbuf, outsize,
                            position, comm);
    }
    ++counters().library_pack_calls;
    return lib.Pack(inbuf, incount, datatype, outbuf, outsize, position, comm);
  }
  const bool host_packed = !device_address(outbuf, &dout);
  const int64_t packed = (int64_t)incount * cn.size;
  if (*position < 0 || (int64_t)*position + packed > (int64_t)outsize) {
    // the reference writes past the end here (pack.cpp:56)
    MPI_Comm_call_errhandler(comm, MPI_ERR_TRUNCATE);
    return MPI_ERR_TRUNCATE;
  }
  if (ensure_runtime()) {
    ++counters().library_pack_calls;
    return lib.Pack(inbuf, incount, datatype, outbuf, outsize, position, comm);
  }
  char *packed_at =
      static_cast<char *>(host_packed ? outbuf : dout) + *position;
  const int err = pack_or_unpack(cn, incount, static_cast<const char *>(din),
                                 packed_at, host_packed, TEMPI_DIR_PACK);
  if (err) {
    TEMPI_LOG(LOG_ERROR, "MPI_Pack: %s", tempi_last_error());
    MPI_Comm_call_errhandler(comm, MPI_ERR_INTERN);
    return MPI_ERR_INTERN;
  }
  *position += (int)packed;
  return MPI_SUCCESS;
}

TEMPI_EXPORT int MPI_Unpack(const void *inbuf, int insize, int *position,
                            void *outbuf, int outcount, MPI_Datatype datatype,
                            MPI_Comm comm) {
  ensure_env();
  const RealMpi &lib = real_mpi();
  if (env().disable) {
    return lib.Unpack(inbuf, insize, position, outbuf, outcount, datatype,
                      comm);
  }
  TraceRange trace("MPI_Unpack");
  Canon cn;
  void *din = nullptr, *dout = nullptr;
  // packed bytes in pageable host memory unpacked into device data: staged
  // on the device, as in MPI_Pack
  if (env().no_pack || outcount <= 0 || !position || !inbuf ||
      !lookup(datatype, &cn) || !cn.ok || !device_address(outbuf, &dout)) {
    if (outcount > 0 && position && inbuf && outbuf &&
        library_needs_staging(inbuf, outbuf)) {
      return library_staged(false, outbuf, outcount, datatype,
                            const_cast<void *>(inbuf), insize, position, comm);
    }
    ++counters().library_unpack_calls;
    return lib.Unpack(inbuf, insize, position, outbuf, outcount, datatype,
                      comm);
  }
  const bool host_packed = !device_address(inbuf, &din);
  const int64_t packed = (int64_t)outcount * cn.size;
  if (*position < 0 || (int64_t)*position + packed > (int64_t)insize) {
    MPI_Comm_call_errhandler(comm, MPI_ERR_TRUNCATE);
    return MPI_ERR_TRUNCATE;
  }
  if (ensure_runtime()) {
    ++counters().library_unpack_calls;
    return lib.Unpack(inbuf, insize, position, outbuf, outcount, datatype,
                      comm);
  }
  char *packed_at =
      static_cast<char *>(host_packed ? const_cast<void *>(inbuf) : din) +
      *position;
  const int err = pack_or_unpack(cn, outcount, static_cast<const char *>(dout),
                                 packed_at, host_packed, TEMPI_DIR_UNPACK);
  if (err) {
    TEMPI_LOG(LOG_ERROR, "MPI_Unpack: %s", tempi_last_error());
    MPI_Comm_call_errhandler(comm, MPI_ERR_INTERN);
    return MPI_ERR_INTERN;
  }
  *position += (int)packed;
  return MPI_SUCCESS;
}

// ---------------------------------------------------------------------------
// tempi_* entry points
// ---------------------------------------------------------------------------

TEMPI_EXPORT int tempi_type_strided_block(MPI_Datatype datatype,
                                          tempi_strided_block *sb,
                                          int64_t *size) {
  Canon cn;
  if (!lookup(datatype, &cn)) {
    return set_error(TEMPI_ERR_ARG,
                     "type was not committed through MPI_Type_commit");
  }
  if (!cn.ok) {
    return set_error(TEMPI_ERR_UNSUPPORTED, "%s", cn.reason.c_str());
  }
  if (sb) {
    *sb = cn.sb;
  }
  if (size) {
    *size = cn.size;
  }
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_type_canonicalize(MPI_Datatype datatype,
                                         tempi_strided_block *sb,
                                         int64_t *size) {
  Canon cn = canonicalize(datatype);
  if (!cn.ok) {
    return set_error(TEMPI_ERR_UNSUPPORTED, "%s", cn.reason.c_str());
  }
  if (sb) {
    *sb = cn.sb;
  }
  if (size) {
    *size = cn.size;
  }
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_num_variants(void) { return kNumVariants; }

TEMPI_EXPORT int tempi_variant_info_get(int variant, tempi_variant_info *info) {
  if (variant < 0 || variant >= kNumVariants || !info) {
    return set_error(TEMPI_ERR_ARG, "variant %d out of range", variant);
  }
  info->lanes_per_row = kVariants[variant].P;
  info->word_bytes = kVariants[variant].W;
  info->unroll = kVariants[variant].U;
  info->nontemporal_stores = kVariants[variant].NTS;
  info->nontemporal_loads = kVariants[variant].NTL;
  return TEMPI_SUCCESS;
}

// packed bytes of `count` objects of `sb`, with the copy's own validation
// of the descriptor (dimension count, signs, 64-bit products)
static int packed_bytes(const tempi_strided_block &sb, int64_t count,
                        int64_t *bytes) {
  Geom g;
  int err = make_geom(sb, count, &g);
  if (err) {
    return err;
  }
  *bytes = g.L * g.rows;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_select_variant(const tempi_strided_block *sb,
                                      int64_t incount, const void *strided_buf,
                                      const void *packed_buf, int direction,
                                      int method, int *variant) {
  if (!sb || !variant) {
    return set_error(TEMPI_ERR_ARG, "null argument");
  }
  int err = check_direction(direction);
  if (err) {
    return err;
  }
  Geom g;
  err = make_geom(*sb, incount, &g);
  if (err) {
    return err;
  }
  if (method == TEMPI_SELECT_AUTOTUNE) {
    err = ensure_runtime();
    if (err) {
      return err;
    }
    void *ds = nullptr, *dp = nullptr;
    if (!device_address(strided_buf, &ds) || !device_address(packed_buf, &dp)) {
      return set_error(TEMPI_ERR_ARG, "autotuning needs device-accessible buffers");
    }
    *variant = choose(g, static_cast<const char *>(ds), static_cast<char *>(dp),
                      direction, method, library_stream(), /*in_place=*/false)
                   .variant;
    return TEMPI_SUCCESS;
  }
  *variant = choose(g, static_cast<const char *>(strided_buf),
                    static_cast<char *>(const_cast<void *>(packed_buf)),
                    direction, method, nullptr, /*in_place=*/false)
                 .variant;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_pack_sb(const tempi_strided_block *sb,
                               const void *inbuf, int64_t incount,
                               void *outbuf, int64_t outsize,
                               int64_t *position, int variant, void *stream) {
  if (!sb || !position || incount < 0 || *position < 0 || outsize < 0) {
    return set_error(TEMPI_ERR_ARG, "bad argument");
  }
  int64_t packed = 0;
  int err = packed_bytes(*sb, incount, &packed);
  if (err) {
    return err;
  }
  if (packed > outsize - *position) {
    return set_error(TEMPI_ERR_TRUNCATE, "need %lld bytes, have %lld",
                     (long long)(*position + packed), (long long)outsize);
  }
  err = ensure_runtime();
  if (err) {
    return err;
  }
  void *din = nullptr, *dout = nullptr;
  if (packed > 0 &&
      (!device_address(inbuf, &din) || !device_address(outbuf, &dout))) {
    return set_error(TEMPI_ERR_ARG, "buffers must be device-accessible");
  }
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : library_stream();
  err = run_copy(*sb, incount, static_cast<const char *>(din),
                 static_cast<char *>(dout) + *position, TEMPI_DIR_PACK,
                 variant, s);
  if (err) {
    return err;
  }
  *position += packed;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_unpack_sb(const tempi_strided_block *sb,
                                 const void *inbuf, int64_t insize,
                                 int64_t *position, void *outbuf,
                                 int64_t outcount, int variant, void *stream) {
  if (!sb || !position || outcount < 0 || *position < 0 || insize < 0) {
    return set_error(TEMPI_ERR_ARG, "bad argument");
  }
  int64_t packed = 0;
  int err = packed_bytes(*sb, outcount, &packed);
  if (err) {
    return err;
  }
  if (packed > insize - *position) {
    return set_error(TEMPI_ERR_TRUNCATE, "need %lld bytes, have %lld",
                     (long long)(*position + packed), (long long)insize);
  }
  err = ensure_runtime();
  if (err) {
    return err;
  }
  void *din = nullptr, *dout = nullptr;
  if (packed > 0 &&
      (!device_address(inbuf, &din) || !device_address(outbuf, &dout))) {
    return set_error(TEMPI_ERR_ARG, "buffers must be device-accessible");
  }
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : library_stream();
  err = run_copy(*sb, outcount, static_cast<const char *>(dout),
                 static_cast<char *>(din) + *position, TEMPI_DIR_UNPACK,
                 variant, s);
  if (err) {
    return err;
  }
  *position += packed;
  return TEMPI_SUCCESS;
}
