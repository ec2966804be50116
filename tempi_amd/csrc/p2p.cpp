// p2p.cpp -- MPI_Send / MPI_Recv / MPI_Isend / MPI_Irecv / MPI_Wait /
// MPI_Waitall / MPI_Test for device buffers whose type has a strided form.
//
// reference: src/send.cpp, src/recv.cpp, src/internal/send.cpp (dispatch);
// src/internal/sender.cpp:88-328 (the DEVICE, ONE_SHOT and STAGED senders and
// the model-based choice); src/isend.cpp, src/irecv.cpp,
// src/internal/isend.cpp, irecv.cpp and async_operation.cpp:34-461 (the
// Isend / Irecv state machines, the request handed to the caller, progress);
// src/wait.cpp.
//
// A non-contiguous message travels as its packed bytes (MPI_PACKED, type-map
// order), staged one of three ways:
//   DEVICE   packed into device memory, handed to the MPI library as a device
//            pointer -- only when the library reads device memory
//            (TEMPI_MPI_GPU_AWARE; MPICH built without GPU support does not);
//   ONESHOT  packed by the kernel straight into pinned host memory mapped
//            into the device address space;
//   STAGED   packed into device memory, then copied to pinned host memory.
// The receive side mirrors each.  TEMPI_DATATYPE_{ONESHOT,DEVICE,STAGED}
// force one; otherwise the measured system model (perf.hip) prices each
// available method and the cheapest is taken, cached per (same node, bytes,
// block length), as the reference does.  Without tables: DEVICE when the
// library is GPU-aware, else ONESHOT.
//
// Isend: pack on the library stream, then (once the event has fired) an
// MPI_Isend of the staged bytes.  Irecv: an MPI_Irecv into the staging buffer
// at once, then (once it has completed) the unpack.  Progress is made in every
// call of this file; before any call that can block in the MPI library, all
// sends still packing are posted first, so that a receive this process
// waits on can be matched by its own pending sends (the reference's MPI_Wait
// only progresses the request waited on, and a self-exchange can stall).
// The caller gets a request handle of its own that the MPI library never
// issues; MPI_Wait / MPI_Waitall / MPI_Test complete it (other completion
// calls are not intercepted, as in the reference).  Calls are expected from
// one thread at a time (MPI_THREAD_SINGLE / FUNNELED, as the reference).
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <memory>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace tempi {

namespace {

int g_override_method = -1; // tempi_set_datatype_method; -1 = environment

// ---------------------------------------------------------------------------
// node map: which world ranks share a node (reference: topology.cpp:191-196)
// ---------------------------------------------------------------------------
std::vector<uint64_t> g_node_of_world_rank;

uint64_t fnv1a(const char *s, int n) {
  uint64_t h = 1469598103934665603ull;
  for (int i = 0; i < n; ++i) {
    h = (h ^ (unsigned char)s[i]) * 1099511628211ull;
  }
  return h;
}

bool colocated(MPI_Comm comm, int peer) {
  int me = -1;
  if (MPI_SUCCESS != MPI_Comm_rank(comm, &me) || peer == me) {
    return true;
  }
  if (peer < 0 || g_node_of_world_rank.empty()) {
    return g_node_of_world_rank.size() <= 1 && peer >= 0;
  }
  MPI_Group g = MPI_GROUP_NULL, w = MPI_GROUP_NULL;
  int ranks[2] = {me, peer}, world[2] = {-1, -1};
  bool same = false;
  if (MPI_SUCCESS == MPI_Comm_group(comm, &g) &&
      MPI_SUCCESS == MPI_Comm_group(MPI_COMM_WORLD, &w) &&
      MPI_SUCCESS == MPI_Group_translate_ranks(g, 2, ranks, w, world) &&
      world[0] >= 0 && world[1] >= 0 &&
      world[0] < (int)g_node_of_world_rank.size() &&
      world[1] < (int)g_node_of_world_rank.size()) {
    same = g_node_of_world_rank[world[0]] == g_node_of_world_rank[world[1]];
  }
  if (g != MPI_GROUP_NULL) {
    MPI_Group_free(&g);
  }
  if (w != MPI_GROUP_NULL) {
    MPI_Group_free(&w);
  }
  return same;
}

// ---------------------------------------------------------------------------
// method choice
// ---------------------------------------------------------------------------
std::map<std::tuple<bool, int64_t, int64_t, int64_t>, int> g_method_cache;
// the tables and method model the cache was built on
uint64_t g_method_cache_tables = 0, g_method_cache_model = 0;

// The staging method for a message of `bytes` packed bytes in rows of
// `block` bytes `stride` apart, to a peer on this node or not: forced by the
// environment or tempi_set_datatype_method; else, for a peer on this node,
// the method model's cheapest (mlp.hip: trained on end-to-end timings of
// each method on this kind of node); else the measured tables' cheapest
// (reference: sender.cpp:128-273); without either, DEVICE when the MPI
// library reads device memory, otherwise ONESHOT.  Cached per message shape.
int choose_method_for(bool near, int64_t bytes, int64_t block,
                      int64_t stride) {
  const int m = forced_datatype_method();
  if (m != TEMPI_METHOD_UNKNOWN) {
    return m;
  }
  if (g_method_cache_tables != perf_generation() ||
      g_method_cache_model != method_model_generation()) {
    // tables or method model replaced since: price again
    g_method_cache.clear();
    g_method_cache_tables = perf_generation();
    g_method_cache_model = method_model_generation();
  }
  const auto key = std::make_tuple(near, bytes, block, stride);
  auto it = g_method_cache.find(key);
  if (it != g_method_cache.end()) {
    return it->second;
  }
  std::vector<int> cands = {TEMPI_METHOD_ONESHOT, TEMPI_METHOD_STAGED};
  if (env().gpu_aware_mpi) {
    cands.insert(cands.begin(), TEMPI_METHOD_DEVICE);
  }
  int best = -1;
  if (near) {
    float f[kModelInputs];
    method_features(bytes, block, stride, near, f);
    best = method_model_pick(f, cands.data(), (int)cands.size());
    if (best > 0) {
      ++counters().method_model_choices;
    }
  }
  if (best <= 0) {
    best = env().gpu_aware_mpi ? TEMPI_METHOD_DEVICE : TEMPI_METHOD_ONESHOT;
    const int64_t bl = std::min<int64_t>(std::max<int64_t>(1, block), 512);
    double best_s = INFINITY;
    for (int c : cands) {
      const double t = perf_model_seconds(c, near, bytes, bl);
      if (t < best_s) {
        best_s = t;
        best = c;
      }
    }
  }
  g_method_cache[key] = best;
  return best;
}

int choose_method(const Canon &cn, int64_t bytes, MPI_Comm comm, int peer) {
  const int64_t block = std::max<int64_t>(1, cn.sb.counts[0]);
  const int64_t stride = cn.sb.ndims >= 2 ? cn.sb.strides[1] : block;
  return choose_method_for(colocated(comm, peer), bytes, block, stride);
}

void count_method(int m) {
  switch (m) {
  case TEMPI_METHOD_DEVICE:
    ++counters().p2p_device;
    break;
  case TEMPI_METHOD_ONESHOT:
    ++counters().p2p_oneshot;
    break;
  default:
    ++counters().p2p_staged;
    break;
  }
}

// ---------------------------------------------------------------------------
// staging: the buffers one transfer uses
// ---------------------------------------------------------------------------
struct Staging {
  int method = TEMPI_METHOD_UNKNOWN;
  int64_t bytes = 0;
  void *host = nullptr;     // pinned, mapped (ONESHOT, STAGED)
  void *host_dev = nullptr; // its device address
  void *dev = nullptr;      // device (DEVICE, STAGED)

  int acquire(int m, int64_t n) {
    method = m;
    bytes = n;
    if (m != TEMPI_METHOD_DEVICE) {
      host = pool_acquire(true, (size_t)n);
      if (!host || !device_address(host, &host_dev)) {
        return set_error(TEMPI_ERR_NOMEM, "host staging buffer");
      }
    }
    if (m != TEMPI_METHOD_ONESHOT) {
      dev = pool_acquire(false, (size_t)n);
      if (!dev) {
        return set_error(TEMPI_ERR_NOMEM, "device staging buffer");
      }
    }
    return TEMPI_SUCCESS;
  }
  void release() {
    pool_release(true, host);
    pool_release(false, dev);
    host = host_dev = dev = nullptr;
  }
  // what the MPI library sends from / receives into
  void *wire() const { return method == TEMPI_METHOD_DEVICE ? dev : host; }
  // where the kernels write packed bytes / read them
  char *kernel_side() const {
    return static_cast<char *>(method == TEMPI_METHOD_ONESHOT ? host_dev : dev);
  }
};

// enqueue pack (+ copy to host) of `count` objects into the staging buffer
int enqueue_pack(const Canon &cn, int count, const void *ubuf_dev, Staging &st,
                 hipStream_t s) {
  int err = run_copy(cn.sb, count, static_cast<const char *>(ubuf_dev),
                     st.kernel_side(), TEMPI_DIR_PACK, -1, s);
  if (!err && st.method == TEMPI_METHOD_STAGED) {
    TEMPI_HIP_CHECK(hipMemcpyAsync(st.host, st.dev, (size_t)st.bytes,
                                   hipMemcpyDeviceToHost, s));
  }
  return err;
}

// enqueue (copy to device +) unpack of `count` objects from the staging buffer
int enqueue_unpack(const Canon &cn, int count, void *ubuf_dev, Staging &st,
                   int64_t bytes, hipStream_t s) {
  if (st.method == TEMPI_METHOD_STAGED && bytes > 0) {
    TEMPI_HIP_CHECK(hipMemcpyAsync(st.dev, st.host, (size_t)bytes,
                                   hipMemcpyHostToDevice, s));
  }
  return run_copy(cn.sb, count, static_cast<const char *>(ubuf_dev),
                  st.kernel_side(), TEMPI_DIR_UNPACK, -1, s);
}

int mpi_error_of(int tempi_err) {
  return tempi_err == TEMPI_SUCCESS ? MPI_SUCCESS : MPI_ERR_INTERN;
}

// objects a received message holds (whole objects only)
int received_objects(const MPI_Status &st, int64_t obj_size, int max_count) {
  int n = 0;
  if (MPI_SUCCESS != MPI_Get_count(&st, MPI_PACKED, &n) || n < 0 ||
      obj_size <= 0) {
    return 0;
  }
  return (int)std::min<int64_t>(max_count, n / obj_size);
}

// ---------------------------------------------------------------------------
// asynchronous operations
// ---------------------------------------------------------------------------
class AsyncOp {
public:
  virtual ~AsyncOp() = default;
  virtual bool is_send() const = 0;
  // advance without blocking; MPI error code
  virtual int wake() = 0;
  // block until the packed bytes are handed to the library (sends)
  virtual int post_blocking() { return MPI_SUCCESS; }
  virtual bool done() const = 0;
  // complete if possible without blocking; *flag = completed
  virtual int test(int *flag, MPI_Status *st) = 0;
  // block until complete
  virtual int wait(MPI_Status *st) = 0;
};

// Typed device data whose type has no strided form, handed to an MPI
// library that cannot read device memory: the library sends from (or
// receives into) a pinned host image of the type map's span, taken at the
// start; when a receive completes, only the bytes the library changed in
// the image are written to the device, so writes to the bytes between the
// type map's blocks made while the receive was pending are kept.
class SpanOp : public AsyncOp {
  bool send_;
  bool done_ = false;
  SpanImage img_;
  MPI_Request lib_req_ = MPI_REQUEST_NULL;
  MPI_Status status_{};
  int error_ = MPI_SUCCESS;

  int finish() {
    done_ = true;
    if (!send_ && error_ == MPI_SUCCESS && img_.put_back()) {
      TEMPI_LOG(LOG_ERROR, "receive into device memory: %s", tempi_last_error());
      error_ = MPI_ERR_INTERN;
    }
    return error_;
  }

public:
  explicit SpanOp(bool send) : send_(send) {}
  ~SpanOp() override {
    if (!done_ && lib_req_ != MPI_REQUEST_NULL) {
      (void)real_mpi().Wait(&lib_req_, MPI_STATUS_IGNORE);
    }
  }
  int start(const void *buf, int count, MPI_Datatype dt, int peer, int tag,
            MPI_Comm comm) {
    if (img_.take(buf, dt, count, /*written=*/!send_)) {
      TEMPI_LOG(LOG_ERROR, "staging device data: %s", tempi_last_error());
      return MPI_ERR_NO_MEM;
    }
    return send_ ? real_mpi().Isend(img_.base(), count, dt, peer, tag, comm,
                                    &lib_req_)
                 : real_mpi().Irecv(img_.base(), count, dt, peer, tag, comm,
                                    &lib_req_);
  }
  bool is_send() const override { return send_; }
  int wake() override {
    if (done_) {
      return MPI_SUCCESS;
    }
    int flag = 0;
    const int err = real_mpi().Test(&lib_req_, &flag, &status_);
    if (err != MPI_SUCCESS) {
      error_ = err;
    }
    return flag || err != MPI_SUCCESS ? finish() : MPI_SUCCESS;
  }
  bool done() const override { return done_; }
  int test(int *flag, MPI_Status *st) override {
    const int err = wake();
    *flag = done_;
    if (done_ && st != MPI_STATUS_IGNORE) {
      *st = status_;
    }
    return err;
  }
  int wait(MPI_Status *st) override {
    if (!done_) {
      const int err = real_mpi().Wait(&lib_req_, &status_);
      if (err != MPI_SUCCESS) {
        error_ = err;
      }
      finish();
    }
    if (st != MPI_STATUS_IGNORE) {
      *st = status_;
    }
    return error_;
  }
};

// Would the MPI library be handed `buf` in device memory it cannot read?
bool library_cannot_read(const void *buf, int count, int peer) {
  return !env().disable && !env().gpu_aware_mpi && count > 0 &&
         peer != MPI_PROC_NULL && buf && device_only(buf);
}

class IsendOp : public AsyncOp {
  enum { PACKING, SENDING, DONE } state_ = PACKING;
  Staging stg_;
  hipEvent_t ev_ = nullptr;
  unsigned long long seq_ = 0; // the pack's completion signal (0: none)
  MPI_Request lib_req_ = MPI_REQUEST_NULL;
  int64_t bytes_;
  int peer_, tag_;
  MPI_Comm comm_;

  int post() {
    state_ = SENDING;
    return real_mpi().Isend(stg_.wire(), (int)bytes_, MPI_PACKED, peer_, tag_,
                            comm_, &lib_req_);
  }

public:
  IsendOp(int64_t bytes, int peer, int tag, MPI_Comm comm)
      : bytes_(bytes), peer_(peer), tag_(tag), comm_(comm) {}
  ~IsendOp() override {
    if (state_ == PACKING) {
      // never completed (an error path): the pack may still be running
      (void)hipStreamSynchronize(library_stream());
    }
    if (ev_) {
      (void)hipEventDestroy(ev_);
    }
    stg_.release();
  }
  bool is_send() const override { return true; }
  int start(const Canon &cn, int count, const void *ubuf_dev, int method) {
    hipStream_t s = library_stream();
    int err = stg_.acquire(method, bytes_);
    if (!err && method != TEMPI_METHOD_STAGED) {
      // the pack is all there is to wait for: it may signal its completion
      err = run_copy_signalled(cn.sb, count, static_cast<const char *>(ubuf_dev),
                               stg_.kernel_side(), TEMPI_DIR_PACK, &seq_);
    } else if (!err) {
      err = enqueue_pack(cn, count, ubuf_dev, stg_, s);
    }
    if (!err && hipSuccess != hipEventCreateWithFlags(&ev_, hipEventDisableTiming)) {
      err = set_error(TEMPI_ERR_HIP, "event creation failed");
    }
    if (!err && hipSuccess != hipEventRecord(ev_, s)) {
      err = set_error(TEMPI_ERR_HIP, "event record failed");
    }
    return err;
  }
  int wake() override {
    if (state_ != PACKING) {
      return MPI_SUCCESS;
    }
    if (signal_reached(seq_)) {
      ++counters().signalled_waits;
      return post();
    }
    const hipError_t q = hipEventQuery(ev_);
    if (q == hipErrorNotReady) {
      return MPI_SUCCESS;
    }
    if (q != hipSuccess) {
      set_error(TEMPI_ERR_HIP, "pack for MPI_Isend failed: %s",
                hipGetErrorString(q));
      return MPI_ERR_INTERN;
    }
    return post();
  }
  int post_blocking() override {
    if (state_ != PACKING) {
      return MPI_SUCCESS;
    }
    if (wait_signal_or_event(seq_, ev_)) {
      TEMPI_LOG(LOG_ERROR, "pack for MPI_Isend: %s", tempi_last_error());
      return MPI_ERR_INTERN;
    }
    return post();
  }
  bool done() const override { return state_ == DONE; }
  int test(int *flag, MPI_Status *st) override {
    *flag = 0;
    int err = wake();
    if (err || state_ != SENDING) {
      return err;
    }
    err = real_mpi().Test(&lib_req_, flag, st);
    if (*flag) {
      state_ = DONE;
    }
    return err;
  }
  int wait(MPI_Status *st) override {
    int err = post_blocking();
    if (err) {
      return err;
    }
    err = real_mpi().Wait(&lib_req_, st);
    state_ = DONE;
    return err;
  }
};

class IrecvOp : public AsyncOp {
  enum { RECEIVING, UNPACKING, DONE } state_ = RECEIVING;
  Staging stg_;
  hipEvent_t ev_ = nullptr;
  unsigned long long seq_ = 0; // the unpack's completion signal (0: none)
  MPI_Request lib_req_ = MPI_REQUEST_NULL;
  MPI_Status status_{};
  Canon cn_;
  int count_;
  void *ubuf_dev_;
  int64_t bytes_;
  int error_ = MPI_SUCCESS;

  int unpack() {
    state_ = UNPACKING;
    if (error_ != MPI_SUCCESS) {
      return error_;
    }
    const int n = received_objects(status_, cn_.size, count_);
    hipStream_t s = library_stream();
    int err = TEMPI_SUCCESS;
    if (stg_.method != TEMPI_METHOD_STAGED) {
      err = run_copy_signalled(cn_.sb, n, static_cast<const char *>(ubuf_dev_),
                               stg_.kernel_side(), TEMPI_DIR_UNPACK, &seq_);
    } else {
      err = enqueue_unpack(cn_, n, ubuf_dev_, stg_, (int64_t)n * cn_.size, s);
    }
    if (!err && hipSuccess != hipEventRecord(ev_, s)) {
      err = set_error(TEMPI_ERR_HIP, "event record failed");
    }
    return mpi_error_of(err);
  }

public:
  IrecvOp(const Canon &cn, int count, void *ubuf_dev, int64_t bytes)
      : cn_(cn), count_(count), ubuf_dev_(ubuf_dev), bytes_(bytes) {}
  ~IrecvOp() override {
    if (state_ == UNPACKING) {
      (void)hipStreamSynchronize(library_stream());
    }
    if (ev_) {
      (void)hipEventDestroy(ev_);
    }
    stg_.release();
  }
  bool is_send() const override { return false; }
  int start(int method, int peer, int tag, MPI_Comm comm) {
    int err = stg_.acquire(method, bytes_);
    if (!err && hipSuccess != hipEventCreateWithFlags(&ev_, hipEventDisableTiming)) {
      err = set_error(TEMPI_ERR_HIP, "event creation failed");
    }
    if (err) {
      return MPI_ERR_INTERN;
    }
    return real_mpi().Irecv(stg_.wire(), (int)bytes_, MPI_PACKED, peer, tag,
                            comm, &lib_req_);
  }
  int wake() override {
    if (state_ == RECEIVING) {
      int flag = 0;
      const int err = real_mpi().Test(&lib_req_, &flag, &status_);
      if (err != MPI_SUCCESS) {
        error_ = err;
      }
      if (flag || err != MPI_SUCCESS) {
        return unpack();
      }
      return MPI_SUCCESS;
    }
    if (state_ == UNPACKING) {
      if (error_ != MPI_SUCCESS) {
        state_ = DONE;
        return error_;
      }
      if (signal_reached(seq_)) {
        ++counters().signalled_waits;
        state_ = DONE;
        return MPI_SUCCESS;
      }
      const hipError_t q = hipEventQuery(ev_);
      if (q == hipSuccess) {
        state_ = DONE;
      } else if (q != hipErrorNotReady) {
        set_error(TEMPI_ERR_HIP, "unpack for MPI_Irecv failed");
        return MPI_ERR_INTERN;
      }
    }
    return MPI_SUCCESS;
  }
  bool done() const override { return state_ == DONE; }
  int test(int *flag, MPI_Status *st) override {
    int err = wake();
    if (!err && state_ == UNPACKING) {
      err = wake();
    }
    *flag = state_ == DONE;
    if (*flag && st != MPI_STATUS_IGNORE) {
      *st = status_;
    }
    return err ? err : error_;
  }
  int wait(MPI_Status *st) override {
    if (state_ == RECEIVING) {
      const int err = real_mpi().Wait(&lib_req_, &status_);
      if (err != MPI_SUCCESS) {
        error_ = err;
      }
      const int uerr = unpack();
      if (uerr != MPI_SUCCESS && error_ == MPI_SUCCESS) {
        error_ = uerr;
      }
    }
    if (state_ == UNPACKING && error_ == MPI_SUCCESS &&
        wait_signal_or_event(seq_, ev_)) {
      TEMPI_LOG(LOG_ERROR, "unpack for MPI_Irecv: %s", tempi_last_error());
      error_ = MPI_ERR_INTERN;
    }
    state_ = DONE;
    if (st != MPI_STATUS_IGNORE) {
      *st = status_;
    }
    return error_;
  }
};

// ---------------------------------------------------------------------------
// the caller-visible requests
// ---------------------------------------------------------------------------
std::unordered_map<int64_t, std::unique_ptr<AsyncOp>> g_active;
uint64_t g_next_id = 0;

int64_t key_of(MPI_Request r) { return (int64_t)(intptr_t)r; }

// A handle the library never issues for a live request.  MPICH-family
// libraries encode the handle kind in the two top bits of an int and issue
// live requests with them set; values with both clear and below
// MPI_REQUEST_NULL's object field are free.  Pointer-handle libraries get
// the operation's own address.
template <class R = MPI_Request> R mint_handle(AsyncOp *op) {
  if constexpr (sizeof(R) == sizeof(int)) {
    for (;;) {
      const R r = (R)(0x01000000 + (int)(g_next_id++ % 0x00f00000));
      if (r != MPI_REQUEST_NULL && !g_active.count(key_of(r))) {
        return r;
      }
    }
  } else {
    return reinterpret_cast<R>(op);
  }
}

AsyncOp *find_op(MPI_Request r) {
  auto it = g_active.find(key_of(r));
  return it == g_active.end() ? nullptr : it->second.get();
}

// advance every operation without blocking
int try_progress() {
  int first_err = MPI_SUCCESS;
  for (auto &kv : g_active) {
    if (!kv.second->done()) {
      const int err = kv.second->wake();
      if (err != MPI_SUCCESS && first_err == MPI_SUCCESS) {
        first_err = err;
      }
    }
  }
  return first_err;
}

// hand every still-packing send to the library (before anything that may
// block inside it)
int post_pending_sends() {
  int first_err = MPI_SUCCESS;
  for (auto &kv : g_active) {
    if (kv.second->is_send() && !kv.second->done()) {
      const int err = kv.second->post_blocking();
      if (err != MPI_SUCCESS && first_err == MPI_SUCCESS) {
        first_err = err;
      }
    }
  }
  return first_err;
}

// modelled times of a contiguous transfer staged through the host and sent
// directly (reference: SendRecv1DStaged::model and SendRecvFallback::model,
// sender.cpp:26-61)
bool staged_beats_direct(int64_t bytes) {
  auto curve = [](int c, int64_t b) {
    int n = 0;
    tempi_perf_get_curve(c, nullptr, 0, &n);
    std::vector<double> t((size_t)std::max(n, 1));
    tempi_perf_get_curve(c, t.data(), n, &n);
    return tempi_interp_time(t.data(), n, b);
  };
  const double send = curve(TEMPI_CURVE_INTRA_GPU_GPU, bytes);
  const double staged =
      curve(TEMPI_CURVE_D2H, bytes) + send + curve(TEMPI_CURVE_H2D, bytes);
  return staged < send;
}

// does this call take the device path?  fills the record and the device
// address of the user buffer, and the method when the path decides it
bool device_path(const void *buf, int count, MPI_Datatype dt, int peer,
                 Canon *cn, void **dev, int *method_out) {
  if (env().disable || count <= 0 || peer == MPI_PROC_NULL || !buf ||
      !type_lookup(dt, cn) || !cn->ok || cn->size <= 0 ||
      !device_address(buf, dev) || ensure_runtime()) {
    return false;
  }
  const bool contiguous = cn->sb.ndims == 1 && cn->extent == cn->size;
  if (contiguous) {
    // contiguous device data goes to the library as it is unless staging is
    // asked for (reference: ContiguousMethod, src/internal/env.cpp:81-86)
    // or the library cannot read device memory
    if (env().contiguous_method == TEMPI_METHOD_STAGED ||
        (!env().gpu_aware_mpi && device_only(buf))) {
      *method_out = TEMPI_METHOD_STAGED;
      return true;
    }
    if (env().contiguous_auto &&
        staged_beats_direct((int64_t)count * cn->size)) {
      // reference: SendRecv1D::send, sender.cpp:63-74
      *method_out = TEMPI_METHOD_STAGED;
      return true;
    }
    return false;
  }
  *method_out = TEMPI_METHOD_UNKNOWN; // chosen by the caller with the peer
  return true;
}

// a non-blocking transfer of device data the MPI library cannot read, from
// or into a host image of its span (SpanOp)
int start_span_op(bool send, const void *buf, int count, MPI_Datatype dt,
                  int peer, int tag, MPI_Comm comm, MPI_Request *request) {
  auto op = std::make_unique<SpanOp>(send);
  const int err = op->start(buf, count, dt, peer, tag, comm);
  if (err != MPI_SUCCESS) {
    return err;
  }
  ++counters().host_staged_calls;
  *request = mint_handle(op.get());
  g_active[key_of(*request)] = std::move(op);
  (void)try_progress();
  return MPI_SUCCESS;
}

int check_count(int count, const Canon &cn, int64_t *bytes) {
  *bytes = (int64_t)count * cn.size;
  if (*bytes > INT32_MAX) {
    set_error(TEMPI_ERR_UNSUPPORTED, "message of %lld packed bytes",
              (long long)*bytes);
    return MPI_ERR_COUNT;
  }
  return MPI_SUCCESS;
}

} // namespace

int forced_datatype_method() {
  int m = g_override_method >= 0 ? g_override_method : env().datatype_method;
  if (m == TEMPI_METHOD_DEVICE && !env().gpu_aware_mpi) {
    static bool warned = false;
    if (!warned) {
      warned = true;
      TEMPI_LOG(LOG_WARN, "DEVICE sends need a GPU-aware MPI library "
                          "(TEMPI_MPI_GPU_AWARE); staging through the host");
    }
    m = TEMPI_METHOD_STAGED;
  }
  return m;
}

bool world_ranks_share_node(int a, int b) {
  if (a < 0 || b < 0 || a >= (int)g_node_of_world_rank.size() ||
      b >= (int)g_node_of_world_rank.size()) {
    return g_node_of_world_rank.size() <= 1;
  }
  return g_node_of_world_rank[a] == g_node_of_world_rank[b];
}

void p2p_init() {
  int size = 1;
  if (MPI_SUCCESS != MPI_Comm_size(MPI_COMM_WORLD, &size) || size <= 1) {
    g_node_of_world_rank.assign(1, 0);
    return;
  }
  char name[MPI_MAX_PROCESSOR_NAME] = {};
  int len = 0;
  MPI_Get_processor_name(name, &len);
  uint64_t mine = fnv1a(name, len);
  g_node_of_world_rank.assign((size_t)size, 0);
  if (MPI_SUCCESS != MPI_Allgather(&mine, 1, MPI_UINT64_T,
                                   g_node_of_world_rank.data(), 1,
                                   MPI_UINT64_T, MPI_COMM_WORLD)) {
    g_node_of_world_rank.clear();
  }
}

void p2p_finalize() {
  if (!g_active.empty()) {
    TEMPI_LOG(LOG_ERROR, "%zu point-to-point operations were not completed "
                         "before MPI_Finalize; completing them now",
              g_active.size());
    (void)post_pending_sends();
    for (auto &kv : g_active) {
      (void)kv.second->wait(MPI_STATUS_IGNORE);
    }
    g_active.clear();
  }
  g_method_cache.clear();
  pool_trim();
}

} // namespace tempi

using namespace tempi;

TEMPI_EXPORT int tempi_p2p_method(int colocated, int64_t bytes, int64_t block,
                                 int64_t stride) {
  if (bytes < 0 || block < 0) {
    return TEMPI_METHOD_UNKNOWN;
  }
  ensure_env();
  return choose_method_for(colocated != 0, bytes, block, stride);
}

TEMPI_EXPORT int tempi_set_datatype_method(int method) {
  if (method != -1 && method != TEMPI_METHOD_UNKNOWN &&
      method != TEMPI_METHOD_DEVICE && method != TEMPI_METHOD_ONESHOT &&
      method != TEMPI_METHOD_STAGED) {
    return set_error(TEMPI_ERR_ARG, "method %d", method);
  }
  ensure_env();
  const int prev = g_override_method;
  g_override_method = method;
  g_method_cache.clear();
  return prev;
}

// reference: src/send.cpp:10-17, src/internal/send.cpp:20-50,
// src/internal/sender.cpp:88-126, 150-165, 206-222
TEMPI_EXPORT int MPI_Send(const void *buf, int count, MPI_Datatype datatype,
                          int dest, int tag, MPI_Comm comm) {
  ensure_env();
  Canon cn;
  void *dev = nullptr;
  int method = TEMPI_METHOD_UNKNOWN;
  if (!device_path(buf, count, datatype, dest, &cn, &dev, &method)) {
    if (!env().disable) {
      ++counters().p2p_library;
      (void)post_pending_sends();
    }
    if (library_cannot_read(buf, count, dest)) {
      SpanImage img;
      if (img.take(buf, datatype, count, /*written=*/false)) {
        TEMPI_LOG(LOG_ERROR, "MPI_Send: %s", tempi_last_error());
        return MPI_ERR_NO_MEM;
      }
      ++counters().host_staged_calls;
      return real_mpi().Send(img.base(), count, datatype, dest, tag, comm);
    }
    return real_mpi().Send(buf, count, datatype, dest, tag, comm);
  }
  int64_t bytes = 0;
  int err = check_count(count, cn, &bytes);
  if (err) {
    return err;
  }
  if (method == TEMPI_METHOD_UNKNOWN) {
    method = choose_method(cn, bytes, comm, dest);
  }
  count_method(method);
  Staging st;
  hipStream_t s = library_stream();
  int terr = st.acquire(method, bytes);
  if (!terr && method != TEMPI_METHOD_STAGED) {
    // the pack is all there is to wait for
    terr = run_copy_wait(cn.sb, count, static_cast<const char *>(dev),
                         st.kernel_side(), TEMPI_DIR_PACK);
  } else if (!terr) {
    terr = enqueue_pack(cn, count, dev, st, s);
    if (!terr && hipSuccess != hipStreamSynchronize(s)) {
      terr = set_error(TEMPI_ERR_HIP, "pack for MPI_Send failed");
    }
  }
  if (terr) {
    st.release();
    TEMPI_LOG(LOG_ERROR, "MPI_Send: %s", tempi_last_error());
    return MPI_ERR_INTERN;
  }
  (void)post_pending_sends();
  err = real_mpi().Send(st.wire(), (int)bytes, MPI_PACKED, dest, tag, comm);
  st.release();
  (void)try_progress();
  return err;
}

// reference: src/recv.cpp:19-52, src/internal/sender.cpp:113-126, 166-177,
// 224-236
TEMPI_EXPORT int MPI_Recv(void *buf, int count, MPI_Datatype datatype,
                          int source, int tag, MPI_Comm comm,
                          MPI_Status *status) {
  ensure_env();
  Canon cn;
  void *dev = nullptr;
  int method = TEMPI_METHOD_UNKNOWN;
  if (!device_path(buf, count, datatype, source, &cn, &dev, &method)) {
    if (!env().disable) {
      ++counters().p2p_library;
      (void)post_pending_sends();
    }
    if (library_cannot_read(buf, count, source)) {
      SpanImage img;
      if (img.take(buf, datatype, count, /*written=*/true)) {
        TEMPI_LOG(LOG_ERROR, "MPI_Recv: %s", tempi_last_error());
        return MPI_ERR_NO_MEM;
      }
      ++counters().host_staged_calls;
      int err = real_mpi().Recv(img.base(), count, datatype, source, tag, comm,
                                status);
      if (err == MPI_SUCCESS && img.put_back()) {
        TEMPI_LOG(LOG_ERROR, "MPI_Recv: %s", tempi_last_error());
        err = MPI_ERR_INTERN;
      }
      return err;
    }
    return real_mpi().Recv(buf, count, datatype, source, tag, comm, status);
  }
  int64_t bytes = 0;
  int err = check_count(count, cn, &bytes);
  if (err) {
    return err;
  }
  if (method == TEMPI_METHOD_UNKNOWN) {
    method = choose_method(cn, bytes, comm, source);
  }
  count_method(method);
  Staging st;
  if (st.acquire(method, bytes)) {
    st.release();
    return MPI_ERR_INTERN;
  }
  (void)post_pending_sends();
  MPI_Status local{};
  err = real_mpi().Recv(st.wire(), (int)bytes, MPI_PACKED, source, tag, comm,
                        &local);
  if (err == MPI_SUCCESS) {
    const int n = received_objects(local, cn.size, count);
    hipStream_t s = library_stream();
    int terr = TEMPI_SUCCESS;
    if (method != TEMPI_METHOD_STAGED) {
      terr = run_copy_wait(cn.sb, n, static_cast<const char *>(dev),
                           st.kernel_side(), TEMPI_DIR_UNPACK);
    } else {
      terr = enqueue_unpack(cn, n, dev, st, (int64_t)n * cn.size, s);
      if (!terr && hipSuccess != hipStreamSynchronize(s)) {
        terr = set_error(TEMPI_ERR_HIP, "unpack for MPI_Recv failed");
      }
    }
    if (terr) {
      TEMPI_LOG(LOG_ERROR, "MPI_Recv: %s", tempi_last_error());
      err = MPI_ERR_INTERN;
    }
  }
  st.release();
  if (status != MPI_STATUS_IGNORE) {
    *status = local;
  }
  (void)try_progress();
  return err;
}

// reference: src/isend.cpp:11-16, src/internal/isend.cpp:15-45,
// src/internal/async_operation.cpp:70-206, 334-395
TEMPI_EXPORT int MPI_Isend(const void *buf, int count, MPI_Datatype datatype,
                           int dest, int tag, MPI_Comm comm,
                           MPI_Request *request) {
  ensure_env();
  Canon cn;
  void *dev = nullptr;
  int method = TEMPI_METHOD_UNKNOWN;
  if (!request ||
      !device_path(buf, count, datatype, dest, &cn, &dev, &method)) {
    if (!env().disable) {
      ++counters().p2p_library;
    }
    if (request && library_cannot_read(buf, count, dest)) {
      return start_span_op(true, buf, count, datatype, dest, tag, comm, request);
    }
    const int err = real_mpi().Isend(buf, count, datatype, dest, tag, comm, request);
    if (!env().disable) {
      (void)try_progress();
    }
    return err;
  }
  int64_t bytes = 0;
  int err = check_count(count, cn, &bytes);
  if (err) {
    return err;
  }
  if (method == TEMPI_METHOD_UNKNOWN) {
    method = choose_method(cn, bytes, comm, dest);
  }
  count_method(method);
  auto op = std::make_unique<IsendOp>(bytes, dest, tag, comm);
  if (op->start(cn, count, dev, method)) {
    TEMPI_LOG(LOG_ERROR, "MPI_Isend: %s", tempi_last_error());
    return MPI_ERR_INTERN;
  }
  *request = mint_handle(op.get());
  g_active[key_of(*request)] = std::move(op);
  (void)try_progress();
  return MPI_SUCCESS;
}

// reference: src/irecv.cpp:11-16, src/internal/irecv.cpp:15-45,
// src/internal/async_operation.cpp:209-331, 397-458
TEMPI_EXPORT int MPI_Irecv(void *buf, int count, MPI_Datatype datatype,
                           int source, int tag, MPI_Comm comm,
                           MPI_Request *request) {
  ensure_env();
  Canon cn;
  void *dev = nullptr;
  int method = TEMPI_METHOD_UNKNOWN;
  if (!request ||
      !device_path(buf, count, datatype, source, &cn, &dev, &method)) {
    if (!env().disable) {
      ++counters().p2p_library;
    }
    if (request && library_cannot_read(buf, count, source)) {
      return start_span_op(false, buf, count, datatype, source, tag, comm,
                           request);
    }
    const int err = real_mpi().Irecv(buf, count, datatype, source, tag, comm, request);
    if (!env().disable) {
      (void)try_progress();
    }
    return err;
  }
  int64_t bytes = 0;
  int err = check_count(count, cn, &bytes);
  if (err) {
    return err;
  }
  if (method == TEMPI_METHOD_UNKNOWN) {
    method = choose_method(cn, bytes, comm, source);
  }
  count_method(method);
  auto op = std::make_unique<IrecvOp>(cn, count, dev, bytes);
  err = op->start(method, source, tag, comm);
  if (err != MPI_SUCCESS) {
    return err;
  }
  *request = mint_handle(op.get());
  g_active[key_of(*request)] = std::move(op);
  (void)try_progress();
  return MPI_SUCCESS;
}

// reference: src/wait.cpp:11-16, src/internal/async_operation.cpp:460-477
TEMPI_EXPORT int MPI_Wait(MPI_Request *request, MPI_Status *status) {
  ensure_env();
  if (env().disable || !request) {
    return real_mpi().Wait(request, status);
  }
  AsyncOp *op = find_op(*request);
  if (!op) {
    (void)post_pending_sends();
    return real_mpi().Wait(request, status);
  }
  (void)post_pending_sends();
  const int err = op->wait(status);
  g_active.erase(key_of(*request));
  *request = MPI_REQUEST_NULL;
  return err;
}

// not intercepted by the reference (its version is disabled,
// async_operation.cpp:479-511); needed here so that the requests this
// library hands out can be completed together with the library's own
TEMPI_EXPORT int MPI_Waitall(int count, MPI_Request reqs[],
                             MPI_Status statuses[]) {
  ensure_env();
  if (env().disable || count <= 0 || !reqs) {
    return real_mpi().Waitall(count, reqs, statuses);
  }
  (void)post_pending_sends();
  int first_err = MPI_SUCCESS;
  std::vector<int> lib_idx;
  for (int i = 0; i < count; ++i) {
    AsyncOp *op = find_op(reqs[i]);
    if (!op) {
      lib_idx.push_back(i);
      continue;
    }
    MPI_Status *st = statuses == MPI_STATUSES_IGNORE ? MPI_STATUS_IGNORE
                                                     : &statuses[i];
    const int err = op->wait(st);
    if (err != MPI_SUCCESS && first_err == MPI_SUCCESS) {
      first_err = err;
    }
    g_active.erase(key_of(reqs[i]));
    reqs[i] = MPI_REQUEST_NULL;
  }
  if (!lib_idx.empty()) {
    std::vector<MPI_Request> lr(lib_idx.size());
    std::vector<MPI_Status> ls(lib_idx.size());
    for (size_t k = 0; k < lib_idx.size(); ++k) {
      lr[k] = reqs[lib_idx[k]];
    }
    const int err = real_mpi().Waitall((int)lr.size(), lr.data(),
                                       statuses == MPI_STATUSES_IGNORE
                                           ? MPI_STATUSES_IGNORE
                                           : ls.data());
    for (size_t k = 0; k < lib_idx.size(); ++k) {
      reqs[lib_idx[k]] = lr[k];
      if (statuses != MPI_STATUSES_IGNORE) {
        statuses[lib_idx[k]] = ls[k];
      }
    }
    if (err != MPI_SUCCESS && first_err == MPI_SUCCESS) {
      first_err = err;
    }
  }
  return first_err;
}

TEMPI_EXPORT int MPI_Test(MPI_Request *request, int *flag, MPI_Status *status) {
  ensure_env();
  if (env().disable || !request || !flag) {
    return real_mpi().Test(request, flag, status);
  }
  AsyncOp *op = find_op(*request);
  (void)try_progress();
  if (!op) {
    return real_mpi().Test(request, flag, status);
  }
  const int err = op->test(flag, status);
  if (*flag) {
    g_active.erase(key_of(*request));
    *request = MPI_REQUEST_NULL;
  }
  return err;
}
