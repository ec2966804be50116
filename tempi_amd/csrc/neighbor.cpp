// neighbor.cpp -- MPI_Neighbor_alltoallw of derived types on device buffers.
//
// reference: src/neighbor_alltoallw.cpp:11-18 and
// src/internal/neighbor_alltoallw.cpp:19-79, which post one MPI_Isend and
// one MPI_Irecv per neighbour (each with its own pack or unpack) on a tag
// reserved from the top of the tag space (src/internal/tags.cpp).
//
// Here the whole exchange is three steps, whatever the neighbour count:
// one batched kernel launch packs every outgoing block into one staging
// buffer, the MPI library moves the bytes with MPI_Neighbor_alltoallv (in
// the communicator's collective context, so no tag is taken from the
// caller), and one batched launch unpacks every incoming block.  Staging
// follows the point-to-point methods: pinned host memory written by the
// kernels (ONESHOT), device memory copied to and from pinned host memory
// (STAGED), or device memory handed to a GPU-aware library (DEVICE).  The
// batches and staging buffers of an argument set are kept (a halo exchange
// repeats the same call every step), up to a few argument sets.
//
// Calls whose types do not all have a strided form in this library, or
// whose buffers are not device-accessible, go to the library unchanged.  A
// kept argument set is dropped when one of its types or its communicator is
// freed (their handles may be reused).
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <list>
#include <memory>
#include <vector>

namespace tempi {

namespace {

struct Side {
  std::vector<int64_t> bytes; // packed bytes per neighbour
  std::vector<int> displs;    // their offsets in the staging buffer
  int64_t total = 0;
};

struct Exchange {
  // the arguments this entry was built for
  MPI_Comm comm;
  const void *sendbuf;
  void *recvbuf;
  std::vector<int> scounts, rcounts;
  std::vector<MPI_Aint> sdispls, rdispls;
  std::vector<MPI_Datatype> stypes, rtypes;
  int method = TEMPI_METHOD_ONESHOT;
  // what was built
  Side out, in;
  void *send_host = nullptr, *send_dev = nullptr; // staging
  void *recv_host = nullptr, *recv_dev = nullptr;
  tempi_batch *pack = nullptr;
  tempi_batch *unpack = nullptr;

  ~Exchange() {
    if (pack) {
      tempi_batch_destroy(pack);
    }
    if (unpack) {
      tempi_batch_destroy(unpack);
    }
    pool_release(true, send_host);
    pool_release(false, send_dev);
    pool_release(true, recv_host);
    pool_release(false, recv_dev);
  }

  bool same(MPI_Comm c, const void *sb, void *rb, const int *sc,
            const MPI_Aint *sd, const MPI_Datatype *st, const int *rc,
            const MPI_Aint *rd, const MPI_Datatype *rt, int outdeg,
            int indeg) const {
    if (c != comm || sb != sendbuf || rb != recvbuf ||
        (int)scounts.size() != outdeg || (int)rcounts.size() != indeg) {
      return false;
    }
    for (int i = 0; i < outdeg; ++i) {
      if (sc[i] != scounts[i] || sd[i] != sdispls[i] || st[i] != stypes[i]) {
        return false;
      }
    }
    for (int i = 0; i < indeg; ++i) {
      if (rc[i] != rcounts[i] || rd[i] != rdispls[i] || rt[i] != rtypes[i]) {
        return false;
      }
    }
    return true;
  }
};

constexpr size_t kMaxEntries = 8;
std::list<std::unique_ptr<Exchange>> g_entries; // most recent first

// the side's records; false when any type has no strided form
bool lower_side(int n, const int *counts, const MPI_Datatype *types,
                std::vector<Canon> *cns, Side *side) {
  int64_t off = 0;
  for (int i = 0; i < n; ++i) {
    Canon cn;
    if (counts[i] < 0 || !type_lookup(types[i], &cn) || !cn.ok ||
        cn.size <= 0) {
      return false;
    }
    const int64_t b = (int64_t)counts[i] * cn.size;
    side->bytes.push_back(b);
    side->displs.push_back((int)off);
    off += b;
    if (off > INT32_MAX) {
      return false;
    }
    cns->push_back(cn);
  }
  side->total = off;
  return true;
}

int pick_method(int64_t bytes, const std::vector<Canon> &cns) {
  const int m = forced_datatype_method();
  if (m != TEMPI_METHOD_UNKNOWN) {
    return m;
  }
  int64_t bl = 512;
  for (const Canon &c : cns) {
    bl = std::min<int64_t>(bl, std::max<int64_t>(1, c.sb.counts[0]));
  }
  int best = env().gpu_aware_mpi ? TEMPI_METHOD_DEVICE : TEMPI_METHOD_ONESHOT;
  double best_s = INFINITY;
  for (int c : {TEMPI_METHOD_DEVICE, TEMPI_METHOD_ONESHOT, TEMPI_METHOD_STAGED}) {
    if (c == TEMPI_METHOD_DEVICE && !env().gpu_aware_mpi) {
      continue;
    }
    const double s = perf_model_seconds(c, true, bytes, bl);
    if (s < best_s) {
      best_s = s;
      best = c;
    }
  }
  return best;
}

int make_batch(int n, const std::vector<Canon> &cns, const char *base,
               const MPI_Aint *displs, const int *counts, const Side &side,
               char *staging, int dir, tempi_batch **out) {
  std::vector<tempi_batch_request> reqs((size_t)std::max(n, 1));
  for (int i = 0; i < n; ++i) {
    tempi_batch_request &r = reqs[i];
    r.sb = cns[i].sb;
    r.strided = base + displs[i];
    r.packed = staging + side.displs[i];
    r.count = counts[i];
    r.direction = dir;
    r.variant = -1;
  }
  return tempi_batch_create(reqs.data(), n, env().select, out);
}

// build the batches and staging buffers of one argument set
int build(Exchange &e, int outdeg, int indeg,
          const std::vector<Canon> &scn, const std::vector<Canon> &rcn) {
  const bool host = e.method != TEMPI_METHOD_DEVICE;
  const bool dev = e.method != TEMPI_METHOD_ONESHOT;
  const size_t sb = (size_t)std::max<int64_t>(e.out.total, 1);
  const size_t rb = (size_t)std::max<int64_t>(e.in.total, 1);
  if (host) {
    e.send_host = pool_acquire(true, sb);
    e.recv_host = pool_acquire(true, rb);
    if (!e.send_host || !e.recv_host) {
      return TEMPI_ERR_NOMEM;
    }
  }
  if (dev) {
    e.send_dev = pool_acquire(false, sb);
    e.recv_dev = pool_acquire(false, rb);
    if (!e.send_dev || !e.recv_dev) {
      return TEMPI_ERR_NOMEM;
    }
  }
  // the kernels write the host buffer itself for ONESHOT, device memory
  // otherwise (the batch code resolves mapped host pointers)
  char *ks = static_cast<char *>(e.method == TEMPI_METHOD_ONESHOT ? e.send_host
                                                                   : e.send_dev);
  char *kr = static_cast<char *>(e.method == TEMPI_METHOD_ONESHOT ? e.recv_host
                                                                   : e.recv_dev);
  int err = make_batch(outdeg, scn, static_cast<const char *>(e.sendbuf),
                       e.sdispls.data(), e.scounts.data(), e.out, ks,
                       TEMPI_DIR_PACK, &e.pack);
  if (!err) {
    err = make_batch(indeg, rcn, static_cast<const char *>(e.recvbuf),
                     e.rdispls.data(), e.rcounts.data(), e.in, kr,
                     TEMPI_DIR_UNPACK, &e.unpack);
  }
  return err;
}

int run(Exchange &e, MPI_Comm comm) {
  hipStream_t s = library_stream();
  int err = tempi_batch_launch(e.pack, s);
  if (!err && e.method == TEMPI_METHOD_STAGED && e.out.total > 0 &&
      hipSuccess != hipMemcpyAsync(e.send_host, e.send_dev,
                                   (size_t)e.out.total, hipMemcpyDeviceToHost,
                                   s)) {
    err = set_error(TEMPI_ERR_HIP, "staging copy failed");
  }
  if (!err && hipSuccess != hipStreamSynchronize(s)) {
    err = set_error(TEMPI_ERR_HIP, "pack for MPI_Neighbor_alltoallw failed");
  }
  if (err) {
    return MPI_ERR_INTERN;
  }
  std::vector<int> sc(e.out.bytes.begin(), e.out.bytes.end());
  std::vector<int> rc(e.in.bytes.begin(), e.in.bytes.end());
  const bool d = e.method == TEMPI_METHOD_DEVICE;
  int merr = MPI_Neighbor_alltoallv(d ? e.send_dev : e.send_host, sc.data(),
                                    e.out.displs.data(), MPI_BYTE,
                                    d ? e.recv_dev : e.recv_host, rc.data(),
                                    e.in.displs.data(), MPI_BYTE, comm);
  if (merr != MPI_SUCCESS) {
    return merr;
  }
  if (e.method == TEMPI_METHOD_STAGED && e.in.total > 0 &&
      hipSuccess != hipMemcpyAsync(e.recv_dev, e.recv_host, (size_t)e.in.total,
                                   hipMemcpyHostToDevice, s)) {
    set_error(TEMPI_ERR_HIP, "staging copy failed");
    return MPI_ERR_INTERN;
  }
  err = tempi_batch_launch(e.unpack, s);
  if (!err && hipSuccess != hipStreamSynchronize(s)) {
    err = set_error(TEMPI_ERR_HIP, "unpack for MPI_Neighbor_alltoallw failed");
  }
  return err ? MPI_ERR_INTERN : MPI_SUCCESS;
}

} // namespace

void neighbor_finalize() { g_entries.clear(); }

void neighbor_forget_type(MPI_Datatype t) {
  g_entries.remove_if([&](const std::unique_ptr<Exchange> &e) {
    for (MPI_Datatype x : e->stypes) {
      if (x == t) {
        return true;
      }
    }
    for (MPI_Datatype x : e->rtypes) {
      if (x == t) {
        return true;
      }
    }
    return false;
  });
}

void neighbor_forget_comm(MPI_Comm c) {
  g_entries.remove_if(
      [&](const std::unique_ptr<Exchange> &e) { return e->comm == c; });
}

} // namespace tempi

using namespace tempi;

TEMPI_EXPORT int MPI_Neighbor_alltoallw(
    const void *sendbuf, const int sendcounts[], const MPI_Aint sdispls[],
    const MPI_Datatype sendtypes[], void *recvbuf, const int recvcounts[],
    const MPI_Aint rdispls[], const MPI_Datatype recvtypes[], MPI_Comm comm) {
  ensure_env();
  auto library = [&]() {
    ++counters().p2p_library;
    return real_mpi().Neighbor_alltoallw(sendbuf, sendcounts, sdispls,
                                         sendtypes, recvbuf, recvcounts,
                                         rdispls, recvtypes, comm);
  };
  if (env().disable) {
    return real_mpi().Neighbor_alltoallw(sendbuf, sendcounts, sdispls,
                                         sendtypes, recvbuf, recvcounts,
                                         rdispls, recvtypes, comm);
  }
  int indeg = 0, outdeg = 0, weighted = 0;
  if (MPI_SUCCESS !=
      MPI_Dist_graph_neighbors_count(comm, &indeg, &outdeg, &weighted)) {
    return library(); // not a distributed-graph communicator
  }
  void *sd = nullptr, *rd = nullptr;
  if ((outdeg > 0 && !device_address(sendbuf, &sd)) ||
      (indeg > 0 && !device_address(recvbuf, &rd)) || ensure_runtime()) {
    if (!env().gpu_aware_mpi && ((outdeg > 0 && device_only(sendbuf)) ||
                                 (indeg > 0 && device_only(recvbuf)))) {
      // one side in device memory, the other not handled here: the MPI
      // library cannot read or write device memory
      set_error(TEMPI_ERR_UNSUPPORTED, "MPI_Neighbor_alltoallw with one buffer "
                                       "in device memory and one in host memory");
      TEMPI_LOG(LOG_ERROR, "%s", tempi_last_error());
      MPI_Comm_call_errhandler(comm, MPI_ERR_BUFFER);
      return MPI_ERR_BUFFER;
    }
    return library();
  }
  // the cached argument set, if this call repeats one
  for (auto it = g_entries.begin(); it != g_entries.end(); ++it) {
    if ((*it)->same(comm, sendbuf, recvbuf, sendcounts, sdispls, sendtypes,
                    recvcounts, rdispls, recvtypes, outdeg, indeg)) {
      g_entries.splice(g_entries.begin(), g_entries, it);
      return run(*g_entries.front(), comm);
    }
  }
  // device buffers the MPI library cannot read: the exchange runs here or
  // fails with an error, never with device pointers in the library
  const bool device_only_bufs =
      !env().gpu_aware_mpi && ((outdeg > 0 && device_only(sendbuf)) ||
                               (indeg > 0 && device_only(recvbuf)));
  auto unsupported = [&](const char *why) {
    set_error(TEMPI_ERR_UNSUPPORTED,
              "MPI_Neighbor_alltoallw of device buffers: %s", why);
    TEMPI_LOG(LOG_ERROR, "%s", tempi_last_error());
    MPI_Comm_call_errhandler(comm, MPI_ERR_TYPE);
    return MPI_ERR_TYPE;
  };
  std::vector<Canon> scn, rcn;
  auto e = std::make_unique<Exchange>();
  if (!lower_side(outdeg, sendcounts, sendtypes, &scn, &e->out) ||
      !lower_side(indeg, recvcounts, recvtypes, &rcn, &e->in)) {
    if (device_only_bufs) {
      return unsupported("a type has no strided form");
    }
    return library();
  }
  e->comm = comm;
  e->sendbuf = sd;
  e->recvbuf = rd;
  e->scounts.assign(sendcounts, sendcounts + outdeg);
  e->rcounts.assign(recvcounts, recvcounts + indeg);
  e->sdispls.assign(sdispls, sdispls + outdeg);
  e->rdispls.assign(rdispls, rdispls + indeg);
  e->stypes.assign(sendtypes, sendtypes + outdeg);
  e->rtypes.assign(recvtypes, recvtypes + indeg);
  std::vector<Canon> all(scn);
  all.insert(all.end(), rcn.begin(), rcn.end());
  e->method = pick_method(std::max(e->out.total, e->in.total), all);
  if (build(*e, outdeg, indeg, scn, rcn)) {
    if (device_only_bufs) {
      return unsupported(tempi_last_error());
    }
    TEMPI_LOG(LOG_WARN, "MPI_Neighbor_alltoallw: %s; library path",
              tempi_last_error());
    return library();
  }
  // the cache is keyed on the caller's addresses
  e->sendbuf = sendbuf;
  e->recvbuf = recvbuf;
  switch (e->method) {
  case TEMPI_METHOD_DEVICE:
    ++counters().p2p_device;
    break;
  case TEMPI_METHOD_ONESHOT:
    ++counters().p2p_oneshot;
    break;
  default:
    ++counters().p2p_staged;
  }
  g_entries.push_front(std::move(e));
  if (g_entries.size() > kMaxEntries) {
    g_entries.pop_back();
  }
  return run(*g_entries.front(), comm);
}

// reference: src/comm_free.cpp (drops the communicator's cached state)
TEMPI_EXPORT int MPI_Comm_free(MPI_Comm *comm) {
  ensure_env();
  if (!env().disable && comm) {
    neighbor_forget_comm(*comm);
  }
  return real_mpi().Comm_free(comm);
}
