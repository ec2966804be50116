// batch.hip -- many strided copies in one persistent kernel launch.
//
// The reference launches one kernel per MPI_Pack and synchronizes
// (src/internal/packer_2d.cu:101-108).  A batch instead resolves every
// request once, on the host, into the same per-copy parameters the single
// kernels take, uploads them as a device-resident table, and splits each
// request into tiles of one workgroup pass (256 threads, four loads in
// flight per lane: 16 KiB for 16-byte words).  A persistent grid (up to 8
// workgroups per CU) drains the tile queue with a static, XCD-aware
// assignment; each tile's request comes from a tile-to-request table, and a
// switch runs that request's kernel variant.  Heterogeneous shapes -- 12-B
// halo rows next to 2-KiB face rows -- share one launch.
//
// A first version handed tiles out with one atomic counter; every tile then
// paid for a device-scope atomic serialized with all the others (about 30 ns
// each on MI355X), which dominated small and large batches alike.  The
// static assignment has no shared state, so launches need no reset and
// replay from a captured hipGraph unchanged.  Requests inside a batch must
// not write bytes another request reads or writes.
#include "copy_core.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

namespace tempi {

namespace {

struct BatchItem {
  int32_t kind;            // 4 * variant + 2 * aligned + direction
  uint32_t units_per_tile; // units of one block-pass: groups per block * K
  uint32_t ntiles;         // tiles of this request
  uint32_t pad;
  KParams p;               // after the kind: both in the item's first lines
};

constexpr uint32_t kXcds = 8;

// One block-pass: tile `l` of a request split into `nt` tiles moves K
// blocks of GROUPS units, nt * GROUPS units apart -- units l*GROUPS + g +
// k*nt*GROUPS, every unit of the request exactly once over the nt tiles.
// Spreading a lane's K units across the request rather than over K
// neighbouring blocks measured about 0.2 us faster per launch on MI355X for
// 12-B rows 2 KiB apart (262 144 rows), the same as the single-copy kernels'
// grid-stride order.
template <int P, int W, int U, bool ALIGNED, int DIR, bool NTS, bool NTL>
__device__ __forceinline__ void copy_tile(const KParams &p, uint32_t l,
                                          uint32_t nt) {
  constexpr uint32_t GROUPS = kBlock / P;
  copy_units<P, W, U, ALIGNED, DIR, NTS, NTL>(p, l * GROUPS + threadIdx.x / P,
                                              nt * GROUPS, p.units,
                                              threadIdx.x % P);
}

__device__ __forceinline__ void run_tile(int kind, const KParams &p,
                                         uint32_t l, uint32_t nt) {
  switch (kind) {
#define TEMPI_BCASE(I, P, W, U, NTS, NTL)                                      \
  case 4 * I + 0:                                                              \
    copy_tile<P, W, U, false, 0, NTS, NTL>(p, l, nt);                          \
    break;                                                                     \
  case 4 * I + 1:                                                              \
    copy_tile<P, W, U, false, 1, NTS, NTL>(p, l, nt);                          \
    break;                                                                     \
  case 4 * I + 2:                                                              \
    copy_tile<P, W, U, true, 0, NTS, NTL>(p, l, nt);                           \
    break;                                                                     \
  case 4 * I + 3:                                                              \
    copy_tile<P, W, U, true, 1, NTS, NTL>(p, l, nt);                           \
    break;
    TEMPI_VARIANTS(TEMPI_BCASE)
#undef TEMPI_BCASE
  default:
    break;
  }
}

// One tile: the request it belongs to and its index within the request.
struct Tile {
  uint32_t item;
  uint32_t tile;
};

// Workgroups are dispatched round-robin over the XCDs (workgroup b runs on
// XCD b % 8).  The host orders the tile table so that XCD x owns the
// contiguous range [xcd_begin[x], xcd_begin[x+1]), which holds an eighth of
// every request's tiles: each XCD gets an equal share of each shape (costly
// 12-B-row tiles and cheap 2-KiB-row tiles alike), and neighbouring tiles --
// whose rows can share partial cache lines -- stay in one XCD's L2.  The
// assignment is static: no atomics, no state carried between launches.
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(5)))
    k_batch(const BatchItem *__restrict__ items,
            const Tile *__restrict__ tiles,
            const uint32_t *__restrict__ xcd_begin, uint32_t ntiles,
            int xcd_split) {
  uint32_t t, end, step;
  if (xcd_split) {
    const uint32_t xcd = blockIdx.x % kXcds;
    step = gridDim.x / kXcds;
    t = xcd_begin[xcd] + blockIdx.x / kXcds;
    end = xcd_begin[xcd + 1];
  } else {
    step = gridDim.x;
    t = blockIdx.x;
    end = ntiles;
  }
  for (; t < end; t += step) {
    // uniform addresses: the tile and its request's parameters are read
    // with scalar loads in place (a private copy of the parameter struct
    // would live in scratch memory)
    const Tile tl = tiles[t];
    const uint32_t r = __builtin_amdgcn_readfirstlane(tl.item);
    const uint32_t l = __builtin_amdgcn_readfirstlane(tl.tile);
    const BatchItem &it = items[r];
    run_tile(it.kind, it.p, l, it.ntiles);
  }
}

// Batches of at most kSmallItems requests pass their tile index in the
// kernel arguments.  k_batch's workgroups wait on three dependent loads
// before their first data load (XCD range, tile record, request parameters);
// here the tile follows from the workgroup index by arithmetic on the
// argument block, leaving one dependent load (the request's parameters).
// The parameters stay in device memory: a back-to-back launch with 2 KiB of
// kernel arguments costs about 0.5 us more than one with a few hundred bytes
// on MI355X, more than the loads it would save.
//
// Padded tile numbering: request r owns tiles [base[r], base[r+1]), a range
// that starts at a multiple of 8 and has 8 * share[r] entries.  Tile T runs
// on XCD T % 8 (the grid is a multiple of 8, so T % 8 == blockIdx.x % 8 on
// every pass), and XCD x takes the request's tiles
// [x * share[r], (x + 1) * share[r]): each XCD gets a contiguous eighth of
// every request, as in k_batch.  Padding entries past the request's last
// tile do nothing.
constexpr int kSmallItems = 8;

struct SmallIndex {
  uint32_t n;                     // requests
  uint32_t end;                   // padded tiles of the whole batch
  uint32_t base[kSmallItems];     // first padded tile of each request
  uint32_t share[kSmallItems];    // tiles of each request per XCD
  uint32_t ntiles[kSmallItems];   // tiles of each request
};

// The index is read through the kernel-argument segment pointer (it is the
// first argument, at offset 0) into registers, every entry unconditionally
// and before the loop: the entries past the last request hold `end`, so no
// test of the request count is needed, and the loads issue together.  (With
// the count tested per entry, or with entries read at a run-time index, the
// compiler chained one scalar load and wait per request in front of the
// first data load, or copied the parameter to scratch memory.)
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(5)))
    k_batch_small(const SmallIndex arg, const BatchItem *__restrict__ items) {
  (void)arg;
  const SmallIndex *ix =
      (const SmallIndex *)__builtin_amdgcn_kernarg_segment_ptr();
  const uint32_t end = ix->end;
  uint32_t base[kSmallItems], share[kSmallItems], ntiles[kSmallItems];
#pragma unroll
  for (int i = 0; i < kSmallItems; ++i) {
    base[i] = ix->base[i];
    share[i] = ix->share[i];
    ntiles[i] = ix->ntiles[i];
  }
  for (uint32_t t = blockIdx.x; t < end; t += gridDim.x) {
    // request r: the last whose range starts at or before t
    uint32_t r = 0, b = base[0], sh = share[0], nt = ntiles[0];
#pragma unroll
    for (int i = 1; i < kSmallItems; ++i) {
      const bool in = t >= base[i];
      r = in ? (uint32_t)i : r;
      b = in ? base[i] : b;
      sh = in ? share[i] : sh;
      nt = in ? ntiles[i] : nt;
    }
    const uint32_t q = t - b;
    const uint32_t l = (q % kXcds) * sh + q / kXcds;
    if (l < nt) {
      const BatchItem &it = items[r];
      run_tile(it.kind, it.p, l, nt);
    }
  }
}

} // namespace

} // namespace tempi

using namespace tempi;

namespace tempi {
namespace {

// What one launch of a batch needs: the grid, the kernel, and where its
// tables sit in device memory.
struct BatchLaunch {
  int grid = 0;
  uint32_t ntiles = 0;
  int xcd_split = 0;
  bool small = false; // k_batch_small with `ix`
  SmallIndex ix{};
  const BatchItem *items = nullptr;
  const Tile *tiles = nullptr;        // tile records, grouped by XCD
  const uint32_t *xcd_begin = nullptr; // first tile of each XCD (kXcds + 1)
};

void launch_kernel(const BatchLaunch &l, hipStream_t s) {
  if (l.small) {
    hipLaunchKernelGGL(k_batch_small, dim3(l.grid), dim3(kBlock), 0, s, l.ix,
                       l.items);
  } else {
    hipLaunchKernelGGL(k_batch, dim3(l.grid), dim3(kBlock), 0, s, l.items,
                       l.tiles, l.xcd_begin, l.ntiles, l.xcd_split);
  }
}

// One request as the batch resolves it.
struct Req {
  Geom g;
  const char *ds = nullptr; // strided buffer origin (device address)
  char *dp = nullptr;       // first packed byte (device address)
  int dir = 0;
  int variant = 0;
};

// The host image of a batch's device tables (request parameters, tile
// records, XCD range bounds; each section 256-B aligned) and its launch
// shape, with the table pointers left as offsets into the image.
struct BatchPlan {
  std::vector<char> image;
  size_t tl_off = 0, xb_off = 0;
  BatchLaunch launch;
  int n = 0;
};

int plan_batch(const std::vector<Req> &rs, int cu_count, BatchPlan *pl) {
  BatchLaunch &L = pl->launch;
  L = BatchLaunch{};
  std::vector<BatchItem> items;
  std::vector<uint32_t> item_tiles; // tiles of each item
  uint64_t tiles = 0;
  for (const Req &r : rs) {
    BatchItem it{};
    bool aligned = false;
    int err = fill_kparams(r.g, r.ds, r.dp, r.dir, r.variant, &it.p, &aligned);
    if (err) {
      return err;
    }
    it.kind = 4 * r.variant + 2 * (aligned ? 1 : 0) + r.dir;
    const VariantDesc &v = kVariants[r.variant];
    it.units_per_tile = (uint32_t)(kBlock / v.P) * units_in_flight(v.U);
    const uint64_t t = ((uint64_t)it.p.units + it.units_per_tile - 1) /
                       it.units_per_tile;
    tiles += t;
    if (tiles >= (1ull << 31)) {
      return set_error(TEMPI_ERR_UNSUPPORTED, "batch exceeds 2^31 tiles");
    }
    it.ntiles = (uint32_t)t;
    item_tiles.push_back((uint32_t)t);
    items.push_back(it);
  }
  pl->n = (int)items.size();
  L.ntiles = (uint32_t)tiles;
  // persistent grid: up to 8 workgroups per CU, a multiple of the XCD count
  // when the tiles are split by XCD
  const int64_t cap = (int64_t)cu_count * 8;
  if (!items.empty() && items.size() <= (size_t)kSmallItems) {
    // tile index in the kernel arguments; only the parameters go to the
    // device table
    SmallIndex &ix = L.ix;
    ix.n = (uint32_t)items.size();
    uint32_t base = 0;
    for (size_t r = 0; r < items.size(); ++r) {
      ix.base[r] = base;
      ix.share[r] = (item_tiles[r] + kXcds - 1) / kXcds;
      ix.ntiles[r] = item_tiles[r];
      base += kXcds * ix.share[r];
    }
    for (size_t r = items.size(); r < (size_t)kSmallItems; ++r) {
      ix.base[r] = base;
    }
    ix.end = base;
    L.small = true;
    L.grid = (int)std::min<int64_t>((int64_t)base, cap / kXcds * kXcds);
  } else if ((int64_t)tiles >= 8 * (int64_t)kXcds) {
    L.xcd_split = 1;
    L.grid = (int)(std::min<int64_t>((int64_t)tiles, cap) / kXcds * kXcds);
  } else {
    L.grid = (int)std::max<int64_t>(1, (int64_t)tiles);
  }
  // tile records grouped by XCD: XCD x takes the x-th eighth of every
  // item's tiles (a single group when the grid is not split)
  const int groups = L.small ? 0 : L.xcd_split ? (int)kXcds : 1;
  std::vector<Tile> tile_list;
  tile_list.reserve(L.small ? 0 : tiles);
  std::vector<uint32_t> xcd_begin(kXcds + 1, (uint32_t)tiles);
  for (int x = 0; x < groups; ++x) {
    xcd_begin[x] = (uint32_t)tile_list.size();
    for (size_t r = 0; r < items.size(); ++r) {
      const uint64_t n_r = item_tiles[r];
      const uint64_t lo = n_r * x / groups, hi = n_r * (x + 1) / groups;
      for (uint64_t l = lo; l < hi; ++l) {
        tile_list.push_back(Tile{(uint32_t)r, (uint32_t)l});
      }
    }
  }
  const size_t items_bytes = items.size() * sizeof(BatchItem);
  pl->tl_off = (items_bytes + 255) / 256 * 256;
  pl->xb_off =
      pl->tl_off + (tile_list.size() * sizeof(Tile) + 255) / 256 * 256;
  pl->image.assign(pl->xb_off + 256, 0);
  if (!items.empty()) {
    memcpy(pl->image.data(), items.data(), items_bytes);
  }
  if (!tile_list.empty()) {
    memcpy(pl->image.data() + pl->tl_off, tile_list.data(),
           tile_list.size() * sizeof(Tile));
  }
  memcpy(pl->image.data() + pl->xb_off, xcd_begin.data(),
         xcd_begin.size() * 4);
  return TEMPI_SUCCESS;
}

// the launch of a plan whose image was copied to device address `dev`
BatchLaunch placed(const BatchPlan &pl, char *dev) {
  BatchLaunch L = pl.launch;
  L.items = reinterpret_cast<const BatchItem *>(dev);
  L.tiles = reinterpret_cast<const Tile *>(dev + pl.tl_off);
  L.xcd_begin = reinterpret_cast<const uint32_t *>(dev + pl.xb_off);
  return L;
}

// Check and resolve the caller's requests: geometry, device addresses and,
// for a request that leaves it to the library (variant -1), the variant
// `pick` returns.  Requests that move nothing are left out; `where` gets
// the caller's index of each resolved request.
template <class Pick>
int resolve_requests(const tempi_batch_request *reqs, int64_t n, Pick &&pick,
                     std::vector<Req> *rs, std::vector<int64_t> *where,
                     int64_t *bytes) {
  // every request is checked before the device is touched
  int err = TEMPI_SUCCESS;
  std::vector<Geom> geoms((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    if ((err = check_direction(reqs[i].direction)) ||
        (err = make_geom(reqs[i].sb, reqs[i].count, &geoms[(size_t)i]))) {
      return err;
    }
  }
  if ((err = ensure_runtime())) {
    return err;
  }
  *bytes = 0;
  for (int64_t i = 0; i < n; ++i) {
    const tempi_batch_request &r = reqs[i];
    Req q;
    q.g = geoms[(size_t)i];
    if (q.g.rows == 0 || q.g.L == 0) {
      continue;
    }
    if (q.g.L > (1ll << 30)) {
      return set_error(TEMPI_ERR_UNSUPPORTED,
                       "request %lld: rows over 1 GiB are not batched",
                       (long long)i);
    }
    void *ds = nullptr, *dp = nullptr;
    if (!device_address(r.strided, &ds) || !device_address(r.packed, &dp)) {
      return set_error(TEMPI_ERR_ARG,
                       "request %lld: buffers must be device-accessible",
                       (long long)i);
    }
    q.ds = static_cast<const char *>(ds);
    q.dp = static_cast<char *>(dp);
    q.dir = r.direction;
    q.variant = r.variant;
    if (q.variant < 0 || q.variant >= kNumVariants) {
      q.variant = pick(q);
    }
    *bytes += q.g.L * q.g.rows;
    rs->push_back(q);
    where->push_back(i);
  }
  return TEMPI_SUCCESS;
}

} // namespace
} // namespace tempi

struct tempi_batch {
  int64_t requests = 0;
  int64_t bytes = 0;
  int n = 0;           // items with work
  char *dev = nullptr; // the tables, one allocation
  BatchLaunch launch;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<int> variants;
};

static int batch_launch_on(tempi_batch *b, hipStream_t s) {
  if (b->launch.ntiles == 0) {
    return TEMPI_SUCCESS;
  }
  launch_kernel(b->launch, s);
  TEMPI_HIP_CHECK(hipGetLastError());
  return TEMPI_SUCCESS;
}

namespace {

// The device tables and launch shape of a batch of resolved requests.
int build_batch(tempi_batch *b, const std::vector<Req> &rs) {
  hipStream_t s = library_stream();
  BatchPlan pl;
  int err = plan_batch(rs, device_cu_count(), &pl);
  if (err) {
    return err;
  }
  b->n = pl.n;
  if (hipSuccess != hipMalloc(&b->dev, pl.image.size())) {
    (void)hipGetLastError();
    b->dev = nullptr;
    return set_error(TEMPI_ERR_NOMEM, "batch table allocation failed");
  }
  b->launch = placed(pl, b->dev);
  // one upload, outside any capture
  if (hipSuccess != hipMemcpyAsync(b->dev, pl.image.data(), pl.image.size(),
                                   hipMemcpyHostToDevice, s) ||
      hipSuccess != hipStreamSynchronize(s)) {
    (void)hipGetLastError();
    (void)hipFree(b->dev);
    b->dev = nullptr;
    return set_error(TEMPI_ERR_HIP, "batch table upload failed");
  }
  return TEMPI_SUCCESS;
}

// Tiles of a resolved request (one workgroup pass each).
uint64_t tiles_of(const Req &q) {
  const VariantDesc &v = kVariants[q.variant];
  const int64_t per_unit = (int64_t)v.P * v.U;
  const int64_t nseg = std::max<int64_t>(1, (q.g.L / v.W + per_unit - 1) / per_unit);
  const uint64_t units = (uint64_t)(q.g.rows * nseg);
  const uint64_t per_tile = (uint64_t)(kBlock / v.P) * units_in_flight(v.U);
  return (units + per_tile - 1) / per_tile;
}

// The default policy's batch rule.  The learned selector ranks a copy's
// variants by the time of the copy alone, where a copy under a megabyte is
// latency-bound and variants that split it into more tiles win.  In a batch
// with more tiles than the grid runs at once, throughput decides instead:
// on MI355X the 1 664-request halo batch (all requests under 1 MiB) packed in
// 129 us with the closed-form rule's variants against 139 us with the
// selector's, and unpacked in 174 us against 267 us, while the 4 096 random
// requests (up to 16 MiB) packed in 8.6 ms with the selector's against 9.2 ms.
// So in a batch that saturates the grid, requests under 1 MiB whose variant
// was left to the library take the rule; larger requests, and small batches
// (the headline's three faces: 9.26 us with the selector's picks against
// 9.91 us with the rule's), keep the selector's.
void saturated_batch_policy(const tempi_batch_request *reqs,
                            const std::vector<int64_t> &where,
                            std::vector<Req> *rs) {
  uint64_t tiles = 0;
  for (const Req &q : *rs) {
    tiles += tiles_of(q);
  }
  if (tiles <= (uint64_t)device_cu_count() * 8) {
    return;
  }
  for (size_t k = 0; k < rs->size(); ++k) {
    Req &q = (*rs)[k];
    const int asked = reqs[where[k]].variant;
    if ((asked < 0 || asked >= kNumVariants) && q.g.L * q.g.rows < (1ll << 20)) {
      q.variant = heuristic_variant(q.g, q.ds, q.dp, q.dir);
    }
  }
}

} // namespace

TEMPI_EXPORT int tempi_batch_create(const tempi_batch_request *reqs, int64_t n,
                                    int method, tempi_batch **out) {
  if (!out || (n > 0 && !reqs) || n < 0) {
    return set_error(TEMPI_ERR_ARG, "bad argument");
  }
  *out = nullptr;
  std::unique_ptr<tempi_batch> owner(new tempi_batch());
  tempi_batch *b = owner.get();
  b->requests = n;
  b->variants.assign((size_t)n, -1);
  std::vector<Req> rs;
  std::vector<int64_t> where;
  // the per-copy policy first (in small batches the learned selector's
  // picks for requests under 1 MiB ran the faces of 128^3 - 256^3 float
  // grids 5-28 % faster than the closed-form rule's), then the batch rule
  // for batches that saturate the grid (saturated_batch_policy).  The
  // selector evaluates all the batch's new shapes in one inference first:
  // one launch per shape took 0.14 s to create the 4 096-request batch.
  int err = resolve_requests(
      reqs, n, [](const Req &) { return -1; }, &rs, &where, &b->bytes);
  if (err) {
    return err;
  }
  if (method == TEMPI_SELECT_MODEL ||
      (method == TEMPI_SELECT_DEFAULT && model_available())) {
    std::vector<CopyShape> shapes;
    for (const Req &q : rs) {
      if (q.variant < 0) {
        shapes.push_back({q.g, q.ds, q.dp, q.dir});
      }
    }
    model_prefetch(shapes);
  }
  for (Req &q : rs) {
    if (q.variant < 0) {
      q.variant = choose(q.g, q.ds, q.dp, q.dir, method, library_stream(),
                         /*in_place=*/false)
                      .variant;
    }
  }
  if (method == TEMPI_SELECT_DEFAULT) {
    saturated_batch_policy(reqs, where, &rs);
  }
  for (size_t k = 0; k < rs.size(); ++k) {
    b->variants[(size_t)where[k]] = rs[k].variant;
  }
  if ((err = build_batch(b, rs))) {
    return err;
  }
  *out = owner.release();
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_batch_launch(tempi_batch *b, void *stream) {
  if (!b) {
    return set_error(TEMPI_ERR_ARG, "null batch");
  }
  TraceRange trace("tempi_batch_launch");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : library_stream();
  int err = batch_launch_on(b, s);
  if (!err) {
    ++counters().batch_launches;
  }
  return err;
}

namespace {

// A graph of the launches of `n` batches, each depending on the one before
// (kernel nodes added directly: no stream capture).  The batches' device
// tables must outlive the graph.
int build_graph(tempi_batch *const *bs, int64_t n, hipGraph_t *graph,
                hipGraphExec_t *exec, int64_t *nodes) {
  *graph = nullptr;
  *exec = nullptr;
  *nodes = 0;
  TEMPI_HIP_CHECK(hipGraphCreate(graph, 0));
  hipGraphNode_t prev = nullptr;
  int err = TEMPI_SUCCESS;
  for (int64_t i = 0; i < n && !err; ++i) {
    const BatchLaunch &l = bs[i]->launch;
    if (l.ntiles == 0) {
      continue;
    }
    hipKernelNodeParams kp{};
    void *args_small[] = {const_cast<SmallIndex *>(&l.ix),
                          const_cast<const BatchItem **>(&l.items)};
    uint32_t ntiles = l.ntiles;
    int xcd_split = l.xcd_split;
    void *args_big[] = {const_cast<const BatchItem **>(&l.items),
                        const_cast<const Tile **>(&l.tiles),
                        const_cast<const uint32_t **>(&l.xcd_begin), &ntiles,
                        &xcd_split};
    kp.func = l.small ? reinterpret_cast<void *>(&k_batch_small)
                      : reinterpret_cast<void *>(&k_batch);
    kp.gridDim = dim3((unsigned)l.grid);
    kp.blockDim = dim3(kBlock);
    kp.sharedMemBytes = 0;
    kp.kernelParams = l.small ? args_small : args_big;
    kp.extra = nullptr;
    hipGraphNode_t node = nullptr;
    const hipError_t e = hipGraphAddKernelNode(&node, *graph, prev ? &prev : nullptr,
                                               prev ? 1 : 0, &kp);
    if (e != hipSuccess) {
      err = set_error(TEMPI_ERR_HIP, "hipGraphAddKernelNode: %s",
                      hipGetErrorString(e));
      break;
    }
    prev = node;
    ++*nodes;
  }
  if (!err) {
    const hipError_t e = hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      err = set_error(TEMPI_ERR_HIP, "hipGraphInstantiate: %s",
                      hipGetErrorString(e));
    }
  }
  if (err) {
    (void)hipGetLastError();
    if (*exec) {
      (void)hipGraphExecDestroy(*exec);
    }
    (void)hipGraphDestroy(*graph);
    *graph = nullptr;
    *exec = nullptr;
  }
  return err;
}

} // namespace

TEMPI_EXPORT int tempi_batch_graph_launch(tempi_batch *b, void *stream) {
  if (!b) {
    return set_error(TEMPI_ERR_ARG, "null batch");
  }
  if (b->launch.ntiles == 0) {
    return TEMPI_SUCCESS;
  }
  TraceRange trace("tempi_batch_graph_launch");
  if (!b->exec) {
    int64_t nodes = 0;
    int err = build_graph(&b, 1, &b->graph, &b->exec, &nodes);
    if (err) {
      return err;
    }
  }
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : library_stream();
  TEMPI_HIP_CHECK(hipGraphLaunch(b->exec, s));
  ++counters().graph_launches;
  return TEMPI_SUCCESS;
}

struct tempi_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t nodes = 0;
};

TEMPI_EXPORT int tempi_graph_create(tempi_batch *const *batches, int64_t n,
                                    tempi_graph **out) {
  if (!out || n < 0 || (n > 0 && !batches)) {
    return set_error(TEMPI_ERR_ARG, "bad argument");
  }
  *out = nullptr;
  for (int64_t i = 0; i < n; ++i) {
    if (!batches[i]) {
      return set_error(TEMPI_ERR_ARG, "null batch %lld", (long long)i);
    }
  }
  int err = ensure_runtime();
  if (err) {
    return err;
  }
  std::unique_ptr<tempi_graph> g(new tempi_graph());
  if ((err = build_graph(batches, n, &g->graph, &g->exec, &g->nodes))) {
    return err;
  }
  *out = g.release();
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_graph_launch(tempi_graph *g, void *stream) {
  if (!g) {
    return set_error(TEMPI_ERR_ARG, "null graph");
  }
  if (g->nodes == 0) {
    return TEMPI_SUCCESS;
  }
  TraceRange trace("tempi_graph_launch");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : library_stream();
  TEMPI_HIP_CHECK(hipGraphLaunch(g->exec, s));
  ++counters().graph_launches;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_graph_destroy(tempi_graph *g) {
  if (!g) {
    return TEMPI_SUCCESS;
  }
  if (g->exec) {
    // the graph may have been launched on any stream
    (void)hipDeviceSynchronize();
    (void)hipGraphExecDestroy(g->exec);
  }
  if (g->graph) {
    (void)hipGraphDestroy(g->graph);
  }
  delete g;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_batch_info(const tempi_batch *b, int64_t *requests,
                                  int64_t *tiles, int64_t *packed_bytes) {
  if (!b) {
    return set_error(TEMPI_ERR_ARG, "null batch");
  }
  if (requests) {
    *requests = b->requests;
  }
  if (tiles) {
    *tiles = b->launch.ntiles;
  }
  if (packed_bytes) {
    *packed_bytes = b->bytes;
  }
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_batch_variant(const tempi_batch *b, int64_t i,
                                     int *variant) {
  if (!b || i < 0 || i >= (int64_t)b->variants.size() || !variant) {
    return set_error(TEMPI_ERR_ARG, "bad argument");
  }
  *variant = b->variants[i];
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_batch_destroy(tempi_batch *b) {
  if (!b) {
    return TEMPI_SUCCESS;
  }
  if (b->exec || b->dev) {
    // the batch may have been launched on any stream
    (void)hipDeviceSynchronize();
  }
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
  }
  if (b->dev) {
    (void)hipFree(b->dev);
  }
  delete b;
  return TEMPI_SUCCESS;
}

// ---------------------------------------------------------------------------
// the batch entry point that allocates nothing
// ---------------------------------------------------------------------------
//
// tempi_batch_run_noalloc resolves and launches a batch in one call, with
// the request and tile tables in a file-scope device array instead of
// memory the library allocates, for callers that cannot allocate device
// memory.  The array's host-side name is only a stand-in for it: host code
// uses the address hipGetSymbolAddress reports for the current device,
// taken once per device on the first call, and never the name itself.

namespace tempi {
namespace {

// largest table image (request parameters, tile records, XCD bounds) a
// call can place: 4 MiB holds about 4 000 requests over 2^19 tiles (the
// 1 664-request halo batch of 64 subdomains of 128^3 needs 0.5 MiB)
constexpr size_t kScratchBytes = 4u << 20;

struct ScratchSlot {
  std::once_flag once;
  int err = TEMPI_SUCCESS;
  char *ptr = nullptr;             // the array's address on this device
  hipEvent_t uploaded = nullptr;   // the last call's table upload
  std::mutex mu;
  std::vector<char> image;         // that upload's host source
};
constexpr int kMaxDevices = 64;
ScratchSlot g_scratch_slots[kMaxDevices];

} // namespace
} // namespace tempi

static __device__ __align__(256) char scratch[tempi::kScratchBytes];

namespace tempi {
namespace {

int scratch_slot(ScratchSlot **out) {
  int dev = 0;
  TEMPI_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    return set_error(TEMPI_ERR_UNSUPPORTED, "device %d", dev);
  }
  ScratchSlot &sl = g_scratch_slots[dev];
  std::call_once(sl.once, [&sl] {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      sl.err = set_error(TEMPI_ERR_HIP, "hipGetSymbolAddress(scratch): %s",
                         hipGetErrorString(e));
      return;
    }
    e = hipEventCreateWithFlags(&sl.uploaded, hipEventDisableTiming);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      sl.err = set_error(TEMPI_ERR_HIP, "event creation: %s",
                         hipGetErrorString(e));
      return;
    }
    sl.ptr = static_cast<char *>(p);
  });
  if (sl.err != TEMPI_SUCCESS) {
    return set_error(sl.err, "static scratch of device %d is unusable", dev);
  }
  *out = &sl;
  return TEMPI_SUCCESS;
}

} // namespace
} // namespace tempi

TEMPI_EXPORT int tempi_batch_run_noalloc(const tempi_batch_request *reqs,
                                         int64_t n, void *stream) {
  if ((n > 0 && !reqs) || n < 0) {
    return set_error(TEMPI_ERR_ARG, "bad argument");
  }
  std::vector<Req> rs;
  std::vector<int64_t> where;
  int64_t bytes = 0;
  // the closed-form rule: the learned selector and trial timings would
  // allocate device memory on first use
  int err = resolve_requests(
      reqs, n,
      [](const Req &q) { return heuristic_variant(q.g, q.ds, q.dp, q.dir); },
      &rs, &where, &bytes);
  if (err) {
    return err;
  }
  if (rs.empty()) {
    return TEMPI_SUCCESS;
  }
  BatchPlan pl;
  if ((err = plan_batch(rs, device_cu_count(), &pl))) {
    return err;
  }
  if (pl.image.size() > kScratchBytes) {
    return set_error(TEMPI_ERR_UNSUPPORTED,
                     "batch tables of %zu B (%zu requests, %u tiles) exceed "
                     "the %zu B of static scratch",
                     pl.image.size(), rs.size(), pl.launch.ntiles,
                     kScratchBytes);
  }
  ScratchSlot *sl = nullptr;
  if ((err = scratch_slot(&sl))) {
    return err;
  }
  TraceRange trace("tempi_batch_run_noalloc");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : library_stream();
  std::lock_guard<std::mutex> lk(sl->mu);
  // the previous upload has read its host image (a copy from pageable
  // memory may finish after the call that enqueued it returned)
  TEMPI_HIP_CHECK(hipEventSynchronize(sl->uploaded));
  sl->image.swap(pl.image);
  TEMPI_HIP_CHECK(hipMemcpyAsync(sl->ptr, sl->image.data(), sl->image.size(),
                                 hipMemcpyHostToDevice, s));
  TEMPI_HIP_CHECK(hipEventRecord(sl->uploaded, s));
  launch_kernel(placed(pl, sl->ptr), s);
  TEMPI_HIP_CHECK(hipGetLastError());
  ++counters().batch_launches;
  return TEMPI_SUCCESS;
}
