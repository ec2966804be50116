/*
 * tempi_amd.h -- C ABI of libtempi_amd.so, an MI355X (gfx950) implementation
 * of cwpearson/tempi's derived-datatype MPI_Pack / MPI_Unpack path.
 *
 * Two layers are exported:
 *
 *   1. The MPI interposition boundary.  libtempi_amd.so defines the MPI
 *      functions below and forwards to the next definition in link order
 *      (dlsym(RTLD_NEXT), falling back to the object that defines
 *      MPI_Type_get_envelope), exactly as the reference does.  Link it before
 *      libmpi or LD_PRELOAD it.
 *
 *   2. tempi_* entry points with plain pointers and sizes: the reference's
 *      internal interfaces on this path (canonicalizer result, Packer,
 *      allocators, streams), plus the batched queue, autotuner and learned
 *      variant selector, which have no reference counterpart.
 *
 * Every function returns 0 (TEMPI_SUCCESS) or a negative TEMPI_ERR_* code
 * unless stated otherwise; tempi_last_error() returns a message for the
 * calling thread's last failure.
 */
#ifndef TEMPI_AMD_H
#define TEMPI_AMD_H

#include <mpi.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* 1. MPI interposition                                                     */
/* ------------------------------------------------------------------------ */

/* reference: src/init.cpp:22 */
int MPI_Init(int *argc, char ***argv);
/* reference: src/init_thread.cpp:22 */
int MPI_Init_thread(int *argc, char ***argv, int required, int *provided);
/* reference: src/finalize.cpp:20 */
int MPI_Finalize(void);
/* reference: src/type_commit.cpp:16 -- canonicalizes the committed type */
int MPI_Type_commit(MPI_Datatype *datatype);
/* reference: src/type_free.cpp:14 -- releases the cached record first */
int MPI_Type_free(MPI_Datatype *datatype);
/* reference: src/pack.cpp:28 -- device/mapped buffers run the gfx950 kernel,
   anything else goes to the underlying library.  Unlike the reference, a
   pack that would overrun `outsize` raises MPI_ERR_TRUNCATE through the
   communicator's error handler. */
int MPI_Pack(const void *inbuf, int incount, MPI_Datatype datatype,
             void *outbuf, int outsize, int *position, MPI_Comm comm);
/* reference: src/unpack.cpp:20 -- bytes outside the type map are never
   written */
int MPI_Unpack(const void *inbuf, int insize, int *position, void *outbuf,
               int outcount, MPI_Datatype datatype, MPI_Comm comm);
/* Point-to-point transfers of device-accessible buffers whose type has a
   non-contiguous strided form travel as packed bytes (MPI_PACKED), staged
   in device memory (DEVICE, needs a GPU-aware MPI library), pinned host
   memory the kernel writes directly (ONESHOT) or device memory copied to
   pinned host memory (STAGED); anything else goes to the library as it is.
   reference: src/send.cpp:10, src/internal/send.cpp:20-50,
   src/internal/sender.cpp:88-328 */
int MPI_Send(const void *buf, int count, MPI_Datatype datatype, int dest,
             int tag, MPI_Comm comm);
/* reference: src/recv.cpp:19-52 */
int MPI_Recv(void *buf, int count, MPI_Datatype datatype, int source, int tag,
             MPI_Comm comm, MPI_Status *status);
/* reference: src/isend.cpp:11, src/internal/isend.cpp:15-45,
   src/internal/async_operation.cpp:70-206 -- the caller receives a request
   handle of this library's, completed by MPI_Wait / MPI_Waitall / MPI_Test */
int MPI_Isend(const void *buf, int count, MPI_Datatype datatype, int dest,
              int tag, MPI_Comm comm, MPI_Request *request);
/* reference: src/irecv.cpp:11, src/internal/irecv.cpp:15-45,
   src/internal/async_operation.cpp:209-331 */
int MPI_Irecv(void *buf, int count, MPI_Datatype datatype, int source, int tag,
              MPI_Comm comm, MPI_Request *request);
/* reference: src/wait.cpp:11, src/internal/async_operation.cpp:460-477 */
int MPI_Wait(MPI_Request *request, MPI_Status *status);
/* no active counterpart in the reference (its version is compiled out,
   async_operation.cpp:479-511): completes this library's requests together
   with the MPI library's own */
int MPI_Waitall(int count, MPI_Request array_of_requests[],
                MPI_Status array_of_statuses[]);
/* no counterpart in the reference: tests this library's requests */
int MPI_Test(MPI_Request *request, int *flag, MPI_Status *status);
/* reference: src/neighbor_alltoallw.cpp:11-18,
   src/internal/neighbor_alltoallw.cpp:19-79 (one Isend/Irecv per
   neighbour).  Here: one batched pack launch for all outgoing blocks, the
   library's MPI_Neighbor_alltoallv of the packed bytes, one batched unpack
   launch; argument sets are kept for repeated calls. */
int MPI_Neighbor_alltoallw(const void *sendbuf, const int sendcounts[],
                           const MPI_Aint sdispls[],
                           const MPI_Datatype sendtypes[], void *recvbuf,
                           const int recvcounts[], const MPI_Aint rdispls[],
                           const MPI_Datatype recvtypes[], MPI_Comm comm);
/* reference: src/comm_free.cpp -- drops state kept for the communicator */
int MPI_Comm_free(MPI_Comm *comm);

/* ------------------------------------------------------------------------ */
/* 2. tempi_* entry points                                                  */
/* ------------------------------------------------------------------------ */

#define TEMPI_SUCCESS 0
#define TEMPI_ERR_ARG (-1)         /* bad argument                          */
#define TEMPI_ERR_UNSUPPORTED (-2) /* type has no strided-block form        */
#define TEMPI_ERR_TRUNCATE (-3)    /* output buffer too small               */
#define TEMPI_ERR_HIP (-4)         /* HIP runtime error                     */
#define TEMPI_ERR_NODEVICE (-5)    /* no usable GPU                         */
#define TEMPI_ERR_NOMEM (-6)
#define TEMPI_ERR_IO (-7)          /* model / cache file problem            */

#define TEMPI_MAX_DIMS 8

/* Canonical form of a committed type: `counts[0]` contiguous bytes form a
   block, dimension d >= 1 repeats everything inside it counts[d] times,
   strides[d] bytes apart; the first byte of object 0 sits `start` bytes
   after the buffer origin, and object i starts i*extent bytes after object
   0.  Blocks are listed in MPI type-map order (dimension ndims-1 slowest).
   reference: include/strided_block.hpp:12-67 (StridedBlock) */
typedef struct tempi_strided_block {
  int64_t start;
  int64_t extent;
  int32_t ndims;
  int32_t reserved;
  int64_t counts[TEMPI_MAX_DIMS];
  int64_t strides[TEMPI_MAX_DIMS];
} tempi_strided_block;

/* library version, 100*major + minor */
int tempi_version(void);
/* message for the calling thread's last TEMPI_ERR_* */
const char *tempi_last_error(void);
/* number of visible GPUs (0 when none; never fails) */
int tempi_device_count(void);
/* value of an MPI handle constant compiled into the library ("MPI_BYTE",
   "MPI_FLOAT", "MPI_COMM_WORLD", ...), as an integer; `ok` is set to 0 for
   unknown names.  Lets ABI-agnostic hosts (Python ctypes) build types. */
int64_t tempi_mpi_constant(const char *name, int *ok);
/* size in bytes of MPI_Datatype / MPI_Comm handles in this MPI ABI */
int tempi_mpi_handle_size(void);

/* Canonical form of a type committed through MPI_Type_commit, and its
   packed size per object.  TEMPI_ERR_UNSUPPORTED when the type has none
   (the MPI calls then go to the library).
   reference: to_strided_block(simplify(traverse(t))),
   src/internal/types.cpp:356-373, :567-614, :654-715 */
int tempi_type_strided_block(MPI_Datatype datatype, tempi_strided_block *sb,
                             int64_t *size);

/* Canonicalize without committing through the interposer (the type must be
   committed in the underlying library). Same output as above. */
int tempi_type_canonicalize(MPI_Datatype datatype, tempi_strided_block *sb,
                            int64_t *size);

/* Kernel variants of the strided copy.  Variant ids are 0..n-1. */
typedef struct tempi_variant_info {
  int32_t lanes_per_row; /* lanes of a wavefront that share one row   */
  int32_t word_bytes;    /* bytes per load/store                      */
  int32_t unroll;        /* words each lane moves per work unit       */
  int32_t nontemporal_stores; /* stores marked nontemporal           */
  int32_t nontemporal_loads;  /* loads marked nontemporal            */
} tempi_variant_info;
int tempi_num_variants(void);
int tempi_variant_info_get(int variant, tempi_variant_info *info);

#define TEMPI_DIR_PACK 0
#define TEMPI_DIR_UNPACK 1

#define TEMPI_SELECT_HEURISTIC 0 /* closed-form rule                    */
#define TEMPI_SELECT_MODEL 1     /* learned selector (MLP)              */
#define TEMPI_SELECT_AUTOTUNE 2  /* exhaustive timing, cached           */
#define TEMPI_SELECT_DEFAULT 3   /* library policy (model if loaded)    */

/* Variant the library would use for this copy.  Under the default policy a
   copy of 64 MiB or more is settled on its first run: the model's three
   best variants and the closed-form rule's pick are each timed once, cold,
   writing a scratch image of the destination (never the caller's buffers),
   and the fastest is kept for the shape; before that run this returns the
   model's pick. */
int tempi_select_variant(const tempi_strided_block *sb, int64_t incount,
                         const void *strided_buf, const void *packed_buf,
                         int direction, int method, int *variant);

/* Pack `incount` objects described by `sb` from `inbuf` to outbuf +
   *position, asynchronously on `stream` (NULL: the library stream), then
   advance *position.  variant < 0 lets the library choose.
   reference: Packer::pack_async, include/packer.hpp:27-35 (with the
   synchronous Packer::pack, :38-42, being this plus a stream sync) */
int tempi_pack_sb(const tempi_strided_block *sb, const void *inbuf,
                  int64_t incount, void *outbuf, int64_t outsize,
                  int64_t *position, int variant, void *stream);
/* reference: Packer::unpack_async, include/packer.hpp:37-40 */
int tempi_unpack_sb(const tempi_strided_block *sb, const void *inbuf,
                    int64_t insize, int64_t *position, void *outbuf,
                    int64_t outcount, int variant, void *stream);

/* Autotuner: no counterpart in the reference, which fixes the kernel by a
   rule at commit (include/pack_kernels.cuh:129-166) and measures only whole
   methods offline (src/internal/measure_system.cu:514-609).
   tempi_time_variant: average device time (us) of one launch of `variant`
   with grid cap index `grid_cap` (0..tempi_num_grid_caps()-1) over `reps`
   back-to-back launches on the library stream; copies of 64 MiB or more
   are timed one launch at a time, each after the caches are emptied (data
   that large is not cache-resident between real calls).  The packed (pack)
   or strided (unpack) buffer is written.
   tempi_autotune: time every variant x grid cap, return the fastest and,
   if per_variant_us is not NULL, the best time of each variant
   (tempi_num_variants() floats).  Setting TEMPI_SELECT=autotune makes
   MPI_Pack / MPI_Unpack tune each new shape on first use and cache it. */
int tempi_num_grid_caps(void);
int tempi_time_variant(const tempi_strided_block *sb, int64_t incount,
                       const void *strided, void *packed, int direction,
                       int variant, int grid_cap, int reps, float *us);
int tempi_autotune(const tempi_strided_block *sb, int64_t incount,
                   const void *strided, void *packed, int direction,
                   int *best_variant, int *best_grid_cap,
                   float *per_variant_us);
/* Buffers of tempi_time_variant and tempi_autotune: the copy runs as a pack
   or unpack on them, so the strided buffer must hold every byte of the
   type map from `strided` (start + extent-walk of `incount` objects, as
   for tempi_pack_sb) and the packed buffer incount * packed-size bytes from
   `packed`.  Both are written: the packed side by a pack, the strided
   side's type-map bytes (only those) by an unpack.  The accesses of a copy
   can be checked on the host with tempi_copy_emulate. */

/* Host-side emulation of a copy: no counterpart in the reference.  Runs the
   launches, grid, 32-bit grid-stride index sequence and per-unit address
   arithmetic the library uses for (variant, grid cap) -- with the kernels'
   own address code -- and records the bytes each launch would load and
   store, without touching memory (the addresses are numbers only).
   `max_blocks` is a workgroup cap (0: the default, 16 per CU) and
   `cu_count` the CU count it assumes.  Copies of more than `unit_limit`
   work units report only the launch plan (emulated = 0). */
typedef struct tempi_copy_emulation {
  int64_t launches;   /* kernel launches of the copy                       */
  int64_t max_units;  /* most work units in one launch (stays < 2^31)      */
  int32_t emulated;   /* 1: every unit emulated, fields below filled       */
  int32_t wrapped;    /* 1: a 32-bit unit index would wrap                 */
  int64_t read_lo, read_hi;   /* bytes loaded: [lo, hi) from the read side's
                                 origin (strided origin for a pack, packed
                                 pointer for an unpack)                     */
  int64_t write_lo, write_hi; /* bytes stored, from the written side's origin */
  int64_t bytes_read, bytes_written; /* totals over all loads and stores   */
  int32_t exact;      /* 1: every byte the copy must write was stored once
                         and no other byte; 0: not so; -1: not checked (the
                         written span exceeds 256 MiB)                      */
  int32_t reserved;
} tempi_copy_emulation;
int tempi_copy_emulate(const tempi_strided_block *sb, int64_t incount,
                       uint64_t strided_addr, uint64_t packed_addr,
                       int direction, int variant, int max_blocks,
                       int cu_count, int64_t unit_limit,
                       tempi_copy_emulation *out);

/* Read `bytes` (<= 0: 1 GiB) of a library-owned buffer so that L2 and the
   Infinity Cache hold only clean, unrelated lines; synchronous.  For
   cold-cache timing. */
int tempi_flush_caches(int64_t bytes);
/* Launch latency of an empty kernel: host-side enqueue cost and device
   time per launch (reference: src/internal/measure_system.cu:36-64). */
int tempi_launch_latency_us(int reps, float *host_us, float *device_us);

/* Batched packing: no counterpart in the reference, which launches one
   kernel per MPI_Pack and synchronizes (src/internal/packer_2d.cu:101-108).
   A batch is a fixed set of pack and/or unpack requests (for example the 26
   faces, edges and corners of a 3D halo), resolved once into a
   device-resident descriptor table; one launch runs a persistent kernel that
   drains all of them (a batch of at most 8 requests passes its tile index
   in the kernel arguments instead of a device tile table).  tempi_batch_graph_launch replays the same launch from
   a captured hipGraph; it is slower than tempi_batch_launch on ROCm 7.2,
   whose hipGraphLaunch costs about 17 us of host time against about 2.7 us
   for a kernel launch, and is for callers whose own step is a graph.
   Requests in one batch must be independent (no
   request writes bytes another reads or writes).  Nothing is synchronized:
   wait on the stream.  `method` (TEMPI_SELECT_*) picks the variant of each
   request that leaves it to the library (variant -1), as for a single copy
   of that request. */
typedef struct tempi_batch_request {
  tempi_strided_block sb; /* layout of one object                          */
  const void *strided;    /* strided buffer origin (pack: in, unpack: out)  */
  void *packed;           /* first packed byte (pack: out, unpack: in)      */
  int64_t count;          /* objects                                        */
  int32_t direction;      /* TEMPI_DIR_PACK or TEMPI_DIR_UNPACK             */
  int32_t variant;        /* kernel variant, -1 to let the library choose   */
} tempi_batch_request;
typedef struct tempi_batch tempi_batch;
int tempi_batch_create(const tempi_batch_request *reqs, int64_t n, int method,
                       tempi_batch **batch);
int tempi_batch_launch(tempi_batch *batch, void *stream);
int tempi_batch_graph_launch(tempi_batch *batch, void *stream);
/* One executable graph of several batches' launches, each after the one
   before (for example the pack batch and the unpack batch of a step): one
   tempi_graph_launch replays them all.  The graph refers to the batches'
   device tables: the batches must outlive it.  No counterpart in the
   reference. */
typedef struct tempi_graph tempi_graph;
int tempi_graph_create(tempi_batch *const *batches, int64_t n,
                       tempi_graph **graph);
int tempi_graph_launch(tempi_graph *graph, void *stream);
int tempi_graph_destroy(tempi_graph *graph);
int tempi_batch_info(const tempi_batch *batch, int64_t *requests,
                     int64_t *tiles, int64_t *packed_bytes);
int tempi_batch_variant(const tempi_batch *batch, int64_t request,
                        int *variant);
int tempi_batch_destroy(tempi_batch *batch);
/* No counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc): resolves `reqs` and launches them as one batch on
   `stream` (NULL: the library stream) like tempi_batch_create +
   tempi_batch_launch, but the batch's request and tile tables go to a
   static device array of 4 MiB that the library holds on each GPU until the
   process ends, not to memory it allocates; nothing is allocated on the
   device and nothing waits for the launch.  Requests with variant -1 take
   the closed-form rule (TEMPI_SELECT_HEURISTIC: the learned selector and
   trial timings would allocate).  Batches whose tables exceed the array
   (about 4 000 requests, or 2^19 tiles) fail with TEMPI_ERR_UNSUPPORTED
   before any device work.  Every call on a GPU overwrites that GPU's
   array: calls on one stream follow each other in stream order, but calls
   that overlap on one GPU (other streams, other threads without ordering)
   race on the tables. */
int tempi_batch_run_noalloc(const tempi_batch_request *reqs, int64_t n,
                            void *stream);

/* Learned variant selector: replaces the reference's measured tables and
   interpolation (src/internal/measure_system.cpp:181-299) and its fixed
   kernel rule (include/pack_kernels.cuh:129-166).  An MLP
   16 -> 64 -> 64 -> tempi_mlp_num_outputs() (ReLU, fp32) predicts the
   relative time (log2, centred) of each kernel variant; the selector takes
   the argmin.  Parameters are laid out [W1 16x64, b1, W2 64x64, b2, W3 64xV,
   b3] (row = input); `norm` is the feature mean[16] then 1/std[16].
   tempi_mlp_features: the 16 features of a copy (log2 row bytes, rows,
     total bytes; row dimensions; log2 of the first two dimensions' counts and
     strides and of the remaining count; log2 alignment (capped at 16 B) of
     the strided start, packed pointer, strides and row length; direction;
     12-byte rows; log2 stride/row ratio).
   tempi_mlp_grad: loss and gradient of one minibatch (n <= 32) of
     normalized features X[n][16], targets T[n][V] and optional mask M, by
     the GPU kernel training uses.
   tempi_mlp_train: Adam over minibatches of 32 for `epochs` epochs, all on
     the GPU; `params` holds the initial values and receives the result.
   tempi_mlp_predict: the inference kernel on raw features.
   tempi_mlp_save / tempi_mlp_load: versioned text file.  The library loads
     $TEMPI_MODEL, else selector_mlp.txt next to itself, on first use; a
     missing or malformed file leaves the closed-form rule in place.
   With a model loaded, MPI_Type_commit evaluates it for the new type. */
int tempi_mlp_num_features(void);
int tempi_mlp_num_outputs(void);
int tempi_mlp_num_params(void);
int tempi_mlp_features(const tempi_strided_block *sb, int64_t incount,
                       const void *strided, const void *packed, int direction,
                       float *features);
int tempi_mlp_grad(const float *params, const float *X, const float *T,
                   const float *M, int n, float *grad, float *loss);
int tempi_mlp_train(const float *X, const float *T, const float *M, int64_t n,
                    int epochs, float lr, uint64_t seed, float *params,
                    float *epoch_loss);
int tempi_mlp_predict(const float *params, const float *norm, const float *X,
                      int64_t n, float *pred, int *best);
int tempi_mlp_save(const char *path, const float *params, const float *norm);
int tempi_mlp_load(const char *path);
int tempi_mlp_loaded(void);

/* The library's non-blocking stream (reference: kernStream,
   src/internal/streams.cpp:25), created on first use. */
void *tempi_stream(void);
int tempi_stream_synchronize(void *stream);

/* Memory.  reference: deviceAllocator / hostAllocator
   (include/allocator_device.hpp:22-70, include/allocator_host.hpp:31-60).
   Mapped host memory is pinned and device-accessible: packing into it is
   the ONE_SHOT method. */
void *tempi_alloc_device(int64_t bytes);
int tempi_free_device(void *p);
void *tempi_alloc_host_mapped(int64_t bytes);
int tempi_free_host_mapped(void *p);
int tempi_memcpy(void *dst, const void *src, int64_t bytes); /* synchronous */
int tempi_memset(void *dst, int value, int64_t bytes);       /* synchronous */
/* 1 if the kernels can address `p` (device, managed or mapped host) */
int tempi_is_device_accessible(const void *p);

/* Events for device-side timing on a stream. */
int tempi_event_create(void **event);
int tempi_event_destroy(void *event);
int tempi_event_record(void *event, void *stream);
int tempi_event_synchronize(void *event);
int tempi_event_elapsed_ms(void *start, void *stop, float *ms);

/* Counters (reference: include/counters.hpp:86-193). */
typedef struct tempi_counters {
  int64_t pack_kernel_launches;
  int64_t unpack_kernel_launches;
  int64_t batch_launches;
  int64_t graph_launches;
  int64_t library_pack_calls;   /* MPI_Pack forwarded to the library    */
  int64_t library_unpack_calls; /* MPI_Unpack forwarded to the library  */
  int64_t commits_canonical;    /* MPI_Type_commit with a strided form  */
  int64_t commits_unsupported;
  int64_t model_inferences;     /* MLP inference kernel launches        */
  int64_t autotune_runs;
  int64_t verified_choices;     /* large copies whose model choice was timed */
  int64_t p2p_device;           /* sends / receives of device data by method */
  int64_t p2p_oneshot;
  int64_t p2p_staged;
  int64_t p2p_library;          /* point-to-point calls forwarded as they are */
  int64_t trial_launches;       /* copies the library timed (not the caller's) */
  int64_t method_model_choices; /* staging methods chosen by the method MLP */
  int64_t host_staged_calls;    /* MPI_Pack / MPI_Unpack with one side in memory
                                   the other side's engine cannot address,
                                   staged through device or pinned host memory */
  int64_t signalled_waits;      /* synchronous copies whose own completion
                                   signal the host waited for */
} tempi_counters;
int tempi_counters_get(tempi_counters *out);
int tempi_counters_reset(void);

/* ---- measured system performance and the send-method model ----------------
   The reference's system model (include/measure_system.hpp:26-106,
   struct Performance): contiguous transfer curves (entry i = time of 2^i
   bytes) and 9x9 pack/unpack tables (entry [i][j] = time of 2^(2i+6) bytes
   in blocks of 2^j bytes at a 512-B stride), in seconds, stored as
   perf.json in the cache directory with the reference's keys
   (src/internal/measure_system.cpp:36-63).  Loaded at MPI_Init
   (src/internal/measure_system.cu:32). */
#define TEMPI_CURVE_D2H 0           /* "d2h"                      */
#define TEMPI_CURVE_H2D 1           /* "h2d"                      */
#define TEMPI_CURVE_INTRA_CPU_CPU 2 /* "intraNodeCpuCpuPingpong"  */
#define TEMPI_CURVE_INTRA_GPU_GPU 3 /* "intraNodeGpuGpuPingpong"  */
#define TEMPI_CURVE_INTER_CPU_CPU 4 /* "interNodeCpuCpuPingpong"  */
#define TEMPI_CURVE_INTER_GPU_GPU 5 /* "interNodeGpuGpuPingpong"  */
#define TEMPI_NUM_CURVES 6
#define TEMPI_TABLE_PACK_DEVICE 0   /* "packDevice"   */
#define TEMPI_TABLE_UNPACK_DEVICE 1 /* "unpackDevice" */
#define TEMPI_TABLE_PACK_HOST 2     /* "packHost"     */
#define TEMPI_TABLE_UNPACK_HOST 3   /* "unpackHost"   */
#define TEMPI_NUM_TABLES 4
#define TEMPI_PERF_TABLE_DIM 9      /* rows (bytes) and columns (block) */
#define TEMPI_PERF_CURVE_LEN 24     /* 1 B .. 8 MiB */
/* send methods for non-contiguous data
   (reference: Performance::SendNonContigNd::Method, measure_system.hpp:92) */
#define TEMPI_METHOD_UNKNOWN 0
#define TEMPI_METHOD_DEVICE 1  /* pack in device memory, send from there   */
#define TEMPI_METHOD_ONESHOT 2 /* pack straight into mapped host memory    */
#define TEMPI_METHOD_STAGED 3  /* pack in device memory, copy to the host  */

/* Interpolated time of `bytes` from a curve of n points (entry i = 2^i B),
   +inf when n == 0.  reference: interp_time / interp_time_opt,
   src/internal/measure_system.cpp:182-214. */
double tempi_interp_time(const double *times, int n, int64_t bytes);
/* Interpolated time from a rows x cols table (row-major; row i = 2^(2i+6)
   bytes, column j = block length 2^j), +inf when rows == 0.
   reference: interp_2d / interp_2d_opt, measure_system.cpp:216-297. */
double tempi_interp_2d(const double *table, int rows, int cols, int64_t bytes,
                       int64_t block_length);

/* The process-wide tables.  get: *n (or *rows, *cols) receives the stored
   size; up to `cap` values are copied.  set: n == 0 clears the entry. */
int tempi_perf_clear(void);
int tempi_perf_get_curve(int curve, double *times, int cap, int *n);
int tempi_perf_set_curve(int curve, const double *times, int n);
int tempi_perf_get_table(int table, double *times, int cap, int *rows,
                         int *cols);
int tempi_perf_set_table(int table, const double *times, int rows, int cols);
double tempi_perf_kernel_launch(void); /* "cudaKernelLaunch", seconds */
int tempi_perf_set_kernel_launch(double seconds);
/* perf.json import / export; NULL = <cache dir>/perf.json.
   reference: import_performance / export_performance,
   measure_system.cpp:140-180. */
int tempi_perf_load(const char *path);
int tempi_perf_save(const char *path);
/* Fill the empty entries this GPU can measure (reference: measure_performance,
   src/internal/measure_system.cu:381-609): bit 0 kernel launch, bit 1 the
   d2h/h2d curves, bit 2 the device pack/unpack tables, bit 3 the mapped-host
   (ONE_SHOT) tables.  Each point runs trials of at least 7 samples and at
   most `seconds_per_point` (<= 0: 0.05 s) or 500 samples.  After each
   trial, tempi_iid_test decides whether its samples are independent and
   identically distributed.  Trials repeat until the test passes or
   $TEMPI_MEASURE_TRIALS trials have run (default 10, the reference's
   RunConfig::maxTrials, include/benchmark.hpp:41-46).  The point's time
   is the last trial's trimean (src/internal/statistics.cpp:30-39), and its
   `iid` flag in perf.json is the test's verdict
   (src/internal/benchmark.cpp:44-83).
   Bit 4 (TEMPI_MEASURE_PINGPONG) is collective over MPI_COMM_WORLD and needs
   at least two ranks: ranks 0 and 1 bounce host messages of 1 B - 8 MiB
   through the MPI library (reference: the CPU-CPU ping-pong of
   bin/measure_system.cpp), 7 samples of a repetition count rank 0 picks for
   about 200 us, one-way time = trimean / 2; the curve, broadcast to every
   rank, fills intraNodeCpuCpuPingpong when the two ranks share a node, else
   interNodeCpuCpuPingpong.  It needs no GPU.  The GPU-GPU curves need a
   GPU-aware MPI library and stay imported. */
#define TEMPI_MEASURE_PINGPONG 16
int tempi_perf_measure(int what, double seconds_per_point);
/* The permutation test of NIST SP 800-90B section 5.1 on `n` samples
   (2 <= n <= 65536, all finite), with the eight statistics the reference
   uses (reference: sp_800_90B, src/internal/iid.cpp:184-245; test/iid.cpp).
   Statistics, in order: excursion, number of directional runs, longest
   directional run, number of increases or decreases, number of runs about
   the median, longest run about the median, average collision, maximum
   collision.  tempi_iid_statistics writes the eight statistics of the
   samples to stats[8].  tempi_iid_test shuffles the samples 10 000 times;
   shuffle k uses a splitmix64 stream seeded with
   seed + k * 0x9E3779B97F4A7C15.  It writes counts[2 s] = C0, the shuffles
   whose statistic s exceeds the samples', and counts[2 s + 1] = C1, those
   equal to it (counts may be NULL).  It returns 1 when no statistic has
   C0 + C1 <= 5 or C0 >= 9995 (independent), 0 when one does, and
   TEMPI_ERR_ARG for bad input. */
int tempi_iid_statistics(const double *samples, int64_t n, double *stats);
int tempi_iid_test(const double *samples, int64_t n, uint64_t seed,
                   int64_t *counts);
/* Modelled time of sending `bytes` of non-contiguous data in blocks of
   `block_length` by `method`; TEMPI_ERR_UNSUPPORTED when an input is missing.
   reference: DeviceND/OneshotND/StagedND::model, src/internal/sender.cpp:
   128-137, 180-191, 238-248. */
int tempi_perf_model(int method, int colocated, int64_t bytes,
                     int64_t block_length, double *seconds);
/* DEVICE or ONE_SHOT, whichever models faster (ties and unknowns: DEVICE);
   cached per (colocated, bytes, block_length).  reference:
   SendRecvND::send, src/internal/sender.cpp:251-273. */
int tempi_choose_method(int colocated, int64_t bytes, int64_t block_length);
/* Force the method of later sends and receives of non-contiguous device data
   (TEMPI_METHOD_DEVICE / ONESHOT / STAGED; TEMPI_METHOD_UNKNOWN: choose by
   the model; -1: back to the TEMPI_DATATYPE_* environment).  Returns the
   previous setting.  The runtime form of the reference's
   TEMPI_DATATYPE_{ONESHOT,DEVICE,AUTO} variables (src/internal/env.cpp:
   73-81), for tools and tests. */
int tempi_set_datatype_method(int method);
/* The staging method the library takes for a message of `bytes` packed
   bytes in rows of `block` bytes `stride` apart (TEMPI_METHOD_*): forced by
   tempi_set_datatype_method or TEMPI_DATATYPE_*; else, for a peer on the
   same node, the method model's cheapest; else the measured tables'
   cheapest (reference: src/internal/sender.cpp:128-273, which always uses
   the tables). */
int tempi_p2p_method(int colocated, int64_t bytes, int64_t block,
                     int64_t stride);

/* Learned staging-method model: no counterpart in the reference, whose
   method choice sums measured table entries (sender.cpp:128-273).  The
   selector's network (tempi_mlp_*) with its own weights: output k predicts
   the time of method k + 1 (DEVICE, ONESHOT, STAGED; log2, relative),
   trained on end-to-end timings of each method by tempi_amd/method_train.py.
   tempi_method_features: the network's 16 inputs for a message (log2 bytes,
     row bytes, rows, stride, stride / row bytes; same node; log2 alignment of
     the row length; the rest 0).
   tempi_method_model_load: weights file (NULL: method_mlp.txt next to the
     library; $TEMPI_METHOD_MODEL overrides the default); the library loads
     it on first use.  tempi_method_model_unload: back to the tables. */
int tempi_method_features(int64_t bytes, int64_t block, int64_t stride,
                          int colocated, float *features);
int tempi_method_model_load(const char *path);
int tempi_method_model_unload(void);
int tempi_method_model_loaded(void);

#ifdef __cplusplus
}
#endif
#endif
