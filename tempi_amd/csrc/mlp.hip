// mlp.hip -- the learned kernel-variant selector.
//
// The reference chooses between methods with measured tables and bilinear
// interpolation (src/internal/measure_system.cpp:181-299, tables from
// src/internal/measure_system.cu:514-609) and picks its kernel by a fixed
// rule (include/pack_kernels.cuh:129-166).  Here a small MLP maps 16
// features of a copy (shape, strides, alignment, direction) to the predicted
// relative time of each kernel variant; the selector takes the argmin.  It
// is trained on the autotuner's own timings, by the forward / backward /
// Adam kernels below, and evaluated at MPI_Type_commit (and on first use of
// a new shape) by the inference kernel.  Without a model file the closed-form
// rule is used; the autotuner stays available as fallback and as the judge
// of the model's regret.
//
// Network: x (16) -> relu(W1 x + b1) (64) -> relu(W2 h1 + b2) (64) ->
// W3 h2 + b3 (one output per variant), fp32.  Parameters are stored
// [W1 (16x64, row = input), b1, W2 (64x64), b2, W3 (64xV), b3].
// Loss: masked mean squared error against the targets.
#include "copy_core.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <dlfcn.h>
#include <deque>
#include <map>
#include <random>
#include <set>
#include <string>
#include <unordered_map>

namespace tempi {

namespace {

#define TEMPI_COUNT_V(I, P, W, U, NTS, NTL) +1
constexpr int kOut = 0 TEMPI_VARIANTS(TEMPI_COUNT_V);
#undef TEMPI_COUNT_V
constexpr int kIn = 16, kH1 = 64, kH2 = 64;
static_assert(kIn == kModelInputs, "one input width for both networks");
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kIn * kH1;
constexpr int kOffW2 = kOffB1 + kH1;
constexpr int kOffB2 = kOffW2 + kH1 * kH2;
constexpr int kOffW3 = kOffB2 + kH2;
constexpr int kOffB3 = kOffW3 + kH2 * kOut;
constexpr int kNParams = kOffB3 + kOut;
constexpr int kB = 32;         // samples per minibatch / inference chunk
constexpr int kThreads = 256;

// activations of one minibatch, in LDS
struct Acts {
  float x[kB][kIn];
  float h1[kB][kH1];
  float h2[kB][kH2];
  float y[kB][kOut]; // outputs; replaced in place by dloss/dy in training
  float d2[kB][kH2];
  float d1[kB][kH1];
};

// parameters and activations share one workgroup's LDS; keep them within the
// 64 KiB every CDNA part grants a workgroup
static_assert(sizeof(Acts) + sizeof(float) * (kNParams + 3) <= 65536,
              "MLP kernels exceed 64 KiB of LDS");

// y = act(b + a W) for a[nb][K], W[K][N] row-major (row = input)
template <int K, int N, bool RELU>
__device__ void dense(const float *__restrict__ P, int offW, int offB,
                      const float (*a)[K], float (*out)[N], int nb) {
  for (int idx = threadIdx.x; idx < nb * N; idx += kThreads) {
    const int b = idx / N, j = idx % N;
    float acc = P[offB + j];
#pragma unroll 8
    for (int i = 0; i < K; ++i) {
      acc = fmaf(a[b][i], P[offW + i * N + j], acc);
    }
    out[b][j] = RELU ? fmaxf(acc, 0.f) : acc;
  }
}

// The parameters into LDS, every thread's loads issued before its first
// wait (blocks of kThreads threads).  As a strided loop the compiler waited
// for each load before the next: 26 round trips to memory at the start of
// every launch, most of an inference's 5-18 us on MI355X.
__device__ __forceinline__ void load_params(float *P,
                                            const float *__restrict__ params) {
  constexpr int kPer = (kNParams + kThreads - 1) / kThreads;
  float v[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int p = (int)threadIdx.x + k * kThreads;
    v[k] = p < kNParams ? params[p] : 0.f;
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int p = (int)threadIdx.x + k * kThreads;
    if (p < kNParams) {
      P[p] = v[k];
    }
  }
}

__device__ void forward(const float *__restrict__ P, Acts &a, int nb) {
  dense<kIn, kH1, true>(P, kOffW1, kOffB1, a.x, a.h1, nb);
  __syncthreads();
  dense<kH1, kH2, true>(P, kOffW2, kOffB2, a.h1, a.h2, nb);
  __syncthreads();
  dense<kH2, kOut, false>(P, kOffW3, kOffB3, a.h2, a.y, nb);
  __syncthreads();
}

// dy = 2 m (y - t) / sum(m), written over y; returns (in *loss, thread 0)
// sum m (y-t)^2 / sum m
__device__ void loss_grad(Acts &a, const float *__restrict__ T,
                          const float *__restrict__ M,
                          const uint32_t *__restrict__ rows, int nb,
                          float *s_sum, float *loss) {
  if (threadIdx.x == 0) {
    s_sum[0] = 0.f;
    s_sum[1] = 0.f;
  }
  __syncthreads();
  float cnt = 0.f;
  for (int idx = threadIdx.x; idx < nb * kOut; idx += kThreads) {
    const int b = idx / kOut, k = idx % kOut;
    const uint32_t r = rows[b];
    cnt += M ? M[(size_t)r * kOut + k] : 1.f;
  }
  atomicAdd(&s_sum[0], cnt);
  __syncthreads();
  const float inv = s_sum[0] > 0.f ? 1.f / s_sum[0] : 0.f;
  float l = 0.f;
  for (int idx = threadIdx.x; idx < nb * kOut; idx += kThreads) {
    const int b = idx / kOut, k = idx % kOut;
    const uint32_t r = rows[b];
    const float m = M ? M[(size_t)r * kOut + k] : 1.f;
    const float e = a.y[b][k] - T[(size_t)r * kOut + k];
    a.y[b][k] = 2.f * m * e * inv; // dy, in place: this thread alone reads y[b][k]
    l += m * e * e * inv;
  }
  atomicAdd(&s_sum[1], l);
  __syncthreads();
  if (threadIdx.x == 0) {
    *loss = s_sum[1];
  }
}

__device__ void backward_hidden(const float *__restrict__ P, Acts &a, int nb) {
  for (int idx = threadIdx.x; idx < nb * kH2; idx += kThreads) {
    const int b = idx / kH2, j = idx % kH2;
    float acc = 0.f;
    if (a.h2[b][j] > 0.f) {
#pragma unroll
      for (int k = 0; k < kOut; ++k) {
        acc = fmaf(a.y[b][k], P[kOffW3 + j * kOut + k], acc);
      }
    }
    a.d2[b][j] = acc;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < nb * kH1; idx += kThreads) {
    const int b = idx / kH1, i = idx % kH1;
    float acc = 0.f;
    if (a.h1[b][i] > 0.f) {
#pragma unroll 8
      for (int j = 0; j < kH2; ++j) {
        acc = fmaf(a.d2[b][j], P[kOffW2 + i * kH2 + j], acc);
      }
    }
    a.d1[b][i] = acc;
  }
  __syncthreads();
}

// gradient of parameter p from the minibatch activations
__device__ float param_grad(const Acts &a, int p, int nb) {
  float g = 0.f;
  if (p < kOffB1) {
    const int i = (p - kOffW1) / kH1, j = (p - kOffW1) % kH1;
    for (int b = 0; b < nb; ++b) g = fmaf(a.x[b][i], a.d1[b][j], g);
  } else if (p < kOffW2) {
    const int j = p - kOffB1;
    for (int b = 0; b < nb; ++b) g += a.d1[b][j];
  } else if (p < kOffB2) {
    const int i = (p - kOffW2) / kH2, j = (p - kOffW2) % kH2;
    for (int b = 0; b < nb; ++b) g = fmaf(a.h1[b][i], a.d2[b][j], g);
  } else if (p < kOffW3) {
    const int j = p - kOffB2;
    for (int b = 0; b < nb; ++b) g += a.d2[b][j];
  } else if (p < kOffB3) {
    const int j = (p - kOffW3) / kOut, k = (p - kOffW3) % kOut;
    for (int b = 0; b < nb; ++b) g = fmaf(a.h2[b][j], a.y[b][k], g);
  } else {
    const int k = p - kOffB3;
    for (int b = 0; b < nb; ++b) g += a.y[b][k];
  }
  return g;
}

__device__ void load_rows(Acts &a, const float *__restrict__ X,
                          const uint32_t *__restrict__ rows, int nb) {
  for (int idx = threadIdx.x; idx < nb * kIn; idx += kThreads) {
    const int b = idx / kIn, i = idx % kIn;
    a.x[b][i] = X[(size_t)rows[b] * kIn + i];
  }
  __syncthreads();
}

// Loss and gradient of one minibatch (nb <= kB samples, rows[] indices).
__global__ void __launch_bounds__(kThreads)
    k_mlp_grad(const float *__restrict__ params, const float *__restrict__ X,
               const float *__restrict__ T, const float *__restrict__ M,
               const uint32_t *__restrict__ rows, int nb, float *grad,
               float *loss) {
  __shared__ float P[kNParams];
  __shared__ Acts a;
  __shared__ float s_sum[2];
  load_params(P, params);
  __syncthreads();
  load_rows(a, X, rows, nb);
  forward(P, a, nb);
  loss_grad(a, T, M, rows, nb, s_sum, loss);
  backward_hidden(P, a, nb);
  for (int p = threadIdx.x; p < kNParams; p += kThreads) {
    grad[p] = param_grad(a, p, nb);
  }
}

// One epoch of Adam over the samples in `order` (n entries), minibatches of
// kB; params / m / v in global memory, kept in LDS / registers meanwhile.
__global__ void __launch_bounds__(kThreads)
    k_mlp_train_epoch(float *params, float *adam_m, float *adam_v,
                      const float *__restrict__ X, const float *__restrict__ T,
                      const float *__restrict__ M,
                      const uint32_t *__restrict__ order, int n, float lr,
                      int step0, float *epoch_loss) {
  __shared__ float P[kNParams];
  __shared__ Acts a;
  __shared__ float s_sum[2];
  __shared__ float s_loss;
  load_params(P, params);
  float loss_acc = 0.f;
  int steps = 0;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  __syncthreads();
  for (int b0 = 0; b0 < n; b0 += kB) {
    const int nb = min(kB, n - b0);
    load_rows(a, X, order + b0, nb);
    forward(P, a, nb);
    loss_grad(a, T, M, order + b0, nb, s_sum, &s_loss);
    backward_hidden(P, a, nb);
    const int t = step0 + steps + 1;
    const float c1 = 1.f - powf(b1, (float)t);
    const float c2 = 1.f - powf(b2, (float)t);
    for (int p = threadIdx.x; p < kNParams; p += kThreads) {
      const float g = param_grad(a, p, nb);
      const float m = b1 * adam_m[p] + (1.f - b1) * g;
      const float v = b2 * adam_v[p] + (1.f - b2) * g * g;
      adam_m[p] = m;
      adam_v[p] = v;
      P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
    }
    __syncthreads();
    loss_acc += s_loss;
    ++steps;
  }
  for (int p = threadIdx.x; p < kNParams; p += kThreads) {
    params[p] = P[p];
  }
  if (threadIdx.x == 0) {
    *epoch_loss = steps ? loss_acc / steps : 0.f;
  }
}

// Predicted relative time of every variant for n raw feature vectors, and
// the argmin of each; workgroup w takes chunks w, w + grid, ... of kB
// vectors.
__global__ void __launch_bounds__(kThreads)
    k_mlp_infer(const float *__restrict__ params,
                const float *__restrict__ norm, const float *__restrict__ X,
                int n, float *pred, int *best) {
  __shared__ float P[kNParams];
  __shared__ Acts a;
  load_params(P, params);
  __syncthreads();
  for (int b0 = (int)blockIdx.x * kB; b0 < n; b0 += (int)gridDim.x * kB) {
    const int nb = min(kB, n - b0);
    for (int idx = threadIdx.x; idx < nb * kIn; idx += kThreads) {
      const int b = idx / kIn, i = idx % kIn;
      a.x[b][i] = (X[(size_t)(b0 + b) * kIn + i] - norm[i]) * norm[kIn + i];
    }
    __syncthreads();
    forward(P, a, nb);
    for (int b = threadIdx.x; b < nb; b += kThreads) {
      int arg = 0;
      for (int k = 0; k < kOut; ++k) {
        if (pred) {
          pred[(size_t)(b0 + b) * kOut + k] = a.y[b][k];
        }
        if (a.y[b][k] < a.y[b][arg]) {
          arg = k;
        }
      }
      if (best) {
        best[b0 + b] = arg;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// Commit-time inferences run without waiting (model_precommit).  A batch of
// up to kSlotTypes committed types takes a slot of pinned host memory that
// the kernel reads and writes through its device mapping -- two feature
// vectors and two picks per type -- until its event has fired and its picks
// are in the shape cache.  A commit that finds an inference of ours still
// running adds its type to the open batch instead of launching one of its
// own; the batch launches when the GPU is free of ours again (at the next
// commit), when it is full, or when a lookup needs one of its shapes.
constexpr int kSlots = 8;
constexpr int kSlotTypes = 16;
constexpr int kSlotVecs = 2 * kSlotTypes; // one inference chunk (kB)
static_assert(kSlotVecs <= kB, "a batch is one inference chunk");
constexpr size_t kSlotPicks = (size_t)kSlotVecs * kIn * 4; // picks' offset
constexpr size_t kSlotBytes = kSlotPicks + (size_t)kSlotVecs * 4;

struct Model {
  std::mutex mu;
  bool loaded = false;
  bool tried = false;
  std::vector<float> params;   // kNParams
  std::vector<float> norm;     // mean[kIn], 1/std[kIn]
  float *d_buf = nullptr;      // params, norm, x, pred, best
  std::unordered_map<std::string, int> cache;
  std::string path;
  struct Pending {
    int slot;
    int n = 0;                     // types in the batch
    bool launched = false;         // else the open batch (always the newest)
    std::vector<std::string> keys; // pack, unpack of each type
  };
  char *h_slots = nullptr;     // pinned: kSlots x kSlotBytes
  hipEvent_t ev[kSlots] = {};
  std::deque<Pending> pending; // oldest first
  std::vector<int> free_slots;
};

Model &model() {
  static Model m;
  return m;
}

float log2p(double v) { return (float)std::log2(std::max(v, 1.0)); }

int align_log2(uint64_t v) {
  if (v == 0) {
    return 4;
  }
  int tz = __builtin_ctzll(v);
  return tz > 4 ? 4 : tz;
}

} // namespace

// 16 features of a copy (see the header for their meaning)
void copy_features(const Geom &g, const char *strided, const char *packed,
                   int dir, float *f) {
  const double L = (double)g.L, rows = (double)g.rows;
  f[0] = log2p(L);
  f[1] = log2p(rows);
  f[2] = log2p(L * rows);
  f[3] = (float)g.nrd;
  f[4] = g.nrd >= 1 ? log2p((double)g.cnt[0]) : 0.f;
  f[5] = g.nrd >= 1 ? log2p((double)std::llabs(g.str[0])) : 0.f;
  f[6] = g.nrd >= 2 ? log2p((double)g.cnt[1]) : 0.f;
  f[7] = g.nrd >= 2 ? log2p((double)std::llabs(g.str[1])) : 0.f;
  double rest = 1;
  for (int k = 2; k < g.nrd; ++k) {
    rest *= (double)g.cnt[k];
  }
  f[8] = log2p(rest);
  f[9] = (float)align_log2((uint64_t)(uintptr_t)(strided + g.start));
  f[10] = (float)align_log2((uint64_t)(uintptr_t)packed);
  uint64_t sg = 0;
  for (int k = 0; k < g.nrd; ++k) {
    sg |= (uint64_t)std::llabs(g.str[k]);
  }
  f[11] = (float)align_log2(sg);
  f[12] = (float)align_log2((uint64_t)g.L);
  f[13] = (float)dir;
  f[14] = (g.L % 12 == 0 && g.L % 16 != 0) ? 1.f : 0.f;
  f[15] = g.nrd >= 1 ? log2p((double)std::llabs(g.str[0]) / L) : 0.f;
}

static int model_device_buffers(Model &m) {
  if (m.d_buf) {
    return TEMPI_SUCCESS;
  }
  const size_t bytes = (kNParams + 2 * kIn) * 4 + kB * kIn * 4 +
                       kB * kOut * 4 + kB * 4 + 1024;
  TEMPI_HIP_CHECK(hipMalloc(&m.d_buf, bytes));
  TEMPI_HIP_CHECK(hipMemcpy(m.d_buf, m.params.data(), kNParams * 4,
                            hipMemcpyHostToDevice));
  TEMPI_HIP_CHECK(hipMemcpy(m.d_buf + kNParams, m.norm.data(), 2 * kIn * 4,
                            hipMemcpyHostToDevice));
  return TEMPI_SUCCESS;
}

// run the inference kernel on n <= kB raw feature vectors
static int model_infer(Model &m, const float *x, int n, int *best,
                       float *pred) {
  int err = model_device_buffers(m);
  if (err) {
    return err;
  }
  float *d_norm = m.d_buf + kNParams;
  float *d_x = d_norm + 2 * kIn;
  float *d_pred = d_x + kB * kIn;
  int *d_best = reinterpret_cast<int *>(d_pred + kB * kOut);
  hipStream_t s = library_stream();
  TEMPI_HIP_CHECK(hipMemcpyAsync(d_x, x, (size_t)n * kIn * 4,
                                 hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mlp_infer, dim3(1), dim3(kThreads), 0, s, m.d_buf,
                     d_norm, d_x, n, d_pred, d_best);
  TEMPI_HIP_CHECK(hipGetLastError());
  TEMPI_HIP_CHECK(hipMemcpyAsync(best, d_best, (size_t)n * 4,
                                 hipMemcpyDeviceToHost, s));
  if (pred) {
    TEMPI_HIP_CHECK(hipMemcpyAsync(pred, d_pred, (size_t)n * kOut * 4,
                                   hipMemcpyDeviceToHost, s));
  }
  TEMPI_HIP_CHECK(hipStreamSynchronize(s));
  ++counters().model_inferences;
  return TEMPI_SUCCESS;
}

// run the inference kernel on n raw feature vectors (any n), one workgroup
// per chunk of kB, in buffers from the device pool
static int model_infer_many(Model &m, const float *x, int n, int *best) {
  int err = model_device_buffers(m);
  if (err) {
    return err;
  }
  const size_t xb = (size_t)n * kIn * 4;
  char *d = static_cast<char *>(pool_acquire(false, xb + (size_t)n * 4));
  if (!d) {
    return set_error(TEMPI_ERR_NOMEM, "selector inputs of %d copies", n);
  }
  hipStream_t s = library_stream();
  float *d_x = reinterpret_cast<float *>(d);
  int *d_best = reinterpret_cast<int *>(d + xb);
  const int grid = std::min((n + kB - 1) / kB, 1024);
  err = TEMPI_SUCCESS;
  if (hipSuccess != hipMemcpyAsync(d_x, x, xb, hipMemcpyHostToDevice, s)) {
    err = set_error(TEMPI_ERR_HIP, "selector input upload failed");
  } else {
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(kThreads), 0, s, m.d_buf,
                       m.d_buf + kNParams, d_x, n, nullptr, d_best);
    if (hipSuccess != hipGetLastError() ||
        hipSuccess != hipMemcpyAsync(best, d_best, (size_t)n * 4,
                                     hipMemcpyDeviceToHost, s)) {
      err = set_error(TEMPI_ERR_HIP, "selector inference failed");
    }
  }
  // the buffers go back to the pool only after the stream is done with them
  if (hipSuccess != hipStreamSynchronize(s) && !err) {
    err = set_error(TEMPI_ERR_HIP, "selector inference failed");
  }
  pool_release(false, d);
  if (!err) {
    ++counters().model_inferences;
  }
  return err;
}

// Launch the open batch (the newest pending one, if not launched yet): one
// inference over its types' feature vectors, read from and answered into
// its pinned slot through the slot's device mapping; its event fires when
// the picks are there.  A batch that cannot be launched is dropped (its
// shapes are evaluated again when first looked up).  The model's lock is
// held.
static void launch_open(Model &m) {
  if (m.pending.empty() || m.pending.back().launched) {
    return;
  }
  Model::Pending &b = m.pending.back();
  char *h = m.h_slots + b.slot * kSlotBytes;
  void *d = nullptr;
  hipStream_t s = library_stream();
  bool ok = device_address(h, &d) != 0;
  if (ok) {
    const float *df = static_cast<const float *>(d);
    int *db = reinterpret_cast<int *>(static_cast<char *>(d) + kSlotPicks);
    hipLaunchKernelGGL(k_mlp_infer, dim3(1), dim3(kThreads), 0, s, m.d_buf,
                       m.d_buf + kNParams, df, 2 * b.n, nullptr, db);
    ok = hipSuccess == hipGetLastError() &&
         hipSuccess == hipEventRecord(m.ev[b.slot], s);
  }
  if (!ok) {
    (void)hipGetLastError();
    (void)hipStreamSynchronize(s); // the slot's buffers may still be in use
    m.free_slots.push_back(b.slot);
    m.pending.pop_back();
    return;
  }
  b.launched = true;
  ++counters().model_inferences;
}

// Move the picks of finished commit-time inferences into the shape cache,
// oldest first; `wait` waits for each, launching the open batch first (else
// stops at the first still running, or at the open batch); `keep` false
// drops the picks, and open batches unlaunched (the weights they came from
// are being replaced).  The model's lock is held.
static void harvest(Model &m, bool wait, bool keep = true) {
  while (!m.pending.empty()) {
    Model::Pending &p = m.pending.front();
    if (!p.launched) {
      if (!keep) {
        m.free_slots.push_back(p.slot);
        m.pending.pop_front();
        continue;
      }
      if (!wait) {
        return;
      }
      launch_open(m);
      continue; // launched, or dropped
    }
    const hipError_t e = wait ? hipEventSynchronize(m.ev[p.slot])
                              : hipEventQuery(m.ev[p.slot]);
    if (e == hipErrorNotReady) {
      return;
    }
    if (e == hipSuccess && keep) {
      const int *picks =
          reinterpret_cast<const int *>(m.h_slots + p.slot * kSlotBytes + kSlotPicks);
      for (int v = 0; v < 2 * p.n; ++v) {
        if (picks[v] >= 0 && picks[v] < kOut) {
          m.cache[p.keys[(size_t)v]] = picks[v];
        }
      }
    } else if (e != hipSuccess) {
      (void)hipGetLastError();
    }
    m.free_slots.push_back(p.slot);
    m.pending.pop_front();
  }
}

// pinned slots and events for commit-time inferences (once per model)
static int model_slots(Model &m) {
  if (m.h_slots) {
    return TEMPI_SUCCESS;
  }
  void *h = nullptr;
  // mapped like the ONE_SHOT staging buffers: the kernel reads and writes
  // the slots through their device address
  TEMPI_HIP_CHECK(hipHostMalloc(&h, kSlots * kSlotBytes,
                                hipHostMallocMapped | hipHostMallocPortable));
  for (int i = 0; i < kSlots; ++i) {
    if (hipSuccess != hipEventCreateWithFlags(&m.ev[i], hipEventDisableTiming)) {
      for (int j = 0; j < i; ++j) {
        (void)hipEventDestroy(m.ev[j]);
        m.ev[j] = nullptr;
      }
      (void)hipHostFree(h);
      return set_error(TEMPI_ERR_HIP, "event creation failed");
    }
  }
  m.h_slots = static_cast<char *>(h);
  m.free_slots.clear();
  for (int i = kSlots - 1; i >= 0; --i) {
    m.free_slots.push_back(i);
  }
  return TEMPI_SUCCESS;
}

// `name` in the directory the library was loaded from
static std::string library_file(const char *name) {
  Dl_info info{};
  if (dladdr(reinterpret_cast<void *>(&library_file), &info) &&
      info.dli_fname) {
    std::string p(info.dli_fname);
    const size_t slash = p.rfind('/');
    return (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) +
           "/" + name;
  }
  return name;
}

static std::string default_model_path() {
  return library_file("selector_mlp.txt");
}

static int load_model_file(Model &m, const std::string &path) {
  FILE *f = fopen(path.c_str(), "r");
  if (!f) {
    return set_error(TEMPI_ERR_IO, "cannot open model %s", path.c_str());
  }
  std::string text;
  char chunk[4096];
  size_t got = 0;
  while ((got = fread(chunk, 1, sizeof chunk, f)) > 0) {
    text.append(chunk, got);
  }
  fclose(f);
  // whitespace-separated tokens; numbers read without the process locale
  std::vector<std::string> tok;
  for (size_t i = 0; i < text.size();) {
    while (i < text.size() && isspace((unsigned char)text[i])) {
      ++i;
    }
    size_t j = i;
    while (j < text.size() && !isspace((unsigned char)text[j])) {
      ++j;
    }
    if (j > i) {
      tok.emplace_back(text, i, j - i);
    }
    i = j;
  }
  size_t at = 0;
  auto word = [&](const char *w) { return at < tok.size() && tok[at++] == w; };
  auto integer = [&](int want) {
    int v = 0;
    return at < tok.size() &&
           parse_number(tok[at].data(), tok[at].data() + tok[at].size(), &v) &&
           ++at && v == want;
  };
  auto real = [&](float *v) {
    return at < tok.size() &&
           parse_number(tok[at].data(), tok[at].data() + tok[at].size(), v) &&
           ++at && std::isfinite(*v);
  };
  bool ok = word("tempi-mlp") && integer(1) && word("in") && integer(kIn) &&
            word("h1") && integer(kH1) && word("h2") && integer(kH2) &&
            word("out") && integer(kOut) && word("norm");
  std::vector<float> norm(2 * kIn), params(kNParams);
  for (int i = 0; ok && i < 2 * kIn; ++i) {
    ok = real(&norm[i]);
  }
  ok = ok && word("params");
  for (int i = 0; ok && i < kNParams; ++i) {
    ok = real(&params[i]);
  }
  ok = ok && at == tok.size();
  if (!ok) {
    return set_error(TEMPI_ERR_IO, "malformed model file %s", path.c_str());
  }
  // inferences of the old weights still in flight read d_buf: wait for
  // them and drop their picks
  harvest(m, /*wait=*/true, /*keep=*/false);
  if (m.d_buf) {
    (void)hipFree(m.d_buf);
    m.d_buf = nullptr;
  }
  m.params = params;
  m.norm = norm;
  m.cache.clear();
  m.loaded = true;
  m.path = path;
  return TEMPI_SUCCESS;
}

// the model, loading it on first use ($TEMPI_MODEL, else selector_mlp.txt
// next to the library); nullptr if there is none or it is unusable
static Model *loaded_model() {
  Model &m = model();
  std::lock_guard<std::mutex> lk(m.mu);
  if (!m.tried) {
    m.tried = true;
    const std::string path =
        env().model_path.empty() ? default_model_path() : env().model_path;
    if (load_model_file(m, path)) {
      TEMPI_LOG(LOG_INFO, "no variant model (%s): closed-form rule",
                tempi_last_error());
    }
  }
  return m.loaded ? &m : nullptr;
}

bool model_available() { return loaded_model() != nullptr; }

// variant from the model for this copy, cached per shape; -1 if unavailable
int model_variant(const Geom &g, const char *strided, const char *packed,
                  int dir, const std::string &key) {
  Model *m = loaded_model();
  if (!m || ensure_runtime()) {
    return -1;
  }
  std::lock_guard<std::mutex> lk(m->mu);
  auto it = m->cache.find(key);
  if (it == m->cache.end() && !m->pending.empty()) {
    // the commit-time inference of this shape may still be in flight, or
    // waiting in the open batch (then launched now)
    harvest(*m, /*wait=*/false);
    it = m->cache.find(key);
    bool pending = false;
    for (const Model::Pending &p : m->pending) {
      pending |= std::find(p.keys.begin(), p.keys.end(), key) != p.keys.end();
    }
    if (it == m->cache.end() && pending) {
      harvest(*m, /*wait=*/true);
      it = m->cache.find(key);
    }
  }
  if (it != m->cache.end()) {
    return it->second;
  }
  float f[kIn];
  copy_features(g, strided, packed, dir, f);
  int best = -1;
  if (model_infer(*m, f, 1, &best, nullptr) || best < 0 || best >= kOut) {
    return -1;
  }
  m->cache[key] = best;
  return best;
}

int model_ranked(const Geom &g, const char *strided, const char *packed,
                 int dir, int k, int *out) {
  Model *m = loaded_model();
  if (!m || k <= 0 || ensure_runtime()) {
    return 0;
  }
  float f[kIn];
  float pred[kOut];
  copy_features(g, strided, packed, dir, f);
  int best = -1;
  {
    std::lock_guard<std::mutex> lk(m->mu);
    if (model_infer(*m, f, 1, &best, pred)) {
      return 0;
    }
  }
  int idx[kOut];
  for (int v = 0; v < kOut; ++v) {
    idx[v] = v;
  }
  std::stable_sort(idx, idx + kOut,
                   [&](int a, int b) { return pred[a] < pred[b]; });
  const int n = std::min(k, kOut);
  for (int i = 0; i < n; ++i) {
    out[i] = idx[i];
  }
  return n;
}

// Commit-time evaluation: the type at incount 1 with 256-B aligned buffers,
// both directions, in an inference launch whose picks seed the shape cache.
// Nothing is waited for: the inputs and picks sit in a pinned slot until the
// first lookup of the shape (or a later commit) finds its batch's event
// fired, so MPI_Type_commit does not wait for the GPU -- nor for copies
// already queued on the library stream.  Commits in quick succession share
// launches (the open batch), so a burst of commits is not paced by the
// inference kernel.
void model_precommit(const tempi_strided_block &sb) {
  Model *m = loaded_model();
  if (!m || !have_device() || ensure_runtime()) {
    return;
  }
  Geom g;
  if (make_geom(sb, 1, &g) || g.rows == 0) {
    return;
  }
  // a 256-aligned stand-in for the buffer origin
  const char *base = reinterpret_cast<const char *>((uintptr_t)1 << 20);
  char *pk = reinterpret_cast<char *>((uintptr_t)1 << 21);
  std::lock_guard<std::mutex> lk(m->mu);
  if (model_device_buffers(*m) || model_slots(*m)) {
    return;
  }
  harvest(*m, /*wait=*/false);
  if (m->pending.empty() || m->pending.back().launched) {
    // open a batch
    if (m->free_slots.empty()) {
      // every slot holds a launched batch: wait for the oldest
      (void)hipEventSynchronize(m->ev[m->pending.front().slot]);
      harvest(*m, /*wait=*/false);
    }
    if (m->free_slots.empty()) {
      return;
    }
    Model::Pending b;
    b.slot = m->free_slots.back();
    m->free_slots.pop_back();
    b.keys.reserve(kSlotVecs);
    m->pending.push_back(std::move(b));
  }
  // this type's two feature vectors (pack, unpack) and their picks, written
  // before the batch's launch (whose doorbell releases them to the device)
  Model::Pending &b = m->pending.back();
  char *h = m->h_slots + b.slot * kSlotBytes;
  float *hf = reinterpret_cast<float *>(h) + (size_t)2 * b.n * kIn;
  copy_features(g, base, pk, 0, hf);
  copy_features(g, base, pk, 1, hf + kIn);
  int *hb = reinterpret_cast<int *>(h + kSlotPicks) + 2 * b.n;
  hb[0] = hb[1] = -1;
  b.keys.push_back(shape_key_of(g, base, pk, 0));
  b.keys.push_back(shape_key_of(g, base, pk, 1));
  ++b.n;
  // launch now when no inference of ours is running (the open batch is the
  // only one pending) or when the batch is full; else the next commit or
  // lookup launches it
  if (m->pending.size() == 1 || b.n == kSlotTypes) {
    launch_open(*m);
  }
}

void model_prefetch(const std::vector<CopyShape> &copies) {
  Model *m = loaded_model();
  if (!m || copies.empty() || ensure_runtime()) {
    return;
  }
  std::lock_guard<std::mutex> lk(m->mu);
  harvest(*m, /*wait=*/false);
  std::vector<float> x;
  std::vector<std::string> keys;
  std::set<std::string> seen;
  for (const CopyShape &c : copies) {
    if (times_cold(c.g)) {
      continue;
    }
    std::string key = shape_key_of(c.g, c.strided, c.packed, c.dir);
    if (m->cache.count(key) || !seen.insert(key).second) {
      continue;
    }
    x.resize(x.size() + kIn);
    copy_features(c.g, c.strided, c.packed, c.dir, x.data() + x.size() - kIn);
    keys.push_back(std::move(key));
  }
  if (keys.empty()) {
    return;
  }
  std::vector<int> best(keys.size(), -1);
  if (model_infer_many(*m, x.data(), (int)keys.size(), best.data())) {
    return; // the per-copy choices evaluate these shapes one by one
  }
  for (size_t i = 0; i < keys.size(); ++i) {
    if (best[i] >= 0 && best[i] < kOut) {
      m->cache[keys[i]] = best[i];
    }
  }
}

// at MPI_Finalize: wait for commit-time inferences still in flight
void model_finalize() {
  Model &m = model();
  std::lock_guard<std::mutex> lk(m.mu);
  harvest(m, /*wait=*/true);
}

// ---------------------------------------------------------------------------
// the staging-method model: the same network, another set of weights
// ---------------------------------------------------------------------------
//
// The reference prices each way of moving a non-contiguous message (DEVICE,
// ONE_SHOT, STAGED) as a sum of measured table entries (pack + transfer +
// unpack, src/internal/sender.cpp:128-273).  Here a second instance of the
// selector's network is trained on end-to-end timings of each method on
// device-resident derived types (tempi_amd/method_train.py): output k
// predicts method k + 1's time (log2, relative to the fastest measured for
// the message), and the cheapest of the methods that can run is taken.
// Features (the rest zero): log2 packed bytes, row bytes, rows, stride and
// stride / row bytes; peer on the same node; log2 alignment of the row
// length.

namespace {
Model &method_model_slot() {
  static Model m;
  return m;
}
std::atomic<uint64_t> g_method_model_generation{0};
} // namespace

uint64_t method_model_generation() { return g_method_model_generation.load(); }

void method_features(int64_t bytes, int64_t block, int64_t stride,
                     bool colocated, float *f) {
  for (int i = 0; i < kIn; ++i) {
    f[i] = 0.f;
  }
  const double b = (double)std::max<int64_t>(bytes, 1);
  const double bl = (double)std::max<int64_t>(block, 1);
  const double st = (double)std::max<int64_t>(stride, block);
  f[0] = log2p(b);
  f[1] = log2p(bl);
  f[2] = log2p(b / bl);
  f[3] = log2p(st);
  f[4] = log2p(st / bl);
  f[5] = colocated ? 1.f : 0.f;
  f[6] = (float)align_log2((uint64_t)std::max<int64_t>(block, 1));
}

// $TEMPI_METHOD_MODEL, else method_mlp.txt next to the library, loaded on
// first use; nullptr if there is none
static Model *loaded_method_model() {
  ensure_env();
  Model &m = method_model_slot();
  std::lock_guard<std::mutex> lk(m.mu);
  if (!m.tried) {
    m.tried = true;
    const std::string path = env().method_model_path.empty()
                                 ? library_file("method_mlp.txt")
                                 : env().method_model_path;
    if (load_model_file(m, path)) {
      TEMPI_LOG(LOG_INFO, "no method model (%s): measured tables",
                tempi_last_error());
    }
  }
  return m.loaded ? &m : nullptr;
}

int method_model_pick(const float *f, const int *candidates, int n) {
  Model *m = loaded_method_model();
  if (!m || n <= 0 || ensure_runtime()) {
    return -1;
  }
  float pred[kOut];
  int best = -1;
  {
    std::lock_guard<std::mutex> lk(m->mu);
    if (model_infer(*m, f, 1, &best, pred)) {
      return -1;
    }
  }
  int pick = -1;
  float lowest = INFINITY;
  for (int i = 0; i < n; ++i) {
    const int k = candidates[i] - 1; // output k: method k + 1
    if (k >= 0 && k < kOut && pred[k] < lowest) {
      lowest = pred[k];
      pick = candidates[i];
    }
  }
  return pick;
}

} // namespace tempi

using namespace tempi;

TEMPI_EXPORT int tempi_mlp_num_features(void) { return kIn; }
TEMPI_EXPORT int tempi_mlp_num_outputs(void) { return kOut; }
TEMPI_EXPORT int tempi_mlp_num_params(void) { return kNParams; }

TEMPI_EXPORT int tempi_mlp_features(const tempi_strided_block *sb,
                                    int64_t incount, const void *strided,
                                    const void *packed, int direction,
                                    float *features) {
  if (!sb || !features) {
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
  copy_features(g, static_cast<const char *>(strided),
                static_cast<const char *>(packed), direction, features);
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_mlp_grad(const float *params, const float *X,
                                const float *T, const float *M, int n,
                                float *grad, float *loss) {
  if (!params || !X || !T || !grad || !loss || n <= 0 || n > kB) {
    return set_error(TEMPI_ERR_ARG, "bad argument (1 <= n <= %d)", kB);
  }
  int err = ensure_runtime();
  if (err) {
    return err;
  }
  const size_t np = kNParams * 4, nx = (size_t)n * kIn * 4,
               nt = (size_t)n * kOut * 4;
  char *d = nullptr;
  TEMPI_HIP_CHECK(hipMalloc(&d, np * 2 + nx + 2 * nt + kB * 4 + 64));
  float *d_p = reinterpret_cast<float *>(d);
  float *d_g = reinterpret_cast<float *>(d + np);
  float *d_x = reinterpret_cast<float *>(d + 2 * np);
  float *d_t = reinterpret_cast<float *>(d + 2 * np + nx);
  float *d_m = reinterpret_cast<float *>(d + 2 * np + nx + nt);
  uint32_t *d_rows = reinterpret_cast<uint32_t *>(d + 2 * np + nx + 2 * nt);
  float *d_loss = reinterpret_cast<float *>(d + 2 * np + nx + 2 * nt + kB * 4);
  std::vector<uint32_t> rows(n);
  for (int i = 0; i < n; ++i) {
    rows[i] = (uint32_t)i;
  }
  hipStream_t s = library_stream();
  err = TEMPI_SUCCESS;
  if (hipMemcpy(d_p, params, np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_x, X, nx, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_t, T, nt, hipMemcpyHostToDevice) != hipSuccess ||
      (M && hipMemcpy(d_m, M, nt, hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(d_rows, rows.data(), n * 4, hipMemcpyHostToDevice) !=
          hipSuccess) {
    err = set_error(TEMPI_ERR_HIP, "upload failed");
  }
  if (!err) {
    hipLaunchKernelGGL(k_mlp_grad, dim3(1), dim3(kThreads), 0, s, d_p, d_x,
                       d_t, M ? d_m : nullptr, d_rows, n, d_g, d_loss);
    if (hipStreamSynchronize(s) != hipSuccess ||
        hipMemcpy(grad, d_g, np, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(loss, d_loss, 4, hipMemcpyDeviceToHost) != hipSuccess) {
      err = set_error(TEMPI_ERR_HIP, "gradient kernel failed");
    }
  }
  (void)hipFree(d);
  return err;
}

TEMPI_EXPORT int tempi_mlp_train(const float *X, const float *T,
                                 const float *M, int64_t n, int epochs,
                                 float lr, uint64_t seed, float *params,
                                 float *epoch_loss) {
  if (!X || !T || !params || n <= 0 || epochs <= 0 || n >= (1ll << 31)) {
    return set_error(TEMPI_ERR_ARG, "bad argument");
  }
  int err = ensure_runtime();
  if (err) {
    return err;
  }
  const size_t np = kNParams * 4, nx = (size_t)n * kIn * 4,
               nt = (size_t)n * kOut * 4, no = (size_t)n * 4;
  char *d = nullptr;
  TEMPI_HIP_CHECK(hipMalloc(&d, 3 * np + nx + 2 * nt + no + 256));
  float *d_p = reinterpret_cast<float *>(d);
  float *d_m1 = reinterpret_cast<float *>(d + np);
  float *d_m2 = reinterpret_cast<float *>(d + 2 * np);
  float *d_x = reinterpret_cast<float *>(d + 3 * np);
  float *d_t = reinterpret_cast<float *>(d + 3 * np + nx);
  float *d_m = reinterpret_cast<float *>(d + 3 * np + nx + nt);
  uint32_t *d_o = reinterpret_cast<uint32_t *>(d + 3 * np + nx + 2 * nt);
  float *d_l = reinterpret_cast<float *>(d + 3 * np + nx + 2 * nt + no);
  std::vector<float> zeros(kNParams, 0.f);
  std::vector<uint32_t> order((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    order[i] = (uint32_t)i;
  }
  std::mt19937_64 rng(seed);
  hipStream_t s = library_stream();
  if (hipMemcpy(d_p, params, np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_m1, zeros.data(), np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_m2, zeros.data(), np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_x, X, nx, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_t, T, nt, hipMemcpyHostToDevice) != hipSuccess ||
      (M && hipMemcpy(d_m, M, nt, hipMemcpyHostToDevice) != hipSuccess)) {
    (void)hipFree(d);
    return set_error(TEMPI_ERR_HIP, "upload failed");
  }
  const int steps_per_epoch = (int)((n + kB - 1) / kB);
  for (int e = 0; e < epochs && !err; ++e) {
    std::shuffle(order.begin(), order.end(), rng);
    if (hipMemcpy(d_o, order.data(), no, hipMemcpyHostToDevice) !=
        hipSuccess) {
      err = set_error(TEMPI_ERR_HIP, "upload failed");
      break;
    }
    hipLaunchKernelGGL(k_mlp_train_epoch, dim3(1), dim3(kThreads), 0, s, d_p,
                       d_m1, d_m2, d_x, d_t, M ? d_m : nullptr, d_o, (int)n,
                       lr, e * steps_per_epoch, d_l);
    float l = 0;
    if (hipStreamSynchronize(s) != hipSuccess ||
        hipMemcpy(&l, d_l, 4, hipMemcpyDeviceToHost) != hipSuccess) {
      err = set_error(TEMPI_ERR_HIP, "training kernel failed");
      break;
    }
    if (epoch_loss) {
      epoch_loss[e] = l;
    }
  }
  if (!err && hipMemcpy(params, d_p, np, hipMemcpyDeviceToHost) != hipSuccess) {
    err = set_error(TEMPI_ERR_HIP, "download failed");
  }
  (void)hipFree(d);
  return err;
}

TEMPI_EXPORT int tempi_mlp_predict(const float *params, const float *norm,
                                   const float *X, int64_t n, float *pred,
                                   int *best) {
  if (!params || !norm || !X || n <= 0 || n >= (1ll << 31)) {
    return set_error(TEMPI_ERR_ARG, "bad argument");
  }
  int err = ensure_runtime();
  if (err) {
    return err;
  }
  const size_t np = kNParams * 4, nn = 2 * kIn * 4, nx = (size_t)n * kIn * 4,
               npd = (size_t)n * kOut * 4, nb = (size_t)n * 4;
  char *d = nullptr;
  TEMPI_HIP_CHECK(hipMalloc(&d, np + nn + nx + npd + nb + 256));
  float *d_p = reinterpret_cast<float *>(d);
  float *d_n = reinterpret_cast<float *>(d + np);
  float *d_x = reinterpret_cast<float *>(d + np + nn);
  float *d_pr = reinterpret_cast<float *>(d + np + nn + nx);
  int *d_b = reinterpret_cast<int *>(d + np + nn + nx + npd);
  hipStream_t s = library_stream();
  if (hipMemcpy(d_p, params, np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_n, norm, nn, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_x, X, nx, hipMemcpyHostToDevice) != hipSuccess) {
    err = set_error(TEMPI_ERR_HIP, "upload failed");
  }
  if (!err) {
    hipLaunchKernelGGL(k_mlp_infer, dim3(1), dim3(kThreads), 0, s, d_p, d_n,
                       d_x, (int)n, d_pr, d_b);
    if (hipStreamSynchronize(s) != hipSuccess ||
        (pred && hipMemcpy(pred, d_pr, npd, hipMemcpyDeviceToHost) !=
                     hipSuccess) ||
        (best && hipMemcpy(best, d_b, nb, hipMemcpyDeviceToHost) !=
                     hipSuccess)) {
      err = set_error(TEMPI_ERR_HIP, "inference kernel failed");
    }
  }
  (void)hipFree(d);
  return err;
}

TEMPI_EXPORT int tempi_mlp_save(const char *path, const float *params,
                                const float *norm) {
  if (!path || !params || !norm) {
    return set_error(TEMPI_ERR_ARG, "null argument");
  }
  FILE *f = fopen(path, "w");
  if (!f) {
    return set_error(TEMPI_ERR_IO, "cannot write %s", path);
  }
  fprintf(f, "tempi-mlp 1\nin %d h1 %d h2 %d out %d\nnorm", kIn, kH1, kH2,
          kOut);
  for (int i = 0; i < 2 * kIn; ++i) {
    fprintf(f, " %s", format_number(norm[i]).c_str());
  }
  fprintf(f, "\nparams");
  for (int i = 0; i < kNParams; ++i) {
    fprintf(f, "%s%s", (i % 16) ? " " : "\n", format_number(params[i]).c_str());
  }
  fprintf(f, "\n");
  const bool ok = 0 == fclose(f);
  return ok ? TEMPI_SUCCESS : set_error(TEMPI_ERR_IO, "write %s failed", path);
}

TEMPI_EXPORT int tempi_mlp_load(const char *path) {
  Model &m = model();
  std::lock_guard<std::mutex> lk(m.mu);
  m.tried = true;
  const int err = load_model_file(m, path ? path : default_model_path());
  if (err) {
    m.loaded = false;
  }
  return err;
}

TEMPI_EXPORT int tempi_mlp_loaded(void) { return model_available() ? 1 : 0; }

TEMPI_EXPORT int tempi_method_features(int64_t bytes, int64_t block,
                                       int64_t stride, int colocated,
                                       float *features) {
  if (!features || bytes < 0 || block < 0) {
    return set_error(TEMPI_ERR_ARG, "bad argument");
  }
  method_features(bytes, block, stride, colocated != 0, features);
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_method_model_load(const char *path) {
  ensure_env();
  Model &m = method_model_slot();
  std::lock_guard<std::mutex> lk(m.mu);
  m.tried = true;
  const int err =
      load_model_file(m, path ? path : library_file("method_mlp.txt"));
  if (err) {
    m.loaded = false;
  }
  ++g_method_model_generation;
  return err;
}

TEMPI_EXPORT int tempi_method_model_unload(void) {
  Model &m = method_model_slot();
  std::lock_guard<std::mutex> lk(m.mu);
  m.tried = true;
  m.loaded = false;
  ++g_method_model_generation;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_method_model_loaded(void) {
  return loaded_method_model() ? 1 : 0;
}
