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
  if (g.rows >= (1ll << 32) || units >= (1ll << 32) || words >= (1ll << 32)) {
    return set_error(TEMPI_ERR_UNSUPPORTED,
                     "copy of %lld rows of %lld B exceeds the 2^32 unit space",
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

static int launch_one(const Geom &g, const char *strided, char *packed,
                      int dir, int variant, hipStream_t stream,
                      int max_blocks) {
  KParams p;
  bool aligned = false;
  int err = fill_kparams(g, strided, packed, dir, variant, &p, &aligned);
  if (err) {
    return err;
  }
  const int64_t groups = (int64_t)(kBlock / kVariants[variant].P) *
                         units_in_flight(kVariants[variant].U);
  int64_t blocks = ((int64_t)p.units + groups - 1) / groups;
  const int64_t cap =
      max_blocks > 0 ? max_blocks : (int64_t)device_cu_count() * 16;
  blocks = std::min<int64_t>(blocks, cap);
  blocks = std::max<int64_t>(blocks, 1);
  KernelFn fn = kKernels[variant][aligned ? 1 : 0][dir];
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
  // rows longer than 1 GiB are moved as 1 GiB pieces, one launch each
  constexpr int64_t kPiece = 1ll << 30;
  if (g.L <= kPiece) {
    return launch_one(g, strided, packed, dir, variant, stream, max_blocks);
  }
  Geom piece = g;
  piece.nrd = 0;
  piece.rows = 1;
  for (int64_t r = 0; r < g.rows; ++r) {
    // strided offset of row r
    int64_t off = g.start, q = r;
    for (int k = 0; k < g.nrd; ++k) {
      off += (q % g.cnt[k]) * g.str[k];
      q /= g.cnt[k];
    }
    for (int64_t x = 0; x < g.L; x += kPiece) {
      piece.L = std::min(kPiece, g.L - x);
      piece.start = off + x;
      int err = launch_one(piece, strided, packed + r * g.L + x, dir, variant,
                           stream, max_blocks);
      if (err) {
        return err;
      }
    }
  }
  return TEMPI_SUCCESS;
}

// exported for the host-side unit test of the divisor arithmetic
extern "C" __attribute__((visibility("default"))) uint32_t
tempi_fastdiv_check(uint32_t n, uint32_t d) {
  return make_fastdiv(d).div(n);
}

} // namespace tempi
