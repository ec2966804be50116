// copy_core.hpp -- the per-unit strided copy shared by the single-copy
// kernels (kernels.hip) and the batched queue kernel (batch.hip).
//
// A copy is `units` work units; unit u is segment (u % nseg) of row
// (u / nseg).  A group of P lanes moves one unit as U words of W bytes per
// lane, the P lanes touching consecutive words; each thread issues the loads
// of K = units_in_flight(U) units before storing any of them.  Stores are always aligned
// to the word's natural unit A (W for powers of two, 4 for 12-byte words);
// loads may sit at any byte address (gfx9 global loads accept that).
#pragma once

#include "internal.hpp"

#include <hip/hip_runtime.h>

namespace tempi {

// n / d for 32-bit n by multiply-high and shift (Granlund and Montgomery,
// "Division by Invariant Integers using Multiplication", PLDI 1994, fig. 4.1):
// with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
// n / d = (mulhi(n, m) + n) >> l, the sum taken in 64 bits.
struct FastDiv {
  uint32_t mul = 1;
  uint32_t shift = 0;
  __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const {
    const uint32_t t = (uint32_t)(((uint64_t)n * mul) >> 32);
    return (uint32_t)(((uint64_t)t + n) >> shift);
  }
};

FastDiv make_fastdiv(uint32_t d);

struct KParams {
  const char *src; // first byte of row 0 on the source side
  char *dst;       // first byte of row 0 on the destination side
  int64_t L;       // bytes per row
  uint32_t units;  // rows * nseg
  uint32_t nseg;   // segments per row
  FastDiv segdiv;
  uint32_t words; // L / W
  int32_t nrd;    // row dimensions
  uint32_t cnt[kMaxRowDims];
  FastDiv div[kMaxRowDims];
  int64_t str[kMaxRowDims];
  uint32_t signal; // nonzero: the launch signals its completion (kernels.hip)
};

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

template <int W> struct WordT;
template <> struct WordT<16> { typedef u32x4 T; };
template <> struct WordT<12> { typedef u32x3 T; };
template <> struct WordT<8> { typedef u32x2 T; };
template <> struct WordT<4> { typedef uint32_t T; };
template <> struct WordT<2> { typedef uint16_t T; };
template <> struct WordT<1> { typedef uint8_t T; };

// alignment unit of a W-byte word
__host__ __device__ constexpr int word_align(int W) {
  return (W & (W - 1)) == 0 ? W : 4;
}

// load W bytes from any byte address; NT marks the load nontemporal (the
// line is filled for this access and is first in line for eviction)
template <int W, bool NT>
__device__ __forceinline__ typename WordT<W>::T load_word(const char *p) {
  typedef typename WordT<W>::T T;
  typedef T Tu __attribute__((aligned(1)));
  if constexpr (NT) {
    return __builtin_nontemporal_load(reinterpret_cast<const Tu *>(p));
  } else {
    return *reinterpret_cast<const Tu *>(p);
  }
}

// store W bytes to an address aligned to word_align(W)
template <int W, bool NT>
__device__ __forceinline__ void store_word(char *p, typename WordT<W>::T v) {
  typedef typename WordT<W>::T T;
  typedef T Ta __attribute__((aligned(word_align(W))));
  if constexpr (NT) {
    __builtin_nontemporal_store(v, reinterpret_cast<Ta *>(p));
  } else {
    *reinterpret_cast<Ta *>(p) = v;
  }
}

// Copy n < 16 bytes to d from s (any alignment) with the widest stores d's
// alignment allows: 1, 2, 4 then 8 bytes on the way up to an aligned
// address, 8, 4, 2 then 1 on the way down from one (at most four stores
// either way).
__device__ __forceinline__ void copy_short(char *d, const char *s,
                                           uint32_t n) {
  while (n) {
    const uintptr_t a = (uintptr_t)d;
    if ((a & 1) || n < 2) {
      *d = *s;
      d += 1;
      s += 1;
      n -= 1;
    } else if ((a & 2) || n < 4) {
      store_word<2, false>(d, load_word<2, false>(s));
      d += 2;
      s += 2;
      n -= 2;
    } else if ((a & 4) || n < 8) {
      store_word<4, false>(d, load_word<4, false>(s));
      d += 4;
      s += 4;
      n -= 4;
    } else {
      store_word<8, false>(d, load_word<8, false>(s));
      d += 8;
      s += 8;
      n -= 8;
    }
  }
}

// byte offset of row `row` on the strided side
__host__ __device__ __forceinline__ int64_t row_offset(const KParams &p,
                                                       uint32_t row) {
  int64_t off = 0;
  uint32_t q = row;
#pragma unroll
  for (int k = 0; k < kMaxRowDims; ++k) {
    if (k >= p.nrd) {
      break;
    }
    if (k == p.nrd - 1) {
      off += (int64_t)q * p.str[k];
      break;
    }
    const uint32_t nq = p.div[k].div(q);
    off += (int64_t)(q - nq * p.cnt[k]) * p.str[k];
    q = nq;
  }
  return off;
}

// How unit_addr moves the head and tail bytes of a misaligned row: the
// kernels copy them (DeviceMem); the host-side emulation of a copy records
// the bytes instead (kernels.hip), so one definition of the address
// arithmetic serves both.
struct DeviceMem {
  __device__ static __forceinline__ void copy_short(char *d, const char *s,
                                                    uint32_t n) {
    ::tempi::copy_short(d, s, n);
  }
  __device__ static __forceinline__ void copy_byte(char *d, const char *s) {
    *d = *s;
  }
};

// Addresses of one work unit, computed before any of the thread's loads.
struct UnitAddr {
  const char *s; // first body byte on the source side
  char *d;       // first body byte on the destination side
  uint32_t w;    // first word index of this lane
  uint32_t n;    // body words of the row (0: nothing to move)
};

// Address phase of work unit `u` (wave-uniform when P == 64).  DIR 0 = pack
// (strided -> packed), 1 = unpack (packed -> strided).  Head and tail bytes
// of a row whose destination is not word-aligned are copied here directly.
template <class M, int P, int W, int U, bool ALIGNED, int DIR>
__host__ __device__ __forceinline__ UnitAddr unit_addr(const KParams &p,
                                                       uint32_t u,
                                                       uint32_t lane) {
  constexpr int A = word_align(W);
  const uint32_t row = p.nseg == 1 ? u : p.segdiv.div(u);
  const uint32_t seg = u - row * p.nseg;
  const int64_t soff = row_offset(p, row);
  const int64_t poff = (int64_t)row * p.L;
  UnitAddr a;
  a.s = p.src + (DIR == 0 ? soff : poff);
  a.d = p.dst + (DIR == 0 ? poff : soff);
  a.w = seg * (uint32_t)(P * U) + lane;
  if constexpr (ALIGNED) {
    a.n = p.words;
  } else {
    const uint32_t mis = (uint32_t)((uintptr_t)a.d & (A - 1));
    uint32_t head = mis ? (uint32_t)(A - mis) : 0u;
    if ((int64_t)head > p.L) {
      head = (uint32_t)p.L;
    }
    a.n = (uint32_t)((p.L - head) / W);
    if (seg == 0) {
      const int64_t t0 = (int64_t)head + (int64_t)a.n * W;
      const uint32_t tail = (uint32_t)(p.L - t0);
      if constexpr (P == 1) {
        // one lane per row: the widest aligned stores, at most four each
        M::copy_short(a.d, a.s, head);
        M::copy_short(a.d + t0, a.s + t0, tail);
      } else {
        for (uint32_t b = lane; b < head; b += P) {
          M::copy_byte(a.d + b, a.s + b);
        }
        for (uint32_t b = lane; b < tail; b += P) {
          M::copy_byte(a.d + t0 + b, a.s + t0 + b);
        }
      }
    }
    a.s += head;
    a.d += head;
  }
  return a;
}

// s_waitcnt immediate for vmcnt(0) alone on gfx9 (expcnt and lgkmcnt at
// their maxima, 7 and 15, so only the vector memory counter is waited on)
constexpr int kWaitLoads = 0x0F70;

// The words of one unit, held in registers between the load and store phases.
template <int W, int U> struct UnitWords {
  typename WordT<W>::T v[U];
};

// units a thread keeps in flight: enough for four loads per lane
__host__ __device__ constexpr int units_in_flight(int U) {
  return U >= 4 ? 1 : 4 / U;
}

// Move units first, first + stride, ... (K of them, those below `end`) in
// three phases: every unit's addresses, then every load, then every store.
// Keeping the loads in one straight run matters: with the address arithmetic
// of later units between them, the compiler put a wait for the earlier loads
// in front of the later ones (register reuse across the grid-stride loop),
// and each unit paid a full memory round trip of its own.  NTS / NTL:
// nontemporal stores / loads of the body words.
template <int P, int W, int U, bool ALIGNED, int DIR, bool NTS, bool NTL>
__device__ __forceinline__ void copy_units(const KParams &p, uint32_t first,
                                           uint32_t stride, uint32_t end,
                                           uint32_t lane) {
  constexpr int K = units_in_flight(U);
  UnitAddr a[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    uint32_t u = first + k * stride;
    if constexpr (P == 64 && K == 1) {
      // wave-uniform: the row arithmetic runs on the scalar unit (with
      // several units in flight their state would overflow the SGPRs)
      u = __builtin_amdgcn_readfirstlane(u);
    }
    if (u < end) {
      a[k] = unit_addr<DeviceMem, P, W, U, ALIGNED, DIR>(p, u, lane);
    } else {
      a[k].s = nullptr;
      a[k].d = nullptr;
      a[k].w = 0;
      a[k].n = 0;
    }
  }
  UnitWords<W, U> v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t w = a[k].w + j * P;
      if (w < a[k].n) {
        v[k].v[j] = load_word<W, NTL>(a[k].s + (size_t)w * W);
      }
    }
  }
  // Every load lands before any store, by one wait in straight-line code.
  // Left to the compiler, the waits sit inside the per-word store
  // conditions; on a path that skips a store its load counts as still
  // pending at the loop back edge, and the next iteration waits for it in
  // the middle of its own loads.
  __builtin_amdgcn_s_waitcnt(kWaitLoads);
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t w = a[k].w + j * P;
      if (w < a[k].n) {
        store_word<W, NTS>(a[k].d + (size_t)w * W, v[k].v[j]);
      }
    }
  }
}

constexpr int kBlock = 256;

// The variant list, as X(index, P, W, U, NTS, NTL): P lanes per row, W-byte
// words, U words per lane per unit, nontemporal stores / loads.
#define TEMPI_VARIANTS(X)                                                      \
  X(0, 64, 16, 1, false, false) /* long rows, 1 KiB per wavefront unit */     \
  X(1, 64, 16, 4, false, false) /* long rows, 4 KiB per unit           */     \
  X(2, 64, 16, 4, true, false)  /* as 1, nontemporal stores            */     \
  X(3, 32, 16, 2, false, false) /* rows around 512 B - 1 KiB           */     \
  X(4, 16, 16, 2, false, false) /* rows around 256 - 512 B             */     \
  X(5, 8, 16, 2, false, false)  /* rows around 128 - 256 B             */     \
  X(6, 4, 16, 2, false, false)  /* rows around 64 - 128 B              */     \
  X(7, 1, 16, 4, false, false)  /* lane per row, 16-B words            */     \
  X(8, 1, 8, 4, false, false)   /* lane per row, 8-B words             */     \
  X(9, 1, 4, 4, false, false)   /* lane per row, 4-B words             */     \
  X(10, 16, 4, 4, false, false) /* 4-B words across 16 lanes           */     \
  X(11, 1, 1, 16, false, false) /* lane per row, bytes                 */     \
  X(12, 1, 12, 2, false, false) /* lane per row, 12-B words (3 floats) */     \
  X(13, 64, 16, 4, false, true) /* as 1, nontemporal loads             */     \
  X(14, 32, 16, 2, false, true) /* as 3, nontemporal loads             */     \
  X(15, 16, 16, 2, false, true) /* as 4, nontemporal loads             */     \
  X(16, 8, 16, 2, false, true)  /* as 5, nontemporal loads             */     \
  X(17, 4, 16, 2, false, true)  /* as 6, nontemporal loads             */     \
  X(18, 1, 8, 4, false, true)   /* as 8, nontemporal loads             */

// Work units of one launch stay below 2^31.  The kernels index units with
// 32-bit counters that step by at most (grid * groups * units in flight)
// < 2^24 past the last unit, so no index wraps; launch_copy splits a larger
// copy along its outer row dimensions.
constexpr int64_t kMaxUnits = 1ll << 31;

// Fill the kernel parameters of one copy with `variant`; a copy of
// kMaxUnits units or more is reported as TEMPI_ERR_UNSUPPORTED.
int fill_kparams(const Geom &g, const char *strided, char *packed, int dir,
                 int variant, KParams *p, bool *aligned);

} // namespace tempi
