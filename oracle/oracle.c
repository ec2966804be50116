/*
 * oracle.c -- CPU oracle for MPI_Pack / MPI_Unpack of derived datatypes.
 *
 * TEST INFRASTRUCTURE ONLY: never linked into or called by the product
 * library.  See oracle.h for what each half restates and how it is pinned.
 */
#include "oracle.h"

#include <string.h>

/* ------------------------------------------------------------------------ */
/* type geometry                                                             */
/* ------------------------------------------------------------------------ */

static int64_t min64(int64_t a, int64_t b) { return a < b ? a : b; }
static int64_t max64(int64_t a, int64_t b) { return a > b ? a : b; }

/* lb and ub of node i.  For the replicating constructors the displacement of
   each copy of the child is affine in the copy indices, so the extremes sit at
   the corner indices (MPI-3.1 section 4.1, the lb/ub definition of a
   typemap).  Subarray is resized to [0, prod(sizes)*extent(old)) by
   definition (MPI-3.1 section 4.1.3), which the reference also relies on
   (src/internal/types.cpp:311-315 reads that extent back). */
static void ot_bounds(const ot_node *nodes, int32_t i, int64_t *lb,
                      int64_t *ub) {
  const ot_node *n = &nodes[i];
  int64_t clb = 0, cub = 0, cext = 0;
  if (n->kind != OT_NAMED) {
    ot_bounds(nodes, n->child, &clb, &cub);
    cext = cub - clb;
  }
  switch (n->kind) {
  case OT_NAMED:
    *lb = 0;
    *ub = n->size;
    return;
  case OT_CONTIGUOUS: {
    if (n->count <= 0) {
      *lb = *ub = 0;
      return;
    }
    int64_t d0 = 0, d1 = (n->count - 1) * cext;
    *lb = min64(d0, d1) + clb;
    *ub = max64(d0, d1) + cub;
    return;
  }
  case OT_VECTOR:
  case OT_HVECTOR: {
    if (n->count <= 0 || n->blocklength <= 0) {
      *lb = *ub = 0;
      return;
    }
    int64_t s = n->kind == OT_VECTOR ? n->stride * cext : n->stride;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    int64_t is[2] = {0, n->count - 1}, js[2] = {0, n->blocklength - 1};
    for (int a = 0; a < 2; ++a) {
      for (int b = 0; b < 2; ++b) {
        int64_t d = is[a] * s + js[b] * cext;
        lo = min64(lo, d + clb);
        hi = max64(hi, d + cub);
      }
    }
    *lb = lo;
    *ub = hi;
    return;
  }
  case OT_SUBARRAY: {
    int64_t e = cext;
    for (int d = 0; d < n->ndims; ++d) {
      e *= n->sizes[d];
    }
    *lb = 0;
    *ub = e;
    return;
  }
  case OT_RESIZED:
    *lb = n->lb;
    *ub = n->lb + n->extent;
    return;
  case OT_HINDEXED: {
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (int64_t k = 0; k < n->count; ++k) {
      if (n->blens[k] <= 0) {
        continue;
      }
      int64_t d0 = n->disps[k], d1 = n->disps[k] + (n->blens[k] - 1) * cext;
      lo = min64(lo, min64(d0, d1) + clb);
      hi = max64(hi, max64(d0, d1) + cub);
    }
    if (lo > hi) {
      lo = hi = 0;
    }
    *lb = lo;
    *ub = hi;
    return;
  }
  }
  *lb = *ub = 0;
}

int64_t ot_lb(const ot_node *nodes, int32_t i) {
  int64_t lb, ub;
  ot_bounds(nodes, i, &lb, &ub);
  return lb;
}

int64_t ot_extent(const ot_node *nodes, int32_t i) {
  int64_t lb, ub;
  ot_bounds(nodes, i, &lb, &ub);
  return ub - lb;
}

int64_t ot_size(const ot_node *nodes, int32_t i) {
  const ot_node *n = &nodes[i];
  switch (n->kind) {
  case OT_NAMED:
    return n->size;
  case OT_CONTIGUOUS:
    return n->count * ot_size(nodes, n->child);
  case OT_VECTOR:
  case OT_HVECTOR:
    return n->count * n->blocklength * ot_size(nodes, n->child);
  case OT_SUBARRAY: {
    int64_t s = ot_size(nodes, n->child);
    for (int d = 0; d < n->ndims; ++d) {
      s *= n->subsizes[d];
    }
    return s;
  }
  case OT_RESIZED:
    return ot_size(nodes, n->child);
  case OT_HINDEXED: {
    int64_t s = 0;
    for (int64_t k = 0; k < n->count; ++k) {
      s += n->blens[k] > 0 ? n->blens[k] : 0;
    }
    return s * ot_size(nodes, n->child);
  }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* typemap walk                                                              */
/* ------------------------------------------------------------------------ */

typedef struct walk {
  /* what to do with each block */
  int mode; /* 0 = collect merged blocks, 1 = pack, 2 = unpack */
  const uint8_t *in;
  uint8_t *out;
  int64_t cursor; /* packed-stream position */
  /* collection */
  int64_t *disp, *len, cap, n;
  int64_t pend_disp, pend_len;
} walk;

static void emit(walk *w, int64_t disp, int64_t len) {
  if (len <= 0) {
    return;
  }
  switch (w->mode) {
  case 0:
    if (w->pend_len > 0 && w->pend_disp + w->pend_len == disp) {
      w->pend_len += len;
    } else {
      if (w->pend_len > 0) {
        if (w->n < w->cap) {
          w->disp[w->n] = w->pend_disp;
          w->len[w->n] = w->pend_len;
        }
        ++w->n;
      }
      w->pend_disp = disp;
      w->pend_len = len;
    }
    break;
  case 1:
    memcpy(w->out + w->cursor, w->in + disp, (size_t)len);
    break;
  case 2:
    memcpy(w->out + disp, w->in + w->cursor, (size_t)len);
    break;
  }
  w->cursor += len;
}

/* Emit the typemap of node i placed at byte displacement `base`, in typemap
   order.  A NAMED child repeated at a stride equal to its size is emitted as
   one run (same bytes, same order). */
static void walk_node(walk *w, const ot_node *nodes, int32_t i, int64_t base) {
  const ot_node *n = &nodes[i];
  switch (n->kind) {
  case OT_NAMED:
    emit(w, base, n->size);
    return;
  case OT_CONTIGUOUS: {
    const ot_node *c = &nodes[n->child];
    int64_t cext = ot_extent(nodes, n->child);
    if (c->kind == OT_NAMED && cext == c->size) {
      emit(w, base, n->count * c->size);
      return;
    }
    for (int64_t k = 0; k < n->count; ++k) {
      walk_node(w, nodes, n->child, base + k * cext);
    }
    return;
  }
  case OT_VECTOR:
  case OT_HVECTOR: {
    const ot_node *c = &nodes[n->child];
    int64_t cext = ot_extent(nodes, n->child);
    int64_t s = n->kind == OT_VECTOR ? n->stride * cext : n->stride;
    int run = (c->kind == OT_NAMED && cext == c->size);
    for (int64_t k = 0; k < n->count; ++k) {
      if (run) {
        emit(w, base + k * s, n->blocklength * c->size);
      } else {
        for (int64_t j = 0; j < n->blocklength; ++j) {
          walk_node(w, nodes, n->child, base + k * s + j * cext);
        }
      }
    }
    return;
  }
  case OT_SUBARRAY: {
    int nd = n->ndims;
    int64_t cext = ot_extent(nodes, n->child);
    /* element stride of each dimension, C or Fortran order */
    int64_t es[OT_MAXD];
    for (int d = 0; d < nd; ++d) {
      es[d] = cext;
      if (n->order == 0) {
        for (int j = d + 1; j < nd; ++j) {
          es[d] *= n->sizes[j];
        }
      } else {
        for (int j = 0; j < d; ++j) {
          es[d] *= n->sizes[j];
        }
      }
    }
    for (int d = 0; d < nd; ++d) {
      if (n->subsizes[d] <= 0) {
        return;
      }
    }
    /* odometer over the subarray, fastest dimension last in C order and
       first in Fortran order */
    int64_t idx[OT_MAXD] = {0};
    for (;;) {
      int64_t disp = base;
      for (int d = 0; d < nd; ++d) {
        disp += (n->starts[d] + idx[d]) * es[d];
      }
      walk_node(w, nodes, n->child, disp);
      int d = n->order == 0 ? nd - 1 : 0;
      for (;;) {
        if (++idx[d] < n->subsizes[d]) {
          break;
        }
        idx[d] = 0;
        d += n->order == 0 ? -1 : 1;
        if (d < 0 || d >= nd) {
          return;
        }
      }
    }
  }
  case OT_RESIZED:
    walk_node(w, nodes, n->child, base);
    return;
  case OT_HINDEXED: {
    int64_t cext = ot_extent(nodes, n->child);
    for (int64_t k = 0; k < n->count; ++k) {
      for (int64_t j = 0; j < n->blens[k]; ++j) {
        walk_node(w, nodes, n->child, base + n->disps[k] + j * cext);
      }
    }
    return;
  }
  }
}

static void walk_objects(walk *w, const ot_node *nodes, int32_t root,
                         int64_t incount) {
  int64_t ext = ot_extent(nodes, root);
  for (int64_t o = 0; o < incount; ++o) {
    walk_node(w, nodes, root, o * ext);
  }
}

int64_t ot_blocks(const ot_node *nodes, int32_t root, int64_t incount,
                  int64_t *disp, int64_t *len, int64_t cap) {
  walk w;
  memset(&w, 0, sizeof(w));
  w.mode = 0;
  w.disp = disp;
  w.len = len;
  w.cap = cap;
  walk_objects(&w, nodes, root, incount);
  if (w.pend_len > 0) {
    if (w.n < w.cap) {
      w.disp[w.n] = w.pend_disp;
      w.len[w.n] = w.pend_len;
    }
    ++w.n;
  }
  return w.n;
}

int64_t ot_pack(const ot_node *nodes, int32_t root, const uint8_t *in,
                int64_t incount, uint8_t *out) {
  walk w;
  memset(&w, 0, sizeof(w));
  w.mode = 1;
  w.in = in;
  w.out = out;
  walk_objects(&w, nodes, root, incount);
  return w.cursor;
}

int64_t ot_unpack(const ot_node *nodes, int32_t root, const uint8_t *in,
                  int64_t outcount, uint8_t *out) {
  walk w;
  memset(&w, 0, sizeof(w));
  w.mode = 2;
  w.in = in;
  w.out = out;
  walk_objects(&w, nodes, root, outcount);
  return w.cursor;
}

/* ------------------------------------------------------------------------ */
/* StridedBlock gather / scatter                                             */
/* ------------------------------------------------------------------------ */

/* Restates include/pack_kernels.cuh:74-90 (pack_2d: dst[z*c1*c0 + y*c0 + x] =
   src[z*extent + y*stride1 + x]) and :372-391 (pack_3d, adds z*stride2) for
   any number of dimensions, with the offsets in 64 bits.  The start offset is
   applied to the strided side as the packers do (src/internal/
   packer_2d.cu:26, :53; packer_3d.cu:202, :220), and the packed side is dense
   in row-major order with the object index outermost. */
static int64_t sb_copy(const sb_desc *sb, const uint8_t *in, int64_t count,
                       uint8_t *out, int unpack) {
  int nd = sb->ndims;
  int64_t bl = sb->counts[0];
  int64_t rows = 1;
  for (int d = 1; d < nd; ++d) {
    rows *= sb->counts[d];
  }
  if (bl <= 0 || rows <= 0 || count <= 0) {
    return 0;
  }
  int64_t packed = 0;
  for (int64_t o = 0; o < count; ++o) {
    int64_t idx[SB_MAXD] = {0};
    for (int64_t r = 0; r < rows; ++r) {
      int64_t off = o * sb->extent + sb->start;
      for (int d = 1; d < nd; ++d) {
        off += idx[d] * sb->strides[d];
      }
      if (unpack) {
        memcpy(out + off, in + packed, (size_t)bl);
      } else {
        memcpy(out + packed, in + off, (size_t)bl);
      }
      packed += bl;
      for (int d = 1; d < nd; ++d) {
        if (++idx[d] < sb->counts[d]) {
          break;
        }
        idx[d] = 0;
      }
    }
  }
  return packed;
}

int64_t sb_pack(const sb_desc *sb, const uint8_t *in, int64_t incount,
                uint8_t *out) {
  return sb_copy(sb, in, incount, out, 0);
}

int64_t sb_unpack(const sb_desc *sb, const uint8_t *in, int64_t outcount,
                  uint8_t *out) {
  return sb_copy(sb, in, outcount, out, 1);
}

/* ------------------------------------------------------------------------ */
/* seeded bytes                                                              */
/* ------------------------------------------------------------------------ */

static uint64_t splitmix64_at(uint64_t seed, uint64_t k) {
  /* state after k+1 increments of the golden-ratio gamma, then the SplitMix64
     finaliser (Steele, Lea, Flood, OOPSLA 2014) */
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void oracle_fill_splitmix(uint8_t *buf, int64_t n, uint64_t seed) {
  int64_t words = n / 8;
  for (int64_t k = 0; k < words; ++k) {
    uint64_t v = splitmix64_at(seed, (uint64_t)k);
    memcpy(buf + 8 * k, &v, 8);
  }
  if (n % 8) {
    uint64_t v = splitmix64_at(seed, (uint64_t)words);
    memcpy(buf + 8 * words, &v, (size_t)(n % 8));
  }
}
