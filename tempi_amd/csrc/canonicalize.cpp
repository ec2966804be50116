// canonicalize.cpp -- commit-time lowering of an MPI derived datatype to a
// strided block (tempi_strided_block) that keeps MPI type-map order.
//
// The reference decodes the type into a tree of Dense/Stream nodes
// (src/internal/types.cpp:42-354), rewrites the tree to a fixpoint
// (:567-614) and reads a StridedBlock off the single-child chain
// (:654-715).  This implementation lowers each constructor directly into a
// list of dimensions over a byte run, so that every step is a statement
// about the type map itself:
//
//   * a constructor appends its repetition dimensions outside the child's
//     dimensions and adds its displacement to `start`;
//   * dimensions are never reordered -- the reference's stream_swap
//     (types.cpp:378-404) changes the packed order and is not reproduced;
//   * a count-1 dimension is dropped but its displacement stays in `start`
//     (the reference's stream_elision2, types.cpp:507-513, loses it);
//   * a dimension is folded into its inner neighbour only when it continues
//     exactly where that neighbour ends (the legal case of the reference's
//     stream_flatten, types.cpp:529-563, and stream_dense_fold, :409-449).
//
// Beyond the reference's contiguous/vector/hvector/subarray/named set, the
// lowering accepts dup, resized, single-block struct and the indexed family
// when its displacements factor into nested arithmetic progressions (which
// covers the reference's make_hi / make_hib row layouts,
// support/type.cpp:113-156).  Anything else is reported unsupported, and the
// MPI calls for that type go to the underlying library.
#include "internal.hpp"

#include <algorithm>

namespace tempi {

namespace {

struct Layout {
  int64_t start = 0;
  std::vector<Dim> dims; // dims[0] = {bytes, 1}
};

struct Decoded {
  int combiner = 0;
  std::vector<int> ints;
  std::vector<MPI_Aint> addrs;
  std::vector<MPI_Datatype> types;
};

bool is_named(MPI_Datatype t) {
  int ni, na, nd, comb;
  if (MPI_SUCCESS != MPI_Type_get_envelope(t, &ni, &na, &nd, &comb)) {
    return true; // do not try to free something we cannot inspect
  }
  return comb == MPI_COMBINER_NAMED;
}

// owns the datatype handles MPI_Type_get_contents hands out
struct Contents {
  Decoded d;
  ~Contents() {
    for (MPI_Datatype &t : d.types) {
      if (!is_named(t)) {
        real_mpi().Type_free(&t);
      }
    }
  }
};

bool decode(MPI_Datatype t, Contents &c) {
  int ni = 0, na = 0, nd = 0, comb = 0;
  if (MPI_SUCCESS != MPI_Type_get_envelope(t, &ni, &na, &nd, &comb)) {
    return false;
  }
  c.d.combiner = comb;
  if (comb == MPI_COMBINER_NAMED) {
    return true;
  }
  c.d.ints.resize(ni);
  c.d.addrs.resize(na);
  c.d.types.resize(nd);
  if (MPI_SUCCESS != MPI_Type_get_contents(t, ni, na, nd, c.d.ints.data(),
                                           c.d.addrs.data(),
                                           c.d.types.data())) {
    c.d.types.clear();
    return false;
  }
  return true;
}

bool extent_of(MPI_Datatype t, int64_t *lb, int64_t *ext) {
  MPI_Count l = 0, e = 0;
  if (MPI_SUCCESS != MPI_Type_get_extent_x(t, &l, &e)) {
    return false;
  }
  *lb = (int64_t)l;
  *ext = (int64_t)e;
  return true;
}

bool size_of(MPI_Datatype t, int64_t *size) {
  MPI_Count s = 0;
  if (MPI_SUCCESS != MPI_Type_size_x(t, &s)) {
    return false;
  }
  *size = (int64_t)s;
  return true;
}

// Factor a displacement sequence into nested arithmetic progressions:
// disp[k] = disp[0] + sum_j idx_j(k) * step_j with idx_0 fastest.  Appends
// the progressions (innermost first) to `out`.
bool factor_progression(const std::vector<int64_t> &disp,
                        std::vector<Dim> &out) {
  std::vector<int64_t> seq(disp.size());
  for (size_t i = 0; i < disp.size(); ++i) {
    seq[i] = disp[i] - disp[0];
  }
  while (seq.size() > 1) {
    const int64_t step = seq[1] - seq[0];
    size_t run = 1;
    while (run < seq.size() && seq[run] == seq[0] + (int64_t)run * step) {
      ++run;
    }
    if (seq.size() % run) {
      return false;
    }
    std::vector<int64_t> next;
    for (size_t g = 0; g < seq.size(); g += run) {
      for (size_t i = 0; i < run; ++i) {
        if (seq[g + i] != seq[g] + (int64_t)i * step) {
          return false;
        }
      }
      next.push_back(seq[g]);
    }
    out.push_back(Dim{(int64_t)run, step});
    if (out.size() > 64) {
      return false;
    }
    seq.swap(next);
  }
  return true;
}

bool lower(MPI_Datatype t, Layout &L, std::string &why, int depth);

// the common shape of vector/hvector/contiguous/indexed-block: `bl` copies
// of `child` at its extent, repeated along `outer` (innermost first)
bool lower_repeat(MPI_Datatype child, int64_t bl,
                  const std::vector<Dim> &outer, int64_t disp0, Layout &L,
                  std::string &why, int depth) {
  if (!lower(child, L, why, depth + 1)) {
    return false;
  }
  int64_t clb, cext;
  if (!extent_of(child, &clb, &cext)) {
    why = "extent query failed";
    return false;
  }
  L.dims.push_back(Dim{bl, cext});
  for (const Dim &d : outer) {
    L.dims.push_back(d);
  }
  L.start += disp0;
  return true;
}

bool lower(MPI_Datatype t, Layout &L, std::string &why, int depth) {
  if (depth > 32) {
    why = "type nesting too deep";
    return false;
  }
  Contents c;
  if (!decode(t, c)) {
    why = "envelope/contents query failed";
    return false;
  }
  const Decoded &d = c.d;
  switch (d.combiner) {
  case MPI_COMBINER_NAMED: {
    // reference: types.cpp:168-181 (named -> dense of its extent).  A named
    // type with a hole (MPI_DOUBLE_INT, ...) is not one byte run.
    int64_t lb, ext, size;
    if (!extent_of(t, &lb, &ext) || !size_of(t, &size)) {
      why = "extent query failed";
      return false;
    }
    if (lb != 0 || size != ext || size <= 0) {
      why = "named type is not a dense byte run";
      return false;
    }
    L.start = 0;
    L.dims.assign(1, Dim{size, 1});
    return true;
  }
  case MPI_COMBINER_DUP:
    return lower(d.types[0], L, why, depth + 1);
  case MPI_COMBINER_RESIZED:
    // same type map, only lb/extent change (read back by the caller)
    return lower(d.types[0], L, why, depth + 1);
  case MPI_COMBINER_CONTIGUOUS: {
    // reference: types.cpp:195-229; contiguous(n) = n copies at the extent
    return lower_repeat(d.types[0], d.ints[0], {}, 0, L, why, depth);
  }
  case MPI_COMBINER_VECTOR: {
    // reference: types.cpp:56-111 (stride in child extents)
    int64_t clb, cext;
    if (!extent_of(d.types[0], &clb, &cext)) {
      why = "extent query failed";
      return false;
    }
    return lower_repeat(d.types[0], d.ints[1],
                        {Dim{d.ints[0], (int64_t)d.ints[2] * cext}}, 0, L,
                        why, depth);
  }
  case MPI_COMBINER_HVECTOR: {
    // reference: types.cpp:113-167; the byte stride stays 64-bit here (the
    // reference narrows it to int at :133)
    return lower_repeat(d.types[0], d.ints[1],
                        {Dim{d.ints[0], (int64_t)d.addrs[0]}}, 0, L, why,
                        depth);
  }
  case MPI_COMBINER_HVECTOR_INTEGER: {
    return lower_repeat(d.types[0], d.ints[1],
                        {Dim{d.ints[0], (int64_t)d.ints[2]}}, 0, L, why,
                        depth);
  }
  case MPI_COMBINER_SUBARRAY: {
    // reference: types.cpp:234-319.  ints = ndims, sizes[], subsizes[],
    // starts[], order.  Element stride of dim i is the product of the sizes
    // of the faster dims; the start offsets stay in `start` even for
    // subsize-1 dimensions.
    const int nd = d.ints[0];
    const int order = d.ints[1 + 3 * nd];
    if (!lower(d.types[0], L, why, depth + 1)) {
      return false;
    }
    int64_t clb, cext;
    if (!extent_of(d.types[0], &clb, &cext)) {
      why = "extent query failed";
      return false;
    }
    std::vector<Dim> fastest_first;
    int64_t start = 0;
    for (int k = 0; k < nd; ++k) {
      // k-th fastest dimension
      const int i = (order == MPI_ORDER_C) ? nd - 1 - k : k;
      int64_t es = cext;
      if (order == MPI_ORDER_C) {
        for (int j = i + 1; j < nd; ++j) {
          es *= d.ints[1 + j];
        }
      } else {
        for (int j = 0; j < i; ++j) {
          es *= d.ints[1 + j];
        }
      }
      const int64_t subsize = d.ints[1 + nd + i];
      const int64_t st = d.ints[1 + 2 * nd + i];
      start += st * es;
      fastest_first.push_back(Dim{subsize, es});
    }
    for (const Dim &x : fastest_first) {
      L.dims.push_back(x);
    }
    L.start += start;
    return true;
  }
  case MPI_COMBINER_STRUCT:
  case MPI_COMBINER_STRUCT_INTEGER: {
    // only the one-block struct (possibly with zero-length companions) is a
    // strided block: bl copies of the type at disp
    const int n = d.ints[0];
    int which = -1;
    for (int k = 0; k < n; ++k) {
      if (d.ints[1 + k] > 0) {
        if (which >= 0) {
          why = "struct with several blocks";
          return false;
        }
        which = k;
      }
    }
    if (which < 0) {
      why = "empty struct";
      return false;
    }
    int64_t disp = d.combiner == MPI_COMBINER_STRUCT
                       ? (int64_t)d.addrs[which]
                       : (int64_t)d.ints[1 + n + which];
    return lower_repeat(d.types[which], d.ints[1 + which], {}, disp, L, why,
                        depth);
  }
  case MPI_COMBINER_INDEXED_BLOCK:
  case MPI_COMBINER_HINDEXED_BLOCK:
  case MPI_COMBINER_INDEXED:
  case MPI_COMBINER_HINDEXED:
  case MPI_COMBINER_HINDEXED_INTEGER: {
    const int n = d.ints[0];
    int64_t clb, cext;
    if (!extent_of(d.types[0], &clb, &cext)) {
      why = "extent query failed";
      return false;
    }
    std::vector<int64_t> bl(n), disp(n);
    for (int k = 0; k < n; ++k) {
      switch (d.combiner) {
      case MPI_COMBINER_INDEXED_BLOCK: // count, bl, disps[] (elements)
        bl[k] = d.ints[1];
        disp[k] = (int64_t)d.ints[2 + k] * cext;
        break;
      case MPI_COMBINER_HINDEXED_BLOCK: // count, bl; addrs = disps
        bl[k] = d.ints[1];
        disp[k] = (int64_t)d.addrs[k];
        break;
      case MPI_COMBINER_INDEXED: // count, bls[], disps[] (elements)
        bl[k] = d.ints[1 + k];
        disp[k] = (int64_t)d.ints[1 + n + k] * cext;
        break;
      case MPI_COMBINER_HINDEXED: // count, bls[]; addrs = disps
        bl[k] = d.ints[1 + k];
        disp[k] = (int64_t)d.addrs[k];
        break;
      default: // HINDEXED_INTEGER: count, bls[], disps[] (bytes)
        bl[k] = d.ints[1 + k];
        disp[k] = (int64_t)d.ints[1 + n + k];
        break;
      }
    }
    // zero-length blocks contribute nothing to the type map
    std::vector<int64_t> kept;
    int64_t common = -1;
    for (int k = 0; k < n; ++k) {
      if (bl[k] == 0) {
        continue;
      }
      if (common >= 0 && bl[k] != common) {
        why = "indexed blocks of different lengths";
        return false;
      }
      common = bl[k];
      kept.push_back(disp[k]);
    }
    if (kept.empty()) {
      why = "indexed type with no data";
      return false;
    }
    std::vector<Dim> outer;
    if (!factor_progression(kept, outer)) {
      why = "indexed displacements are not a nested progression";
      return false;
    }
    return lower_repeat(d.types[0], common, outer, kept[0], L, why, depth);
  }
  default:
    why = "unsupported combiner";
    return false;
  }
}

} // namespace

void normalize_dims(std::vector<Dim> &dims) {
  bool changed = true;
  while (changed) {
    changed = false;
    // a count-1 dimension repeats nothing; its offset is already in start
    for (size_t i = 1; i < dims.size(); ++i) {
      if (dims[i].count == 1) {
        dims.erase(dims.begin() + i);
        changed = true;
        break;
      }
    }
    if (changed) {
      continue;
    }
    // dims[i] continues exactly where dims[i-1] ends: one longer dimension.
    // For the byte run (stride 1) this is stride == bytes.
    for (size_t i = 1; i < dims.size(); ++i) {
      int64_t end = 0;
      if (!__builtin_mul_overflow(dims[i - 1].count, dims[i - 1].stride, &end) &&
          dims[i].stride == end) {
        dims[i - 1].count *= dims[i].count;
        dims.erase(dims.begin() + i);
        changed = true;
        break;
      }
    }
  }
}

Canon canonicalize(MPI_Datatype dt) {
  Canon c;
  Layout L;
  std::string why;
  if (!lower(dt, L, why, 0)) {
    c.reason = why;
    return c;
  }
  if (!extent_of(dt, &c.lb, &c.extent) || !size_of(dt, &c.size)) {
    c.reason = "extent query failed";
    return c;
  }
  for (const Dim &x : L.dims) {
    if (x.count <= 0) {
      c.reason = "zero-sized type";
      return c;
    }
  }
  normalize_dims(L.dims);
  if (L.dims.size() > (size_t)TEMPI_MAX_DIMS) {
    c.reason = "too many dimensions";
    return c;
  }
  // the lowering must account for every byte MPI says the type holds
  int64_t total = 1;
  for (const Dim &x : L.dims) {
    total *= x.count;
  }
  if (total != c.size) {
    c.reason = "lowered size disagrees with MPI_Type_size";
    return c;
  }
  c.sb.start = L.start;
  c.sb.extent = c.extent;
  c.sb.ndims = (int32_t)L.dims.size();
  for (size_t i = 0; i < L.dims.size(); ++i) {
    c.sb.counts[i] = L.dims[i].count;
    c.sb.strides[i] = L.dims[i].stride;
  }
  c.ok = true;
  return c;
}

} // namespace tempi
