/*
 * oracle.h -- CPU oracle for the derived-datatype pack/unpack path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library (libtempi_amd.so)
 * links, loads or calls this code.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, and only as the checker.
 *
 * Two independent restatements live here:
 *
 *  1. A typemap walker (ot_*) over a tree of MPI type-constructor nodes.  It
 *     enumerates (displacement, length) blocks straight from the MPI-3.1
 *     constructor definitions (contiguous, vector, hvector, subarray, resized,
 *     hindexed), the semantics the reference decodes in
 *     src/internal/types.cpp:42-354 (Type::from_mpi_datatype).  Packing is the
 *     concatenation of those blocks for each of `incount` objects placed
 *     `extent` bytes apart, as MPI_Pack defines it.
 *
 *  2. A StridedBlock gather/scatter (sb_*) restating the reference's kernel
 *     index math, include/pack_kernels.cuh:64-120 (2D) and :350-433 (3D):
 *         dst[i*c_n..c_1*c0 + ... + y*c0 + x] = src[i*extent + start + ... + y*s1 + x]
 *     generalised to any number of dimensions and 64-bit offsets.
 *
 * Pinning (see tests/test_oracle.py): both restatements are checked against
 * the committed MPICH golden vectors (tests/golden/), against live MPICH
 * MPI_Pack where libmpi is present, against closed forms (packed size,
 * extent), against known answers probed from MPICH (vector(3,2,4,BYTE) ->
 * 0 1 4 5 8 9, ...) and against each other on seeded random types.
 */
#ifndef TEMPI_ORACLE_H
#define TEMPI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum ot_kind {
  OT_NAMED = 0,       /* predefined type of `size` contiguous bytes         */
  OT_CONTIGUOUS = 1,  /* count copies of child                               */
  OT_VECTOR = 2,      /* count blocks of blocklength children, stride in ext */
  OT_HVECTOR = 3,     /* same, stride in bytes                               */
  OT_SUBARRAY = 4,    /* ndims, sizes, subsizes, starts, order               */
  OT_RESIZED = 5,     /* child with new lb / extent                          */
  OT_HINDEXED = 6,    /* count blocks: blens[k] children at disps[k] bytes   */
};

#define OT_MAXD 8

typedef struct ot_node {
  int32_t kind;
  int32_t child; /* index of the old type in the node array, -1 for NAMED */
  int64_t count;
  int64_t blocklength;
  int64_t stride; /* VECTOR: in child extents; HVECTOR: bytes */
  int64_t size;   /* NAMED: bytes */
  int32_t ndims;
  int32_t order; /* SUBARRAY: 0 = C (last dim fastest), 1 = Fortran */
  int64_t sizes[OT_MAXD];
  int64_t subsizes[OT_MAXD];
  int64_t starts[OT_MAXD];
  int64_t lb; /* RESIZED */
  int64_t extent;
  const int64_t *blens; /* HINDEXED */
  const int64_t *disps; /* HINDEXED, bytes */
} ot_node;

/* lower bound / extent / size (bytes of data) of node `i` */
int64_t ot_lb(const ot_node *nodes, int32_t i);
int64_t ot_extent(const ot_node *nodes, int32_t i);
int64_t ot_size(const ot_node *nodes, int32_t i);

/* Write up to `cap` merged (disp, len) blocks of `incount` objects of node
   `root` into disp/len; returns the total number of merged blocks (which may
   exceed cap, in which case only the first cap are written). */
int64_t ot_blocks(const ot_node *nodes, int32_t root, int64_t incount,
                  int64_t *disp, int64_t *len, int64_t cap);

/* MPI_Pack / MPI_Unpack semantics over raw memory.  `in` (pack) and `out`
   (unpack) point at the origin of object 0; displacements may be negative.
   Returns bytes packed / consumed. */
int64_t ot_pack(const ot_node *nodes, int32_t root, const uint8_t *in,
                int64_t incount, uint8_t *out);
int64_t ot_unpack(const ot_node *nodes, int32_t root, const uint8_t *in,
                  int64_t outcount, uint8_t *out);

/* StridedBlock (include/strided_block.hpp:12-67): the first byte of object
   0 is at `start`; counts[0] contiguous bytes form a block; dimension d >= 1
   repeats the inner part counts[d] times at strides[d] bytes. */
#define SB_MAXD 8
typedef struct sb_desc {
  int64_t start;
  int64_t extent;
  int32_t ndims;
  int32_t pad_;
  int64_t counts[SB_MAXD];
  int64_t strides[SB_MAXD];
} sb_desc;

int64_t sb_pack(const sb_desc *sb, const uint8_t *in, int64_t incount,
                uint8_t *out);
int64_t sb_unpack(const sb_desc *sb, const uint8_t *in, int64_t outcount,
                  uint8_t *out);

/* Seeded byte stream: word k (little endian) of the stream is the SplitMix64
   output after k+1 steps from `seed` (Steele, Lea, Flood 2014). */
void oracle_fill_splitmix(uint8_t *buf, int64_t n, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
