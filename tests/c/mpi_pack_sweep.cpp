// The reference's own MPI benchmarks as a plain MPI program (it links only
// the MPI library and the HIP runtime; run it with libtempi_amd.so preloaded
// to measure this library, or without it to measure the MPI library alone):
//
//   mpi_pack_sweep pack <device|mapped|host> <iterations>
//     per-call MPI_Pack / MPI_Unpack wall time over the 2D sweep of
//     bin/bench_mpi_pack.cpp:127-143: packed sizes 1 KiB, 1 MiB, 4 MiB,
//     incount 1 and 2, blocks of 1-512 B at a 512-B stride, each as
//     MPI_Type_vector, MPI_Type_create_hvector and MPI_Type_create_subarray
//     of MPI_BYTE (support/type.cpp:216-242).  "device": both buffers from
//     hipMalloc; "mapped": the packed buffer in pinned mapped host memory
//     (the reference's stage = true, bin/bench_mpi_pack.cpp:49-53); "host":
//     both buffers from malloc.
//   mpi_pack_sweep commit <iterations>
//     MPI_Type_commit wall time of the 3D types of bin/bench_type_commit.cpp:
//     28-53 (five layouts of support/type.cpp over 24 extents in a 1024^3
//     byte allocation).
//
// One JSON line per case; times are the trimean (Q1 + 2 median + Q3) / 4 of
// the iterations, the reference's statistic (src/internal/statistics.cpp:
// 30-39), in seconds, from std::chrono::steady_clock (MPI_Wtime, which the
// reference reads, has microsecond ticks in some MPI builds: too coarse for
// the smallest host packs).
#include <hip/hip_runtime.h>
#include <mpi.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#define CHECK_HIP(x)                                                           \
  do {                                                                         \
    if ((x) != hipSuccess) {                                                   \
      fprintf(stderr, "HIP error %s at line %d\n", #x, __LINE__);              \
      exit(2);                                                                 \
    }                                                                          \
  } while (0)

static double now() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

static double trimean(std::vector<double> v) {
  std::sort(v.begin(), v.end());
  auto q = [&](double f) {
    const double x = f * (double)(v.size() - 1);
    const size_t i = (size_t)x;
    const double r = x - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - r) + v[i + 1] * r : v[i];
  };
  return (q(0.25) + 2 * q(0.5) + q(0.75)) / 4;
}

// ---- 2D byte layouts: numBlocks blocks of blockLength bytes, stride apart
static MPI_Datatype make_2d(int kind, int nb, int bl, int stride) {
  MPI_Datatype t;
  if (kind == 0) {
    MPI_Type_vector(nb, bl, stride, MPI_BYTE, &t);
  } else if (kind == 1) {
    MPI_Type_create_hvector(nb, bl, stride, MPI_BYTE, &t);
  } else {
    int sizes[2] = {nb, stride}, subsizes[2] = {nb, bl}, starts[2] = {0, 0};
    MPI_Type_create_subarray(2, sizes, subsizes, starts, MPI_ORDER_C, MPI_BYTE,
                             &t);
  }
  return t;
}
static const char *k2d[] = {"2d_byte_vector", "2d_byte_hvector",
                            "2d_byte_subarray"};

static int run_pack(const std::string &mode, int iters) {
  const int targets[] = {1024, 1024 * 1024, 4 * 1024 * 1024};
  const int blocks[] = {1, 2, 4, 8, 32, 64, 128, 256, 512};
  const int stride = 512;
  for (int target : targets) {
    for (int count = 1; count <= 2; ++count) {
      for (int bl : blocks) {
        const int nb = target / bl;
        for (int kind = 0; kind < 3; ++kind) {
          MPI_Datatype t = make_2d(kind, nb, bl, stride);
          MPI_Type_commit(&t);
          MPI_Aint lb, extent;
          MPI_Type_get_extent(t, &lb, &extent);
          int size = 0, packed = 0;
          MPI_Type_size(t, &size);
          MPI_Pack_size(count, t, MPI_COMM_WORLD, &packed);
          const size_t span = (size_t)extent * count;
          char *src = nullptr, *dst = nullptr;
          if (mode == "host") {
            // not filled: a pack's cost does not depend on the values, and
            // the untimed first calls fault the pages in
            src = (char *)malloc(span);
            dst = (char *)malloc(packed);
          } else {
            CHECK_HIP(hipMalloc(&src, span));
            CHECK_HIP(hipMemset(src, 1, span));
            if (mode == "mapped") {
              CHECK_HIP(hipHostMalloc(&dst, packed, hipHostMallocMapped));
            } else {
              CHECK_HIP(hipMalloc(&dst, packed));
            }
            CHECK_HIP(hipDeviceSynchronize());
          }
          // untimed first calls: two through this library (the first call of
          // a shape may settle its kernel variant), one for the MPI library
          const int untimed = mode == "host" ? 1 : 2;
          std::vector<double> tp, tu;
          for (int it = -untimed; it < iters; ++it) {
            int pos = 0;
            double t0 = now();
            MPI_Pack(src, count, t, dst, packed, &pos, MPI_COMM_WORLD);
            double t1 = now();
            pos = 0;
            MPI_Unpack(dst, packed, &pos, src, count, t, MPI_COMM_WORLD);
            double t2 = now();
            if (it >= 0) {
              tp.push_back(t1 - t0);
              tu.push_back(t2 - t1);
            }
          }
          printf("{\"layout\": \"%s\", \"count\": %d, \"target\": %d, "
                 "\"block\": %d, \"stride\": %d, \"mode\": \"%s\", "
                 "\"bytes\": %lld, \"iterations\": %d, \"pack_s\": %.9g, "
                 "\"unpack_s\": %.9g}\n",
                 k2d[kind], count, target, bl, stride, mode.c_str(),
                 (long long)size * count, iters, trimean(tp), trimean(tu));
          fflush(stdout);
          if (mode == "host") {
            free(src);
            free(dst);
          } else {
            CHECK_HIP(hipFree(src));
            if (mode == "mapped") {
              CHECK_HIP(hipHostFree(dst));
            } else {
              CHECK_HIP(hipFree(dst));
            }
          }
          MPI_Type_free(&t);
        }
      }
    }
  }
  return 0;
}

// ---- 3D byte layouts of a copy extent (x bytes, y rows, z planes) in an
// allocation of ax * ay * az bytes
struct E3 {
  int x, y, z;
};
static MPI_Datatype make_3d(int kind, E3 c, E3 a) {
  MPI_Datatype row, plane, cube;
  switch (kind) {
  case 0: { // one 3D subarray
    int sizes[3] = {a.z, a.y, a.x}, sub[3] = {c.z, c.y, c.x}, st[3] = {0, 0, 0};
    MPI_Type_create_subarray(3, sizes, sub, st, MPI_ORDER_C, MPI_BYTE, &cube);
    return cube;
  }
  case 1: { // 2D subarray planes, repeated by a vector
    int sizes[2] = {a.y, a.x}, sub[2] = {c.y, c.x}, st[2] = {0, 0};
    MPI_Type_create_subarray(2, sizes, sub, st, MPI_ORDER_C, MPI_BYTE, &plane);
    MPI_Type_vector(c.z, 1, 1, plane, &cube);
    MPI_Type_free(&plane);
    return cube;
  }
  case 2: // a one-block vector row, hvector planes, hvector cube
    MPI_Type_vector(1, c.x, a.x, MPI_BYTE, &row);
    break;
  case 3: // a row of single bytes
    MPI_Type_vector(c.x, 1, 1, MPI_BYTE, &row);
    break;
  default: // a vector plane, hvector cube
    MPI_Type_vector(c.y, c.x, a.x, MPI_BYTE, &plane);
    MPI_Type_create_hvector(c.z, 1, (MPI_Aint)a.x * a.y, plane, &cube);
    MPI_Type_free(&plane);
    return cube;
  }
  MPI_Type_create_hvector(c.y, 1, a.x, row, &plane);
  MPI_Type_create_hvector(c.z, 1, (MPI_Aint)a.x * a.y, plane, &cube);
  MPI_Type_free(&row);
  MPI_Type_free(&plane);
  return cube;
}
static const char *k3d[] = {"subarray", "subarray_v", "byte_v1_hv_hv",
                            "byte_vn_hv_hv", "byte_v_hv"};

static int run_commit(int iters) {
  const E3 alloc{1024, 1024, 1024};
  const E3 dims[] = {{1, 1024, 1024},  {2, 1024, 512},  {4, 1024, 256},
                     {8, 1024, 128},   {16, 1024, 64},  {32, 1024, 32},
                     {64, 1024, 16},   {128, 1024, 8},  {256, 1024, 4},
                     {512, 1024, 2},   {1024, 1024, 1}, {1, 1024, 1},
                     {2, 1024, 1},     {4, 1024, 1},    {8, 1024, 1},
                     {16, 1024, 1},    {32, 1024, 1},   {64, 1024, 1},
                     {128, 1024, 1},   {256, 1024, 1},  {512, 1024, 1},
                     {12, 512, 512},   {512, 3, 512},   {512, 512, 3}};
  for (const E3 &d : dims) {
    for (int kind = 0; kind < 5; ++kind) {
      std::vector<double> tc, tm;
      for (int it = -2; it < iters; ++it) {
        double t0 = now();
        MPI_Datatype t = make_3d(kind, d, alloc);
        double t1 = now();
        MPI_Type_commit(&t);
        double t2 = now();
        MPI_Type_free(&t);
        if (it >= 0) {
          tm.push_back(t1 - t0);
          tc.push_back(t2 - t1);
        }
      }
      printf("{\"layout\": \"%s\", \"x\": %d, \"y\": %d, \"z\": %d, "
             "\"iterations\": %d, \"create_s\": %.9g, \"commit_s\": %.9g}\n",
             k3d[kind], d.x, d.y, d.z, iters, trimean(tm), trimean(tc));
      fflush(stdout);
    }
  }
  return 0;
}

int main(int argc, char **argv) {
  MPI_Init(&argc, &argv);
  int rc = 1;
  if (argc >= 4 && std::string(argv[1]) == "pack") {
    rc = run_pack(argv[2], atoi(argv[3]));
  } else if (argc >= 3 && std::string(argv[1]) == "commit") {
    rc = run_commit(atoi(argv[2]));
  } else {
    fprintf(stderr, "usage: %s pack <device|mapped|host> <iters> | "
                    "commit <iters>\n", argv[0]);
  }
  MPI_Finalize();
  return rc;
}
