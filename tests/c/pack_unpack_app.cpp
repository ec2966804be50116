// An unmodified-style MPI program, run with libtempi_amd.so preloaded:
// for each type it packs a host copy of seeded bytes (the interposer sends
// host buffers to the MPI library: the expected bytes) and the same bytes in
// device memory (the gfx950 kernels), compares, unpacks on the device into a
// 0xAB-filled buffer and compares that with the library's own unpack.
// Case list after the reference's test/pack_unpack.cpp:160-302.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <mpi.h>

#include "tempi_amd.h" // the counters' layout only: nothing is linked from it

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define CHECK_HIP(x)                                                           \
  do {                                                                         \
    if ((x) != hipSuccess) {                                                   \
      fprintf(stderr, "HIP error %s at %d\n", #x, __LINE__);                   \
      exit(2);                                                                 \
    }                                                                          \
  } while (0)

static int failures = 0;

static void fill(std::vector<unsigned char> &v, uint64_t seed) {
  uint64_t s = seed;
  for (auto &b : v) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    b = (unsigned char)(s >> 56);
  }
}

static void test_type(const char *name, MPI_Datatype t, int count) {
  MPI_Type_commit(&t);
  MPI_Aint lb, extent, tlb, text;
  MPI_Type_get_extent(t, &lb, &extent);
  MPI_Type_get_true_extent(t, &tlb, &text);
  int packed = 0;
  MPI_Pack_size(count, t, MPI_COMM_WORLD, &packed);
  const size_t span = (size_t)(tlb + text + (count - 1) * extent) + 64;
  std::vector<unsigned char> src(span), exp(packed), got(packed);
  fill(src, 7 + span);
  // expected: host buffers go to the MPI library
  int pe = 0;
  MPI_Pack(src.data(), count, t, exp.data(), packed, &pe, MPI_COMM_WORLD);
  // device
  unsigned char *dsrc, *dpk, *ddst;
  CHECK_HIP(hipMalloc(&dsrc, span));
  CHECK_HIP(hipMalloc(&dpk, packed + 1));
  CHECK_HIP(hipMalloc(&ddst, span));
  CHECK_HIP(hipMemcpy(dsrc, src.data(), span, hipMemcpyHostToDevice));
  int pg = 0;
  MPI_Pack(dsrc, count, t, dpk, packed, &pg, MPI_COMM_WORLD);
  CHECK_HIP(hipMemcpy(got.data(), dpk, packed, hipMemcpyDeviceToHost));
  bool ok = pg == pe && 0 == memcmp(got.data(), exp.data(), packed);
  // unpack into 0xAB on both sides
  std::vector<unsigned char> uexp(span, 0xAB), ugot(span);
  int p = 0;
  MPI_Unpack(exp.data(), packed, &p, uexp.data(), count, t, MPI_COMM_WORLD);
  CHECK_HIP(hipMemset(ddst, 0xAB, span));
  // the library works on its own non-blocking stream, as the reference's
  // kernStream: the caller finishes producing its buffers first
  CHECK_HIP(hipDeviceSynchronize());
  p = 0;
  MPI_Unpack(dpk, packed, &p, ddst, count, t, MPI_COMM_WORLD);
  CHECK_HIP(hipMemcpy(ugot.data(), ddst, span, hipMemcpyDeviceToHost));
  ok = ok && p == pe && 0 == memcmp(ugot.data(), uexp.data(), span);
  printf("%s %s count=%d packed=%d\n", ok ? "PASS" : "FAIL", name, count,
         packed);
  failures += ok ? 0 : 1;
  CHECK_HIP(hipFree(dsrc));
  CHECK_HIP(hipFree(dpk));
  CHECK_HIP(hipFree(ddst));
  MPI_Type_free(&t);
}

static MPI_Datatype vec(int n, int bl, int st) {
  MPI_Datatype t;
  MPI_Type_vector(n, bl, st, MPI_BYTE, &t);
  return t;
}

static MPI_Datatype sub2(int nb, int bl, int st) {
  MPI_Datatype t;
  int sizes[2] = {nb, st}, subs[2] = {nb, bl}, starts[2] = {0, 0};
  MPI_Type_create_subarray(2, sizes, subs, starts, MPI_ORDER_C, MPI_BYTE, &t);
  return t;
}

static MPI_Datatype off_sub(int px, int py, int pz, int ax, int ay, int az,
                            int ox, int oy, int oz) {
  MPI_Datatype t;
  int sizes[3] = {ax, ay, az}, subs[3] = {px, py, pz}, starts[3] = {ox, oy, oz};
  MPI_Type_create_subarray(3, sizes, subs, starts, MPI_ORDER_C, MPI_BYTE, &t);
  return t;
}

static MPI_Datatype v_hv(int cx, int cy, int cz, int ax, int ay) {
  MPI_Datatype plane, full;
  MPI_Type_vector(cy, cx, ax, MPI_BYTE, &plane);
  MPI_Type_create_hvector(cz, 1, (MPI_Aint)ax * ay, plane, &full);
  MPI_Type_free(&plane);
  return full;
}

int main(int argc, char **argv) {
  MPI_Init(&argc, &argv);
  MPI_Datatype c;
  MPI_Type_contiguous(10, MPI_BYTE, &c);
  test_type("contiguous_contiguous(10)", c, 30);
  for (int count : {1, 2}) {
    test_type("2d_byte_vector(2,3,4)", vec(2, 3, 4), count);
    test_type("2d_byte_subarray(2,3,4)", sub2(2, 3, 4), count);
    test_type("off_subarray", off_sub(2, 3, 4, 16, 16, 16, 1, 1, 4), count);
    test_type("byte_v_hv(2,3,4)/(10,10,10)", v_hv(2, 3, 4, 10, 10), count);
  }
  test_type("byte_v_hv(10,10,1)/(10,10,10)", v_hv(10, 10, 1, 10, 10), 1);
  test_type("byte_v_hv(4,3,4)/(200,200,200)", v_hv(4, 3, 4, 200, 200), 1);
  for (int count : {1, 3}) {
    test_type("byte_v_hv(100,100,100)/(200,200,200)",
              v_hv(100, 100, 100, 200, 200), count);
  }
  test_type("vector(1024,256,512)", vec(1024, 256, 512), 1);
  // the preloaded library's counters, if it is there: kernels launched
  typedef int (*counters_fn)(tempi_counters *);
  if (counters_fn f = (counters_fn)dlsym(RTLD_DEFAULT, "tempi_counters_get")) {
    tempi_counters c{};
    f(&c);
    printf("pack kernels %lld unpack kernels %lld library packs %lld\n",
           (long long)c.pack_kernel_launches, (long long)c.unpack_kernel_launches,
           (long long)c.library_pack_calls);
  }
  MPI_Finalize();
  printf("%d failures\n", failures);
  return failures ? 1 : 0;
}
