// pool.cpp -- cached staging buffers for point-to-point transfers.
//
// The reference keeps freed device and pinned-host buffers in power-of-two
// slabs so that repeated sends do not pay for cudaMalloc / cudaHostRegister
// each time (include/allocator_slab.hpp:17-198, allocator_host.hpp:31-60).
// Here: one free list per size class (powers of two from 4 KiB) for device
// memory (hipMalloc) and one for pinned host memory mapped into the device
// address space (hipHostMalloc, Mapped | Portable: the kernels store packed
// bytes into it directly).  Buffers return to their list on release (up to
// 2 GiB of free buffers per kind; past that a released buffer is freed) and
// the lists are freed at MPI_Finalize.
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <map>
#include <unordered_map>

namespace tempi {

namespace {

struct Pool {
  std::mutex mu;
  std::multimap<size_t, void *> free_list; // size class -> buffer
  std::unordered_map<void *, size_t> size_of;
  size_t free_bytes = 0; // held in free_list
};

// free buffers kept per kind; beyond this a released buffer is freed
constexpr size_t kKeepBytes = 2ull << 30;

Pool &pool(bool host) {
  static Pool device_pool, host_pool;
  return host ? host_pool : device_pool;
}

size_t size_class(size_t bytes) {
  size_t c = 4096;
  while (c < bytes) {
    c <<= 1;
  }
  return c;
}

} // namespace

void *pool_acquire(bool host, size_t bytes) {
  Pool &p = pool(host);
  const size_t c = size_class(bytes);
  {
    std::lock_guard<std::mutex> lk(p.mu);
    auto it = p.free_list.find(c);
    if (it != p.free_list.end()) {
      void *b = it->second;
      p.free_list.erase(it);
      p.free_bytes -= c;
      return b;
    }
  }
  void *b = nullptr;
  const hipError_t e =
      host ? hipHostMalloc(&b, c, hipHostMallocMapped | hipHostMallocPortable)
           : hipMalloc(&b, c);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error(TEMPI_ERR_NOMEM, "staging buffer of %zu B: %s", c,
              hipGetErrorString(e));
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(p.mu);
  p.size_of[b] = c;
  return b;
}

void pool_release(bool host, void *b) {
  if (!b) {
    return;
  }
  Pool &p = pool(host);
  std::lock_guard<std::mutex> lk(p.mu);
  auto it = p.size_of.find(b);
  if (it == p.size_of.end()) {
    return;
  }
  if (p.free_bytes + it->second > kKeepBytes) {
    (void)(host ? hipHostFree(b) : hipFree(b));
    p.size_of.erase(it);
    return;
  }
  p.free_bytes += it->second;
  p.free_list.emplace(it->second, b);
}

void pool_trim() {
  for (bool host : {false, true}) {
    Pool &p = pool(host);
    std::lock_guard<std::mutex> lk(p.mu);
    for (auto &kv : p.free_list) {
      (void)(host ? hipHostFree(kv.second) : hipFree(kv.second));
      p.size_of.erase(kv.second);
    }
    p.free_list.clear();
    p.free_bytes = 0;
  }
}

} // namespace tempi
