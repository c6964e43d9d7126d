// pool.cpp — memory pools of the C ABI (sdr_pool_*): device and pinned host blocks kept
// for reuse, so a serving slot runs request after request without hipMalloc / hipFree.
//
// A request is served from the smallest cached block of at least its size class (the
// next power of two, 4 KiB at least) and at most four times that, else by a fresh
// allocation of its class: requests of similar size (scenarios of neighbouring grid
// sizes) share blocks, at the price of up to 4x the bytes in a reused block.  An
// allocation that fails frees the pool's cached blocks and tries once more.
#include <hip/hip_runtime.h>

#include <memory>

#include "../../include/sdr_amd.h"
#include "host.hpp"

namespace sdr {

thread_local DevicePool* g_pool = nullptr;

namespace {
size_t size_class(size_t b) {
  size_t c = 4096;
  while (c < b) c <<= 1;
  return c;
}
}  // namespace

void* DevicePool::get(size_t bytes, bool host) {
  const size_t want = size_class(bytes);
  {
    std::lock_guard<std::mutex> g(mu);
    auto& fl = host ? free_host : free_dev;
    auto it = fl.lower_bound(want);
    if (it != fl.end() && it->first <= 4 * want) {
      void* p = it->second;
      cached -= it->first;
      live += it->first;
      fl.erase(it);
      ++hits;
      return p;
    }
    ++misses;
  }
  void* p = nullptr;
  auto alloc = [&]() -> hipError_t {
    return host ? hipHostMalloc(&p, want, hipHostMallocDefault) : hipMalloc(&p, want);
  };
  hipError_t e = alloc();
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim();
    e = alloc();
  }
  HIPCHK(e);
  std::lock_guard<std::mutex> g(mu);
  (host ? host_cap : dev_cap)[p] = want;
  live += want;
  return p;
}

void DevicePool::put(void* p, bool host) {
  std::lock_guard<std::mutex> g(mu);
  auto& caps = host ? host_cap : dev_cap;
  auto it = caps.find(p);
  if (it == caps.end()) {  // not this pool's block (cannot happen through Buf): free it
    if (host) (void)hipHostFree(p);
    else (void)hipFree(p);
    return;
  }
  (host ? free_host : free_dev).emplace(it->second, p);
  live -= it->second;
  cached += it->second;
}

void DevicePool::trim() {
  std::lock_guard<std::mutex> g(mu);
  for (auto& kv : free_dev) {
    (void)hipFree(kv.second);
    dev_cap.erase(kv.second);
  }
  for (auto& kv : free_host) {
    (void)hipHostFree(kv.second);
    host_cap.erase(kv.second);
  }
  free_dev.clear();
  free_host.clear();
  cached = 0;
}

DevicePool::~DevicePool() { trim(); }

}  // namespace sdr

struct sdr_pool {
  sdr::DevicePool pool;
};

using sdr::Error;
using sdr::guard;

extern "C" {

int sdr_pool_new(int32_t device, sdr_pool** out) {
  return guard([&] {
    if (!out) throw Error(SDR_E_INVALID, "null argument");
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) throw Error(SDR_E_NODEV, "no HIP device");
    if (device < 0 || device >= cnt) throw Error(SDR_E_INVALID, "no such device");
    auto p = std::make_unique<sdr_pool>();
    p->pool.device = device;
    *out = p.release();
  });
}

int sdr_pool_bind(sdr_pool* p) {
  return guard([&] {
    if (p) HIPCHK(hipSetDevice(p->pool.device));
    sdr::g_pool = p ? &p->pool : nullptr;
  });
}

int sdr_pool_get_stats(const sdr_pool* p, sdr_pool_stats* out) {
  return guard([&] {
    if (!p || !out) throw Error(SDR_E_INVALID, "null argument");
    auto& pl = const_cast<sdr_pool*>(p)->pool;
    std::lock_guard<std::mutex> g(pl.mu);
    out->hits = pl.hits;
    out->misses = pl.misses;
    out->cached_bytes = pl.cached;
    out->live_bytes = pl.live;
    out->cached_blocks = pl.free_dev.size() + pl.free_host.size();
    out->live_blocks = pl.dev_cap.size() + pl.host_cap.size() - out->cached_blocks;
  });
}

int sdr_pool_trim(sdr_pool* p) {
  return guard([&] {
    if (!p) throw Error(SDR_E_INVALID, "null argument");
    HIPCHK(hipSetDevice(p->pool.device));
    p->pool.trim();
  });
}

int sdr_pool_free(sdr_pool* p) {
  return guard([&] {
    if (!p) return;
    {
      std::lock_guard<std::mutex> g(p->pool.mu);
      if (p->pool.live != 0) throw Error(SDR_E_STATE, "the pool still has live blocks (free its solvers first)");
    }
    if (sdr::g_pool == &p->pool) sdr::g_pool = nullptr;
    (void)hipSetDevice(p->pool.device);
    delete p;
  });
}

}  // extern "C"
