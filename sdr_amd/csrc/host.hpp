// host.hpp — host-side plumbing shared by the library's C++ files: the error type every
// entry point reports through, HIP status checks, device and pinned host buffers, and
// the guard that turns an exception into an SDR_E_* code plus sdr_last_error() text.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>

#include "../../include/sdr_amd.h"

namespace sdr {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                            \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess)                                                                    \
      throw Error(e_ == hipErrorOutOfMemory ? SDR_E_OOM : SDR_E_HIP,                         \
                  std::string(#x) + ": " + hipGetErrorString(e_));                           \
  } while (0)

// ------------------------------------------------------------------ memory pools
// A pool keeps the device and pinned host blocks its buffers release and hands them to
// later allocations of a similar size, so a serving slot runs request after request
// without hipMalloc / hipFree (hipFree waits for the device).  A thread allocates from
// the pool bound to it (sdr_pool_bind); a released block returns to the pool it came
// from.  One solver at a time per pool: a block is reused only by work queued after the
// stream that used it last was synchronised or on that same stream.
struct DevicePool {
  int device = 0;
  std::mutex mu;
  std::multimap<size_t, void*> free_dev, free_host;  // by capacity
  std::unordered_map<void*, size_t> dev_cap, host_cap;  // every block, by address
  uint64_t hits = 0, misses = 0;
  size_t cached = 0, live = 0;
  void* get(size_t bytes, bool host);
  void put(void* p, bool host);
  void trim();  // free every cached block
  ~DevicePool();
};
extern thread_local DevicePool* g_pool;  // the calling thread's bound pool (null: none)

// ------------------------------------------------------------------ device buffer
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  bool owned = true;  // false: a view of memory the buffer does not own (never freed)
  DevicePool* pool = nullptr;  // where an owned block goes back to (null: hipFree)
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() { release(); }
  void release() {
    if (p && owned) {
      if (pool) pool->put(p, false);
      else (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
    owned = true;
    pool = nullptr;
  }
  void alloc(size_t b) {
    release();
    if (b == 0) b = 256;
    if (g_pool) {
      p = g_pool->get(b, false);
      pool = g_pool;
    } else {
      HIPCHK(hipMalloc(&p, b));
    }
    bytes = b;
  }
  void view(void* q, size_t b) {
    release();
    p = q;
    bytes = b;
    owned = false;
  }
  void ensure(size_t b) {
    if (b > bytes) alloc(b);
  }
  // exchange the blocks of two buffers (growth keeps the old one alive for a copy)
  void swap(Buf& o) {
    std::swap(p, o.p), std::swap(bytes, o.bytes), std::swap(owned, o.owned), std::swap(pool, o.pool);
  }
  template <class U>
  U* as() const {
    return (U*)p;
  }
};

struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevicePool* pool = nullptr;
  ~PinnedBuf() { release(); }
  void release() {
    if (p) {
      if (pool) pool->put(p, true);
      else (void)hipHostFree(p);
    }
    p = nullptr;
    bytes = 0;
    pool = nullptr;
  }
  void ensure(size_t b) {
    if (b <= bytes) return;
    release();
    if (g_pool) {
      p = g_pool->get(b, true);
      pool = g_pool;
    } else {
      HIPCHK(hipHostMalloc(&p, b, hipHostMallocDefault));
    }
    bytes = b;
  }
};

// ------------------------------------------------------------------ C ABI plumbing
extern thread_local std::string g_err;  // sdr_last_error()


template <class F>
int guard(F&& f) {
  try {
    f();
    g_err.clear();
    return SDR_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return SDR_E_OOM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return SDR_E_INVALID;
  }
}

}  // namespace sdr
