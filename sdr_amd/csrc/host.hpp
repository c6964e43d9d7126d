// host.hpp — host-side plumbing shared by the library's C++ files: the error type every
// entry point reports through, HIP status checks, device and pinned host buffers, and
// the guard that turns an exception into an SDR_E_* code plus sdr_last_error() text.
#pragma once

#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>

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

// ------------------------------------------------------------------ device buffer
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  bool owned = true;  // false: a view of memory the buffer does not own (never freed)
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() { release(); }
  void release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    owned = true;
  }
  void alloc(size_t b) {
    release();
    if (b == 0) b = 256;
    HIPCHK(hipMalloc(&p, b));
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
  template <class U>
  U* as() const {
    return (U*)p;
  }
};

struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  void ensure(size_t b) {
    if (b <= bytes) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    HIPCHK(hipHostMalloc(&p, b, hipHostMallocDefault));
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
