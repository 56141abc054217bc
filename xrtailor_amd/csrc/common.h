// Shared host helpers: error reporting macros and an owning device buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/xrtailor_amd.h"

namespace xrt {

void SetError(const std::string& msg);

#define XRT_CHECK_HIP(expr)                                                       \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      ::xrt::SetError(std::string(#expr) + ": " + hipGetErrorString(_e));        \
      return XRT_ERR_HIP;                                                         \
    }                                                                             \
  } while (0)

#define XRT_REQUIRE(cond, code, msg) \
  do {                               \
    if (!(cond)) {                   \
      ::xrt::SetError(msg);          \
      return (code);                 \
    }                                \
  } while (0)

template <typename T>
struct DevBuf {
  T* ptr = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { Free(); }
  void Free() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  hipError_t Alloc(size_t count) {
    Free();
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void**>(&ptr), count * sizeof(T));
  }
  // The copy has completed when this returns: callers upload from temporaries (set-up
  // vectors that die at the end of their scope), and HIP does not promise that an
  // asynchronous copy from pageable memory has read its source when the call returns
  // (the runtime may queue it to a worker thread). Never called under stream capture.
  hipError_t Upload(const T* host, size_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(ptr, host, count * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
  }
  hipError_t Upload(const std::vector<T>& host, hipStream_t s) {
    hipError_t e = Alloc(host.size());
    if (e != hipSuccess) return e;
    return Upload(host.data(), host.size(), s);
  }
};

}  // namespace xrt
