// Solver context: host mirror of every instance plus the device buffers (one HIP stream).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/xrtailor_amd.h"
#include "device.h"

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
  hipError_t Upload(const T* host, size_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    return hipMemcpyAsync(ptr, host, count * sizeof(T), hipMemcpyHostToDevice, s);
  }
  hipError_t Upload(const std::vector<T>& host, hipStream_t s) {
    hipError_t e = Alloc(host.size());
    if (e != hipSuccess) return e;
    return Upload(host.data(), host.size(), s);
  }
};

struct HostInstance {
  int nv = 0, nt = 0;
  std::vector<float> rest_pos;  // 3*nv
  std::vector<int32_t> tris;    // 3*nt
  std::vector<float> inv_mass;  // nv
  std::vector<int32_t> edges;   // 2*ne, original order
  std::vector<float> rest_len, s_comp;
  std::vector<int32_t> bends;  // 4*nb
  std::vector<float> theta, b_comp;
  std::vector<int32_t> slots;  // LRA slot vertices
  std::vector<std::vector<double>> slot_dist;  // geodesic distance per slot
  std::vector<float> x0_init, v_init;          // optional initial state (3*nv)
  // built by xrt_start
  std::vector<int32_t> s_color, b_color;
  int s_ncol = 0, b_ncol = 0;
  std::vector<int32_t> s_order, b_order;    // sorted position -> original index
  std::vector<int32_t> s_coloff, b_coloff;  // local color offsets (ncol + 1)
  std::vector<int32_t> lra_ptr, lra_slot;
  std::vector<float> lra_rest;
  InstDesc desc{};
  int sbeg = 0, bbeg = 0, kbeg = 0;
  int ne() const { return static_cast<int>(edges.size() / 2); }
  int nb() const { return static_cast<int>(bends.size() / 4); }
};

}  // namespace xrt

struct xrt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  xrt_params params{};
  std::vector<xrt::HostInstance> inst;
  bool started = false;
  // device properties
  int cu_count = 0;
  size_t max_lds = 0;
  int wave_size = 64;
  int64_t l2_bytes = 0;
  // device buffers
  xrt::DevBuf<float4> x0, v, x;
  xrt::DevBuf<xrt::StretchC> sc;
  xrt::DevBuf<float> slam;
  xrt::DevBuf<int4> bidx;
  xrt::DevBuf<float2> bpar;
  xrt::DevBuf<int32_t> s_coloff, b_coloff;
  xrt::DevBuf<int32_t> slot_vert, lra_ptr, lra_slot;
  xrt::DevBuf<float> lra_rest;
  xrt::DevBuf<float4> slot_pos;  // plain-path LRA snapshot
  xrt::DevBuf<xrt::InstDesc> desc;
  xrt::DevBuf<int32_t> order;
  // batched launch configuration
  int batched_block = 1024;
  size_t batched_lds = 0;
  int lds_verts = 0;
  int max_slots = 0;
  // graph of one batched frame
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  // timing events
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};
