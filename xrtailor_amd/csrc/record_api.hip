// Per-frame cache of batched frames (ClothSolverUniversal::SaveCache,
// xrtailor/src/pipeline/impl/universal/ClothSolver.cpp:107-144, with the per-frame normal
// update of SimulateSwift, ClothSolverBase.cpp:242-244).
//
// The reference copies each frame's positions and normals to the host synchronously after
// the frame.  Here each frame's positions (and the vertex normals of every cloth, the
// reference's UpdateNormals formula over one vertex -> corner table of all instances) are
// packed to float3 in one of two device slots; a second stream (default priority) copies the
// slot into pinned host memory while the next frame computes, and a helper thread moves it
// into the caller's array while the calling thread keeps launching frames (a large frame
// graph must keep being enqueued, or the device runs dry).  Events order the device slot
// reuse (repacked only after its copy ended); a pinned slot is refilled only after its
// helper thread has joined.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "context.h"

namespace xrt {
namespace {
constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void PackFloat3Kernel(const float4* __restrict__ in, int n,
                                                           float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = in[i];
  out[3 * i] = p.x;
  out[3 * i + 1] = p.y;
  out[3 * i + 2] = p.z;
}

hipError_t Pack(hipStream_t s, const float4* in, int n, float* out) {
  if (n <= 0) return hipSuccess;
  PackFloat3Kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(in, n, out);
  return hipGetLastError();
}

// the vertex -> corner table of every instance as one mesh (global vertex ids)
int SetupNormals(xrt_ctx* ctx) {
  RecordData& R = ctx->record;
  if (R.vptr.ptr) return XRT_OK;
  const int N = ctx->wide.n_total;
  std::vector<int4> tris;
  for (const auto& h : ctx->inst)
    for (int f = 0; f < h.nt; ++f)
      tris.push_back(make_int4(h.desc.vbeg + h.tris[3 * f], h.desc.vbeg + h.tris[3 * f + 1],
                               h.desc.vbeg + h.tris[3 * f + 2], 0));
  const int nt = static_cast<int>(tris.size());
  std::vector<int32_t> vptr(N + 1, 0), corner(3 * static_cast<size_t>(nt));
  for (int f = 0; f < nt; ++f) {
    vptr[tris[f].x + 1]++;
    vptr[tris[f].y + 1]++;
    vptr[tris[f].z + 1]++;
  }
  for (int v = 0; v < N; ++v) vptr[v + 1] += vptr[v];
  std::vector<int32_t> fill(vptr.begin(), vptr.end() - 1);
  for (int f = 0; f < nt; ++f) {  // ascending face order per vertex
    corner[fill[tris[f].x]++] = 3 * f;
    corner[fill[tris[f].y]++] = 3 * f + 1;
    corner[fill[tris[f].z]++] = 3 * f + 2;
  }
  XRT_CHECK_HIP(R.tris.Upload(tris, ctx->stream));
  XRT_CHECK_HIP(R.vptr.Upload(vptr, ctx->stream));
  XRT_CHECK_HIP(R.corner.Upload(corner, ctx->stream));
  XRT_CHECK_HIP(R.normals.Alloc(std::max(N, 1)));
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return XRT_OK;
}

int Setup(xrt_ctx* ctx, bool normals) {
  RecordData& R = ctx->record;
  const size_t N = static_cast<size_t>(ctx->wide.n_total);
  const size_t per = 3 * N * (normals ? 2 : 1);  // floats per slot
  if (!R.copy_stream) {
    XRT_CHECK_HIP(hipStreamCreateWithFlags(&R.copy_stream, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
      XRT_CHECK_HIP(hipEventCreateWithFlags(&R.packed[k], hipEventDisableTiming));
      XRT_CHECK_HIP(hipEventCreateWithFlags(&R.copied[k], hipEventDisableTiming));
    }
  }
  if (R.slot_floats < per) {
    XRT_CHECK_HIP(R.dev.Alloc(2 * per));
    if (R.host) XRT_CHECK_HIP(hipHostFree(R.host));
    R.host = nullptr;
    XRT_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&R.host), sizeof(float) * 2 * per,
                                hipHostMallocDefault));
    R.slot_floats = per;
    R.fresh[0] = R.fresh[1] = true;
  }
  if (normals) return SetupNormals(ctx);
  return XRT_OK;
}
}  // namespace

void RecordData::Release() {
  if (copy_stream) (void)hipStreamSynchronize(copy_stream);
  for (int k = 0; k < 2; ++k) {
    if (packed[k]) (void)hipEventDestroy(packed[k]);
    if (copied[k]) (void)hipEventDestroy(copied[k]);
    packed[k] = copied[k] = nullptr;
  }
  if (copy_stream) (void)hipStreamDestroy(copy_stream);
  copy_stream = nullptr;
  if (host) (void)hipHostFree(host);
  host = nullptr;
  slot_floats = 0;
}

}  // namespace xrt

using namespace xrt;

extern "C" {

int xrt_record_frames(xrt_ctx* ctx, int32_t nframes, float* positions, float* normals) {
  XRT_REQUIRE(ctx && ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(nframes > 0 && positions, XRT_ERR_INVALID_ARG, "nframes > 0 and positions non-null");
  int rc = Setup(ctx, normals != nullptr);
  if (rc != XRT_OK) return rc;
  RecordData& R = ctx->record;
  hipStream_t s = ctx->stream;
  const int N = ctx->wide.n_total;
  const size_t frame_floats = 3 * static_cast<size_t>(N);
  std::thread mover[2];
  struct Joiner {  // every helper thread has ended when this returns, also on errors
    std::thread* t;
    ~Joiner() {
      for (int k = 0; k < 2; ++k)
        if (t[k].joinable()) t[k].join();
    }
  } joiner{mover};
  for (int f = 0; f <= nframes; ++f) {
    const int slot = f & 1;
    if (f < nframes) {
      rc = xrt_frames_batched(ctx, 1);
      if (rc != XRT_OK) return rc;
      float* dslot = R.dev.ptr + slot * R.slot_floats;
      if (!R.fresh[slot]) XRT_CHECK_HIP(hipStreamWaitEvent(s, R.copied[slot], 0));  // slot free
      XRT_CHECK_HIP(Pack(s, ctx->x0.ptr, N, dslot));
      if (normals) {
        XRT_CHECK_HIP(VertexNormals(ctx->x0.ptr, R.tris.ptr, R.vptr.ptr, R.corner.ptr, N,
                                    R.normals.ptr, s));
        XRT_CHECK_HIP(Pack(s, R.normals.ptr, N, dslot + frame_floats));
      }
      XRT_CHECK_HIP(hipEventRecord(R.packed[slot], s));
      if (mover[slot].joinable()) mover[slot].join();  // pinned slot moved out
      XRT_CHECK_HIP(hipStreamWaitEvent(R.copy_stream, R.packed[slot], 0));
      XRT_CHECK_HIP(hipMemcpyAsync(R.host + slot * R.slot_floats, dslot,
                                   sizeof(float) * frame_floats * (normals ? 2 : 1),
                                   hipMemcpyDeviceToHost, R.copy_stream));
      XRT_CHECK_HIP(hipEventRecord(R.copied[slot], R.copy_stream));
      R.fresh[slot] = false;
    }
    if (f >= 1) {  // frame f - 1 is in pinned memory once its copy has ended
      const int ps = (f - 1) & 1;
      XRT_CHECK_HIP(hipEventSynchronize(R.copied[ps]));
      const float* h = R.host + ps * R.slot_floats;
      float* dp = positions + frame_floats * (f - 1);
      float* dn = normals ? normals + frame_floats * (f - 1) : nullptr;
      mover[ps] = std::thread([h, dp, dn, frame_floats]() {
        std::memcpy(dp, h, sizeof(float) * frame_floats);
        if (dn) std::memcpy(dn, h + frame_floats, sizeof(float) * frame_floats);
      });
    }
  }
  return XRT_OK;
}

}  // extern "C"
