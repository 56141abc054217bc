// Reverse-mode kernels of the body-cloth contact solves, replaying the forward's contact
// records (built with -ffp-contract=off like contact.hip: the activity tests below are the
// forward's expressions and decide identically on the taped input).
//
// VF (CCD offset 0, then DCD offset 1e-3): x' = x - min(n.(x - q) - off, 0) n with the
//   record (face, q) fixed; n = FaceNormal(pred face), so gx' -> gx - n (n.gx) when active.
//   Exact under replay: q lies in the face plane, n.q does not move with x.
// EE: delta_i += -(1 - s) lam fn, delta_j += -s lam fn with lam = (c1 - c2).fn - 1e-3,
//   c1 = (1 - s) x_i + s x_j, then x += relax delta / count.  The record (s, t, fn and the
//   obstacle edge) is held fixed; counts are recounted from the taped state.
// The adjoints are FP64 (double4 gx); the activity tests run in FP32 on the taped state.
#include <hip/hip_runtime.h>

#include "adjoint.h"
#include "bvh_device.h"
#include "contact_math.h"

namespace xrt {
namespace {
using namespace cm;
constexpr int kBs = 256;
inline int Blocks(int n) { return (n + kBs - 1) / kBs; }

__device__ __forceinline__ bool VfActive(BvhView t, int f, const float4 q, float offset, F3 x,
                                         F3& nrm, float& lam) {
  const int4 tr = t.tris[f];
  nrm = FaceNormal(L3(t.pred[tr.x]), L3(t.pred[tr.y]), L3(t.pred[tr.z]));
  lam = D(nrm, S(x, L3(q))) - offset;
  return lam < 0;
}

// the EE record of edge e is active at this state (SolveEEDCD's tests)
__device__ __forceinline__ bool EeActive(const EeRec& r, F3 xi, F3 xj) {
  if (r.a.x > 1e6f) return false;
  const float s = r.a.y, t = r.a.z;
  const F3 c1 = A(M(xi, 1 - s), M(xj, s));
  const F3 c2 = A(M(L3(r.p2), 1 - t), M(L3(r.q2), t));
  const F3 d = S(c1, c2);
  if (Len(d) < kEpsC) return false;
  return !(D(d, L3(r.fn)) - 1e-3f > 0);
}
}  // namespace

__global__ __launch_bounds__(kBs) void VfAdjKernel(const float4* __restrict__ x_pc, int n, BvhView t,
                                                   const int32_t* __restrict__ rt_face,
                                                   const float4* __restrict__ rt_q,
                                                   const int32_t* __restrict__ vf_face,
                                                   const float4* __restrict__ vf_q,
                                                   double4* __restrict__ gx) {
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (i >= n) return;
  // forward replay: CCD solve, then DCD solve on its result
  float4 xi = x_pc[i];
  F3 nc{0, 0, 0}, nd{0, 0, 0};
  bool ac = false, ad = false;
  float lam;
  const int fc = rt_face[i];
  if (fc >= 0 && VfActive(t, fc, rt_q[i], 0.0f, L3(xi), nc, lam)) {
    ac = true;
    xi.x += -lam * nc.x;
    xi.y += -lam * nc.y;
    xi.z += -lam * nc.z;
  }
  const int fd = vf_face[i];
  if (fd >= 0) ad = VfActive(t, fd, vf_q[i], 1e-3f, L3(xi), nd, lam);
  if (!ac && !ad) return;
  double4 g = gx[i];
  if (ad) {
    const double p = nd.x * g.x + nd.y * g.y + nd.z * g.z;
    g.x -= p * nd.x; g.y -= p * nd.y; g.z -= p * nd.z;
  }
  if (ac) {
    const double p = nc.x * g.x + nc.y * g.y + nc.z * g.z;
    g.x -= p * nc.x; g.y -= p * nc.y; g.z -= p * nc.z;
  }
  gx[i] = g;
}

__global__ __launch_bounds__(kBs) void EeAdjCountKernel(const float4* __restrict__ x,
                                                        const int2* __restrict__ edges, int ne,
                                                        const EeRec* __restrict__ rec,
                                                        int32_t* __restrict__ cnt) {
  const int e = blockIdx.x * kBs + threadIdx.x;
  if (e >= ne) return;
  const int2 ed = edges[e];
  if (!EeActive(rec[e], L3(x[ed.x]), L3(x[ed.y]))) return;
  atomicAdd(cnt + ed.x, 1);
  atomicAdd(cnt + ed.y, 1);
}

// per edge: the adjoint of its multiplier lam (0 for an inactive edge)
__global__ __launch_bounds__(kBs) void EeAdjKernel(const float4* __restrict__ x,
                                                   const int2* __restrict__ edges, int ne,
                                                   const EeRec* __restrict__ rec, float relax,
                                                   const double4* __restrict__ gx,
                                                   const int32_t* __restrict__ cnt,
                                                   double* __restrict__ gl_out) {
  const int e = blockIdx.x * kBs + threadIdx.x;
  if (e >= ne) return;
  const int2 ed = edges[e];
  const EeRec r = rec[e];
  if (!EeActive(r, L3(x[ed.x]), L3(x[ed.y]))) {
    gl_out[e] = 0.0;
    return;
  }
  const double s = r.a.y;
  const double fx = r.fn.x, fy = r.fn.y, fz = r.fn.z;
  const double4 gi = gx[ed.x], gj = gx[ed.y];
  const double ci = static_cast<double>(relax) / cnt[ed.x];
  const double cj = static_cast<double>(relax) / cnt[ed.y];
  // adjoint of lam (the corrections are -(1 - s) lam fn and -s lam fn, each scaled by
  // relax / count at its vertex)
  gl_out[e] = -(1 - s) * ci * (fx * gi.x + fy * gi.y + fz * gi.z) -
              s * cj * (fx * gj.x + fy * gj.y + fz * gj.z);
}

// per vertex with an active edge: the adjoint of c1 = (1 - s) x_i + s x_j through lam =
// fn . c1 - offset, summed over the vertex's edges in ascending edge order
__global__ __launch_bounds__(kBs) void EeAdjApplyKernel(double4* __restrict__ gx, int n,
                                                        int32_t* __restrict__ cnt,
                                                        const double* __restrict__ gl,
                                                        const int2* __restrict__ edges,
                                                        const EeRec* __restrict__ rec,
                                                        const int32_t* __restrict__ ve_ptr,
                                                        const int32_t* __restrict__ ve_edge) {
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (i >= n || cnt[i] <= 0) return;
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int q = ve_ptr[i]; q < ve_ptr[i + 1]; ++q) {
    const int e = ve_edge[q];
    const double g = gl[e];
    if (g == 0.0) continue;
    const EeRec r = rec[e];
    const double s = r.a.y;
    const double w = edges[e].x == i ? 1.0 - s : s;
    ax += w * r.fn.x * g;
    ay += w * r.fn.y * g;
    az += w * r.fn.z * g;
  }
  double4 v = gx[i];
  v.x += ax;
  v.y += ay;
  v.z += az;
  gx[i] = v;
  cnt[i] = 0;
}

hipError_t LaunchVfAdj(hipStream_t s, const float4* x_pc, int n, const int4* obs_tris,
                       const float4* obs_pred, const int32_t* rt_face, const float4* rt_q,
                       const int32_t* vf_face, const float4* vf_q, double4* gx) {
  if (n <= 0) return hipSuccess;
  BvhView t;
  t.tris = obs_tris;
  t.pred = const_cast<float4*>(obs_pred);
  VfAdjKernel<<<Blocks(n), kBs, 0, s>>>(x_pc, n, t, rt_face, rt_q, vf_face, vf_q, gx);
  return hipGetLastError();
}

hipError_t LaunchEeAdj(hipStream_t s, const float4* x_pe, const int2* edges, int ne, int nv,
                       const EeRec* rec, float relax, double4* gx, int32_t* cnt, double* gl,
                       const int32_t* ve_ptr, const int32_t* ve_edge) {
  if (ne <= 0 || nv <= 0) return hipSuccess;
  EeAdjCountKernel<<<Blocks(ne), kBs, 0, s>>>(x_pe, edges, ne, rec, cnt);
  EeAdjKernel<<<Blocks(ne), kBs, 0, s>>>(x_pe, edges, ne, rec, relax, gx, cnt, gl);
  EeAdjApplyKernel<<<Blocks(nv), kBs, 0, s>>>(gx, nv, cnt, gl, edges, rec, ve_ptr, ve_edge);
  return hipGetLastError();
}

}  // namespace xrt
