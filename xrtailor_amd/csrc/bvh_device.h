// Device-side LBVH helpers shared by the build / query kernels and the contact kernels
// (header-only: every translation unit inlines its own copy).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bvh.h"

namespace xrt {

constexpr int kMaxHits = 256;  // LBVH_MAX_BUFFER_SIZE (xrtailor/include/xrtailor/core/Common.cuh:24)

// MathFunctions::min / max (xrtailor/include/xrtailor/math/MathFunctions.cuh:24-51)
__device__ __forceinline__ float Mn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float Mx(float a, float b) { return a > b ? a : b; }

__device__ __forceinline__ Box MergeBox(const Box& a, const Box& b) {
  return Box{make_float4(Mn(a.lo.x, b.lo.x), Mn(a.lo.y, b.lo.y), Mn(a.lo.z, b.lo.z), 0.0f),
             make_float4(Mx(a.hi.x, b.hi.x), Mx(a.hi.y, b.hi.y), Mx(a.hi.z, b.hi.z), 0.0f)};
}

// Bounds::Overlap(p): strict containment (xrtailor/src/physics/broad_phase/Bounds.cu:130-136)
__device__ __forceinline__ bool OverlapPoint(const Box& b, float4 p) {
  return p.x > b.lo.x && p.x < b.hi.x && p.y > b.lo.y && p.y < b.hi.y && p.z > b.lo.z &&
         p.z < b.hi.z;
}

__device__ __forceinline__ bool GetIntersection(float d1, float d2, float4 p1, float4 p2,
                                                float4& hit) {  // Bounds.cu:139-147
  if ((d1 * d2) >= 0.0f) return false;
  if (d1 == d2) return false;
  const float s = -d1 / (d2 - d1);
  hit = make_float4(p1.x + (p2.x - p1.x) * s, p1.y + (p2.y - p1.y) * s, p1.z + (p2.z - p1.z) * s,
                    0.0f);
  return true;
}

__device__ __forceinline__ bool InBox(float4 h, const Box& b, int axis) {  // Bounds.cu:149-158
  if (axis == 1 && h.z > b.lo.z && h.z < b.hi.z && h.y > b.lo.y && h.y < b.hi.y) return true;
  if (axis == 2 && h.z > b.lo.z && h.z < b.hi.z && h.x > b.lo.x && h.x < b.hi.x) return true;
  if (axis == 3 && h.x > b.lo.x && h.x < b.hi.x && h.y > b.lo.y && h.y < b.hi.y) return true;
  return false;
}

// Bounds::Overlap(p1, p2, hit): segment vs box (Bounds.cu:160-191)
__device__ __forceinline__ bool OverlapSegment(const Box& b, float4 p1, float4 p2) {
  if (p2.x < b.lo.x && p1.x < b.lo.x) return false;
  if (p2.x > b.hi.x && p1.x > b.hi.x) return false;
  if (p2.y < b.lo.y && p1.y < b.lo.y) return false;
  if (p2.y > b.hi.y && p1.y > b.hi.y) return false;
  if (p2.z < b.lo.z && p1.z < b.lo.z) return false;
  if (p2.z > b.hi.z && p1.z > b.hi.z) return false;
  if (OverlapPoint(b, p1)) return true;
  float4 h;
  return (GetIntersection(p1.x - b.lo.x, p2.x - b.lo.x, p1, p2, h) && InBox(h, b, 1)) ||
         (GetIntersection(p1.y - b.lo.y, p2.y - b.lo.y, p1, p2, h) && InBox(h, b, 2)) ||
         (GetIntersection(p1.z - b.lo.z, p2.z - b.lo.z, p1, p2, h) && InBox(h, b, 3)) ||
         (GetIntersection(p1.x - b.hi.x, p2.x - b.hi.x, p1, p2, h) && InBox(h, b, 1)) ||
         (GetIntersection(p1.y - b.hi.y, p2.y - b.hi.y, p1, p2, h) && InBox(h, b, 2)) ||
         (GetIntersection(p1.z - b.hi.z, p2.z - b.hi.z, p1, p2, h) && InBox(h, b, 3));
}

// Bounds::Overlap(b, thickness) = Overlap(b.Dilate(thickness)) (Bounds.cu:96-128)
__device__ __forceinline__ bool OverlapDilated(const Box& rhs, const Box& lhs, float th) {
  const float lx = lhs.lo.x - th, ly = lhs.lo.y - th, lz = lhs.lo.z - th;
  const float hx = lhs.hi.x + th, hy = lhs.hi.y + th, hz = lhs.hi.z + th;
  if (rhs.lo.x > hx || rhs.hi.x < lx) return false;
  if (rhs.lo.y > hy || rhs.hi.y < ly) return false;
  if (rhs.lo.z > hz || rhs.hi.z < lz) return false;
  return true;
}

// QueryDeviceStackless (xrtailor/src/physics/broad_phase/lbvh/BVH.cu:957-1021): visits the
// leaves whose boxes contain p (segment == false) or meet segment p->q, in traversal order,
// calling f(object) for at most kMaxHits of them.
template <bool kSegment, typename F>
__device__ __forceinline__ int StacklessQuery(const BvhView& t, float4 p, float4 q, F&& f) {
  int node = 0, found = 0;
  const int max_steps = 2 * t.nf;  // a stackless pass visits each node at most once
  for (int step = 0; node != -1 && step < max_steps; ++step) {
    if (found >= kMaxHits) break;
    const Box bx = t.box[node];
    const bool hit = kSegment ? OverlapSegment(bx, p, q) : OverlapPoint(bx, p);
    const int4 nd = t.nodes[node];
    if (hit) {
      if (nd.w >= 0) {
        f(nd.w);
        ++found;
        node = t.escape[node];
      } else {
        node = nd.y;
      }
    } else {
      node = t.escape[node];
    }
  }
  return found;
}

}  // namespace xrt
