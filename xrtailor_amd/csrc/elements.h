// Per-element device functions of the XPBD substep, shared by the plain, batched and
// wide (multi-instance) launch structures and by the adjoint's replay.  P is any
// indexable float4 store (global or LDS pointer).
#pragma once
#include "device.h"

namespace xrt {

// ------------------------------------------------------------------ elements
// SolveStretch_Kernel body (xrtailor/src/physics/dynamics/BasicSolverHelper.cu:80-133) on
// loaded endpoint values; returns whether a and b changed.  Every launch structure calls
// this one function, so all of them round identically.
__device__ __forceinline__ bool StretchSolve(float4& a, float4& b, const StretchC& c, float& lam,
                                             float inv_dt2) {
  // no FMA contraction: inlined into different launch structures (and unroll variants) the
  // compiler would otherwise fuse different products, and the paths would round differently
#pragma clang fp contract(off)
  const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
  const float w1 = a.w, w2 = b.w;
  if (!(dist != c.rest && w1 + w2 > 0.0f)) return false;
  const float inv = 1.0f / (dist + kEps);
  const float at = c.compliance * inv_dt2;
  const float dl = (-(dist - c.rest) - at * lam) / (w1 + w2 + at);
  lam += dl;
  const float s = dl * inv;
  a.x += w1 * s * dx; a.y += w1 * s * dy; a.z += w1 * s * dz;
  b.x -= w2 * s * dx; b.y -= w2 * s * dy; b.z -= w2 * s * dz;
  return true;
}

template <typename P>
__device__ __forceinline__ void StretchProject(P X, const StretchC c, float* __restrict__ lam_ptr,
                                               float inv_dt2, int iter) {
  float lam = (iter == 0) ? 0.0f : *lam_ptr;
  float4 a = X[c.i0];
  float4 b = X[c.i1];
  if (StretchSolve(a, b, c, lam, inv_dt2)) {
    X[c.i0] = a;
    X[c.i1] = b;
  }
  *lam_ptr = lam;
}

// (no FMA contraction here either: these are inlined into the solve bodies)
__device__ __forceinline__ float3 Cross(float3 a, float3 b) {
#pragma clang fp contract(off)
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float Dot(float3 a, float3 b) {
#pragma clang fp contract(off)
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ float3 Scale(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 Add(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 Sub(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }

// SolveBend_Kernel body (xrtailor/src/physics/dynamics/BasicSolverHelper.cu:8-72): Mueller
// dihedral bend with d = clamp(n1.n2, lo, 1) and no multiplier accumulation.  BendMode
// selects the reference formula (lo 0, rest theta0, +q) or the consistent one
// (lo -1, rest pi - theta0, -q); both cost the same.
// On loaded vertex values; returns whether they changed.
__device__ __forceinline__ bool BendSolve(float4& x1, float4& x2, float4& x3, float4& x4,
                                          const float2 par, float inv_dt2, const BendMode m) {
#pragma clang fp contract(off)  // see StretchSolve
  const float3 p2 = make_float3(x2.x - x1.x, x2.y - x1.y, x2.z - x1.z);
  const float3 p3 = make_float3(x3.x - x1.x, x3.y - x1.y, x3.z - x1.z);
  const float3 p4 = make_float3(x4.x - x1.x, x4.y - x1.y, x4.z - x1.z);
  const float3 c23 = Cross(p2, p3), c24 = Cross(p2, p4);
  const float l23r = sqrtf(Dot(c23, c23)), l24r = sqrtf(Dot(c24, c24));
  const float3 n1 = Scale(c23, 1.0f / l23r), n2 = Scale(c24, 1.0f / l24r);
  float d = Dot(n1, n2);
  if (d != d) return false;  // NaN from a degenerate triangle
  d = fminf(fmaxf(d, m.dlo), 1.0f);
  const float ang = acosf(d);
  if (ang < kEps) return false;
  const float i23 = 1.0f / (l23r + kEps), i24 = 1.0f / (l24r + kEps);
  const float3 q3 = Scale(Add(Cross(p2, n2), Scale(Cross(n1, p2), d)), i23);
  const float3 q4 = Scale(Add(Cross(p2, n1), Scale(Cross(n2, p2), d)), i24);
  const float3 q2 = Sub(Scale(Add(Cross(p3, n2), Scale(Cross(n1, p3), d)), -i23),
                        Scale(Add(Cross(p4, n1), Scale(Cross(n2, p4), d)), i24));
  const float3 q1 = Scale(Add(Add(q2, q3), q4), -1.0f);
  const float w1 = x1.w, w2 = x2.w, w3 = x3.w, w4 = x4.w;
  const float denom = par.y * inv_dt2 +
                      (w1 * Dot(q1, q1) + w2 * Dot(q2, q2) + w3 * Dot(q3, q3) + w4 * Dot(q4, q4));
  if (denom < kEps) return false;
  const float lam = m.sgn * sqrtf(1.0f - d * d) * (ang - (m.th0 + m.th1 * par.x)) / denom;
  x1.x += w1 * lam * q1.x; x1.y += w1 * lam * q1.y; x1.z += w1 * lam * q1.z;
  x2.x += w2 * lam * q2.x; x2.y += w2 * lam * q2.y; x2.z += w2 * lam * q2.z;
  x3.x += w3 * lam * q3.x; x3.y += w3 * lam * q3.y; x3.z += w3 * lam * q3.z;
  x4.x += w4 * lam * q4.x; x4.y += w4 * lam * q4.y; x4.z += w4 * lam * q4.z;
  return true;
}

template <typename P>
__device__ __forceinline__ void BendProject(P X, const int4 id, const float2 par, float inv_dt2,
                                            const BendMode m) {
  float4 x1 = X[id.x], x2 = X[id.y], x3 = X[id.z], x4 = X[id.w];
  if (BendSolve(x1, x2, x3, x4, par, inv_dt2, m)) {
    X[id.x] = x1; X[id.y] = x2; X[id.z] = x3; X[id.w] = x4;
  }
}

// One color of stretch constraints [beg, end) by the threads of a workgroup, U per thread
// per round: the records and endpoint positions of all U are loaded before any is solved
// (they share no vertex within a color), so U gathers are in flight at once.
template <int U, typename P>
__device__ __forceinline__ void StretchColor(P X, const StretchC* __restrict__ sc,
                                             float* __restrict__ slam, int beg, int end,
                                             float inv_dt2, int iter, int tid, int bs) {
  for (int e0 = beg + tid; e0 < end; e0 += U * bs) {
    StretchC c[U];
    float lam[U];
    float4 a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * bs;
      if (e < end) {
        c[u] = sc[e];
        lam[u] = iter == 0 ? 0.0f : slam[e];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (e0 + u * bs < end) {
        a[u] = X[c[u].i0];
        b[u] = X[c[u].i1];
      }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * bs;
      if (e >= end) continue;
      if (StretchSolve(a[u], b[u], c[u], lam[u], inv_dt2)) {
        X[c[u].i0] = a[u];
        X[c[u].i1] = b[u];
      }
      slam[e] = lam[u];
    }
  }
}

// One color of bend elements [beg, end), U per thread per round (see StretchColor).
template <int U, typename P>
__device__ __forceinline__ void BendColor(P X, const int4* __restrict__ bidx,
                                          const float2* __restrict__ bpar, int beg, int end,
                                          float inv_dt2, const BendMode m, int tid, int bs) {
  for (int e0 = beg + tid; e0 < end; e0 += U * bs) {
    int4 id[U];
    float2 par[U];
    float4 x1[U], x2[U], x3[U], x4[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * bs;
      if (e < end) {
        id[u] = bidx[e];
        par[u] = bpar[e];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (e0 + u * bs < end) {
        x1[u] = X[id[u].x]; x2[u] = X[id[u].y]; x3[u] = X[id[u].z]; x4[u] = X[id[u].w];
      }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (e0 + u * bs >= end) continue;
      if (BendSolve(x1[u], x2[u], x3[u], x4[u], par[u], inv_dt2, m)) {
        X[id[u].x] = x1[u]; X[id[u].y] = x2[u]; X[id[u].z] = x3[u]; X[id[u].w] = x4[u];
      }
    }
  }
}

// SolveGeodesicLRA_Kernel (xrtailor/src/physics/dynamics/LongRangeAttachments.cu:5-26),
// one target vertex at a time; slot positions come from the pre-pass snapshot `sp`.
template <typename PS>
__device__ __forceinline__ float4 LraVertex(float4 xt, PS sp, const int32_t* __restrict__ slot,
                                            const float* __restrict__ rest, int kb, int ke,
                                            float stretchiness) {
#pragma clang fp contract(off)  // see StretchSolve
  for (int k = kb; k < ke; ++k) {
    const float target = rest[k] * stretchiness;
    if (xt.w < kEps && target > 0.0f) continue;
    const float4 s = sp[slot[k]];
    const float dx = xt.x - s.x, dy = xt.y - s.y, dz = xt.z - s.z;
    const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
    if (dist > target) {
      const float f = target / dist;
      xt.x += -dx + dx * f; xt.y += -dy + dy * f; xt.z += -dz + dz * f;
    }
  }
  return xt;
}

// plane y = ground_y with ComputeFriction (xrtailor/src/physics/ClothSolverHelper.cu:156-169)
__device__ __forceinline__ float4 GroundVertex(float4 x, float4 x0, float ground_y, float mu) {
#pragma clang fp contract(off)  // see StretchSolve
  if (x.w <= kEps || x.y >= ground_y) return x;
  const float cl = ground_y - x.y;
  float fx = 0.0f, fz = 0.0f;
  if (mu > 0.0f && cl > 0.0f) {
    const float tx = x.x - x0.x, tz = x.z - x0.z;  // tangential part of (x - x0) w.r.t. +y
    const float tl = sqrtf(tx * tx + tz * tz);
    const float mx = cl * mu;
    const float s = (tl > 0.0f) ? fminf(mx / tl, 1.0f) : 1.0f;
    fx = -tx * s; fz = -tz * s;
  }
  x.x += fx; x.y += cl; x.z += fz;
  return x;
}

}  // namespace xrt
