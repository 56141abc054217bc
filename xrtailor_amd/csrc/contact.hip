// Contact kernels (built with -ffp-contract=off like the LBVH, so that the queries'
// float predicates round as in the FP32 oracle).
#include <hip/hip_runtime.h>

#include "bvh_device.h"
#include "contact.h"
#include "contact_math.h"

namespace xrt {
namespace {
using namespace cm;
constexpr int kBlock = 256;
constexpr float kFix = 4294967296.0f;     // 2^32: 32.32 fixed point
inline int Blocks(int n) { return (n + kBlock - 1) / kBlock; }

// Ericson ClosestPtPointTriangle (BasicPrimitiveTests.cu:6-100)
__device__ float NearestPoint(F3 p, F3 a, F3 b, F3 c, F3& q) {
  const F3 e1 = S(b, a), e2 = S(c, a), e3 = S(c, b);
  const F3 v1 = S(p, a);
  const float d1 = D(e1, v1), d2 = D(e2, v1);
  if (d1 <= 0 && d2 <= 0) { q = a; return Len(S(p, q)); }
  const F3 v2 = S(p, b);
  const float d3 = D(e1, v2), d4 = D(e2, v2);
  if (d3 >= 0 && d4 <= d3) { q = b; return Len(S(p, q)); }
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) { q = A(a, M(e1, d1 / (d1 - d3))); return Len(S(p, q)); }
  const F3 v3 = S(p, c);
  const float d5 = D(e1, v3), d6 = D(e2, v3);
  if (d6 >= 0 && d5 <= d6) { q = c; return Len(S(p, q)); }
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) { q = A(a, M(e2, d2 / (d2 - d6))); return Len(S(p, q)); }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    q = A(b, M(e3, (d4 - d3) / ((d4 - d3) + (d5 - d6))));
    return Len(S(p, q));
  }
  const float dd = 1.0f / (va + vb + vc);
  q = A(a, A(M(e1, vb * dd), M(e2, vc * dd)));
  return Len(S(p, q));
}

// Moller-Trumbore, hit if 0 < s < length (BasicPrimitiveTests.cu:242-284)
__device__ float RayIntersect(F3 o, F3 d, float length, F3 a, F3 b, F3 c) {
  const F3 e1 = S(b, a), e2 = S(c, a);
  const F3 n = C(d, e2);
  const float det = D(e1, n);
  if (fabsf(det) < kEpsC) return INFINITY;
  const float inv = 1.0f / det;
  const F3 t = S(o, a);
  const float u = D(t, n) * inv;
  if (u < 0 || u > 1) return INFINITY;
  const F3 q = C(t, e1);
  const float v = D(d, q) * inv;
  if (v < 0 || v + u > 1) return INFINITY;
  const float s = D(e2, q) * inv;
  if (s > 0 && s < length) return s;
  return INFINITY;
}

// Ericson ClosestPtSegmentSegment with EPSILON, t < EPS -> 0 (BasicPrimitiveTests.cu:303-361)
__device__ float ClosestSegSeg(F3 p1, F3 q1, F3 p2, F3 q2, float& s, float& t) {
  const F3 d1 = S(q1, p1), d2 = S(q2, p2), r = S(p1, p2);
  const float a = D(d1, d1), e = D(d2, d2), f = D(d2, r);
  if (a <= kEpsC && e <= kEpsC) {
    s = t = 0;
  } else if (a <= kEpsC) {
    s = 0;
    t = Cl(f / e, 0.0f, 1.0f);
  } else {
    const float c = D(d1, r);
    if (e <= kEpsC) {
      t = 0;
      s = Cl(-c / a, 0.0f, 1.0f);
    } else {
      const float b = D(d1, d2), den = a * e - b * b;
      s = den > kEpsC ? Cl((b * f - c * e) / den, 0.0f, 1.0f) : 0.0f;
      t = (b * s + f) / e;
      if (t < kEpsC) { t = 0; s = Cl(-c / a, 0.0f, 1.0f); }
      else if (t > 1) { t = 1; s = Cl((b - c) / a, 0.0f, 1.0f); }
    }
  }
  return Len(S(A(p1, M(d1, s)), A(p2, M(d2, t))));
}

// SignedVertexFaceDistance (BasicPrimitiveTests.cu:369-386); w zeroed when degenerate
__device__ __forceinline__ float SignedVF(F3 x, F3 y0, F3 y1, F3 y2, F3& n, float* w) {
  n = C(Nz(S(y1, y0)), Nz(S(y2, y0)));
  if (D(n, n) < 1e-6f) { w[0] = w[1] = w[2] = w[3] = 0; return INFINITY; }
  n = Nz(n);
  const float h = D(S(x, y0), n);
  const float b0 = Mixed(S(y1, x), S(y2, x), n);
  const float b1 = Mixed(S(y2, x), S(y0, x), n);
  const float b2 = Mixed(S(y0, x), S(y1, x), n);
  w[0] = 1.0f;
  w[1] = -b0 / (b0 + b1 + b2);
  w[2] = -b1 / (b0 + b1 + b2);
  w[3] = -b2 / (b0 + b1 + b2);
  return h;
}

// SignedEdgeEdgeDistance (BasicPrimitiveTests.cu:420-436): n is left unnormalized (the
// reference discards glm::normalize(n))
__device__ __forceinline__ float SignedEE(F3 x0, F3 x1, F3 y0, F3 y1, F3& n, float* w) {
  n = C(Nz(S(x1, x0)), Nz(S(y1, y0)));
  if (D(n, n) < 1e-6f) { w[0] = w[1] = w[2] = w[3] = 0; return INFINITY; }
  const float h = D(S(x0, y0), n);
  const float a0 = Mixed(S(y1, x1), S(y0, x1), n), a1 = Mixed(S(y0, x0), S(y1, x0), n);
  const float b0 = Mixed(S(x0, y1), S(x1, y1), n), b1 = Mixed(S(x1, y0), S(x0, y0), n);
  w[0] = a0 / (a0 + a1);
  w[1] = a1 / (a0 + a1);
  w[2] = -b0 / (b0 + b1);
  w[3] = -b1 / (b0 + b1);
  return h;
}

__device__ __forceinline__ void AddFixed(long long* dq, int32_t* cnt, int i, F3 v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(dq + 3 * i),
            static_cast<unsigned long long>(__float2ll_rn(v.x * kFix)));
  atomicAdd(reinterpret_cast<unsigned long long*>(dq + 3 * i + 1),
            static_cast<unsigned long long>(__float2ll_rn(v.y * kFix)));
  atomicAdd(reinterpret_cast<unsigned long long*>(dq + 3 * i + 2),
            static_cast<unsigned long long>(__float2ll_rn(v.z * kFix)));
  atomicAdd(cnt + i, 1);
}

// ------------------------------------------------------------------ bodies
// Each body handles one vertex / edge / pair; the single-instance kernels below and the
// multi-instance kernels (one launch for every instance) call the same body.

// PhysicsMesh::Interpolate (xrtailor/src/physics/PhysicsMesh.cu:610-620)
__device__ __forceinline__ void InterpolateBody(const float4* __restrict__ x0,
                                                const float4* __restrict__ x,
                                                float4* __restrict__ x1, int i, float alpha) {
  const float4 a = x0[i], b = x[i];
  const float r = 1.0f - alpha;
  x1[i] = make_float4(r * a.x + alpha * b.x, r * a.y + alpha * b.y, r * a.z + alpha * b.z, a.w);
}

// GenerateVF_DCD_Kernel (PBDCollision.cu:17-43)
// the obstacle query of the generate bodies: over the packed records when they are current
// (the batched collision pass packs them after its refit), else over the node arrays
template <bool kSegment, bool kPk, class F>
__device__ __forceinline__ void ObsQuery(const BvhView& t, float4 p, float4 q, F&& f) {
  if constexpr (kPk) StacklessQueryPacked<kSegment>(t.packed, t.nf, p, q, f);
  else StacklessQuery<kSegment>(t, p, q, f);
}

template <bool kPk = false>
__device__ __forceinline__ void VfDcdGenerateBody(const float4* __restrict__ x, int i,
                                                  const BvhView& t, int32_t* __restrict__ face,
                                                  float4* __restrict__ qout) {
  const float4 p4 = x[i];
  const F3 p = L3(p4);
  float best = kScalarMax;
  int bi = -1;
  F3 bq{0, 0, 0};
  ObsQuery<false, kPk>(t, p4, p4, [&](int f) {
    const int4 tr = t.tris[f];
    F3 q;
    const float d = NearestPoint(p, L3(t.pred[tr.x]), L3(t.pred[tr.y]), L3(t.pred[tr.z]), q);
    if (d < best) { best = d; bi = f; bq = q; }
  });
  face[i] = bi;
  qout[i] = make_float4(bq.x, bq.y, bq.z, 0.0f);
}

// GenerateVFCCD_Kernel (PBDCollision.cu:236-270)
template <bool kPk = false>
__device__ __forceinline__ void VfCcdGenerateBody(const float4* __restrict__ x0,
                                                  const float4* __restrict__ x, int i,
                                                  const BvhView& t, int32_t* __restrict__ face,
                                                  float4* __restrict__ qout) {
  face[i] = -1;
  qout[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float4 a4 = x0[i], b4 = x[i];
  const F3 p1 = L3(a4);
  F3 d = S(L3(b4), p1);
  const float len = Len(d);
  if (len < kEpsC) return;
  d = M(d, 1.0f / len);
  float best = kScalarMax;
  ObsQuery<true, kPk>(t, a4, b4, [&](int f) {
    const int4 tr = t.tris[f];
    const float s = RayIntersect(p1, d, len, L3(t.pred[tr.x]), L3(t.pred[tr.y]), L3(t.pred[tr.z]));
    if (s < best) {
      best = s;
      face[i] = f;
      const F3 q = A(p1, M(d, s));
      qout[i] = make_float4(q.x, q.y, q.z, 0.0f);
    }
  });
}

// GenerateEE_DCD_Kernel (PBDCollision.cu:100-150)
template <bool kPk = false>
__device__ __forceinline__ void EeDcdGenerateBody(const float4* __restrict__ x,
                                                  const int2* __restrict__ edges, int e,
                                                  const BvhView& t,
                                                  const uint32_t* __restrict__ rtri,
                                                  EeRec* __restrict__ rec,
                                                  uint8_t* __restrict__ none) {
  const int2 ed = edges[e];
  const float4 pc4 = x[ed.x], qc4 = x[ed.y];
  const F3 pc = L3(pc4), qc = L3(qc4);
  EeRec r;
  r.a = make_float4(kScalarMax, 0.0f, 0.0f, 0.0f);
  r.p2 = r.q2 = r.fn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  ObsQuery<true, kPk>(t, pc4, qc4, [&](int f) {
    const int4 tr = t.tris[f];
    const uint32_t info = rtri[f];
    const F3 ov[3] = {L3(t.pred[tr.x]), L3(t.pred[tr.y]), L3(t.pred[tr.z])};
    for (int k = 0; k < 3; ++k) {
      if (!((info >> (3 + k)) & 1u)) continue;
      const int j = (k + 1) % 3;
      float s, tt;
      const float dist = ClosestSegSeg(pc, qc, ov[k], ov[j], s, tt);
      if (dist == 0.0f) continue;  // the reference uses the distance as the success flag
      if (dist < r.a.x) {
        r.a = make_float4(dist, s, tt, 0.0f);
        r.p2 = make_float4(ov[k].x, ov[k].y, ov[k].z, 0.0f);
        r.q2 = make_float4(ov[j].x, ov[j].y, ov[j].z, 0.0f);
        const F3 fn = FaceNormal(ov[0], ov[1], ov[2]);
        r.fn = make_float4(fn.x, fn.y, fn.z, 0.0f);
      }
    }
  });
  rec[e] = r;
  none[e] = r.a.x > 1e6f ? 1 : 0;
}

// SolveVFDCD_Kernel / SolveVFCCD_Kernel (PBDCollision.cu:54-76,280-300)
// VF solve: push along the pred face normal to the contact point (+offset), then, with
// fr.mu > 0, ComputeFriction (ClothSolverHelper.cu:156-169) on that correction: the part of
// the vertex's substep displacement (x - x0) relative to the obstacle face's (its vertices'
// mean frame displacement / substeps) that is tangential to the normal is removed, at most
// mu x |correction|.  The reference applies no friction on mesh contacts; mu = 0 is its
// behaviour.
__device__ __forceinline__ void VfSolveBody(float4* __restrict__ x, int i, const BvhView& t,
                                            const int32_t* __restrict__ face,
                                            const float4* __restrict__ q, float offset,
                                            const VfFriction& fr) {
  const int f = face[i];
  if (f < 0) return;
  const int4 tr = t.tris[f];
  const F3 nrm = FaceNormal(L3(t.pred[tr.x]), L3(t.pred[tr.y]), L3(t.pred[tr.z]));
  float4 xi = x[i];
  const float lam = D(nrm, S(L3(xi), L3(q[i]))) - offset;
  if (lam < 0) {
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (fr.mu > 0.0f) {
      const float4 a0 = fr.o_x0[tr.x], b0 = fr.o_x0[tr.y], c0 = fr.o_x0[tr.z];
      const float4 a1 = fr.o_x[tr.x], b1 = fr.o_x[tr.y], c1 = fr.o_x[tr.z];
      const float ox = ((a1.x - a0.x) + (b1.x - b0.x) + (c1.x - c0.x)) / 3.0f * fr.inv_ns;
      const float oy = ((a1.y - a0.y) + (b1.y - b0.y) + (c1.y - c0.y)) / 3.0f * fr.inv_ns;
      const float oz = ((a1.z - a0.z) + (b1.z - b0.z) + (c1.z - c0.z)) / 3.0f * fr.inv_ns;
      const float4 p0 = fr.x0[i];
      const float rx = (xi.x - p0.x) - ox, ry = (xi.y - p0.y) - oy, rz = (xi.z - p0.z) - oz;
      const float rn = rx * nrm.x + ry * nrm.y + rz * nrm.z;
      const float tx = rx - nrm.x * rn, ty = ry - nrm.y * rn, tz = rz - nrm.z * rn;
      const float tl = sqrtf(tx * tx + ty * ty + tz * tz);
      if (tl > 0.0f) {
        const float s = fminf(-lam * fr.mu / tl, 1.0f);
        fx = -tx * s;
        fy = -ty * s;
        fz = -tz * s;
      }
    }
    xi.x += -lam * nrm.x + fx;
    xi.y += -lam * nrm.y + fy;
    xi.z += -lam * nrm.z + fz;
    x[i] = xi;
  }
}

// SolveEEDCD_Kernel (PBDCollision.cu:155-214)
// none[e] mirrors `rec[e].a.x > 1e6f` (no record), one byte per edge, so that the edges
// without a record, most of them, cost no read of their 64-byte record
__device__ __forceinline__ void EeSolveBody(const float4* __restrict__ x,
                                            const int2* __restrict__ edges, int e,
                                            const EeRec* __restrict__ rec,
                                            const uint8_t* __restrict__ none,
                                            long long* __restrict__ dq, int32_t* __restrict__ cnt) {
  if (none[e]) return;
  const EeRec r = rec[e];
  if (r.a.x > 1e6f) return;
  const int2 ed = edges[e];
  const float s = r.a.y, t = r.a.z;
  const F3 fn = L3(r.fn);
  const F3 c1 = A(M(L3(x[ed.x]), 1 - s), M(L3(x[ed.y]), s));
  const F3 c2 = A(M(L3(r.p2), 1 - t), M(L3(r.q2), t));
  const F3 d = S(c1, c2);
  if (Len(d) < kEpsC) return;
  const float lam = D(d, fn) - 1e-3f;
  if (lam > 0) return;
  AddFixed(dq, cnt, ed.x, M(fn, -(1 - s) * lam));
  AddFixed(dq, cnt, ed.y, M(fn, -s * lam));
}

// ApplyDeltas_Kernel / ApplyPBDImpulse_Kernel: x += delta / count * relax, then reset
__device__ __forceinline__ void ApplyFixedBody(float4* __restrict__ x, int i,
                                               long long* __restrict__ dq,
                                               int32_t* __restrict__ cnt, float relax) {
  const int c = cnt[i];
  if (c <= 0) return;
  const double sc = static_cast<double>(relax) / (static_cast<double>(c) * 4294967296.0);
  float4 xi = x[i];
  xi.x += static_cast<float>(static_cast<double>(dq[3 * i]) * sc);
  xi.y += static_cast<float>(static_cast<double>(dq[3 * i + 1]) * sc);
  xi.z += static_cast<float>(static_cast<double>(dq[3 * i + 2]) * sc);
  x[i] = xi;
  dq[3 * i] = dq[3 * i + 1] = dq[3 * i + 2] = 0;
  cnt[i] = 0;
}

// ------------------------------------------------------------- repulsion
struct RMesh {
  MeshTopo t;
  MeshState s;
  __device__ __forceinline__ F3 X(int i) const { return L3(s.x[i]); }
  __device__ __forceinline__ F3 X0(int i) const { return L3(s.x0[i]); }
  __device__ __forceinline__ float W(int i) const { return s.fixed_w >= 0.0f ? s.fixed_w : s.x0[i].w; }
  __device__ __forceinline__ bool Free(int i) const { return s.fixed_w < 0.0f && s.x0[i].w > 0.0f; }
};

__device__ F3 EdgeNormal(const RMesh& m, int e) {  // Edge::ComputeNormal (Edge.cu:36-50)
  F3 sum{0, 0, 0};
  int cnt = 0;
  for (int k = m.t.ef_ptr[e]; k < m.t.ef_ptr[e + 1] && cnt < 4; ++k, ++cnt) {
    const int4 tr = m.t.tris[m.t.ef[k]];
    sum = A(sum, FaceNormal(m.X(tr.x), m.X(tr.y), m.X(tr.z)));
  }
  return cnt ? M(sum, 1.0f / static_cast<float>(cnt)) : sum;
}

__device__ bool InEdge(float w, const RMesh& m0, int e0, const RMesh& m1, int e1) {
  const int2 a = m0.t.edges[e0];
  const F3 x = A(M(m0.X(a.x), 1 - w), M(m0.X(a.y), w));  // BasicPrimitiveTests.cu:286-301
  bool in = true;
  const int2 b = m1.t.edges[e1];
  const F3 n0 = m1.X(b.x), n1 = m1.X(b.y);
  int cnt = 0;
  for (int k = m1.t.ef_ptr[e1]; k < m1.t.ef_ptr[e1 + 1] && cnt < 4; ++k, ++cnt) {
    const F3 fn = L3(m1.t.face_n[m1.t.ef[k]]);
    in &= Mixed(S(n1, n0), fn, S(x, n0)) >= 0;
  }
  return in;
}

// AccumulateRepulsions (PBDRepulsion.cu:264-277): nodes [0, kSplit) belong to m0, the
// rest to m1.  With `out` (the adjoint's replay) the accumulated corners go to out->mask
// and out->ids, and the sums only where dq / cnt are given (a null dq adds counts only).
template <int kSplit>
__device__ __forceinline__ void Accumulate(const RMesh& m0, const RMesh& m1, const int* ids,
                                           const F3* corr, long long* dq, int32_t* cnt,
                                           RepOut* out = nullptr) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const RMesh& m = i < kSplit ? m0 : m1;
    if (Len(corr[i]) < kEpsC) continue;
    if (!m.Free(ids[i])) continue;
    if (out) {
      out->mask |= 1 << i;
      if (cnt) atomicAdd(cnt + ids[i], 1);
    } else {
      AddFixed(dq, cnt, ids[i], corr[i]);
    }
  }
  if (out) {
#pragma unroll
    for (int i = 0; i < 4; ++i) out->ids[i] = ids[i];
  }
}

// The inside test of EvaluateVFProximity on the current state, which rejects most pairs;
// it is evaluated before the start-of-substep side (both are pure, so the order does not
// change any result).  It is SignedVF's weights with an early exit: min(-w1, -w2, -w3) >=
// 1e-6 needs b0, b1, b2 of one strict sign (a NaN or a zero weight cannot pass either), so
// that is checked before the divisions, on the same expressions.  VfSigns is the part
// before the divisions (num receives b0, b1, b2), VfWeights the divisions and the bound.
__device__ __forceinline__ bool VfSigns(int node, int4 tr, bool self, F3 x, F3 y0, F3 y1, F3 y2,
                                        float4& num) {
  if (self && (node == tr.x || node == tr.y || node == tr.z)) return false;
  F3 n = C(Nz(S(y1, y0)), Nz(S(y2, y0)));
  if (D(n, n) < 1e-6f) return false;  // degenerate: weights 0
  n = Nz(n);
  const float b0 = Mixed(S(y1, x), S(y2, x), n);
  const float b1 = Mixed(S(y2, x), S(y0, x), n);
  const float b2 = Mixed(S(y0, x), S(y1, x), n);
  if (!((b0 > 0 && b1 > 0 && b2 > 0) || (b0 < 0 && b1 < 0 && b2 < 0))) return false;
  num = make_float4(b0, b1, b2, 0.0f);
  return true;
}

__device__ __forceinline__ bool VfWeights(float4 num) {
  const float b0 = num.x, b1 = num.y, b2 = num.z;
  const float w1 = -b0 / (b0 + b1 + b2), w2 = -b1 / (b0 + b1 + b2), w3 = -b2 / (b0 + b1 + b2);
  return fminf(-w1, fminf(-w2, -w3)) >= 1e-6f;
}

// The rest of EvaluateVFProximity (PBDRepulsion.cu:329-409) for a pair that passed the
// inside test: the start-of-substep side and the constraint, accumulated into dq / cnt
__device__ void VfResponse(const RMesh& c, int node, const RMesh& o, int4 tr, float h,
                           long long* dq, int32_t* cnt, RepOut* out) {
  float w0[4];
  F3 pn;
  const float d0 = SignedVF(c.X0(node), o.X0(tr.x), o.X0(tr.y), o.X0(tr.z), pn, w0);
  const bool face_free = o.Free(tr.x) && o.Free(tr.y) && o.Free(tr.z);
  if (d0 < 0.0f && face_free) pn = M(pn, -1.0f);
  const F3 p1 = o.X(tr.x);
  F3 n = FaceNormal(p1, o.X(tr.y), o.X(tr.z));
  const float side = D(n, pn) > 0.0f ? 1.0f : -1.0f;
  const F3 q = S(c.X(node), p1), p2 = S(o.X(tr.y), p1), p3 = S(o.X(tr.z), p1);
  n = M(C(p2, p3), side);
  const float c23 = Len(n);
  n = Nz(n);
  const float Cv = D(q, n) - h;
  if (Cv > 0) return;
  const F3 dcq = n;
  const F3 dcp2 = M(A(C(p3, q), M(C(n, p3), D(n, q))), 1.0f / c23);
  const F3 dcp3 = M(A(C(p2, q), M(C(n, p2), D(n, q))), -1.0f / c23);
  const F3 dcp1 = S(S(M(dcq, -1.0f), dcp2), dcp3);
  const float wq = c.W(node), w1 = o.W(tr.x), w2 = o.W(tr.y), w3 = o.W(tr.z);
  const float den = wq * D(dcq, dcq) + w1 * D(dcp1, dcp1) + w2 * D(dcp2, dcp2) + w3 * D(dcp3, dcp3);
  if (den < kEpsC) return;
  const float s = Cv / den;
  const F3 corr[4] = {M(dcq, -wq * s), M(dcp1, -w1 * s), M(dcp2, -w2 * s), M(dcp3, -w3 * s)};
  const int ids[4] = {node, tr.x, tr.y, tr.z};
  if (out) out->side = side;
  Accumulate<1>(c, o, ids, corr, dq, cnt, out);
}

// EvaluateVFProximity (PBDRepulsion.cu:329-409) of vertex `node` against the face with
// vertices tr (the face id itself is not needed)
__device__ void EvalVFt(const RMesh& c, int node, const RMesh& o, int4 tr, bool self, float h,
                        long long* dq, int32_t* cnt, RepOut* out = nullptr) {
  float4 num;
  if (!VfSigns(node, tr, self, c.X(node), o.X(tr.x), o.X(tr.y), o.X(tr.z), num)) return;
  if (!VfWeights(num)) return;
  VfResponse(c, node, o, tr, h, dq, cnt, out);
}

__device__ __forceinline__ void EvalVF(const RMesh& c, int node, const RMesh& o, int face,
                                       bool self, float h, long long* dq, int32_t* cnt) {
  EvalVFt(c, node, o, o.t.tris[face], self, h, dq, cnt);
}

// EvaluateEEProximity (PBDRepulsion.cu:411-468) of the edges with vertices a (cloth) and
// b (obstacle or cloth); ua, ub: their unit directions in the current state,
// Nz(x_a.y - x_a.x).  The edge ids, ids() = (e0, e1), are needed only by the pairs that
// pass the inside test.
// The inside test of EvaluateEEProximity (PBDRepulsion.cu:411-468) on the current state of
// the edges with vertices a (cloth) and b (obstacle or cloth): x0, x1, y0, y1 are their
// positions, ua, ub their unit directions Nz(x_a.y - x_a.x).  True, with SignedEE's weights
// in bw, when the barycentric part passes; InEdge and the response follow in EEResponse.
// The weights take early exits: an accepted pair has every weight > 0 (a NaN weight fails
// InEdge or leaves another weight <= 0), which needs a0, a1 and b0, b1 each of one strict
// sign, checked before the divisions on the same expressions.  EESigns is the part before
// the divisions (num receives a0, a1, b0, b1), EEWeights the divisions and the lower bound.
__device__ __forceinline__ bool EESigns(int2 a, int2 b, F3 x0, F3 x1, F3 y0, F3 y1, bool self,
                                        F3 ua, F3 ub, float4& num) {
  if (self && (a.x == b.x || a.x == b.y || a.y == b.x || a.y == b.y)) return false;
  const F3 ncur = C(ua, ub);
  if (D(ncur, ncur) < 1e-6f) return false;  // degenerate: weights 0
  const F3 d00 = S(y0, x0), d01 = S(y1, x0), d10 = S(y0, x1), d11 = S(y1, x1);
  const float a0 = Mixed(d11, d10, ncur), a1 = Mixed(d00, d01, ncur);
  if (!((a0 > 0 && a1 > 0) || (a0 < 0 && a1 < 0))) return false;
  // b0 = Mixed(S(x0, y1), S(x1, y1), ncur) and b1 = Mixed(S(x1, y0), S(x0, y0), ncur): both
  // differences negated, which leaves every nonzero result the same bits (rounding is
  // symmetric; only the sign of an exact zero can differ, and zero fails the test either way)
  const float b0 = Mixed(d01, d11, ncur), b1 = Mixed(d10, d00, ncur);
  if (!((b0 > 0 && b1 > 0) || (b0 < 0 && b1 < 0))) return false;
  num = make_float4(a0, a1, b0, b1);
  return true;
}

__device__ __forceinline__ bool EEWeights(float4 num, float4& bw) {
  bw = make_float4(num.x / (num.x + num.y), num.y / (num.x + num.y), -num.z / (num.z + num.w),
                   -num.w / (num.z + num.w));
  const float mn = fminf(fminf(bw.x, bw.y), fminf(-bw.z, -bw.w));
  return mn >= 1e-6f;
}

__device__ __forceinline__ bool EEInside(int2 a, int2 b, F3 x0, F3 x1, F3 y0, F3 y1, bool self,
                                         F3 ua, F3 ub, float4& bw) {
  float4 num;
  return EESigns(a, b, x0, x1, y0, y1, self, ua, ub, num) && EEWeights(num, bw);
}

// The rest of EvaluateEEProximity for a pair that passed EEInside with weights w: the
// InEdge tests, the start-of-substep side and the constraint, accumulated into dq / cnt
__device__ void EEResponse(const RMesh& c, int2 a, const RMesh& o, int2 b, F3 x0, F3 x1, F3 y0,
                           F3 y1, int e0, int e1, float h, float4 w4, long long* dq,
                           int32_t* cnt, RepOut* out = nullptr) {
  float bw[4] = {w4.x, w4.y, w4.z, w4.w};
  if (!(InEdge(bw[1], c, e0, o, e1) && InEdge(-bw[3], o, e1, c, e0))) return;
  float w[4];
  F3 pn;
  const float d0 = SignedEE(c.X0(a.x), c.X0(a.y), o.X0(b.x), o.X0(b.y), pn, w);
  const bool e1_free = o.Free(b.x) && o.Free(b.y);
  if (!e1_free) pn = EdgeNormal(o, e1);
  else if (d0 < 0.0f) pn = M(pn, -1.0f);
  const F3 dd = A(A(M(x0, bw[0]), M(x1, bw[1])), A(M(y0, bw[2]), M(y1, bw[3])));
  const float Cv = D(dd, pn) - h;
  if (Cv > 0) return;
  bw[2] *= -1.0f;
  bw[3] *= -1.0f;
  const F3 g[4] = {M(pn, bw[0]), M(pn, bw[1]), M(pn, -bw[2]), M(pn, -bw[3])};
  const float wq[4] = {c.W(a.x), c.W(a.y), o.W(b.x), o.W(b.y)};
  float den = 0.0f;
  for (int i = 0; i < 4; ++i) den += wq[i] * bw[i] * bw[i];
  if (den < kEpsC) return;
  const float s = Cv / den;
  F3 corr[4];
  for (int i = 0; i < 4; ++i) corr[i] = M(g[i], -s * wq[i]);
  const int ids4[4] = {a.x, a.y, b.x, b.y};
  if (out) {
    out->pn_x0 = e1_free ? 1 : 0;
    out->pn_sign = d0 < 0.0f ? -1.0f : 1.0f;
    out->pn = make_float3(pn.x, pn.y, pn.z);
  }
  Accumulate<2>(c, o, ids4, corr, dq, cnt, out);
}

__device__ __forceinline__ void EvalEE(const RMesh& c, int e0, const RMesh& o, int e1, bool self,
                                       float h, F3 ua, F3 ub, long long* dq, int32_t* cnt) {
  const int2 a = c.t.edges[e0], b = o.t.edges[e1];
  const F3 x0 = c.X(a.x), x1 = c.X(a.y), y0 = o.X(b.x), y1 = o.X(b.y);
  float4 bw;
  if (EEInside(a, b, x0, x1, y0, y1, self, ua, ub, bw))
    EEResponse(c, a, o, b, x0, x1, y0, y1, e0, e1, h, bw, dq, cnt);
}

// EvaluateRep_Kernel body (PBDRepulsion.cu:470-501) for pair p: cloth face pr.x against
// obstacle face pr.y (p < nobs) or cloth face pr.y.  The VF and EE halves are separate
// bodies so that the multi-instance path can run them as separate launches.
__device__ __forceinline__ bool PairValid(const RMesh& c, const RMesh& other, int2 pr) {
  return pr.x >= 0 && pr.x < c.t.nf && pr.y >= 0 && pr.y < other.t.nf;
}

__device__ __forceinline__ void RepulsionPairVfBody(const RMesh& c, const RMesh& o, int2 pr,
                                                    bool self, float h,
                                                    long long* __restrict__ dq,
                                                    int32_t* __restrict__ cnt) {
  const RMesh& other = self ? c : o;
  if (!PairValid(c, other, pr)) return;
  const uint32_t r0 = c.t.rtri[pr.x];
  const int4 f0 = c.t.tris[pr.x];
  const int v0[3] = {f0.x, f0.y, f0.z};
  for (int i = 0; i < 3; ++i)
    if ((r0 >> i) & 1u) EvalVF(c, v0[i], other, pr.y, self, h, dq, cnt);
}

__device__ __forceinline__ void RepulsionPairEeBody(const RMesh& c, const RMesh& o, int2 pr,
                                                    bool self, float h,
                                                    long long* __restrict__ dq,
                                                    int32_t* __restrict__ cnt) {
  const RMesh& other = self ? c : o;
  if (!PairValid(c, other, pr)) return;
  const uint32_t r0 = c.t.rtri[pr.x], r1 = other.t.rtri[pr.y];
  if (!((r0 >> 3) & 7u) || !((r1 >> 3) & 7u)) return;  // no owned edge pair
  const int4 fe0 = c.t.face_edges[pr.x], fe1 = other.t.face_edges[pr.y];
  const int e0[3] = {fe0.x, fe0.y, fe0.z}, e1[3] = {fe1.x, fe1.y, fe1.z};
  // unit directions of the owned edges, computed at most once per pair and only for edge
  // pairs that share no vertex (the same values EvalEE would compute itself)
  F3 u0[3], u1[3];
  bool h0[3] = {false, false, false}, h1[3] = {false, false, false};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (!((r0 >> (3 + i)) & 1u)) continue;
    const int2 ea = c.t.edges[e0[i]];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (!((r1 >> (3 + j)) & 1u)) continue;
      const int2 eb = other.t.edges[e1[j]];
      if (self && (ea.x == eb.x || ea.x == eb.y || ea.y == eb.x || ea.y == eb.y)) continue;
      if (!h0[i]) { u0[i] = Nz(S(c.X(ea.y), c.X(ea.x))); h0[i] = true; }
      if (!h1[j]) { u1[j] = Nz(S(other.X(eb.y), other.X(eb.x))); h1[j] = true; }
      EvalEE(c, e0[i], other, e1[j], self, h, u0[i], u1[j], dq, cnt);
    }
  }
}

__device__ __forceinline__ void RepulsionPairBody(const RMesh& c, const RMesh& o, int2 pr,
                                                  bool self, float h, long long* __restrict__ dq,
                                                  int32_t* __restrict__ cnt) {
  RepulsionPairVfBody(c, o, pr, self, h, dq, cnt);
  RepulsionPairEeBody(c, o, pr, self, h, dq, cnt);
}

// Test items of one pair as a mask: bit i (0..2) for the VF test of vertex i of face pr.x,
// bit 3 + 3i + j for the EE test of edge i of face pr.x against edge j of face pr.y.  These
// are the owned tests that RepulsionPairVfBody / RepulsionPairEeBody evaluate past their
// adjacency checks: edge l of a face joins its vertices l and l + 1 (face_edges), so the
// shared-vertex tests need only the two faces' vertices.
__device__ __forceinline__ uint32_t PairItemMask(int4 f0, uint32_t r0, int4 f1, uint32_t r1,
                                                 bool self) {
  const int v0[3] = {f0.x, f0.y, f0.z}, v1[3] = {f1.x, f1.y, f1.z};
  uint32_t m = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (((r0 >> i) & 1u) && !(self && (v0[i] == v1[0] || v0[i] == v1[1] || v0[i] == v1[2])))
      m |= 1u << i;
  if (!((r0 >> 3) & 7u) || !((r1 >> 3) & 7u)) return m;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (!((r0 >> (3 + i)) & 1u)) continue;
    const int a0 = v0[i], a1 = v0[(i + 1) % 3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (!((r1 >> (3 + j)) & 1u)) continue;
      const int b0 = v1[j], b1 = v1[(j + 1) % 3];
      if (self && (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1)) continue;
      m |= 1u << (3 + 3 * i + j);
    }
  }
  return m;
}

constexpr int kSelfBit = static_cast<int>(0x80000000u);  // item.y: cloth-vs-cloth item

// ------------------------------------------------------- single-instance kernels
__global__ __launch_bounds__(kBlock) void InterpolateKernel(const float4* __restrict__ x0,
                                                            const float4* __restrict__ x,
                                                            float4* __restrict__ x1, int n,
                                                            float alpha) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) InterpolateBody(x0, x, x1, i, alpha);
}

__global__ __launch_bounds__(kBlock) void VfDcdGenerateKernel(const float4* __restrict__ x, int n,
                                                              BvhView t, int32_t* __restrict__ face,
                                                              float4* __restrict__ qout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) VfDcdGenerateBody(x, i, t, face, qout);
}

__global__ __launch_bounds__(kBlock) void VfCcdGenerateKernel(const float4* __restrict__ x0,
                                                              const float4* __restrict__ x, int n,
                                                              BvhView t, int32_t* __restrict__ face,
                                                              float4* __restrict__ qout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) VfCcdGenerateBody(x0, x, i, t, face, qout);
}

__global__ __launch_bounds__(kBlock) void EeDcdGenerateKernel(const float4* __restrict__ x,
                                                              const int2* __restrict__ edges, int ne,
                                                              BvhView t,
                                                              const uint32_t* __restrict__ rtri,
                                                              EeRec* __restrict__ rec,
                                                              uint8_t* __restrict__ none) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < ne) EeDcdGenerateBody(x, edges, e, t, rtri, rec, none);
}

__global__ __launch_bounds__(kBlock) void VfSolveKernel(float4* __restrict__ x, int n, BvhView t,
                                                        const int32_t* __restrict__ face,
                                                        const float4* __restrict__ q, float offset,
                                                        VfFriction fr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) VfSolveBody(x, i, t, face, q, offset, fr);
}

__global__ __launch_bounds__(kBlock) void EeSolveKernel(const float4* __restrict__ x,
                                                        const int2* __restrict__ edges, int ne,
                                                        const EeRec* __restrict__ rec,
                                                        const uint8_t* __restrict__ none,
                                                        long long* __restrict__ dq,
                                                        int32_t* __restrict__ cnt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < ne) EeSolveBody(x, edges, e, rec, none, dq, cnt);
}

__global__ __launch_bounds__(kBlock) void ApplyFixedKernel(float4* __restrict__ x, int n,
                                                           long long* __restrict__ dq,
                                                           int32_t* __restrict__ cnt, float relax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ApplyFixedBody(x, i, dq, cnt, relax);
}

// a grid-stride loop over the device-side pair count
__global__ __launch_bounds__(kBlock) void RepulsionKernel(RMesh c, RMesh o,
                                                          const int2* __restrict__ pairs,
                                                          const int32_t* __restrict__ n_obs,
                                                          const int32_t* __restrict__ n_total,
                                                          int cap, float h,
                                                          long long* __restrict__ dq,
                                                          int32_t* __restrict__ cnt) {
  const int total = min(*n_total, cap);
  const int nobs = min(*n_obs, total);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x)
    RepulsionPairBody(c, o, pairs[p], p >= nobs, h, dq, cnt);
}

__global__ __launch_bounds__(kBlock) void FindPairsAtKernel(BvhView l, BvhView r, float th,
                                                            const uint32_t* __restrict__ cnt,
                                                            const uint32_t* __restrict__ off,
                                                            const int32_t* __restrict__ base,
                                                            int2* __restrict__ pairs, int cap) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= l.nf) return;
  FindPairsBody(l, r, th, id, static_cast<int64_t>(*base) + off[id], cnt[id], pairs, cap);
}

// a pair total above the capacity: raise the overflow word to the pairs per face needed
__device__ __forceinline__ void NotePairOverflow(int total, int cap, int nf, int32_t* overflow) {
  if (total > cap && overflow) atomicMax(overflow, (total + max(nf, 1) - 1) / max(nf, 1));
}

__global__ void AddTotalsKernel(const int32_t* a, const int32_t* b, int32_t* out, int cap, int nf,
                                int32_t* overflow) {
  if (threadIdx.x == 0) {
    *out = *a + *b;
    NotePairOverflow(*out, cap, nf, overflow);
  }
}

// An EE item that passed the sign part of the inside test, queued for the rest (the weights'
// divisions and lower bound) and its response: the cloth edge, the other edge (bit 31: a
// cloth edge) and the weights' numerators.  The divisions run at response time, 64 queued
// items at a time, instead of in every wave that holds one survivor.
struct EeQueued {
  int e0, e1;
  float4 num;
};

// a load through the global address space: pointers read from instance records would
// otherwise be loaded through flat addressing, which also waits on LDS traffic.  Device
// allocations have the same address in the flat and the global aperture (the integer route
// keeps the compiler from folding the cast away).
template <class T>
__device__ __forceinline__ T Gld(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return *(const __attribute__((address_space(1))) T*)(reinterpret_cast<uintptr_t>(p));
#else
  return *p;  // the host pass of this device function (never executed)
#endif
}

// the sign part of the inside test of EE item `it` (RepulsionEeItemBody up to its weights);
// ent receives the queued form when it passes
__device__ __forceinline__ bool EeItemInside(const RMesh& c, const RMesh& o, int2 it,
                                             EeQueued& ent) {
  if (it.x < 0 || it.x >= c.t.ne) return false;
  const bool self = it.y < 0;
  const int e1 = it.y & 0x7fffffff;
  if (e1 >= (self ? c.t.ne : o.t.ne)) return false;
  const int2* oe = self ? c.t.edges : o.t.edges;  // pointers, not meshes, are selected
  const float4* ox = self ? c.s.x : o.s.x;
  const int2 a = Gld(c.t.edges + it.x), b = Gld(oe + e1);
  const F3 x0 = L3(Gld(c.s.x + a.x)), x1 = L3(Gld(c.s.x + a.y)), y0 = L3(Gld(ox + b.x)),
           y1 = L3(Gld(ox + b.y));
  const F3 ua = Nz(S(x1, x0)), ub = Nz(S(y1, y0));
  ent.e0 = it.x;
  ent.e1 = it.y;
  return EESigns(a, b, x0, x1, y0, y1, self, ua, ub, ent.num);
}

// the rest of the inside test and the response of a queued EE item (positions reloaded: the
// pass does not move them)
__device__ __forceinline__ void EeItemRespond(const RMesh& c, const RMesh& o, const EeQueued& e,
                                              float h, long long* __restrict__ dq,
                                              int32_t* __restrict__ cnt, RepOut* out = nullptr) {
  float4 bw;
  if (!EEWeights(e.num, bw)) return;
  const int e1 = e.e1 & 0x7fffffff;
  const int2 a = c.t.edges[e.e0];
  if (e.e1 < 0) {
    const int2 b = c.t.edges[e1];
    EEResponse(c, a, c, b, c.X(a.x), c.X(a.y), c.X(b.x), c.X(b.y), e.e0, e1, h, bw, dq, cnt,
               out);
  } else {
    const int2 b = o.t.edges[e1];
    EEResponse(c, a, o, b, c.X(a.x), c.X(a.y), o.X(b.x), o.X(b.y), e.e0, e1, h, bw, dq, cnt,
               out);
  }
}

// A VF item that passed the sign part of the inside test, queued for the rest (the weights'
// divisions and lower bound) and its response: the cloth vertex, the face (bit 31: a cloth
// face) and the weights' numerators b0, b1, b2
struct VfQueued {
  int node, f;
  float4 num;
};

// the sign part of the inside test of VF item `it` (cloth vertex it.x against face it.y of
// the obstacle or the cloth); ent receives the queued form when it passes
__device__ __forceinline__ bool VfItemInside(const RMesh& c, const RMesh& o, int2 it,
                                             VfQueued& ent) {
  if (it.x < 0 || it.x >= c.t.nv) return false;
  const bool self = it.y < 0;
  const int f = it.y & 0x7fffffff;
  if (f >= (self ? c.t.nf : o.t.nf)) return false;
  const int4* otris = self ? c.t.tris : o.t.tris;  // pointers, not meshes, are selected
  const float4* ox = self ? c.s.x : o.s.x;
  const int4 tr = Gld(otris + f);
  ent.node = it.x;
  ent.f = it.y;
  return VfSigns(it.x, tr, self, L3(Gld(c.s.x + it.x)), L3(Gld(ox + tr.x)), L3(Gld(ox + tr.y)),
                 L3(Gld(ox + tr.z)), ent.num);
}

// the rest of the inside test and the response of a queued VF item (the two meshes are
// separate inlined calls: selecting one of two structs by reference would keep both in
// scratch memory)
__device__ __forceinline__ void VfItemRespond(const RMesh& c, const RMesh& o, const VfQueued& e,
                                              float h, long long* __restrict__ dq,
                                              int32_t* __restrict__ cnt) {
  if (!VfWeights(e.num)) return;
  const int f = e.f & 0x7fffffff;
  if (e.f < 0) VfResponse(c, e.node, c, c.t.tris[f], h, dq, cnt, nullptr);
  else VfResponse(c, e.node, o, o.t.tris[f], h, dq, cnt, nullptr);
}

// One wave's share of an item list: 64 inside tests per step (item q of the wave's
// stride), the survivors queued in LDS (qw: 128 entries); every 64 queued (and at the
// end) all lanes run responses together, so the rare response path does not idle most of
// a wave.  Same tests on the same values; the sums are fixed point, so the order of the
// responses changes nothing.
template <class Q, class Inside, class Respond>
__device__ __forceinline__ void QueuedItems(int cnt, int base, int stride, Q* qw, Inside&& inside,
                                            Respond&& respond) {
  const int lane = threadIdx.x & 63;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int qn = 0;  // queued entries (the same in every lane)
  while (true) {
    const bool have = base < cnt;  // uniform over the wave
    if (have) {
      const int qi = base + lane;
      Q ent;
      const bool pass = qi < cnt && inside(qi, ent);
      const uint64_t m = __ballot(pass);
      if (pass) qw[qn + __popcll(m & below)] = ent;
      qn += __popcll(m);
      base += stride;
      __builtin_amdgcn_wave_barrier();
    }
    if (qn >= 64 || (!have && qn > 0)) {
      const int take = qn < 64 ? qn : 64;
      if (lane < take) respond(qw[qn - take + lane]);
      qn -= take;
      __builtin_amdgcn_wave_barrier();
    }
    if (!have && qn == 0) break;
  }
}

// ------------------------------------------------ adjoint replay of the PBD repulsion
// Repulsion item `it` (VF: is_ee = 0, EE: 1) of instance c on the state (x, x0): the forward's
// tests and response, with the response's corners and decisions in `out` instead of sums;
// cnt (indexed like x) receives the Jacobi counts when given.
__device__ __forceinline__ void ReplayRepItem(const InstContactDev& c, const float4* x,
                                              const float4* x0, int is_ee, int2 it, float h,
                                              int32_t* cnt, RepOut& out) {
  const RMesh cm{c.cloth, MeshState{x + c.vbeg, x0 + c.vbeg, -1.0f}};
  const RMesh om{c.obs, MeshState{c.o_x, c.o_x0, 1e-3f}};  // obstacle nodes: inv mass 1e-3
  int32_t* cn = cnt ? cnt + c.vbeg : nullptr;
  if (!is_ee) {
    const int f = it.y & 0x7fffffff;
    if (it.x < 0 || it.x >= c.cloth.nv) return;
    if (it.y < 0) {
      if (f < c.cloth.nf) EvalVFt(cm, it.x, cm, c.cloth.tris[f], true, h, nullptr, cn, &out);
    } else if (c.has_obs && f < c.obs.nf) {
      EvalVFt(cm, it.x, om, c.obs.tris[f], false, h, nullptr, cn, &out);
    }
    return;
  }
  if (it.y >= 0 && !c.has_obs) return;
  EeQueued ent;
  if (EeItemInside(cm, om, it, ent)) EeItemRespond(cm, om, ent, h, nullptr, cn, &out);
}

// dual numbers (value, one tangent) in FP64 for the responses' Jacobians
struct Dd {
  double v, d;
};
__device__ __forceinline__ Dd DK(double v) { return Dd{v, 0.0}; }
__device__ __forceinline__ Dd operator+(Dd a, Dd b) { return Dd{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dd operator-(Dd a, Dd b) { return Dd{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dd operator*(Dd a, Dd b) { return Dd{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dd operator/(Dd a, Dd b) {
  return Dd{a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)};
}
__device__ __forceinline__ Dd DSqrt(Dd a) {
  const double s = sqrt(a.v);
  return Dd{s, a.v > 0.0 ? a.d / (2.0 * s) : 0.0};
}
struct D3 {
  Dd x, y, z;
};
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator*(D3 a, Dd s) { return D3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ Dd Dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 Crs3(D3 a, D3 b) {
  return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ D3 Nrm3(D3 a) { return a * (DK(1.0) / DSqrt(Dot3(a, a))); }
__device__ __forceinline__ D3 In3(const double* x, int k, int dir) {
  return D3{Dd{x[k], dir == k ? 1.0 : 0.0}, Dd{x[k + 1], dir == k + 1 ? 1.0 : 0.0},
            Dd{x[k + 2], dir == k + 2 ? 1.0 : 0.0}};
}

// VF response (EvalVFt after its tests) on in = node, y0, y1, y2 (12 values), tangent
// direction dir: the four corrections' tangents dot the output adjoint yb
__device__ double VfRespDir(const double* in, int dir, const RepOut& r, const double* w, double h,
                            const double* yb) {
  const D3 p1 = In3(in, 3, dir);
  const D3 q = In3(in, 0, dir) - p1, p2 = In3(in, 6, dir) - p1, p3 = In3(in, 9, dir) - p1;
  D3 n = Crs3(p2, p3) * DK(r.side);
  const Dd c23 = DSqrt(Dot3(n, n));
  n = n * (DK(1.0) / c23);
  const Dd C = Dot3(q, n) - DK(h);
  const Dd nq = Dot3(n, q);
  D3 dc[4];
  dc[0] = n;
  dc[2] = (Crs3(p3, q) + Crs3(n, p3) * nq) * (DK(1.0) / c23);
  dc[3] = (Crs3(p2, q) + Crs3(n, p2) * nq) * (DK(-1.0) / c23);
  dc[1] = dc[0] * DK(-1.0) - dc[2] - dc[3];
  Dd den = DK(0.0);
  for (int i = 0; i < 4; ++i) den = den + DK(w[i]) * Dot3(dc[i], dc[i]);
  const Dd s = C / den;
  double acc = 0.0;
  for (int i = 0; i < 4; ++i) {
    const D3 corr = dc[i] * (DK(-w[i]) * s);
    acc += corr.x.d * yb[3 * i] + corr.y.d * yb[3 * i + 1] + corr.z.d * yb[3 * i + 2];
  }
  return acc;
}

__device__ __forceinline__ Dd Mixed3(D3 a, D3 b, D3 c) { return Dot3(a, Crs3(b, c)); }

// EE response (EEResponse after its tests) on in = a0, a1, b0, b1 (current) then the same at
// the start of the substep (24 values)
__device__ double EeRespDir(const double* in, int dir, const RepOut& r, const double* w, double h,
                            const double* yb) {
  const D3 x0 = In3(in, 0, dir), x1 = In3(in, 3, dir), y0 = In3(in, 6, dir), y1 = In3(in, 9, dir);
  D3 pn;
  if (r.pn_x0)
    pn = Crs3(Nrm3(In3(in, 15, dir) - In3(in, 12, dir)), Nrm3(In3(in, 21, dir) - In3(in, 18, dir))) *
         DK(r.pn_sign);
  else
    pn = D3{DK(r.pn.x), DK(r.pn.y), DK(r.pn.z)};
  const D3 n = Crs3(Nrm3(x1 - x0), Nrm3(y1 - y0));
  const Dd a0 = Mixed3(y1 - x1, y0 - x1, n), a1 = Mixed3(y0 - x0, y1 - x0, n);
  const Dd b0 = Mixed3(x0 - y1, x1 - y1, n), b1 = Mixed3(x1 - y0, x0 - y0, n);
  const Dd wt[4] = {a0 / (a0 + a1), a1 / (a0 + a1), DK(0.0) - b0 / (b0 + b1),
                    DK(0.0) - b1 / (b0 + b1)};
  const D3 dd = x0 * wt[0] + x1 * wt[1] + y0 * wt[2] + y1 * wt[3];
  const Dd C = Dot3(dd, pn) - DK(h);
  Dd den = DK(0.0);
  for (int i = 0; i < 4; ++i) den = den + DK(w[i]) * wt[i] * wt[i];
  const Dd s = C / den;
  double acc = 0.0;
  for (int i = 0; i < 4; ++i) {
    const D3 corr = pn * (wt[i] * (DK(-w[i]) * s));
    acc += corr.x.d * yb[3 * i] + corr.y.d * yb[3 * i + 1] + corr.z.d * yb[3 * i + 2];
  }
  return acc;
}

constexpr double kAdjFix = 4294967296.0;  // 2^32: adjoint sums in 32.32 fixed point

__device__ __forceinline__ void AddFix3(unsigned long long* p, double x, double y, double z) {
  atomicAdd(p, static_cast<unsigned long long>(__double2ll_rn(x * kAdjFix)));
  atomicAdd(p + 1, static_cast<unsigned long long>(__double2ll_rn(y * kAdjFix)));
  atomicAdd(p + 2, static_cast<unsigned long long>(__double2ll_rn(z * kAdjFix)));
}

// -------------------------------------------------------- multi-instance kernels

#define XRT_MULTI_INDEX(kind, count_expr)                                  \
  const int lblk = XcdBlock(off, n);                                       \
  if (lblk < 0) return;                                                    \
  const int k = InstOfBlock(off, n, lblk);                                 \
  const InstContactDev& c = ic[k];                                         \
  const int i = (lblk - off[k]) * kBlock + threadIdx.x;                    \
  if (i >= (count_expr)) return;

__global__ __launch_bounds__(kBlock) void MultiInterpolateKernel(const InstContactDev* ic,
                                                                 const int32_t* off, int n,
                                                                 float alpha) {
  XRT_MULTI_INDEX(kObsVert, c.obs_nv)
  InterpolateBody(c.o_x0, c.o_x, c.o_x1, i, alpha);
}

// midstep normals and the contact points pred = x1 + tol n1
__global__ __launch_bounds__(kBlock) void MultiNormalPredKernel(const InstContactDev* ic,
                                                                const int32_t* off, int n,
                                                                float tol) {
  XRT_MULTI_INDEX(kObsVert, c.obs_nv)
  c.o_n1[i] = VertexNormalBody(c.o_x1, c.obs.tris, c.o_vptr, c.o_corner, i);
  RefitPredBody(c.o_x1, c.o_n1, tol, i, c.obvh.pred);
}

__global__ __launch_bounds__(kBlock) void MultiObsGrowKernel(const InstContactDev* ic,
                                                             const int32_t* off, int n) {
  XRT_MULTI_INDEX(kObsFace, c.obs_nf)
  GrowLeafBody(c.obvh.nodes, c.obvh.tris, c.obvh.pred, nullptr, c.obs_nf, i, c.obvh.box,
               c.obvh.flags);
}

// one workgroup per instance merges its obstacle tree
__global__ __launch_bounds__(1024) void MultiObsBottomUpKernel(const InstContactDev* ic) {
  const InstContactDev& c = ic[blockIdx.x];
  if (c.has_obs && c.obs_nf > 1)
    LevelBottomUpWG(c.obvh.nodes, c.obvh.box, c.obvh.lvl, c.obvh.flags, c.obs_nf, 0);
}

__global__ __launch_bounds__(kBlock) void MultiVfGenerateKernel(const InstContactDev* ic,
                                                                const int32_t* off, int n,
                                                                const float4* __restrict__ x0,
                                                                const float4* __restrict__ x) {
  XRT_MULTI_INDEX(kVert, c.nv)
  if (!c.has_obs) return;
  VfDcdGenerateBody<true>(x + c.vbeg, i, c.obvh, c.vf_face, c.vf_q);
  VfCcdGenerateBody<true>(x0 + c.vbeg, x + c.vbeg, i, c.obvh, c.rt_face, c.rt_q);
}

__global__ __launch_bounds__(kBlock) void MultiEeGenerateKernel(const InstContactDev* ic,
                                                                const int32_t* off, int n,
                                                                const float4* __restrict__ x) {
  XRT_MULTI_INDEX(kEdge, c.ne)
  if (!c.has_obs) return;
  EeDcdGenerateBody<true>(x + c.vbeg, c.cloth.edges, i, c.obvh, c.obs.rtri, c.ee, c.ee_none);
}

// cloth LBVH Update(ccd): pred = x, then leaves grow by bounds(x, x0)
__global__ __launch_bounds__(kBlock) void MultiClothPredKernel(const InstContactDev* ic,
                                                               const int32_t* off, int n,
                                                               const float4* __restrict__ x) {
  XRT_MULTI_INDEX(kVert, c.nv)
  RefitPredBody(x + c.vbeg, nullptr, 0.0f, i, c.cbvh.pred);
}

__global__ __launch_bounds__(kBlock) void MultiClothGrowKernel(const InstContactDev* ic,
                                                               const int32_t* off, int n,
                                                               const float4* __restrict__ x,
                                                               const float4* __restrict__ x0) {
  XRT_MULTI_INDEX(kFace, c.nf)
  GrowLeafBody(c.cbvh.nodes, c.cbvh.tris, x + c.vbeg, x0 + c.vbeg, c.nf, i, c.cbvh.box,
               c.cbvh.flags);
}

// one workgroup per instance merges its cloth tree (grow-only)
__global__ __launch_bounds__(1024) void MultiClothBottomUpKernel(const InstContactDev* ic) {
  const InstContactDev& c = ic[blockIdx.x];
  if (c.scal && c.nf > 1)
    LevelBottomUpWG(c.cbvh.nodes, c.cbvh.box, c.cbvh.lvl, c.cbvh.flags, c.nf, 1);
}

// frame start: one workgroup per tree merges the freshly built trees of every instance
// (workgroups [0, n): cloth trees when `cloth`, [n, 2n): obstacle trees)
__global__ __launch_bounds__(1024) void MultiBuildMergeKernel(const InstContactDev* ic, int n,
                                                              int cloth) {
  const bool obs = blockIdx.x >= static_cast<unsigned>(n);
  const InstContactDev& c = ic[obs ? blockIdx.x - n : blockIdx.x];
  if (!c.scal) return;
  if (obs) {
    if (c.has_obs && c.obs_nf > 1)
      LevelBottomUpWG(c.obvh.nodes, c.obvh.box, c.obvh.lvl, c.obvh.flags, c.obs_nf, 0);
  } else if (cloth && c.nf > 1) {
    LevelBottomUpWG(c.cbvh.nodes, c.cbvh.box, c.cbvh.lvl, c.cbvh.flags, c.nf, 0);
  }
}


// traversal records of every cloth tree (thread i: leaf nf - 1 + i and internal node i)
__global__ __launch_bounds__(kBlock) void MultiPackClothKernel(const InstContactDev* ic,
                                                               const int32_t* off, int n) {
  XRT_MULTI_INDEX(kFace, c.nf)
  PackNodeBody(c.cbvh, c.nf - 1 + i);
  if (i < c.nf - 1) PackNodeBody(c.cbvh, i);
  c.cbvh.leaf_face[i] = c.cloth.trir[c.cbvh.nodes[c.nf - 1 + i].w];
}

// traversal records of every obstacle tree
__global__ __launch_bounds__(kBlock) void MultiPackObsKernel(const InstContactDev* ic,
                                                             const int32_t* off, int n) {
  XRT_MULTI_INDEX(kObsFace, c.obs_nf)
  if (!c.has_obs) return;
  PackNodeBody(c.obvh, c.obs_nf - 1 + i);
  if (i < c.obs_nf - 1) PackNodeBody(c.obvh, i);
  c.obvh.leaf_face[i] = c.obs.trir[c.obvh.nodes[c.obs_nf - 1 + i].w];
}

// pairs per cloth leaf against the obstacle (self = 0) or the cloth itself (self = 1),
// traversing the packed records, with the repulsion item counts of those pairs when the
// item lists are on; the first kPairStage hits of leaf i are kept at stage[j * nf + i] as
// rhs object | item mask << kStageObjBits (when the instance has a stage)
// exclusive scan of three counters over the 256 threads of a block; tot = the block sums
__device__ __forceinline__ void BlockScan3(uint32_t v[3], uint32_t tot[3]) {
  __shared__ uint32_t wsum[kBlock / 64][3];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc[3] = {v[0], v[1], v[2]};
#pragma unroll
  for (int d = 1; d < 64; d <<= 1)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const uint32_t t = __shfl_up(inc[q], d, 64);
      if (lane >= d) inc[q] += t;
    }
  if (lane == 63)
    for (int q = 0; q < 3; ++q) wsum[w][q] = inc[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    uint32_t before = 0, all = 0;
    for (int j = 0; j < kBlock / 64; ++j) {
      if (j < w) before += wsum[j][q];
      all += wsum[j][q];
    }
    v[q] = before + inc[q] - v[q];
    tot[q] = all;
  }
}

__global__ __launch_bounds__(kBlock) void MultiCountPairsKernel(const InstContactDev* ic,
                                                                const int32_t* off, int n,
                                                                float th, int self) {
  const int lblk = XcdBlock(off, n);
  if (lblk < 0) return;
  const int k = InstOfBlock(off, n, lblk);
  const InstContactDev& c = ic[k];
  const int lb = lblk - off[k];
  const int i = lb * kBlock + threadIdx.x;  // every thread reaches the block scan
  uint32_t v = 0, nvf = 0, nee = 0;
  if (i < c.nf && (self || c.has_obs)) {
    const Box lb = c.cbvh.packed[c.nf - 1 + i];
    const int lobj = -__float_as_int(lb.lo.w) - 1;
    const bool items = c.ee_cap > 0;
    int4 f0 = make_int4(0, 0, 0, 0);
    if (items) f0 = c.cloth.trir[lobj];
    const uint32_t r0 = static_cast<uint32_t>(f0.w);
    uint32_t* st = c.stage + i;
    const int stride = c.nf;
    TraversePackedFace(self ? c.cbvh.packed : c.obvh.packed,
                       self ? c.cbvh.leaf_face : c.obvh.leaf_face, self ? c.nf : c.obs_nf, lb, th,
                       [&](int robj, int4 f1) {
                     uint32_t m = 0;
                     if (items) {
                       m = PairItemMask(f0, r0, f1, static_cast<uint32_t>(f1.w), self != 0);
                       nvf += __popc(m & 7u);
                       nee += __popc(m >> 3);
                     }
                     if (c.stage && v < static_cast<uint32_t>(kPairStage))
                       st[static_cast<int64_t>(v) * stride] =
                           static_cast<uint32_t>(robj) | (m << kStageObjBits);
                     ++v;
                   });
  }
  // block-local exclusive prefixes per leaf, block sums for MultiScanKernel
  uint32_t e[3] = {v, nvf, nee}, tot[3];
  BlockScan3(e, tot);
  if (i < c.nf) {
    c.cbvh.counts[i] = v;
    c.cbvh.offsets[i] = e[0];
    if (c.ee_cap > 0) {
      c.ivf[i] = e[1];
      c.iee[i] = e[2];
      c.nitem[i] = make_uint2(nvf, nee);
    }
  }
  if (threadIdx.x == 0) {
    const int nb = (c.nf + kBlock - 1) / kBlock;
    c.bsum[lb] = tot[0];
    c.bsum[nb + lb] = tot[1];
    c.bsum[2 * nb + lb] = tot[2];
  }
}

// one workgroup per instance: exclusive scan of the count pass's block sums (the per-leaf
// offsets are these plus the block-local prefixes); the pair sum goes to scal[1] (obstacle
// pass) or to the tree's total (cloth pass), the item sums to scal[5], scal[6] (obstacle
// pass) or scal[7], scal[8]
__global__ __launch_bounds__(1024) void MultiScanKernel(const InstContactDev* ic, int self) {
  const InstContactDev& c = ic[blockIdx.x];
  if (c.nf <= 0 || !c.scal) return;
  const int nb = (c.nf + kBlock - 1) / kBlock;
  BlockExclusiveScan(c.bsum, nb, self ? c.cbvh.total : c.scal + 1);
  __syncthreads();  // the scans share their workgroup buffer
  BlockExclusiveScan(c.bsum + nb, nb, c.scal + (self ? 7 : 5));
  __syncthreads();
  BlockExclusiveScan(c.bsum + 2 * nb, nb, c.scal + (self ? 8 : 6));
}

// the item lists of instance c are complete and used for VF (is_ee = 0) or EE tests: lists
// on, every pair stored and every item within the list (valid once the totals of both
// traversal passes are in scal)
__device__ __forceinline__ bool ItemsUsable(const InstContactDev& c, int is_ee) {
  return c.ee_cap > 0 && c.scal[2] <= c.pair_cap &&
         c.scal[is_ee ? 4 : 3] <= (is_ee ? c.ee_cap : c.vf_cap);
}

// wave-wide exclusive prefix sum over the 64 lanes (all active); *total: the wave's sum
__device__ __forceinline__ uint32_t WaveExclusiveSum(uint32_t v, uint32_t* total) {
  const int lane = threadIdx.x & 63;
  uint32_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(inc, d, 64);
    if (lane >= d) inc += t;
  }
  *total = __shfl(inc, 63, 64);
  return inc - v;
}

__device__ __forceinline__ uint32_t WaveMax(uint32_t v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = max(v, static_cast<uint32_t>(__shfl_xor(v, d, 64)));
  return v;
}

__device__ __forceinline__ int Pick3(int4 v, int q) { return q == 0 ? v.x : (q == 1 ? v.y : v.z); }

// FindPairs of every instance's leaves (pairs of leaf i fill [lo, lo + n) backwards, slots
// below the capacity only): copied from the count pass's stage when the leaf had at most
// kPairStage hits, else found by a second traversal.  With the item lists on, the items of
// the pairs go to the lists, the cloth-pair items after all obstacle-pair items; only
// slots below the list capacities are written.  The cloth-pair pass runs after the totals:
// when both lists will be used, nothing reads the cloth pairs and they are not stored.
// Item layout: the 64 leaves of a wave own the contiguous range of their item counts from
// the first leaf's offset.  The staged leaves' items fill its front in (hit j, lane, bit)
// order: per hit index the wave takes an exclusive sum of its lanes' item counts, so each
// store instruction writes neighbouring slots instead of 64 scattered ones (one thread per
// leaf writing its own run took 2.0 ms per cloth pass, 0.9 ms without those stores).  The
// traversed leaves' items follow, each leaf's in its hit order.  The lists hold the same
// items as the reference's per-pair tests; the repulsion sums are fixed point, so the order
// of the items cannot change a result bit.
__global__ __launch_bounds__(kBlock) void MultiEmitPairsKernel(const InstContactDev* ic,
                                                               const int32_t* off, int n, float th,
                                                               int self) {
  const int lblk = XcdBlock(off, n);
  if (lblk < 0) return;
  const int k = InstOfBlock(off, n, lblk);
  const InstContactDev& c = ic[k];
  if (!self && !c.has_obs) return;
  const int blk = lblk - off[k];
  const int i = blk * kBlock + static_cast<int>(threadIdx.x);  // every lane stays for the
  const int i0 = i - static_cast<int>(threadIdx.x & 63);        // wave sums; i0: lane 0's leaf
  if (i0 >= c.nf) return;                                         // (uniform over the wave)
  const bool leaf = i < c.nf;
  const bool store = !self || !(ItemsUsable(c, 0) && ItemsUsable(c, 1));
  const int64_t cap = store ? c.pair_cap : 0;
  const int nb = (c.nf + kBlock - 1) / kBlock;
  const bool items = c.ee_cap > 0;
  const MeshTopo& ot = self ? c.cloth : c.obs;
  const int tag = self ? kSelfBit : 0;
  uint32_t cnt = 0;
  int64_t lo = 0;
  int lobj = 0;
  int4 f0 = make_int4(0, 0, 0, 0), fe0 = make_int4(0, 0, 0, 0);
  uint32_t r0 = 0;
  uint2 nit = make_uint2(0, 0);
  if (leaf) {
    cnt = c.cbvh.counts[i];
    lo = static_cast<int64_t>(c.scal[self ? 1 : 0]) + c.bsum[blk] + c.cbvh.offsets[i];
    lobj = -__float_as_int(c.cbvh.packed[c.nf - 1 + i].lo.w) - 1;
    if (items) {
      f0 = c.cloth.tris[lobj];
      fe0 = c.cloth.face_edges[lobj];
      r0 = c.cloth.rtri[lobj];
      nit = c.nitem[i];
    }
  }
  const bool trav = leaf && (!c.stage || cnt > static_cast<uint32_t>(kPairStage));
  const bool staged = leaf && !trav;
  // the wave's item ranges: staged leaves' items first, then the traversed leaves'
  int64_t pv = 0, pe = 0, tv = 0, te = 0;
  if (items) {
    pv = (self ? static_cast<int64_t>(c.scal[5]) : 0) + c.bsum[nb + blk] + c.ivf[i0];
    pe = (self ? static_cast<int64_t>(c.scal[6]) : 0) + c.bsum[2 * nb + blk] + c.iee[i0];
    uint32_t sv, se, unused;
    WaveExclusiveSum(staged ? nit.x : 0u, &sv);
    WaveExclusiveSum(staged ? nit.y : 0u, &se);
    tv = pv + sv + WaveExclusiveSum(trav ? nit.x : 0u, &unused);
    te = pe + se + WaveExclusiveSum(trav ? nit.y : 0u, &unused);
  }
  const uint32_t jn = WaveMax(staged ? cnt : 0u);
  for (uint32_t j = 0; j < jn; ++j) {
    const bool have = staged && j < cnt;
    uint32_t se = 0;
    if (have) se = c.stage[static_cast<int64_t>(j) * c.nf + i];
    const int robj = static_cast<int>(se & ((1u << kStageObjBits) - 1u));
    if (have) {
      const int64_t idx = lo + cnt - 1 - j;
      if (idx < cap) c.pairs[idx] = make_int2(lobj, robj);
    }
    if (!items) continue;
    const uint32_t m = se >> kStageObjBits;
    uint32_t tot;
    const uint32_t pre = WaveExclusiveSum(__popc(m & 7u) | (__popc(m >> 3) << 16), &tot);
    int64_t qv = pv + (pre & 0xffffu), qe = pe + (pre >> 16);
    pv += tot & 0xffffu;
    pe += tot >> 16;
    if (!m) continue;
#pragma unroll
    for (int q = 0; q < 3; ++q)
      if ((m >> q) & 1u) {
        if (qv < c.vf_cap) c.rep_vf[qv] = make_int2(Pick3(f0, q), robj | tag);
        ++qv;
      }
    if (!(m >> 3)) continue;
    const int4 fe1 = ot.face_edges[robj];
#pragma unroll
    for (int q = 0; q < 9; ++q)
      if ((m >> (3 + q)) & 1u) {
        if (qe < c.ee_cap) c.rep_ee[qe] = make_int2(Pick3(fe0, q / 3), Pick3(fe1, q % 3) | tag);
        ++qe;
      }
  }
  if (!trav) return;
  // a leaf with more hits than the stage holds (or no stage: self contact came on after
  // start, or a mesh too large for the stage's object field) traverses again, in the count
  // pass's hit order
  int64_t index = lo + cnt - 1;
  const int v0[3] = {f0.x, f0.y, f0.z};
  TraversePacked(self ? c.cbvh.packed : c.obvh.packed, self ? c.nf : c.obs_nf,
                 c.cbvh.packed[c.nf - 1 + i], th, [&](int robj) {
                   if (index >= lo && index < cap) c.pairs[index] = make_int2(lobj, robj);
                   --index;
                   if (!items) return;
                   const int4 f1 = ot.trir[robj];
                   const uint32_t m =
                       PairItemMask(f0, r0, f1, static_cast<uint32_t>(f1.w), self != 0);
#pragma unroll
                   for (int q = 0; q < 3; ++q)
                     if ((m >> q) & 1u) {
                       if (tv < c.vf_cap) c.rep_vf[tv] = make_int2(v0[q], robj | tag);
                       ++tv;
                     }
                   if (!(m >> 3)) return;
                   const int4 fe1 = ot.face_edges[robj];
#pragma unroll
                   for (int q = 0; q < 9; ++q)
                     if ((m >> (3 + q)) & 1u) {
                       if (te < c.ee_cap)
                         c.rep_ee[te] = make_int2(Pick3(fe0, q / 3), Pick3(fe1, q % 3) | tag);
                       ++te;
                     }
                 });
}

__global__ __launch_bounds__(kBlock) void MultiAddTotalsKernel(const InstContactDev* ic, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || !ic[k].scal) return;
  ic[k].scal[2] = ic[k].scal[1] + *ic[k].cbvh.total;
  NotePairOverflow(ic[k].scal[2], ic[k].pair_cap, ic[k].nf, ic[k].overflow);
  if (ic[k].ee_cap > 0) {
    ic[k].scal[3] = ic[k].scal[5] + ic[k].scal[7];
    ic[k].scal[4] = ic[k].scal[6] + ic[k].scal[8];
  }
}


// repulsion tests of every instance, one VF (is_ee = 0) or EE item per thread; instances
// whose list is not usable (off, overflowed, or pairs beyond the pair capacity) are left to
// MultiRepPairsKernel.  Eight waves per SIMD (64 VGPRs; the spills fall in the response
// path): the tests wait on dependent gathers (item -> edges -> positions), and the EE launch
// measured 1.70 ms at 86 VGPRs (5 waves), 1.50 at 80, 1.37 at 72, 1.31 ms at 64.
template <int is_ee>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void
MultiRepItemKernel(const InstContactDev* ic, const int32_t* off, int n,
                   const float4* __restrict__ x, const float4* __restrict__ x0, float h) {
  const int lblk = XcdBlock(off, n);
  if (lblk < 0) return;
  const int k = InstOfBlock(off, n, lblk);
  const InstContactDev& c = ic[k];
  if (!ItemsUsable(c, is_ee)) return;
  const int cnt = c.scal[is_ee ? 4 : 3];
  const int nblk = off[k + 1] - off[k];
  const RMesh cm{c.cloth, MeshState{x + c.vbeg, x0 + c.vbeg, -1.0f}};
  const RMesh om{c.obs, MeshState{c.o_x, c.o_x0, 1e-3f}};  // obstacle nodes: inv mass 1e-3
  const int2* items = is_ee ? c.rep_ee : c.rep_vf;
  const int w = threadIdx.x >> 6;
  const int base0 = (lblk - off[k]) * kBlock + w * 64, stride = nblk * kBlock;
  if (!is_ee) {
    __shared__ VfQueued vq[kBlock / 64][128];
    QueuedItems(cnt, base0, stride, vq[w],
                [&](int qi, VfQueued& e) { return VfItemInside(cm, om, Gld(items + qi), e); },
                [&](const VfQueued& e) { VfItemRespond(cm, om, e, h, c.dq, c.cnt); });
    return;
  }
  // EE: the same queue, written out (the shared helper compiled to a 13 % slower EE launch)
  __shared__ EeQueued queue[kBlock / 64][128];
  const int lane = threadIdx.x & 63;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  EeQueued* qw = queue[w];
  int qn = 0;  // queued entries (the same in every lane)
  int base = base0;
  while (true) {
    const bool have = base < cnt;  // uniform over the wave
    if (have) {
      const int qi = base + lane;
      EeQueued ent;
      const bool pass = qi < cnt && EeItemInside(cm, om, Gld(items + qi), ent);
      const uint64_t m = __ballot(pass);
      if (pass) qw[qn + __popcll(m & below)] = ent;
      qn += __popcll(m);
      base += stride;
      __builtin_amdgcn_wave_barrier();
    }
    if (qn >= 64 || (!have && qn > 0)) {
      const int take = qn < 64 ? qn : 64;
      if (lane < take) EeItemRespond(cm, om, qw[qn - take + lane], h, c.dq, c.cnt);
      qn -= take;
      __builtin_amdgcn_wave_barrier();
    }
    if (!have && qn == 0) break;
  }
}

// Adjoint tape of the repulsion: the items of every instance that produced a response on
// the state (x, x0), appended to pool as (instance, is_ee, item) at *count (cap entries;
// flags bit 0: the pool overflowed; bit 1: an instance evaluated its pairs without item
// lists, which the reverse cannot replay)
template <int is_ee>
__global__ __launch_bounds__(kBlock) void MultiRepRecordKernel(
    const InstContactDev* ic, const int32_t* off, int n, const float4* __restrict__ x,
    const float4* __restrict__ x0, float h, int4* __restrict__ pool,
    unsigned long long* __restrict__ count, unsigned long long cap, int32_t* __restrict__ flags) {
  const int lblk = XcdBlock(off, n);
  if (lblk < 0) return;
  const int k = InstOfBlock(off, n, lblk);
  const InstContactDev& c = ic[k];
  if (!c.scal) return;
  if (!ItemsUsable(c, is_ee)) {
    if (threadIdx.x == 0 && lblk == off[k] && c.scal[2] > 0) atomicOr(flags, 2);
    return;
  }
  const int total = c.scal[is_ee ? 4 : 3];
  const int nblk = off[k + 1] - off[k];
  const int2* items = is_ee ? c.rep_ee : c.rep_vf;
  for (int q = (lblk - off[k]) * kBlock + threadIdx.x; q < total; q += nblk * kBlock) {
    const int2 it = items[q];
    RepOut out;
    ReplayRepItem(c, x, x0, is_ee, it, h, nullptr, out);
    if (!out.mask) continue;
    const unsigned long long slot = atomicAdd(count, 1ull);
    if (slot < cap)
      pool[slot] = make_int4(k, is_ee, it.x, it.y);
    else
      atomicOr(flags, 1);
  }
}

__global__ void RepMarkKernel(const unsigned long long* __restrict__ count,
                              unsigned long long* __restrict__ mark) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *mark = *count;
}

// reverse, step 1: the Jacobi counts of the recorded responses in [range[0], range[1])
__global__ __launch_bounds__(kBlock) void RepAdjCountKernel(
    const InstContactDev* ic, const int4* __restrict__ pool,
    const unsigned long long* __restrict__ range, const float4* __restrict__ x,
    const float4* __restrict__ x0, float h, unsigned long long cap, int32_t* __restrict__ cnt) {
  const unsigned long long b = range[0], e = min(range[1], cap);
  for (unsigned long long q = b + blockIdx.x * kBlock + threadIdx.x; q < e;
       q += static_cast<unsigned long long>(gridDim.x) * kBlock) {
    const int4 r = pool[q];
    RepOut out;
    ReplayRepItem(ic[r.x], x, x0, r.y, make_int2(r.z, r.w), h, cnt, out);
  }
}

// reverse, step 2: each response's corrections reversed (Jacobian by dual numbers, one
// input direction at a time) with output adjoint relax / count * gx at its accumulated
// corners; cloth inputs add to add_x, the start-of-substep positions of EE normals to
// add_x0 (32.32 fixed point, so the sums do not depend on their order)
__global__ __launch_bounds__(kBlock) void RepAdjVjpKernel(
    const InstContactDev* ic, const int4* __restrict__ pool,
    const unsigned long long* __restrict__ range, const float4* __restrict__ x,
    const float4* __restrict__ x0, float h, float relax, unsigned long long cap,
    const int32_t* __restrict__ cnt, const double4* __restrict__ gx,
    unsigned long long* __restrict__ add_x, unsigned long long* __restrict__ add_x0) {
  const unsigned long long b = range[0], e = min(range[1], cap);
  for (unsigned long long q = b + blockIdx.x * kBlock + threadIdx.x; q < e;
       q += static_cast<unsigned long long>(gridDim.x) * kBlock) {
    const int4 r = pool[q];
    const InstContactDev& c = ic[r.x];
    RepOut out;
    ReplayRepItem(c, x, x0, r.y, make_int2(r.z, r.w), h, nullptr, out);
    if (!out.mask) continue;
    const bool self = r.w < 0;
    // corner i on the cloth: VF node / EE edge a always; the other side when self contact
    bool cl[4];
    cl[0] = true;
    cl[1] = r.y == 1 || self;
    cl[2] = cl[3] = self;
    double in[24], w[4], yb[12];
    for (int i = 0; i < 4; ++i) {
      const int id = out.ids[i];
      const float4 p = cl[i] ? x[c.vbeg + id] : c.o_x[id];
      const float4 p0 = cl[i] ? x0[c.vbeg + id] : c.o_x0[id];
      in[3 * i] = p.x; in[3 * i + 1] = p.y; in[3 * i + 2] = p.z;
      in[12 + 3 * i] = p0.x; in[12 + 3 * i + 1] = p0.y; in[12 + 3 * i + 2] = p0.z;
      w[i] = cl[i] ? static_cast<double>(p0.w) : static_cast<double>(1e-3f);
      const bool acc = (out.mask >> i) & 1;
      const double f = acc ? static_cast<double>(relax) / cnt[c.vbeg + id] : 0.0;
      const double4 g = acc ? gx[c.vbeg + id] : make_double4(0, 0, 0, 0);
      yb[3 * i] = f * g.x; yb[3 * i + 1] = f * g.y; yb[3 * i + 2] = f * g.z;
    }
    const int nin = r.y ? 24 : 12;
    double xb[24];
    for (int d = 0; d < nin; ++d)
      xb[d] = r.y ? EeRespDir(in, d, out, w, h, yb) : VfRespDir(in, d, out, w, h, yb);
    for (int i = 0; i < 4; ++i) {
      if (!cl[i]) continue;
      const int v = c.vbeg + out.ids[i];
      AddFix3(add_x + 3 * static_cast<int64_t>(v), xb[3 * i], xb[3 * i + 1], xb[3 * i + 2]);
      if (r.y)
        AddFix3(add_x0 + 3 * static_cast<int64_t>(v), xb[12 + 3 * i], xb[12 + 3 * i + 1],
                xb[12 + 3 * i + 2]);
    }
  }
}

// reverse, step 3: gx += add_x, gx0 += add_x0; sums and counts cleared
__global__ __launch_bounds__(kBlock) void RepAdjApplyKernel(int n, double4* __restrict__ gx,
                                                           double4* __restrict__ gx0,
                                                           unsigned long long* __restrict__ add_x,
                                                           unsigned long long* __restrict__ add_x0,
                                                           int32_t* __restrict__ cnt) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int64_t* ax = reinterpret_cast<const int64_t*>(add_x) + 3 * static_cast<int64_t>(i);
  const int64_t* a0 = reinterpret_cast<const int64_t*>(add_x0) + 3 * static_cast<int64_t>(i);
  if (ax[0] | ax[1] | ax[2]) {
    double4 g = gx[i];
    g.x += static_cast<double>(ax[0]) / kAdjFix;
    g.y += static_cast<double>(ax[1]) / kAdjFix;
    g.z += static_cast<double>(ax[2]) / kAdjFix;
    gx[i] = g;
  }
  if (a0[0] | a0[1] | a0[2]) {
    double4 g = gx0[i];
    g.x += static_cast<double>(a0[0]) / kAdjFix;
    g.y += static_cast<double>(a0[1]) / kAdjFix;
    g.z += static_cast<double>(a0[2]) / kAdjFix;
    gx0[i] = g;
  }
  for (int j = 0; j < 3; ++j) {
    add_x[3 * static_cast<int64_t>(i) + j] = 0ull;
    add_x0[3 * static_cast<int64_t>(i) + j] = 0ull;
  }
  cnt[i] = 0;
}

// the per-pair VF and / or EE halves of the instances whose item lists are not usable
__global__ __launch_bounds__(kBlock) void MultiRepPairsKernel(const InstContactDev* ic,
                                                              const int32_t* off, int n,
                                                              const float4* __restrict__ x,
                                                              const float4* __restrict__ x0,
                                                              float h) {
  const int lblk = XcdBlock(off, n);
  if (lblk < 0) return;
  const int k = InstOfBlock(off, n, lblk);
  const InstContactDev& c = ic[k];
  const bool vf = !ItemsUsable(c, 0);
  const bool ee = !ItemsUsable(c, 1);
  if (!vf && !ee) return;
  const int nblk = off[k + 1] - off[k];
  const RMesh cm{c.cloth, MeshState{x + c.vbeg, x0 + c.vbeg, -1.0f}};
  const RMesh om{c.obs, MeshState{c.o_x, c.o_x0, 1e-3f}};
  const int total = min(c.scal[2], c.pair_cap);
  const int nobs = min(c.scal[1], total);
  for (int p = (lblk - off[k]) * kBlock + threadIdx.x; p < total; p += nblk * kBlock) {
    if (vf) RepulsionPairVfBody(cm, om, c.pairs[p], p >= nobs, h, c.dq, c.cnt);
    if (ee) RepulsionPairEeBody(cm, om, c.pairs[p], p >= nobs, h, c.dq, c.cnt);
  }
}

__global__ __launch_bounds__(kBlock) void MultiApplyKernel(const InstContactDev* ic,
                                                           const int32_t* off, int n,
                                                           float4* __restrict__ x, float relax) {
  XRT_MULTI_INDEX(kVert, c.nv)
  ApplyFixedBody(x + c.vbeg, i, c.dq, c.cnt, relax);
}

__global__ __launch_bounds__(kBlock) void MultiVfSolveKernel(const InstContactDev* ic,
                                                             const int32_t* off, int n,
                                                             float4* __restrict__ x,
                                                             const float4* __restrict__ x0,
                                                             float mu, float inv_ns) {
  XRT_MULTI_INDEX(kVert, c.nv)
  if (!c.has_obs) return;
  const VfFriction fr{x0 + c.vbeg, c.o_x0, c.o_x, inv_ns, mu};
  VfSolveBody(x + c.vbeg, i, c.obvh, c.rt_face, c.rt_q, 0.0f, fr);
  VfSolveBody(x + c.vbeg, i, c.obvh, c.vf_face, c.vf_q, 1e-3f, fr);
}

__global__ __launch_bounds__(kBlock) void MultiEeSolveKernel(const InstContactDev* ic,
                                                             const int32_t* off, int n,
                                                             const float4* __restrict__ x) {
  XRT_MULTI_INDEX(kEdge, c.ne)
  if (!c.has_obs) return;
  EeSolveBody(x + c.vbeg, c.cloth.edges, i, c.ee, c.ee_none, c.dq, c.cnt);
}

__global__ __launch_bounds__(kBlock) void MultiObsCopyKernel(const InstContactDev* ic,
                                                             const int32_t* off, int n) {
  XRT_MULTI_INDEX(kObsVert, c.obs_nv)
  c.o_x0[i] = c.o_x[i];
}
#undef XRT_MULTI_INDEX
}  // namespace

hipError_t LaunchInterpolate(hipStream_t s, const float4* x0, const float4* x, float4* x1, int n,
                             float alpha) {
  if (n <= 0) return hipSuccess;
  InterpolateKernel<<<Blocks(n), kBlock, 0, s>>>(x0, x, x1, n, alpha);
  return hipGetLastError();
}
hipError_t LaunchVfDcdGenerate(hipStream_t s, const float4* x, int n, const Lbvh& obs,
                               int32_t* face, float4* q) {
  if (n <= 0 || !obs.active) return hipSuccess;
  VfDcdGenerateKernel<<<Blocks(n), kBlock, 0, s>>>(x, n, obs.view(), face, q);
  return hipGetLastError();
}
hipError_t LaunchVfCcdGenerate(hipStream_t s, const float4* x0, const float4* x, int n,
                               const Lbvh& obs, int32_t* face, float4* q) {
  if (n <= 0 || !obs.active) return hipSuccess;
  VfCcdGenerateKernel<<<Blocks(n), kBlock, 0, s>>>(x0, x, n, obs.view(), face, q);
  return hipGetLastError();
}
hipError_t LaunchEeDcdGenerate(hipStream_t s, const float4* x, const int2* edges, int ne,
                               const Lbvh& obs, const uint32_t* obs_rtri, EeRec* rec,
                               uint8_t* none) {
  if (ne <= 0 || !obs.active) return hipSuccess;
  EeDcdGenerateKernel<<<Blocks(ne), kBlock, 0, s>>>(x, edges, ne, obs.view(), obs_rtri, rec,
                                                    none);
  return hipGetLastError();
}
hipError_t LaunchVfSolve(hipStream_t s, float4* x, int n, const Lbvh& obs, const int32_t* face,
                         const float4* q, float offset, const VfFriction& fr) {
  if (n <= 0 || !obs.active) return hipSuccess;
  VfSolveKernel<<<Blocks(n), kBlock, 0, s>>>(x, n, obs.view(), face, q, offset, fr);
  return hipGetLastError();
}
hipError_t LaunchEeSolve(hipStream_t s, const float4* x, const int2* edges, int ne,
                         const EeRec* rec, const uint8_t* none, long long* dq, int32_t* cnt) {
  if (ne <= 0) return hipSuccess;
  EeSolveKernel<<<Blocks(ne), kBlock, 0, s>>>(x, edges, ne, rec, none, dq, cnt);
  return hipGetLastError();
}
hipError_t LaunchApplyFixedDeltas(hipStream_t s, float4* x, int n, long long* dq, int32_t* cnt,
                                  float relax) {
  if (n <= 0) return hipSuccess;
  ApplyFixedKernel<<<Blocks(n), kBlock, 0, s>>>(x, n, dq, cnt, relax);
  return hipGetLastError();
}
hipError_t LaunchRepulsion(hipStream_t s, const MeshTopo& cloth, const MeshState& cs,
                           const MeshTopo& obs, const MeshState& os, const int2* pairs,
                           const int32_t* n_obs_pairs, const int32_t* n_total, int cap, float h,
                           long long* dq, int32_t* cnt) {
  if (cap <= 0) return hipSuccess;
  RMesh c{cloth, cs}, o{obs, os};
  const int grid = std::min(Blocks(cap), 4096);
  RepulsionKernel<<<grid, kBlock, 0, s>>>(c, o, pairs, n_obs_pairs, n_total, cap, h, dq, cnt);
  return hipGetLastError();
}
hipError_t LbvhFindPairsAt(Lbvh& lhs, const Lbvh& rhs, float thickness, int2* pairs,
                           const int32_t* base, int cap, hipStream_t s) {
  if (!lhs.active || !rhs.active) return hipSuccess;
  FindPairsAtKernel<<<Blocks(lhs.nf), kBlock, 0, s>>>(
      lhs.view(), rhs.view(), thickness, reinterpret_cast<const uint32_t*>(lhs.counts.ptr),
      reinterpret_cast<const uint32_t*>(lhs.offsets.ptr), base, pairs, cap);
  return hipGetLastError();
}
hipError_t LaunchAddTotals(hipStream_t s, const int32_t* a, const int32_t* b, int32_t* out, int cap,
                           int nf, int32_t* overflow) {
  AddTotalsKernel<<<1, 64, 0, s>>>(a, b, out, cap, nf, overflow);
  return hipGetLastError();
}

// ------------------------------------------------------------ multi-instance launchers
namespace {
using MK = MultiMaps::Kind;
inline bool Empty(const MultiMaps& m, MK k) { return m.n <= 0 || m.blocks[k] <= 0; }
}  // namespace

hipError_t MultiObstacleRefit(hipStream_t s, const InstContactDev* ic, const MultiMaps& m,
                              float alpha, float tol) {
  if (Empty(m, MK::kObsVert)) return hipSuccess;
  MultiInterpolateKernel<<<XcdGrid(m.blocks[MK::kObsVert]), kBlock, 0, s>>>(ic, m.off[MK::kObsVert], m.n, alpha);
  MultiNormalPredKernel<<<XcdGrid(m.blocks[MK::kObsVert]), kBlock, 0, s>>>(ic, m.off[MK::kObsVert], m.n, tol);
  MultiObsGrowKernel<<<XcdGrid(m.blocks[MK::kObsFace]), kBlock, 0, s>>>(ic, m.off[MK::kObsFace], m.n);
  MultiObsBottomUpKernel<<<m.n, 1024, 0, s>>>(ic);
  // the generate pass that follows queries the packed records
  MultiPackObsKernel<<<XcdGrid(m.blocks[MK::kObsFace]), kBlock, 0, s>>>(ic, m.off[MK::kObsFace], m.n);
  return hipGetLastError();
}

hipError_t MultiGenerate(hipStream_t s, const InstContactDev* ic, const MultiMaps& m,
                         const float4* x0, const float4* x) {
  if (Empty(m, MK::kVert)) return hipSuccess;
  MultiVfGenerateKernel<<<XcdGrid(m.blocks[MK::kVert]), kBlock, 0, s>>>(ic, m.off[MK::kVert], m.n, x0, x);
  if (!Empty(m, MK::kEdge))
    MultiEeGenerateKernel<<<XcdGrid(m.blocks[MK::kEdge]), kBlock, 0, s>>>(ic, m.off[MK::kEdge], m.n, x);
  return hipGetLastError();
}

hipError_t MultiPairs(hipStream_t s, const InstContactDev* ic, const MultiMaps& m,
                      const float4* x, const float4* x0, float thickness) {
  if (Empty(m, MK::kFace)) return hipSuccess;
  const int fb = m.blocks[MK::kFace];
  const int32_t* fo = m.off[MK::kFace];
  MultiClothPredKernel<<<XcdGrid(m.blocks[MK::kVert]), kBlock, 0, s>>>(ic, m.off[MK::kVert], m.n, x);
  MultiClothGrowKernel<<<XcdGrid(fb), kBlock, 0, s>>>(ic, fo, m.n, x, x0);
  MultiClothBottomUpKernel<<<m.n, 1024, 0, s>>>(ic);
  MultiPackClothKernel<<<XcdGrid(fb), kBlock, 0, s>>>(ic, fo, m.n);
  if (!Empty(m, MK::kObsFace))
    MultiPackObsKernel<<<XcdGrid(m.blocks[MK::kObsFace]), kBlock, 0, s>>>(ic, m.off[MK::kObsFace], m.n);
  // obstacle pass, then the cloth pass, whose emit follows the totals (it knows whether the
  // item lists will be used)
  for (int self = 0; self < 2; ++self) {
    MultiCountPairsKernel<<<XcdGrid(fb), kBlock, 0, s>>>(ic, fo, m.n, thickness, self);
    MultiScanKernel<<<m.n, 1024, 0, s>>>(ic, self);
    if (self) MultiAddTotalsKernel<<<Blocks(m.n), kBlock, 0, s>>>(ic, m.n);
    MultiEmitPairsKernel<<<XcdGrid(fb), kBlock, 0, s>>>(ic, fo, m.n, thickness, self);
  }
  return hipGetLastError();
}

hipError_t MultiRepulsion(hipStream_t s, const InstContactDev* ic, const MultiMaps& m, float4* x,
                          const float4* x0, float h, float relax, hipEvent_t ee_begin,
                          hipEvent_t ee_end) {
  if (Empty(m, MK::kPair)) return hipSuccess;
  MultiRepItemKernel<0><<<XcdGrid(m.blocks[MK::kPair]), kBlock, 0, s>>>(ic, m.off[MK::kPair], m.n, x, x0, h);
  if (ee_begin) {
    const hipError_t e = hipEventRecord(ee_begin, s);
    if (e != hipSuccess) return e;
  }
  MultiRepItemKernel<1><<<XcdGrid(m.blocks[MK::kPair]), kBlock, 0, s>>>(ic, m.off[MK::kPair], m.n, x, x0, h);
  if (ee_end) {
    const hipError_t e = hipEventRecord(ee_end, s);
    if (e != hipSuccess) return e;
  }
  // the fallback runs on the smaller per-vertex grid: most of its blocks exit at once
  MultiRepPairsKernel<<<XcdGrid(m.blocks[MK::kVert]), kBlock, 0, s>>>(ic, m.off[MK::kVert], m.n, x, x0, h);
  MultiApplyKernel<<<XcdGrid(m.blocks[MK::kVert]), kBlock, 0, s>>>(ic, m.off[MK::kVert], m.n, x, relax);
  return hipGetLastError();
}

hipError_t MultiVfSolve(hipStream_t s, const InstContactDev* ic, const MultiMaps& m, float4* x,
                        const float4* x0, float mu, float inv_ns) {
  if (Empty(m, MK::kVert)) return hipSuccess;
  MultiVfSolveKernel<<<XcdGrid(m.blocks[MK::kVert]), kBlock, 0, s>>>(ic, m.off[MK::kVert], m.n, x,
                                                                     x0, mu, inv_ns);
  return hipGetLastError();
}

hipError_t MultiEeSolve(hipStream_t s, const InstContactDev* ic, const MultiMaps& m, float4* x,
                        float relax) {
  if (Empty(m, MK::kEdge)) return hipSuccess;
  MultiEeSolveKernel<<<XcdGrid(m.blocks[MK::kEdge]), kBlock, 0, s>>>(ic, m.off[MK::kEdge], m.n, x);
  MultiApplyKernel<<<XcdGrid(m.blocks[MK::kVert]), kBlock, 0, s>>>(ic, m.off[MK::kVert], m.n, x, relax);
  return hipGetLastError();
}

hipError_t MultiBuildMerge(hipStream_t s, const InstContactDev* ic, const MultiMaps& m, bool cloth) {
  if (m.n <= 0) return hipSuccess;
  MultiBuildMergeKernel<<<2 * m.n, 1024, 0, s>>>(ic, m.n, cloth ? 1 : 0);
  return hipGetLastError();
}

hipError_t MultiObstacleFrameEnd(hipStream_t s, const InstContactDev* ic, const MultiMaps& m) {
  if (Empty(m, MK::kObsVert)) return hipSuccess;
  MultiObsCopyKernel<<<XcdGrid(m.blocks[MK::kObsVert]), kBlock, 0, s>>>(ic, m.off[MK::kObsVert], m.n);
  return hipGetLastError();
}


hipError_t MultiRepRecord(hipStream_t s, const InstContactDev* ic, const MultiMaps& m,
                          const float4* x, const float4* x0, float h, int4* pool,
                          unsigned long long* count, unsigned long long cap, int32_t* flags,
                          unsigned long long* mark_begin, unsigned long long* mark_end) {
  RepMarkKernel<<<1, 64, 0, s>>>(count, mark_begin);
  if (!Empty(m, MK::kPair)) {
    MultiRepRecordKernel<0><<<XcdGrid(m.blocks[MK::kPair]), kBlock, 0, s>>>(
        ic, m.off[MK::kPair], m.n, x, x0, h, pool, count, cap, flags);
    MultiRepRecordKernel<1><<<XcdGrid(m.blocks[MK::kPair]), kBlock, 0, s>>>(
        ic, m.off[MK::kPair], m.n, x, x0, h, pool, count, cap, flags);
  }
  RepMarkKernel<<<1, 64, 0, s>>>(count, mark_end);
  return hipGetLastError();
}

hipError_t RepulsionAdj(hipStream_t s, const InstContactDev* ic, int n_total, const int4* pool,
                        const unsigned long long* range, unsigned long long cap, const float4* x,
                        const float4* x0, float h, float relax, int32_t* cnt, double4* gx,
                        double4* gx0, unsigned long long* add_x, unsigned long long* add_x0) {
  if (n_total <= 0) return hipSuccess;
  constexpr int kGrid = 2048;  // grid-stride over the range (its length is on the device)
  RepAdjCountKernel<<<kGrid, kBlock, 0, s>>>(ic, pool, range, x, x0, h, cap, cnt);
  RepAdjVjpKernel<<<kGrid, kBlock, 0, s>>>(ic, pool, range, x, x0, h, relax, cap, cnt, gx, add_x,
                                           add_x0);
  RepAdjApplyKernel<<<(n_total + kBlock - 1) / kBlock, kBlock, 0, s>>>(n_total, gx, gx0, add_x,
                                                                     add_x0, cnt);
  return hipGetLastError();
}

}  // namespace xrt
