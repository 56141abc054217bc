// Body-cloth PBD collision and PBD-repulsion self contact (gfx950).
//
// Per-mesh contact topology (instance-local ids):
//   tris int4[F], edges int2[E], face_edges int4[F] (edge of nodes i,i+1), ef_ptr/ef CSR
//   (faces per edge, ascending, <= 4 used), face_n float4[F] rest face normals,
//   rtri u32[F] representative-triangle bits (vertex i: bit i, edge i: bit 3+i)
// Contact records (per cloth vertex / edge):
//   vf_face int, vf_q float4 (VF DCD); rt_face int, rt_q float4 (VF CCD);
//   ee: float4 a = (dist, s, t, -), p2, q2, fn (EE DCD)
// Jacobi deltas accumulate in 32.32 fixed point (int64 atomics): the sum, and therefore the
// result, does not depend on the order in which contributions arrive.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bvh.h"

namespace xrt {

struct MeshTopo {
  int nv = 0, nf = 0, ne = 0;
  const int4* tris = nullptr;
  const int2* edges = nullptr;
  const int4* face_edges = nullptr;
  const int32_t* ef_ptr = nullptr;
  const int32_t* ef = nullptr;
  const float4* face_n = nullptr;
  const uint32_t* rtri = nullptr;
};

struct MeshState {      // positions of one mesh (w of x / x0 = inverse mass)
  const float4* x = nullptr;
  const float4* x0 = nullptr;
  float fixed_w = -1.0f;  // >= 0: every node has this inverse mass and is not free (obstacle)
};

struct EeRec {
  float4 a, p2, q2, fn;
};

hipError_t LaunchInterpolate(hipStream_t s, const float4* x0, const float4* x, float4* x1, int n,
                             float alpha);
hipError_t LaunchVfDcdGenerate(hipStream_t s, const float4* x, int n, const Lbvh& obs,
                               int32_t* face, float4* q);
hipError_t LaunchVfCcdGenerate(hipStream_t s, const float4* x0, const float4* x, int n,
                               const Lbvh& obs, int32_t* face, float4* q);
hipError_t LaunchEeDcdGenerate(hipStream_t s, const float4* x, const int2* edges, int ne,
                               const Lbvh& obs, const uint32_t* obs_rtri, EeRec* rec);
hipError_t LaunchVfSolve(hipStream_t s, float4* x, int n, const Lbvh& obs, const int32_t* face,
                         const float4* q, float offset);
hipError_t LaunchEeSolve(hipStream_t s, const float4* x, const int2* edges, int ne,
                         const EeRec* rec, long long* dq, int32_t* cnt);
hipError_t LaunchApplyFixedDeltas(hipStream_t s, float4* x, int n, long long* dq, int32_t* cnt,
                                  float relax);
// pairs of (cloth face, other face): entries < n_obs_pairs (read on device) are cloth vs
// obstacle, the rest cloth vs cloth; pairs beyond min(total, cap) are ignored
hipError_t LaunchRepulsion(hipStream_t s, const MeshTopo& cloth, const MeshState& cs,
                           const MeshTopo& obs, const MeshState& os, const int2* pairs,
                           const int32_t* n_obs_pairs, const int32_t* n_total, int cap, float h,
                           long long* dq, int32_t* cnt);
// FindPairs into pairs + *base (device offset), only slots < cap
hipError_t LbvhFindPairsAt(Lbvh& lhs, const Lbvh& rhs, float thickness, int2* pairs,
                           const int32_t* base, int cap, hipStream_t s);
hipError_t LaunchAddTotals(hipStream_t s, const int32_t* a, const int32_t* b, int32_t* out);

}  // namespace xrt
