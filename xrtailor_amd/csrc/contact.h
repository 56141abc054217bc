// Body-cloth PBD collision and PBD-repulsion self contact (gfx950).
//
// Per-mesh contact topology (instance-local ids):
//   tris int4[F], edges int2[E], face_edges int4[F] (edge of nodes i,i+1), ef_ptr/ef CSR
//   (faces per edge, ascending, <= 4 used), face_n float4[F] rest face normals,
//   rtri u32[F] representative-triangle bits (vertex i: bit i, edge i: bit 3+i),
//   trir int4[F] = (tris.xyz, rtri)
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
  const int4* trir = nullptr;  // (vertices, rtri bits) per face: one load for the item masks
};

struct MeshState {      // positions of one mesh (w of x / x0 = inverse mass)
  const float4* x = nullptr;
  const float4* x0 = nullptr;
  float fixed_w = -1.0f;  // >= 0: every node has this inverse mass and is not free (obstacle)
};

struct EeRec {
  float4 a, p2, q2, fn;
};

// One PBD repulsion response as the adjoint replays it: the corners whose corrections were
// accumulated (mask, bit i = corner i) and their node ids; VF: the orientation taken from
// the start-of-substep state; EE: the normal used (pn), and whether it is the
// start-of-substep edge-edge normal times pn_sign (pn_x0, a free other edge) or the other
// edge's fixed normal.
struct RepOut {
  int mask = 0;
  int ids[4] = {0, 0, 0, 0};
  float side = 1.0f;
  int pn_x0 = 0;
  float pn_sign = 1.0f;
  float3 pn = make_float3(0.0f, 0.0f, 0.0f);
};

hipError_t LaunchInterpolate(hipStream_t s, const float4* x0, const float4* x, float4* x1, int n,
                             float alpha);
hipError_t LaunchVfDcdGenerate(hipStream_t s, const float4* x, int n, const Lbvh& obs,
                               int32_t* face, float4* q);
hipError_t LaunchVfCcdGenerate(hipStream_t s, const float4* x0, const float4* x, int n,
                               const Lbvh& obs, int32_t* face, float4* q);
// none[e] = 1 where edge e got no record (rec[e].a.x > 1e6), read by the EE solve first
hipError_t LaunchEeDcdGenerate(hipStream_t s, const float4* x, const int2* edges, int ne,
                               const Lbvh& obs, const uint32_t* obs_rtri, EeRec* rec,
                               uint8_t* none);
// friction of the VF solves (mu = 0: none, the reference's behaviour): x0 the cloth's
// start-of-substep positions, o_x0 / o_x the obstacle's frame-start / frame-end positions,
// inv_ns = 1 / substeps (the obstacle moves (o_x - o_x0) inv_ns per substep)
struct VfFriction {
  const float4* x0;
  const float4* o_x0;
  const float4* o_x;
  float inv_ns;
  float mu;
};
hipError_t LaunchVfSolve(hipStream_t s, float4* x, int n, const Lbvh& obs, const int32_t* face,
                         const float4* q, float offset, const VfFriction& fr);
hipError_t LaunchEeSolve(hipStream_t s, const float4* x, const int2* edges, int ne,
                         const EeRec* rec, const uint8_t* none, long long* dq, int32_t* cnt);
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
// *out = *a + *b; a total above cap (an instance of nf faces) raises *overflow to the
// pairs-per-face figure the instance would need
hipError_t LaunchAddTotals(hipStream_t s, const int32_t* a, const int32_t* b, int32_t* out, int cap,
                           int nf, int32_t* overflow);

// ----------------------------------------------------------------- all instances at once
// Device view of one instance's contact state for the multi-instance kernels.  Every
// pointer refers to that instance's own buffers; cloth ids are instance-local and the
// cloth state arrays passed to the launchers are offset by vbeg.
struct InstContactDev {
  int vbeg = 0, nv = 0, nf = 0, ne = 0;
  int has_obs = 0;  // obstacle mesh present
  int obs_nv = 0, obs_nf = 0;
  int pair_cap = 0;
  MeshTopo cloth, obs;
  BvhView cbvh, obvh;
  float4* o_x0 = nullptr;
  float4* o_x = nullptr;
  float4* o_x1 = nullptr;
  float4* o_n1 = nullptr;
  const int32_t* o_vptr = nullptr;
  const int32_t* o_corner = nullptr;
  int32_t* vf_face = nullptr;
  int32_t* rt_face = nullptr;
  float4* vf_q = nullptr;
  float4* rt_q = nullptr;
  EeRec* ee = nullptr;
  uint8_t* ee_none = nullptr;
  long long* dq = nullptr;
  int32_t* cnt = nullptr;
  int2* pairs = nullptr;
  // [1] obstacle pairs, [2] all pairs, [3] / [4] VF / EE items, [5] / [6] VF / EE items of
  // the obstacle pairs, [7] / [8] of the cloth pairs
  int32_t* scal = nullptr;
  // repulsion test items rebuilt at each pair refresh (capacities 0: lists off): each wave of
  // 64 cloth leaves owns the range its leaves' counts add up to, from its first leaf's
  // offset; ivf / iee: per cloth leaf, the block-local exclusive prefix of the item counts of
  // the current traversal pass; nitem: the leaf's own counts (VF, EE)
  int vf_cap = 0, ee_cap = 0;
  int2* rep_vf = nullptr;
  int2* rep_ee = nullptr;
  uint32_t* ivf = nullptr;
  uint32_t* iee = nullptr;
  uint2* nitem = nullptr;
  // pair refresh: the first kPairStage hits of each cloth leaf, hit j of leaf i at
  // stage[j * nf + i] as rhs object | item mask << kStageObjBits (null: the pairs are found
  // by a second traversal; also for meshes of 2^kStageObjBits faces or more)
  uint32_t* stage = nullptr;
  // pair refresh: per 256-leaf block of the count pass, its pair / VF / EE item sums
  // (3 x Blocks(nf)), then their exclusive prefix
  uint32_t* bsum = nullptr;
  // the context's pair-overflow word: the largest pairs-per-face figure of an instance
  // whose pairs exceeded pair_cap since it was last read (pairs beyond pair_cap are lost)
  int32_t* overflow = nullptr;
};

constexpr int kPairStage = 128;
constexpr int kStageObjBits = 20;  // the item mask (12 bits) takes the rest of a stage entry

// block -> instance maps: per kind, the prefix sums (n + 1 entries, device) of the
// 256-thread blocks each instance needs, followed by the instance of each block (device),
// and the block totals (host)
struct MultiMaps {
  enum Kind { kVert, kEdge, kFace, kObsVert, kObsFace, kPair, kKinds };
  const int32_t* off[kKinds] = {};
  int blocks[kKinds] = {};
  int n = 0;
};

// collision pass: obstacle x1 = lerp(x0, x, alpha), midstep normals, RefitMidstep(tol)
hipError_t MultiObstacleRefit(hipStream_t s, const InstContactDev* ic, const MultiMaps& m,
                              float alpha, float tol);
// collision pass: VF-DCD, VF-CCD and EE-DCD records of every instance with an obstacle
hipError_t MultiGenerate(hipStream_t s, const InstContactDev* ic, const MultiMaps& m,
                         const float4* x0, const float4* x);
// pair refresh: cloth LBVH Update(x, x0, ccd), Traverse(cloth, obstacle) ++
// Traverse(cloth, cloth) into each instance's pair list (scal[1] obstacle pairs, scal[2]
// total), then the pairs' repulsion test items
hipError_t MultiPairs(hipStream_t s, const InstContactDev* ic, const MultiMaps& m,
                      const float4* x, const float4* x0, float thickness);
// PBD repulsion over every instance's pairs (one VF or EE test item per thread), applied
// x relax; events (if not null) are recorded around the EE kernel
hipError_t MultiRepulsion(hipStream_t s, const InstContactDev* ic, const MultiMaps& m, float4* x,
                          const float4* x0, float h, float relax, hipEvent_t ee_begin = nullptr,
                          hipEvent_t ee_end = nullptr);
// Adjoint tape of a repulsion pass on the state (x, x0), before MultiRepulsion: the items of
// every instance that produce a response go to pool (instance, is_ee, item) at *count (at
// most cap); *mark_begin / *mark_end receive the count before and after.  flags |= 1 when
// the pool overflowed, |= 2 when an instance evaluated its pairs without item lists.
hipError_t MultiRepRecord(hipStream_t s, const InstContactDev* ic, const MultiMaps& m,
                          const float4* x, const float4* x0, float h, int4* pool,
                          unsigned long long* count, unsigned long long cap, int32_t* flags,
                          unsigned long long* mark_begin, unsigned long long* mark_end);
// Reverse of a repulsion pass + Jacobi apply from its recorded responses pool[range[0],
// range[1]) on the pass's input state (x, x0 = start of the substep): gx (every instance's
// nodes, the adjoint after the apply) becomes the adjoint before the pass, gx0 += through
// the start-of-substep edge-edge normals.  cnt (int[n_total]) and add_x / add_x0
// (3 n_total each) are zero on entry and on return.
hipError_t RepulsionAdj(hipStream_t s, const InstContactDev* ic, int n_total, const int4* pool,
                        const unsigned long long* range, unsigned long long cap, const float4* x,
                        const float4* x0, float h, float relax, int32_t* cnt, double4* gx,
                        double4* gx0, unsigned long long* add_x, unsigned long long* add_x0);
// VF-CCD (offset 0) then VF-DCD (offset 1e-3) solves
hipError_t MultiVfSolve(hipStream_t s, const InstContactDev* ic, const MultiMaps& m, float4* x,
                        const float4* x0, float mu, float inv_ns);
// EE-DCD deltas applied x relax
hipError_t MultiEeSolve(hipStream_t s, const InstContactDev* ic, const MultiMaps& m, float4* x,
                        float relax);
// frame start: merge the internal boxes of trees built with merge = false (cloth trees
// when `cloth`, obstacle trees always)
hipError_t MultiBuildMerge(hipStream_t s, const InstContactDev* ic, const MultiMaps& m, bool cloth);
// frame end: obstacle x0 = x
hipError_t MultiObstacleFrameEnd(hipStream_t s, const InstContactDev* ic, const MultiMaps& m);

}  // namespace xrt
