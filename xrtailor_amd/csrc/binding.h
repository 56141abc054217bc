// The steps either side of the substep in the garment pipeline: linear blend skinning of
// the obstacle (body.cu:39-58), the nearest obstacle face of every cloth vertex
// (QueryNearestBarycentricClothToObstacle, ClothSolverHelper.cu:25-66) and the binding
// constraint between cloth vertices and body points (BindingConstraint.cu:20-99).
//
// HBM layout: binding items of every instance concatenated (instance k at bbeg); the
// nearest-face records of the cloth vertices at the state's vertex offsets (int4 face
// vertex ids + face, float4 barycentric weights + distance); skinning weights per skinned
// obstacle as CSR and the joint transforms of every skinned obstacle in one array.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "blockmap.h"
#include "bvh.h"

namespace xrt {

// BindingItem (BindingConstraint.cuh:9-19): cloth vertex, body face vertices, barycentric
// weights, rest distance L, stiffness
struct BindItem {
  int4 idx;    // cloth vertex (instance-local), body vertices u, v, w
  float4 uvwl; // u, v, w, L
  float k;
  float pad[3];
};

struct BindInstDev {
  int32_t bbeg, nb, vbeg, pad;
  const float4* o_x;  // obstacle positions the binding pulls towards (the frame's target)
};

struct SkinInstDev {
  int32_t nv, nj, toff, pad;  // obstacle vertices, joints, offset of its transforms
  const int32_t* ptr;         // weights CSR over the vertices (nv + 1)
  const int32_t* joint;
  const float* weight;
  const float4* rest;         // rest (shaped) vertices
  float4* out;                // skinned vertices (the obstacle's target positions)
};

constexpr int kMaxJoints = 256;

// nearest face of every query point (corner points of the tree, QueryDevice order); *flag
// |= 1 when a search would overflow its 64-entry stack (that point's record is face -1)
hipError_t LaunchNearestBary(hipStream_t s, const BvhView& t, const float4* q, int n, int4* idx,
                             float4* uvw, int32_t* flag);
// BindingConstraint::Generate of one instance: items[i] from cloth vertex vert[i]'s nearest
// record, its rest distance dist[i] (< EPSILON: the current distance) and stiffness[i]
hipError_t LaunchBindingGenerate(hipStream_t s, const int32_t* vert, const float* stiffness,
                                 const float* dist, int n, const float4* cloth_x0,
                                 const int4* skin_idx, const float4* skin_uvw,
                                 const float4* obs_x0, BindItem* items);
// BindingConstraint::Solve of the map's instances
hipError_t LaunchBindingSolve(hipStream_t s, const BindInstDev* inst, const BindItem* items,
                              const DevMap& m, float4* x);
// linear blend skinning of the map's obstacles; T holds every skinned obstacle's joint
// transforms (3x4 row-major, 12 floats per joint)
hipError_t LaunchLbs(hipStream_t s, const SkinInstDev* inst, const float* T, const DevMap& m);

}  // namespace xrt
