// Linear BVH over triangle faces (Karras 2012) for gfx950.
//
// Device layout of one tree with F faces (2F-1 nodes; internal nodes [0, F-1), leaves
// [F-1, 2F-1) in Morton order, as in xrtailor/src/physics/broad_phase/lbvh/BVH.cu:489-502):
//   nodes  int4[2F-1]  (parent, left, right, object); -1 = none
//   box    Box[2F-1]   lower / upper corner (w unused)
//   escape int[2F-1]   stackless "next node on miss" (BVH.cu:212-273), node-indexed
//   morton u32[F], order int[F]   sorted codes and the face id of each leaf
//   v, pred float4[nv] corner points x0 + tol n and x + tol n (BVH.cu:382-421), per vertex
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.h"

namespace xrt {

struct Box {
  float4 lo, hi;
};

struct BvhView {
  int nf = 0, nv = 0;
  const int4* tris = nullptr;
  float4* v = nullptr;
  float4* pred = nullptr;
  uint32_t* morton = nullptr;
  int32_t* order = nullptr;
  int4* nodes = nullptr;
  Box* box = nullptr;
  int32_t* escape = nullptr;
  uint32_t* root = nullptr;  // 6 order-preserving encoded floats (lower xyz, upper xyz)
  int32_t* flags = nullptr;     // bottom-up arrival counters (internal nodes)
  uint32_t* counts = nullptr;   // pair traversal: pairs per lhs leaf
  uint32_t* offsets = nullptr;  //                 their exclusive prefix
  int32_t* total = nullptr;     //                 and their sum (1 int)
  Box* packed = nullptr;        // traversal records (PackNodeBody), batched pair refresh
  int32_t* lvl = nullptr;       // LevelBottomUpWG's depth and child lists (4 nf ints)
  int4* leaf_face = nullptr;    // per leaf slot: its face's record (batched pair refresh)
};

// Owning storage of one LBVH plus its build scratch.
struct Lbvh {
  int nf = 0, nv = 0, nblocks = 0;
  bool active = false;
  DevBuf<int4> tris;
  DevBuf<float4> v, pred;
  DevBuf<Box> leaf;  // leaf boxes in face order
  DevBuf<uint32_t> morton, morton_tmp;
  DevBuf<int32_t> order, order_tmp;
  DevBuf<uint64_t> code;
  DevBuf<int4> nodes;
  DevBuf<Box> box;
  DevBuf<int32_t> escape, flags;
  DevBuf<uint32_t> root;
  DevBuf<uint32_t> hist;  // radix histogram / offsets
  DevBuf<int32_t> counts, offsets, total;  // pair traversal scratch (per lhs leaf)
  DevBuf<Box> packed;                      // traversal records of the nodes
  DevBuf<int32_t> lvl;                     // depth lists of the level-by-level merge
  DevBuf<int4> leaf_face;                  // per leaf slot: its face's record
  BvhView view() const;
  // (re)allocate for a mesh; uploads faces (host int32[3*nf])
  hipError_t Setup(const int32_t* tris_host, int nf, int nv, hipStream_t s);
};

// One tree of a multi-tree build: its buffers and the points it is built from.
struct TreeBuildDev {
  const float4* x0 = nullptr;
  int nv = 0, nf = 0, nblocks = 0;
  const int4* tris = nullptr;
  float4* v = nullptr;
  float4* pred = nullptr;
  Box* leaf = nullptr;
  uint32_t* morton = nullptr;
  uint32_t* morton_tmp = nullptr;
  int32_t* order = nullptr;
  int32_t* order_tmp = nullptr;
  uint64_t* code = nullptr;
  int4* nodes = nullptr;
  Box* box = nullptr;
  int32_t* escape = nullptr;
  uint32_t* root = nullptr;
  uint32_t* hist = nullptr;
  int32_t* flags = nullptr;
};

// A set of trees rebuilt together, every build step of every tree in one launch (the
// frame-start build: BVH::UpdateData + BVH::Construct from x0, tolerance 0, no CCD).
// Per kind of work the device maps hold the n + 1 prefix sums of the trees' block counts,
// then the tree of every block.
struct TreeSet {
  enum Kind { kVert, kFace, kTile, kNode, kKinds };
  int n = 0;
  std::vector<Lbvh*> trees;
  DevBuf<TreeBuildDev> dev;
  DevBuf<int32_t> maps;
  const int32_t* off[kKinds] = {};
  int blocks[kKinds] = {};
  // trees with faces only; x0[i] is tree i's point array (nv float4)
  hipError_t Setup(const std::vector<Lbvh*>& trees, const std::vector<const float4*>& x0,
                   hipStream_t s);
};

// every tree of the set as LbvhBuild(tree, x0, null, null, 0, 0, 0, s, merge = false)
hipError_t MultiLbvhBuild(TreeSet& set, hipStream_t s);

// BVH::UpdateData + BVH::Construct; x0/x/n are float4 per-vertex arrays (n, x may be null).
// merge = false leaves the internal boxes to a later multi-tree merge (the batched frame
// start merges every instance's trees in one launch).
hipError_t LbvhBuild(Lbvh& b, const float4* x0, const float4* x, const float4* n, float tol,
                     int ccd, float scr, hipStream_t s, bool merge = true);
// BVH::RefitMidstep: pred = x1 + tol n1, leaves grow, parents re-merged
hipError_t LbvhRefitMidstep(Lbvh& b, const float4* x1, const float4* n1, float tol, hipStream_t s);
// BVH::Update(ccd): pred = x, leaves grow by bounds(x[, x0]), ancestors grow
hipError_t LbvhUpdate(Lbvh& b, const float4* x, const float4* x0, int ccd, hipStream_t s);
// stackless point (p1 null) / segment queries, <= 256 hits each
hipError_t LbvhQuery(const Lbvh& b, const float4* p0, const float4* p1, int n, int32_t* hits,
                     int32_t* counts, hipStream_t s);
// Traverse(lhs, rhs, thickness): counts pairs per lhs leaf and their exclusive offsets;
// *total (device) receives the pair count
hipError_t LbvhCountPairs(Lbvh& lhs, const Lbvh& rhs, float thickness, hipStream_t s);
// fills (lhs object, rhs object) pairs; writes only slots < cap
hipError_t LbvhFindPairs(Lbvh& lhs, const Lbvh& rhs, float thickness, int2* pairs, int cap,
                         hipStream_t s);
// per-vertex normals summed in ascending face order over a vertex->corner CSR
hipError_t VertexNormals(const float4* x, const int4* tris, const int32_t* vptr,
                         const int32_t* corner, int nv, float4* n, hipStream_t s);

float DecodeOrdered(uint32_t u);

}  // namespace xrt
