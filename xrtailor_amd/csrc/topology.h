// Host-side setup of a cloth instance: stretch edges, bend elements, rest state,
// constraint coloring and geodesic rest lengths.  Runs once per scene (the reference
// runs the same steps in AddCloth / SetupInternalDynamics).
#pragma once
#include <cstdint>
#include <vector>

namespace xrt {

// Unique edges in order of first appearance over faces, local edges (0,1),(1,2),(2,0),
// oriented as in the first face holding the edge (edge order of
// PrepareBasicStretchData, xrtailor/src/physics/ClothSolverHelper.cpp:90-109, with the
// vcglib enumeration replaced by this deterministic one).
std::vector<int32_t> BuildEdges(const int32_t* tris, int nt);

// Dihedral elements (v0,v1,v2,v3) per pair of faces on an edge
// (GenerateBendingElements, xrtailor/src/physics/ClothSolverHelper.cpp:4-88).
std::vector<int32_t> BuildBends(const int32_t* tris, int nt, const std::vector<int32_t>& edges);

std::vector<float> RestLengths(const float* pos, const std::vector<int32_t>& edges);
std::vector<float> RestAngles(const float* pos, const std::vector<int32_t>& bends);

// Constraint coloring, conflict = shared vertex; returns compact colors 0..k-1.
//  kColorGreedy : first fit in constraint index order (seed unused)
//  kColorPalette: GraphColoring::Paint (xrtailor/src/physics/graph_coloring/Base.cu:12-317)
//                 made deterministic (SplitMix64 streams, start-of-round resolution, hash
//                 priority instead of the bare index for ranking conflicts)
constexpr int kColorGreedy = 0;
constexpr int kColorPalette = 1;
std::vector<int32_t> ColorConstraints(const int32_t* cons, int arity, int n, int nv,
                                      uint64_t seed, int mode, int* n_colors);

// Dijkstra over the edge graph (Geodesic::ComputeGeodesicDistance,
// xrtailor/src/utils/Geodesic.cpp:6-79).
std::vector<double> Geodesic(const float* pos, int nv, const std::vector<int32_t>& edges, int src);

uint64_t SplitMix64(uint64_t* state);

// Contact topology of a mesh: face edge ids (edge of nodes i, i+1; Face::SetEdge,
// xrtailor/src/memory/Face.cu:13-25), faces per edge ascending (BuildEFAdjacency,
// xrtailor/src/physics/ClothSolverHelper.cpp:111-125), representative-triangle bits
// (RTriBuild_Kernel, xrtailor/src/physics/representative_triangle/RTriangle.cu:19-85) and
// rest face normals (Face::Update, xrtailor/src/memory/Face.cu:55-59).
struct ContactTopo {
  std::vector<int32_t> face_edges, ef_ptr, ef;
  std::vector<float> face_n;
  std::vector<uint32_t> rtri;
};
ContactTopo BuildContactTopo(const float* pos, int nv, const int32_t* tris, int nt,
                             const std::vector<int32_t>& edges);

}  // namespace xrt
