// Device-side data layout of the cloth substep (shared by kernels and host code).
//
// HBM layout (all instances concatenated, CSR-indexed by InstDesc):
//   x0, v, x : float4[N]   SoA state; x0.w = x.w = inverse mass, v.w unused
//   sc       : StretchC[E] color-sorted per instance, instance-local vertex ids
//   slam     : float[E]    XPBD multipliers of the stretch constraints
//   bidx     : int4[B]     color-sorted bend elements, instance-local ids
//   bpar     : float2[B]   (rest angle, compliance)
//   s_coloff / b_coloff    color offsets (absolute indices into sc / bidx)
//   slot_vert, lra_ptr, lra_slot, lra_rest   geodesic long-range attachments
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace xrt {

constexpr float kEps = 1e-6f;  // EPSILON (xrtailor/include/xrtailor/core/Scalar.hpp:37)

// Multi-instance launches map B logical blocks (off[n] = B; off holds the n + 1 prefix
// sums of the instances' block counts, then the instance of every block) onto a grid of
// XcdGrid(B) blocks.  The hardware deals consecutive workgroups to the 8 XCDs in turn, so
// physical block b runs logical block (b % 8) * ceil(B / 8) + b / 8: each XCD walks a
// contiguous range of logical blocks -- a few instances at a time, whose data then stay
// in that XCD's L2.  -1: idle block.
inline int XcdGrid(int blocks) { return (blocks + 7) / 8 * 8; }

__device__ __forceinline__ int XcdBlock(const int32_t* __restrict__ off, int n) {
  const int total = off[n];
  const int l = static_cast<int>(blockIdx.x & 7u) * ((total + 7) >> 3) + static_cast<int>(blockIdx.x >> 3);
  return l < total ? l : -1;
}

// the instance of logical block b of such a map
__device__ __forceinline__ int InstOfBlock(const int32_t* __restrict__ off, int n, int b) {
  return off[n + 1 + b];
}

// Maps built by AppendBlockMap (blockmap.h) list entries instead: InstOfBlock gives the
// entry of a block and this the instance of an entry.
__device__ __forceinline__ int InstOfEntry(const int32_t* __restrict__ off, int n, int j) {
  return off[n + 1 + off[n] + j];
}

struct StretchC {  // BasicConstraint::StretchConstraint without lambda (BasicConstraint.cuh:8-14)
  int32_t i0, i1;
  float rest, compliance;
};

struct InstDesc {
  int32_t vbeg, nv;      // vertices [vbeg, vbeg+nv)
  int32_t s_cbeg, s_ncol;  // stretch colors: s_coloff[s_cbeg + c] .. [s_cbeg + c + 1]
  int32_t b_cbeg, b_ncol;  // bend colors
  int32_t slot_beg, nslots;  // slot_vert[slot_beg ..)
  int32_t lra_vbeg;        // lra_ptr[lra_vbeg + i], i in [0, nv]
  int32_t pad[3];
};

// per-substep constants shared by every kernel (the SimParams subset they read)
struct StepConsts {
  float gx, gy, gz;
  float dt;
  float inv_dt2;        // 1/dt^2 for XPBD alpha-tilde
  float damp_factor;    // 1 - damping * dt
  float stretchiness;   // long_range_stretchiness
  float ground_y;
  float friction;
  float relaxation;
  int32_t iterations;
  int32_t solve_bending;
  int32_t enable_ground;
  int32_t pad;
  // bend model (see xrt_params::bend_model): d clamped to [bend_dlo, 1], rest angle
  // bend_th0 + bend_th1 * theta0, multiplier scaled by bend_sgn
  float bend_dlo, bend_th0, bend_th1, bend_sgn;
};

struct BendMode {
  float dlo, th0, th1, sgn;
};

struct StrainC {  // FEM::StrainConstraint (FEMConstraint.cuh:8-19) without the material
  int32_t i0, i1, i2;
  float area;
  float inv[4];  // glm::inverse of the rest matrix: [0][0], [0][1], [1][0], [1][1]
};

struct FemConsts {
  float c00, c01, c10, c11, c22;  // orthotropic elasticity entries (FEMHelper.cu:103-111)
  float iso_at;                   // 1 / (bending stiffness dt^2)
};

struct BatchArgs {
  const InstDesc* inst;
  const int32_t* order;  // workgroup -> instance (largest first)
  float4* x0;
  float4* v;
  float4* x;
  const StretchC* sc;
  float* slam;
  const int4* bidx;
  const float2* bpar;
  const int32_t* s_coloff;
  const int32_t* b_coloff;
  const int32_t* slot_vert;
  const int32_t* lra_ptr;
  const int32_t* lra_slot;
  const float* lra_rest;
  int32_t lds_verts;   // instances with nv <= lds_verts keep x in LDS
  int32_t max_slots;   // LDS staging for LRA slot positions
  // quality-mode FEM constraints in place of stretch and bend when fem != 0: per instance
  // (first face color offset, face colors), the global face color offsets, the strain
  // elements, the isometric bending data in the bend elements' order
  int32_t fem;
  const int2* fcol;
  const int32_t* f_coloff;
  const StrainC* fsc;
  const float4* iso_k;
  const float* iso_c;
  float* iso_lam;
  FemConsts fk;
};

}  // namespace xrt
