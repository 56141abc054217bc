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
};

}  // namespace xrt
