// Quality-mode cloth constraints (xrtailor/src/physics/dynamics/FEMHelper.cu:10-392): the
// orthotropic Green-strain triangle constraint and the isometric (quadratic) bending
// constraint, projected per color like the XPBD stretch and bend constraints.
//
// HBM layout: strain elements color-sorted per instance (32 B each: face vertices, rest
// area, the inverse rest matrix in glm order); isometric bending in the bend elements' color
// order (the bends' int4 indices, the weights K as float4 and coeff, the multiplier).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.h"

namespace xrt {

struct StrainC {  // FEM::StrainConstraint (FEMConstraint.cuh:8-19) without the material
  int32_t i0, i1, i2;
  float area;
  float inv[4];  // glm::inverse of the rest matrix: [0][0], [0][1], [1][0], [1][1]
};

struct FemConsts {
  float c00, c01, c10, c11, c22;  // orthotropic elasticity entries (FEMHelper.cu:103-111)
  float iso_at;                   // 1 / (bending stiffness dt^2)
};

struct FemArgs {
  const InstDesc* inst;
  const int2* fcol;       // per instance: first face color offset, face colors
  const int32_t* f_coloff;
  const StrainC* sc;
  const int32_t* b_coloff;
  const int4* bidx;
  const float4* iso_k;
  const float* iso_c;
  float* iso_lam;
  float4* x;
};

// one color of one instance
hipError_t LaunchStrain(hipStream_t s, const StrainC* sc, float4* x, int cnt, const FemConsts& k);
hipError_t LaunchIsoBend(hipStream_t s, const int4* bidx, const float4* iso_k, const float* iso_c,
                         float* lam, float4* x, int cnt, int iter, const FemConsts& k);
// one color of every instance (per-color block map: n_inst + 1 prefix sums, then the
// instance of every block)
hipError_t LaunchWideStrain(hipStream_t s, const FemArgs& a, int n_inst, int color,
                            const int32_t* map, int blocks, const FemConsts& k);
hipError_t LaunchWideIsoBend(hipStream_t s, const FemArgs& a, int n_inst, int color,
                             const int32_t* map, int blocks, int iter, const FemConsts& k);

}  // namespace xrt
