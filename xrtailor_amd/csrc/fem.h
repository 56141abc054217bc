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

// wide launches: 256-thread blocks; the strain launch takes two faces per thread
// (kFemPerBlock per block), the isometric bending launch one hinge per thread and the
// bend elements' block maps
constexpr int kFemBlock = 256;
constexpr int kFemPerBlock = 2 * kFemBlock;

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

__device__ __forceinline__ float FemDot3(float ax, float ay, float az, float bx, float by,
                                         float bz) {
#pragma clang fp contract(off)
  return ax * bx + ay * by + az * bz;
}

// The two projections are evaluated without FMA contraction whichever kernel inlines them
// (the plain, wide and persistent launches then agree bit for bit).
//
// SolveStrain_Kernel + ComputeStrainCorrection (FEMHelper.cu:94-212), glm's column-major
// products written index by index: F = [p0-p2, p1-p2] inv, Green strain e, stress S = C e,
// P = S F, constraint gradients H = inv^T P area, PBD step energy / sum w |grad|^2.
// X: the instance's positions (global memory or LDS).
template <typename P>
__device__ __forceinline__ void StrainProject(P X, const StrainC& c, const FemConsts& k) {
#pragma clang fp contract(off)
  float4 a = X[c.i0], b = X[c.i1], d = X[c.i2];
  const float p13[3] = {a.x - d.x, a.y - d.y, a.z - d.z};
  const float p23[3] = {b.x - d.x, b.y - d.y, b.z - d.z};
  float F0[3], F1[3];  // F[c][0], F[c][1]
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    F0[q] = p13[q] * c.inv[0] + p23[q] * c.inv[2];
    F1[q] = p13[q] * c.inv[1] + p23[q] * c.inv[3];
  }
  const float e00 = 0.5f * (F0[0] * F0[0] + F0[1] * F0[1] + F0[2] * F0[2] - 1.0f);
  const float e11 = 0.5f * (F1[0] * F1[0] + F1[1] * F1[1] + F1[2] * F1[2] - 1.0f);
  const float e01 = 0.5f * (F0[0] * F1[0] + F0[1] * F1[1] + F0[2] * F1[2]);
  const float s00 = k.c00 * e00 + k.c01 * e11;
  const float s11 = k.c10 * e00 + k.c11 * e11;
  const float s01 = k.c22 * e01;
  const float psi = 0.5f * (e00 * s00 + e01 * s01 + e01 * s01 + e11 * s11);
  const float energy = c.area * psi;
  float g0[3], g1[3], g2[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const float pk0 = s00 * F0[q] + s01 * F1[q];
    const float pk1 = s01 * F0[q] + s11 * F1[q];
    g0[q] = (c.inv[0] * pk0 + c.inv[1] * pk1) * c.area;
    g1[q] = (c.inv[2] * pk0 + c.inv[3] * pk1) * c.area;
    g2[q] = -g0[q] - g1[q];
  }
  float sum = a.w * FemDot3(g0[0], g0[1], g0[2], g0[0], g0[1], g0[2]);
  sum += b.w * FemDot3(g1[0], g1[1], g1[2], g1[0], g1[1], g1[2]);
  sum += d.w * FemDot3(g2[0], g2[1], g2[2], g2[0], g2[1], g2[2]);
  if (!(fabsf(sum) > kEps)) return;
  const float s = energy / sum;
  const float fa = -(s * a.w), fb = -(s * b.w), fd = -(s * d.w);
  if (a.w > kEps) { a.x += fa * g0[0]; a.y += fa * g0[1]; a.z += fa * g0[2]; X[c.i0] = a; }
  if (b.w > kEps) { b.x += fb * g1[0]; b.y += fb * g1[1]; b.z += fb * g1[2]; X[c.i1] = b; }
  if (d.w > kEps) { d.x += fd * g2[0]; d.y += fd * g2[1]; d.z += fd * g2[2]; X[c.i2] = d; }
}

// SolveIsometricBending_Kernel (FEMHelper.cu:317-388) with Q = coeff K K^T formed on the
// fly (entries coeff * (K_i K_j), the reference's Q).  The quadratic form and Q x are
// evaluated on positions relative to the hinge's first vertex: the rows of Q sum to zero,
// so this is the same energy and gradient without the cancellation of O(1 m) coordinates
// against centimetre-sized bending in FP32.  The multiplier is reset at iteration 0.
template <typename P>
__device__ __forceinline__ void IsoProject(P X, const int4 id, const float4 K4, const float coeff,
                                           float* lam_ptr, int iter, const FemConsts& k) {
#pragma clang fp contract(off)
  float4 x[4] = {X[id.x], X[id.y], X[id.z], X[id.w]};
  const float K[4] = {K4.x, K4.y, K4.z, K4.w};
  float y[4][3];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    y[j][0] = x[j].x - x[0].x;
    y[j][1] = x[j].y - x[0].y;
    y[j][2] = x[j].z - x[0].z;
  }
  const float lam = iter == 0 ? 0.0f : *lam_ptr;
  float energy = 0.0f;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      energy += coeff * (K[q] * K[j]) *
                FemDot3(y[q][0], y[q][1], y[q][2], y[j][0], y[j][1], y[j][2]);
  energy *= 0.5f;
  if (iter == 0) *lam_ptr = 0.0f;
  if (fabsf(energy) < kEps) return;
  float d[4][3];
#pragma unroll
  for (int j = 0; j < 4; ++j) d[j][0] = d[j][1] = d[j][2] = 0.0f;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float q = coeff * (K[i] * K[j]);
      d[j][0] += q * y[i][0];
      d[j][1] += q * y[i][1];
      d[j][2] += q * y[i][2];
    }
  float div = 0.0f;
#pragma unroll
  for (int j = 0; j < 4; ++j)
    div += x[j].w * FemDot3(d[j][0], d[j][1], d[j][2], d[j][0], d[j][1], d[j][2]);
  const float dl = (-energy - k.iso_at * lam) / (div + k.iso_at);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float f = dl * x[j].w;
    x[j].x += f * d[j][0];
    x[j].y += f * d[j][1];
    x[j].z += f * d[j][2];
  }
  X[id.x] = x[0];
  X[id.y] = x[1];
  X[id.z] = x[2];
  X[id.w] = x[3];
  *lam_ptr = lam + dl;
}

// One color of strain elements [beg, end) inside a workgroup (the persistent kernel): U
// elements per thread per round, their element records loaded before any is projected so
// that the loads are in flight together (elements of a color are independent, so the
// results do not depend on U).
template <int U, typename P>
__device__ __forceinline__ void StrainColor(P X, const StrainC* __restrict__ sc, int beg, int end,
                                            const FemConsts& k, int tid, int bs) {
  for (int e0 = beg + tid; e0 < end; e0 += U * bs) {
    StrainC c[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (e0 + u * bs < end) c[u] = sc[e0 + u * bs];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (e0 + u * bs < end) StrainProject(X, c[u], k);
  }
}

// One color of isometric bending elements [beg, end), U per thread per round (as above).
template <int U, typename P>
__device__ __forceinline__ void IsoColor(P X, const int4* __restrict__ bidx,
                                         const float4* __restrict__ iso_k,
                                         const float* __restrict__ iso_c, float* __restrict__ lam,
                                         int beg, int end, int iter, const FemConsts& k, int tid,
                                         int bs) {
  for (int e0 = beg + tid; e0 < end; e0 += U * bs) {
    int4 id[U];
    float4 K[U];
    float cf[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * bs;
      if (e < end) {
        id[u] = bidx[e];
        K[u] = iso_k[e];
        cf[u] = iso_c[e];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (e0 + u * bs < end) IsoProject(X, id[u], K[u], cf[u], lam + e0 + u * bs, iter, k);
  }
}

// one color of one instance
hipError_t LaunchStrain(hipStream_t s, const StrainC* sc, float4* x, int cnt, const FemConsts& k);
hipError_t LaunchIsoBend(hipStream_t s, const int4* bidx, const float4* iso_k, const float* iso_c,
                         float* lam, float4* x, int cnt, int iter, const FemConsts& k);
// one color of every instance (per-color block map: n_inst + 1 prefix sums, then the
// instance of every block; kFemPerBlock faces or kFemBlock hinges per block)
hipError_t LaunchWideStrain(hipStream_t s, const FemArgs& a, int n_inst, int color,
                            const int32_t* map, int blocks, const FemConsts& k);
hipError_t LaunchWideIsoBend(hipStream_t s, const FemArgs& a, int n_inst, int color,
                             const int32_t* map, int blocks, int iter, const FemConsts& k);

}  // namespace xrt
