// Small 3-vector helpers of the contact kernels.  Every file that includes this is built
// with -ffp-contract=off and correctly rounded division / square root, so the same
// expression gives the same bits in the forward kernels and in their adjoint replays.
#pragma once
#include <hip/hip_runtime.h>

namespace xrt {
namespace cm {

constexpr float kEpsC = 1e-6f;                 // EPSILON
constexpr float kScalarMax = 3.402823466e38f;  // SCALAR_MAX

struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 L3(float4 a) { return F3{a.x, a.y, a.z}; }
__device__ __forceinline__ F3 A(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 S(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 M(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float D(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 C(F3 a, F3 b) {
  return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float Len(F3 a) { return sqrtf(D(a, a)); }
__device__ __forceinline__ F3 Nz(F3 a) { return M(a, 1.0f / Len(a)); }
__device__ __forceinline__ float Mixed(F3 a, F3 b, F3 c) { return D(a, C(b, c)); }
__device__ __forceinline__ float Cl(float x, float a, float b) { return x < a ? a : (x > b ? b : x); }
__device__ __forceinline__ F3 FaceNormal(F3 a, F3 b, F3 c) { return Nz(C(S(b, a), S(c, a))); }

}  // namespace cm
}  // namespace xrt
