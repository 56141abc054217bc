// Quality-mode FEM constraint kernels for gfx950: one thread per element of a color (the
// elements of a color share no vertex, so the corrections are written directly).
#include <hip/hip_runtime.h>

#include "fem.h"

namespace xrt {

namespace {
constexpr int kBlock = kFemBlock;

__global__ __launch_bounds__(kBlock) void StrainKernel(const StrainC* __restrict__ sc,
                                                       float4* __restrict__ x, int cnt, FemConsts k) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < cnt) StrainProject(x, sc[e], k);
}

__global__ __launch_bounds__(kBlock) void IsoBendKernel(const int4* __restrict__ bidx,
                                                        const float4* __restrict__ iso_k,
                                                        const float* __restrict__ iso_c,
                                                        float* __restrict__ lam,
                                                        float4* __restrict__ x, int cnt, int iter,
                                                        FemConsts k) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < cnt) IsoProject(x, bidx[e], iso_k[e], iso_c[e], lam + e, iter, k);
}

// the wide launches' block of a per-color map (XCD-aware like every multi-instance map)
__device__ __forceinline__ bool MapBlock(const int32_t* __restrict__ map, int n_inst, int& inst,
                                         int& blk) {
  const int l = XcdBlock(map, n_inst);
  if (l < 0) return false;
  inst = InstOfBlock(map, n_inst, l);
  blk = l - map[inst];
  return true;
}

__global__ __launch_bounds__(kBlock) void WideStrainKernel(FemArgs a, int color,
                                                           const int32_t* __restrict__ map,
                                                           int n_inst, FemConsts k) {
  int inst, blk;
  if (!MapBlock(map, n_inst, inst, blk)) return;
  const int2 fc = a.fcol[inst];
  if (color >= fc.y) return;
  const int beg = a.f_coloff[fc.x + color], end = a.f_coloff[fc.x + color + 1];
  // two faces per thread (kFemPerBlock faces per block), both element records loaded
  // first: the faces of a color are independent, so the results do not depend on the grouping
  const int e = beg + blk * kFemPerBlock + threadIdx.x;
  if (e >= end) return;
  float4* X = a.x + a.inst[inst].vbeg;
  const StrainC c0 = a.sc[e];
  const bool two = e + kBlock < end;
  StrainC c1;
  if (two) c1 = a.sc[e + kBlock];
  StrainProject(X, c0, k);
  if (two) StrainProject(X, c1, k);
}

__global__ __launch_bounds__(kBlock) void WideIsoBendKernel(FemArgs a, int color,
                                                            const int32_t* __restrict__ map,
                                                            int n_inst, int iter, FemConsts k) {
  int inst, blk;
  if (!MapBlock(map, n_inst, inst, blk)) return;
  const InstDesc d = a.inst[inst];
  if (color >= d.b_ncol) return;
  const int beg = a.b_coloff[d.b_cbeg + color], end = a.b_coloff[d.b_cbeg + color + 1];
  // one hinge per thread (two per thread measured slower: 23.0 -> 25.4 us per launch)
  const int e = beg + blk * kBlock + threadIdx.x;
  if (e >= end) return;
  IsoProject(a.x + d.vbeg, a.bidx[e], a.iso_k[e], a.iso_c[e], a.iso_lam + e, iter, k);
}
}  // namespace

hipError_t LaunchStrain(hipStream_t s, const StrainC* sc, float4* x, int cnt, const FemConsts& k) {
  if (cnt <= 0) return hipSuccess;
  StrainKernel<<<(cnt + kBlock - 1) / kBlock, kBlock, 0, s>>>(sc, x, cnt, k);
  return hipGetLastError();
}

hipError_t LaunchIsoBend(hipStream_t s, const int4* bidx, const float4* iso_k, const float* iso_c,
                         float* lam, float4* x, int cnt, int iter, const FemConsts& k) {
  if (cnt <= 0) return hipSuccess;
  IsoBendKernel<<<(cnt + kBlock - 1) / kBlock, kBlock, 0, s>>>(bidx, iso_k, iso_c, lam, x, cnt,
                                                               iter, k);
  return hipGetLastError();
}

hipError_t LaunchWideStrain(hipStream_t s, const FemArgs& a, int n_inst, int color,
                            const int32_t* map, int blocks, const FemConsts& k) {
  if (blocks <= 0) return hipSuccess;
  WideStrainKernel<<<XcdGrid(blocks), kBlock, 0, s>>>(a, color, map, n_inst, k);
  return hipGetLastError();
}

hipError_t LaunchWideIsoBend(hipStream_t s, const FemArgs& a, int n_inst, int color,
                             const int32_t* map, int blocks, int iter, const FemConsts& k) {
  if (blocks <= 0) return hipSuccess;
  WideIsoBendKernel<<<XcdGrid(blocks), kBlock, 0, s>>>(a, color, map, n_inst, iter, k);
  return hipGetLastError();
}

}  // namespace xrt
