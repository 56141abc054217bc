// XPBD dynamics kernels for gfx950: predict, colored stretch / bend projection,
// geodesic long-range attachments, ground plane, finalize.
//
// Two launch structures share the same per-element device functions:
//  * plain   -- one launch per kernel per color per instance (the reference's structure,
//               ClothSolverBase::SimulateSwift, xrtailor/src/pipeline/base/ClothSolverBase.cpp:190-240)
//  * batched -- one persistent workgroup per instance runs a whole substep (predict, all
//               iterations and colors, finalize) with workgroup barriers between color
//               batches; the iterate x lives in LDS when the instance fits.
#include "dynamics.h"
#include "elements.h"
#include "fem.h"

namespace xrt {

// PredictPositions (ClothSolverHelper.cu:125-131): v += g dt for free vertices, x = x0 + v dt,
// each written as one fused multiply-add so that every kernel that inlines it rounds the
// same way (a product hoisted out of a loop would otherwise round separately)
__device__ __forceinline__ float4 PredictVertex(const float4 p, float4& vi, const StepConsts& k) {
  if (p.w > kEps) {
    vi.x = fmaf(k.gx, k.dt, vi.x);
    vi.y = fmaf(k.gy, k.dt, vi.y);
    vi.z = fmaf(k.gz, k.dt, vi.z);
  }
  return make_float4(fmaf(vi.x, k.dt, p.x), fmaf(vi.y, k.dt, p.y), fmaf(vi.z, k.dt, p.z), p.w);
}

// ------------------------------------------------------------- plain kernels
__global__ __launch_bounds__(256) void PredictKernel(const float4* __restrict__ x0,
                                                     float4* __restrict__ v,
                                                     float4* __restrict__ x, int n, StepConsts k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 vi = v[i];
  x[i] = PredictVertex(x0[i], vi, k);
  v[i] = vi;
}

__global__ __launch_bounds__(256) void StretchKernel(const StretchC* __restrict__ sc,
                                                     float* __restrict__ lam,
                                                     float4* __restrict__ x, int cnt,
                                                     float inv_dt2, int iter) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cnt) return;
  StretchProject(x, sc[k], lam + k, inv_dt2, iter);
}

__global__ __launch_bounds__(256) void BendKernel(const int4* __restrict__ bidx,
                                                  const float2* __restrict__ bpar,
                                                  float4* __restrict__ x, int cnt, float inv_dt2,
                                                  BendMode m) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cnt) return;
  BendProject(x, bidx[k], bpar[k], inv_dt2, m);
}

__global__ __launch_bounds__(256) void GatherSlotsKernel(const float4* __restrict__ x,
                                                         const int32_t* __restrict__ slot_vert,
                                                         float4* __restrict__ slot_pos, int ns) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < ns) slot_pos[j] = x[slot_vert[j]];
}

__global__ __launch_bounds__(256) void LraKernel(float4* __restrict__ x,
                                                 const float4* __restrict__ slot_pos,
                                                 const int32_t* __restrict__ ptr,
                                                 const int32_t* __restrict__ slot,
                                                 const float* __restrict__ rest, int n,
                                                 float stretchiness) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int kb = ptr[i], ke = ptr[i + 1];
  if (kb == ke) return;
  x[i] = LraVertex(x[i], slot_pos, slot, rest, kb, ke, stretchiness);
}

__global__ __launch_bounds__(256) void GroundKernel(float4* __restrict__ x,
                                                    const float4* __restrict__ x0, int n,
                                                    float ground_y, float mu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = GroundVertex(x[i], x0[i], ground_y, mu);
}

__global__ __launch_bounds__(256) void FinalizeKernel(float4* __restrict__ x0,
                                                      float4* __restrict__ v,
                                                      const float4* __restrict__ x, int n,
                                                      float inv_dt, float damp_factor) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = x[i];
  const float4 q = x0[i];
  v[i] = make_float4((p.x - q.x) * inv_dt * damp_factor, (p.y - q.y) * inv_dt * damp_factor,
                     (p.z - q.z) * inv_dt * damp_factor, 0.0f);
  x0[i] = p;
}

// ------------------------------------------------------------ batched kernel
// Constraints per thread per round of the persistent kernel (kLds: iterate in LDS; global:
// more gathers in flight), by unroll variant: 0 (default), 1 (light), 2 (heavy).  The variant
// changes only how many independent constraints of one color a thread loads before solving
// them, never the result.
template <bool kLds, int kVariant>
struct Unroll {
  static constexpr int s = kVariant == 1 ? (kLds ? 1 : 2) : kVariant == 2 ? (kLds ? 4 : 8) : (kLds ? 2 : 4);
  static constexpr int b = kVariant == 1 ? 1 : kVariant == 2 ? (kLds ? 2 : 4) : (kLds ? 1 : 2);
  static constexpr int f = kVariant == 1 ? 2 : kVariant == 2 ? 8 : 4;
};

template <bool kLds, int kVariant>
__device__ __forceinline__ void SubstepBody(const BatchArgs& a, const InstDesc& d, int inst,
                                            const StepConsts& k, float4* lds_x,
                                            float4* lds_slots) {
  const int tid = threadIdx.x, bs = blockDim.x;
  float4* __restrict__ gx0 = a.x0 + d.vbeg;
  float4* __restrict__ gv = a.v + d.vbeg;
  float4* __restrict__ gx = a.x + d.vbeg;
  float4* X = kLds ? lds_x : gx;
  const int nv = d.nv;

  // predict
  for (int i = tid; i < nv; i += bs) {
    float4 vi = gv[i];
    X[i] = PredictVertex(gx0[i], vi, k);
    gv[i] = vi;
  }
  __syncthreads();

  const int32_t* sco = a.s_coloff + d.s_cbeg;
  const int32_t* bco = a.b_coloff + d.b_cbeg;
  const int32_t* lptr = a.lra_ptr + d.lra_vbeg;
  // constraints per thread per round: more gathers in flight where x lives in global memory
  constexpr int kUs = Unroll<kLds, kVariant>::s, kUb = Unroll<kLds, kVariant>::b,
                kUf = Unroll<kLds, kVariant>::f;
  const int2 fc = a.fem ? a.fcol[inst] : make_int2(0, 0);
  for (int it = 0; it < k.iterations; ++it) {
    if (a.fem) {  // quality mode: FEM strain colors, then isometric bending colors
      for (int c = 0; c < fc.y; ++c) {
        StrainColor<kUf>(X, a.fsc, a.f_coloff[fc.x + c], a.f_coloff[fc.x + c + 1], a.fk, tid, bs);
        __syncthreads();
      }
      if (k.solve_bending)
        for (int c = 0; c < d.b_ncol; ++c) {
          IsoColor<kUf>(X, a.bidx, a.iso_k, a.iso_c, a.iso_lam, bco[c], bco[c + 1], it, a.fk, tid,
                        bs);
          __syncthreads();
        }
    }
    for (int c = 0; c < (a.fem ? 0 : d.s_ncol); ++c) {
      StretchColor<kUs>(X, a.sc, a.slam, sco[c], sco[c + 1], k.inv_dt2, it, tid, bs);
      __syncthreads();
    }
    if (k.solve_bending && !a.fem) {
      const BendMode m{k.bend_dlo, k.bend_th0, k.bend_th1, k.bend_sgn};
      for (int c = 0; c < d.b_ncol; ++c) {
        BendColor<kUb>(X, a.bidx, a.bpar, bco[c], bco[c + 1], k.inv_dt2, m, tid, bs);
        __syncthreads();
      }
    }
    if (d.nslots > 0) {
      for (int j = tid; j < d.nslots; j += bs) lds_slots[j] = X[a.slot_vert[d.slot_beg + j]];
      __syncthreads();
    }
    if (d.nslots > 0 || k.enable_ground) {
      for (int i = tid; i < nv; i += bs) {
        float4 xi = X[i];
        if (d.nslots > 0) {
          const int kb = lptr[i], ke = lptr[i + 1];
          if (kb != ke) xi = LraVertex(xi, lds_slots, a.lra_slot, a.lra_rest, kb, ke, k.stretchiness);
        }
        if (k.enable_ground) xi = GroundVertex(xi, gx0[i], k.ground_y, k.friction);
        X[i] = xi;
      }
      __syncthreads();
    }
  }

  // finalize
  const float inv_dt = 1.0f / k.dt;
  for (int i = tid; i < nv; i += bs) {
    const float4 p = X[i];
    const float4 q = gx0[i];
    gv[i] = make_float4((p.x - q.x) * inv_dt * k.damp_factor, (p.y - q.y) * inv_dt * k.damp_factor,
                        (p.z - q.z) * inv_dt * k.damp_factor, 0.0f);
    gx0[i] = p;
    if (kLds) gx[i] = p;
  }
}

template <int kVariant>
__global__ __launch_bounds__(1024) void BatchedSubstepKernel(BatchArgs a, StepConsts k) {
  extern __shared__ float4 lds[];
  const int inst = a.order[blockIdx.x];
  const InstDesc d = a.inst[inst];
  float4* lds_slots = lds + a.lds_verts;
  if (d.nv <= a.lds_verts)
    SubstepBody<true, kVariant>(a, d, inst, k, lds, lds_slots);
  else
    SubstepBody<false, kVariant>(a, d, inst, k, lds, lds_slots);
}

// ------------------------------------------------------------------ launchers
static inline int Blocks(int n, int bs) { return (n + bs - 1) / bs; }

hipError_t LaunchPredict(hipStream_t s, const float4* x0, float4* v, float4* x, int n,
                         const StepConsts& k) {
  if (n <= 0) return hipSuccess;
  PredictKernel<<<Blocks(n, 256), 256, 0, s>>>(x0, v, x, n, k);
  return hipGetLastError();
}
hipError_t LaunchStretch(hipStream_t s, const StretchC* sc, float* lam, float4* x, int cnt,
                         float inv_dt2, int iter) {
  if (cnt <= 0) return hipSuccess;
  StretchKernel<<<Blocks(cnt, 256), 256, 0, s>>>(sc, lam, x, cnt, inv_dt2, iter);
  return hipGetLastError();
}
hipError_t LaunchBend(hipStream_t s, const int4* bidx, const float2* bpar, float4* x, int cnt,
                      float inv_dt2, const BendMode& m) {
  if (cnt <= 0) return hipSuccess;
  BendKernel<<<Blocks(cnt, 256), 256, 0, s>>>(bidx, bpar, x, cnt, inv_dt2, m);
  return hipGetLastError();
}
hipError_t LaunchLra(hipStream_t s, float4* x, const int32_t* slot_vert, float4* slot_pos, int ns,
                     const int32_t* ptr, const int32_t* slot, const float* rest, int n,
                     float stretchiness) {
  if (ns <= 0 || n <= 0) return hipSuccess;
  GatherSlotsKernel<<<Blocks(ns, 256), 256, 0, s>>>(x, slot_vert, slot_pos, ns);
  LraKernel<<<Blocks(n, 256), 256, 0, s>>>(x, slot_pos, ptr, slot, rest, n, stretchiness);
  return hipGetLastError();
}
// every instance's LRA pass in two launches: the slot snapshot of all instances, then every
// target vertex (the same bodies and slot ordinals as LaunchLra per instance)
__global__ __launch_bounds__(256) void WideGatherSlotsKernel(const InstDesc* __restrict__ inst,
                                                             const float4* __restrict__ x,
                                                             const int32_t* __restrict__ slot_vert,
                                                             float4* __restrict__ slot_pos) {
  const InstDesc d = inst[blockIdx.y];
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j < d.nslots) slot_pos[d.slot_beg + j] = x[d.vbeg + slot_vert[d.slot_beg + j]];
}

__global__ __launch_bounds__(256) void WideLraKernel(const InstDesc* __restrict__ inst,
                                                     float4* __restrict__ x,
                                                     const float4* __restrict__ slot_pos,
                                                     const int32_t* __restrict__ ptr,
                                                     const int32_t* __restrict__ slot,
                                                     const float* __restrict__ rest,
                                                     float stretchiness) {
  const InstDesc d = inst[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (d.nslots == 0 || i >= d.nv) return;
  const int kb = ptr[d.lra_vbeg + i], ke = ptr[d.lra_vbeg + i + 1];
  if (kb == ke) return;
  x[d.vbeg + i] = LraVertex(x[d.vbeg + i], slot_pos + d.slot_beg, slot, rest, kb, ke, stretchiness);
}

hipError_t LaunchWideLra(hipStream_t s, const InstDesc* inst, int n_inst, int max_slots, int max_nv,
                         float4* x, const int32_t* slot_vert, float4* slot_pos_all,
                         const int32_t* ptr, const int32_t* slot, const float* rest,
                         float stretchiness) {
  if (n_inst <= 0 || max_slots <= 0 || max_nv <= 0) return hipSuccess;
  WideGatherSlotsKernel<<<dim3(Blocks(max_slots, 256), n_inst), 256, 0, s>>>(inst, x, slot_vert,
                                                                            slot_pos_all);
  WideLraKernel<<<dim3(Blocks(max_nv, 256), n_inst), 256, 0, s>>>(inst, x, slot_pos_all, ptr, slot,
                                                                  rest, stretchiness);
  return hipGetLastError();
}

hipError_t LaunchGround(hipStream_t s, float4* x, const float4* x0, int n, float ground_y,
                        float mu) {
  if (n <= 0) return hipSuccess;
  GroundKernel<<<Blocks(n, 256), 256, 0, s>>>(x, x0, n, ground_y, mu);
  return hipGetLastError();
}
hipError_t LaunchFinalize(hipStream_t s, float4* x0, float4* v, const float4* x, int n,
                          const StepConsts& k) {
  if (n <= 0) return hipSuccess;
  FinalizeKernel<<<Blocks(n, 256), 256, 0, s>>>(x0, v, x, n, 1.0f / k.dt, k.damp_factor);
  return hipGetLastError();
}
hipError_t LaunchBatchedSubstep(hipStream_t s, const BatchArgs& a, const StepConsts& k,
                                int n_instances, int block, size_t lds_bytes, int unroll) {
  if (n_instances <= 0) return hipSuccess;
  if (block < 64 || block > 1024 || block % 64 != 0 || unroll < 0 || unroll > 2)
    return hipErrorInvalidValue;
  if (unroll == 1)
    BatchedSubstepKernel<1><<<n_instances, block, lds_bytes, s>>>(a, k);
  else if (unroll == 2)
    BatchedSubstepKernel<2><<<n_instances, block, lds_bytes, s>>>(a, k);
  else
    BatchedSubstepKernel<0><<<n_instances, block, lds_bytes, s>>>(a, k);
  return hipGetLastError();
}
hipError_t ConfigureBatchedKernel(size_t lds_bytes) {
  const void* fns[3] = {reinterpret_cast<const void*>(&BatchedSubstepKernel<0>),
                        reinterpret_cast<const void*>(&BatchedSubstepKernel<1>),
                        reinterpret_cast<const void*>(&BatchedSubstepKernel<2>)};
  for (const void* f : fns) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             static_cast<int>(lds_bytes));
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace xrt
