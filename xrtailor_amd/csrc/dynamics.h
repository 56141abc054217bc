// Launchers of the XPBD dynamics kernels (dynamics.hip).
#pragma once
#include "device.h"

namespace xrt {

hipError_t LaunchPredict(hipStream_t s, const float4* x0, float4* v, float4* x, int n,
                         const StepConsts& k);
hipError_t LaunchStretch(hipStream_t s, const StretchC* sc, float* lam, float4* x, int cnt,
                         float inv_dt2, int iter);
hipError_t LaunchBend(hipStream_t s, const int4* bidx, const float2* bpar, float4* x, int cnt,
                      float inv_dt2, const BendMode& m);
// LRA of every instance (slot_pos_all: one slot snapshot per slot of every instance)
hipError_t LaunchWideLra(hipStream_t s, const InstDesc* inst, int n_inst, int max_slots, int max_nv,
                         float4* x, const int32_t* slot_vert, float4* slot_pos_all,
                         const int32_t* ptr, const int32_t* slot, const float* rest,
                         float stretchiness);
hipError_t LaunchLra(hipStream_t s, float4* x, const int32_t* slot_vert, float4* slot_pos, int ns,
                     const int32_t* ptr, const int32_t* slot, const float* rest, int n,
                     float stretchiness);
hipError_t LaunchGround(hipStream_t s, float4* x, const float4* x0, int n, float ground_y,
                        float mu);
hipError_t LaunchFinalize(hipStream_t s, float4* x0, float4* v, const float4* x, int n,
                          const StepConsts& k);
// unroll: 0 default, 1 light, 2 heavy (constraints per thread per round; same results)
hipError_t LaunchBatchedSubstep(hipStream_t s, const BatchArgs& a, const StepConsts& k,
                                int n_instances, int block, size_t lds_bytes, int unroll);
hipError_t ConfigureBatchedKernel(size_t lds_bytes);

}  // namespace xrt
