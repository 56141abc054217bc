// Wide (all instances per launch) forward kernels with an optional tape, and the
// reverse-mode (adjoint) kernels of the substep.
//
// A wide launch processes one color of every instance: grid (ceil(max_cnt / 256), n_inst),
// blockIdx.y = instance, so a whole substep is a fixed sequence of launches whatever the
// number of instances.  The forward replays bit for bit: the tape-writing and the plain
// wide forward are the same kernel.
//
// Tape of one substep (float4 units, all instances concatenated):
//   x0s                 N        start-of-substep x0
//   per iteration:
//     stretch           2 E      (x_i.xyz, lambda before), (x_j.xyz, -) per constraint
//     bend              3 B      x1..x4 packed (12 floats) per element
//     lra               N        x before the long-range attachments (when any instance
//                                has slots)
//     ground            N        x before the ground plane (when enabled)
//     contact           N        x before the VF solves (when obstacles are in contact)
//     ee                N        x before the EE solve
// Adjoint state (FP64): gx0, gv, gx double4[N] (w unused); glam double[E]; gcomp double[E],
// gbcomp double[B] accumulate d/d compliance.
#pragma once
#include <hip/hip_runtime.h>

#include "bvh.h"
#include "contact.h"
#include "device.h"

namespace xrt {

struct WideArgs {
  const InstDesc* inst;
  const int32_t* s_coloff;
  const int32_t* b_coloff;
  const StretchC* sc;
  float* slam;
  const int4* bidx;
  const float2* bpar;
  float4* x;         // iterate
  const float4* xw;  // a state array whose .w is the inverse mass (reverse kernels)
};

// ------------------------------------------------------------------ forward
// map (optional): this color's block map -- n_inst + 1 prefix sums of the instances' block
// counts, then the instance of every block (`blocks` in total); without it the grid is
// (blocks of the largest color, n_inst); per_thread 2 (untaped, with a map): two elements per
// thread in half-size blocks (same results)
hipError_t LaunchWideStretch(hipStream_t s, const WideArgs& a, int n_inst, int color, int max_cnt,
                             float inv_dt2, int iter, float4* tape, const int32_t* map = nullptr,
                             int blocks = 0, int per_thread = 1);
hipError_t LaunchWideBend(hipStream_t s, const WideArgs& a, int n_inst, int color, int max_cnt,
                          float inv_dt2, const BendMode& m, float4* tape,
                          const int32_t* map = nullptr, int blocks = 0, int per_thread = 1);
hipError_t LaunchCopy4(hipStream_t s, float4* dst, const float4* src, int n);
hipError_t LaunchZero4(hipStream_t s, float4* p, int n);
hipError_t LaunchZero1(hipStream_t s, float* p, int n);

// ------------------------------------------------------------------ reverse
// geodesic LRA pass reversed (SolveGeodesicLRA, LongRangeAttachments.cu:5-26, as the
// forward's per-target slot sequence): xpre = the tape's pre-LRA state of every instance;
// each target replays its slots into xs (one float4 per constraint) and walks them back,
// leaving each constraint's slot adjoint in gslot; then every slot's adjoints are summed in
// constraint order (slot_ptr / slot_cons: the constraints of each global slot) into the slot
// vertex's gx.
hipError_t LaunchWideLraAdj(hipStream_t s, const InstDesc* inst, int n_inst, int max_nv,
                            int total_slots, const float4* xpre, const int32_t* slot_vert,
                            const int32_t* ptr, const int32_t* slot, const float* rest,
                            float stretchiness, float4* xs, double4* gslot,
                            const int32_t* slot_inst, const int32_t* slot_ptr,
                            const int32_t* slot_cons, double4* gx);
hipError_t LaunchZeroD4(hipStream_t s, double4* p, int n);
hipError_t LaunchZeroD1(hipStream_t s, double* p, int n);
// finalize: gx = gx0 + gv * f ; gx0 = -gv * f      (f = damp_factor / dt)
hipError_t LaunchFinalizeAdj(hipStream_t s, double4* gx0, const double4* gv, double4* gx, int n,
                             double f);
// predict: gx0 += gx ; gv = gx * dt
hipError_t LaunchPredictAdj(hipStream_t s, double4* gx0, double4* gv, const double4* gx, int n,
                            double dt);
hipError_t LaunchGroundAdj(hipStream_t s, const float4* xpre, const float4* x0s, double4* gx,
                           double4* gx0, int n, float ground_y, float mu);
hipError_t LaunchWideStretchAdj(hipStream_t s, const WideArgs& a, int n_inst, int color,
                                int max_cnt, float inv_dt2, int iter, const float4* tape,
                                double4* gx, double* glam, double* gcomp);
hipError_t LaunchWideBendAdj(hipStream_t s, const WideArgs& a, int n_inst, int color, int max_cnt,
                             float inv_dt2, const BendMode& m, const float4* tape, double4* gx,
                             double* gbcomp);
// contacts (adjoint_contact.hip, built FP-strict like the forward contact kernels so that
// the replayed activity tests decide as the forward did); one instance per call
// obs_pred: the obstacle's contact positions (x1 + tol n1) of the pass that made the records
hipError_t LaunchVfAdj(hipStream_t s, const float4* x_pc, int n, const int4* obs_tris,
                       const float4* obs_pred, const int32_t* rt_face, const float4* rt_q,
                       const int32_t* vf_face, const float4* vf_q, double4* gx);
// cnt (int[nv]) is zero on entry and on return; gl (double[ne]) is scratch; ve_ptr / ve_edge:
// the edges of each vertex (CSR, ascending edge id), over which each vertex gathers its
// adjoint terms in a fixed order (no floating-point atomics: the reverse is bitwise
// reproducible)
hipError_t LaunchEeAdj(hipStream_t s, const float4* x_pe, const int2* edges, int ne, int nv,
                       const EeRec* rec, float relax, double4* gx, int32_t* cnt, double* gl,
                       const int32_t* ve_ptr, const int32_t* ve_edge);

}  // namespace xrt
