// Swift-mode particle self-collision: spatial hash + neighbor cache + Jacobi collision
// (SpatialHashGPU, xrtailor/src/physics/broad_phase/spatial_hash/SpatialHashGPU.cu:9-176;
// CollideParticles, xrtailor/src/physics/ClothSolverHelper.cu:156-219).
//
// HBM layout (instances concatenated like the state, instance k at vertex offset vbeg):
//   keys, vals, keys_tmp, vals_tmp : uint32 / int32 [N]   hash sort ping-pong
//   cell_start, cell_end           : uint32 [2N]          cell table of instance k at 2 vbeg
//   nb                             : uint32 [N * max_nb]  neighbor lists of instance k at
//                                    vbeg * max_nb, slot s of particle i at s * nv + i
//   orig                           : float4 [N]           positions at xrt_start
//   delta                          : float4 [N]           collision deltas, w = count
//   hist                           : uint32               radix histograms (256 per tile)
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "blockmap.h"

namespace xrt {

struct ParticleInstDev {
  int32_t vbeg, nv;
  int32_t table;     // HashParams::table_size = 2 nv
  int32_t max_nb;    // HashParams::max_num_neighbors
  int32_t nblocks;   // radix tiles
  int32_t pad;
  int64_t nbeg;      // vbeg * max_nb
  int64_t hbeg;      // histogram offset (256 * nblocks entries)
  float diameter, spacing, spacing2, diameter2;  // HashParams (FP32 products, host)
};

struct ParticleBufs {
  const ParticleInstDev* inst;
  uint32_t *keys, *keys_tmp, *cell_start, *cell_end, *nb, *hist;
  int32_t *vals, *vals_tmp;
  const float4* orig;
  float4* delta;
};

// SpatialHashGPU::Hash of the instances of `vmap` / `tmap` (per-vertex and per-radix-tile
// maps over the same instances): hash x, stable radix sort (passes of 8 bits, an even
// count), cell table, neighbor lists.
hipError_t LaunchParticleHash(hipStream_t s, const ParticleBufs& b, const DevMap& vmap,
                              const DevMap& tmap, int passes, const float4* x);
// CollideParticles: Jacobi deltas from the neighbor lists, then ApplyDeltas x relax;
// events (optional) are recorded around the collision kernel alone
hipError_t LaunchCollideParticles(hipStream_t s, const ParticleBufs& b, const DevMap& vmap,
                                  float4* x, const float4* x0, float mu, float relax,
                                  hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);
// adds the number of list entries (up to each sentinel) of the map's instances to *total
hipError_t LaunchCountNeighbors(hipStream_t s, const ParticleBufs& b, const DevMap& vmap,
                                unsigned long long* total);

}  // namespace xrt
