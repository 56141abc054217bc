// Swift-mode particle self-collision kernels for gfx950 (built with -ffp-contract=off and
// correctly rounded division / square root: the hash cells and the neighbor tests are
// integer outputs of FP32 arithmetic and agree bit for bit with the FP32 oracle).
//
// Every kernel runs the instances of a block map (all instances for the batched substep,
// one for the plain path), a few instances per XCD.  The hash is a stable LSD radix sort
// of (cell hash, particle) per instance -- the order of the reference's stable radix sort
// -- so the cell table and the neighbor lists are deterministic.  The neighbor search
// walks particles in sorted (cell) order, so the lanes of a wave share their 27 cells and
// the cells' particles stay in L2; the collision pass walks particles in index order and
// reads the column-major lists with unit stride across lanes.
#include <hip/hip_runtime.h>

#include "bvh_device.h"
#include "device.h"
#include "particle.h"

namespace xrt {

namespace {
constexpr int kBlock = 256;

// HashCoords (SpatialHashGPU.cu:13-16): 32-bit wrapping products, C remainder, abs
__device__ __forceinline__ int32_t HashCoords(int32_t x, int32_t y, int32_t z, int32_t table) {
  const uint32_t h = (static_cast<uint32_t>(x) * 92837111u) ^
                     (static_cast<uint32_t>(y) * 689287499u) ^
                     (static_cast<uint32_t>(z) * 283923481u);
  const int32_t r = static_cast<int32_t>(h) % table;
  return r < 0 ? -r : r;
}

// ComputeIntCoord (SpatialHashGPU.cu:9-11)
__device__ __forceinline__ int32_t IntCoord(float v, float spacing) {
  return static_cast<int32_t>(floorf(v / spacing));
}

__device__ __forceinline__ float Len2(float dx, float dy, float dz) {
  return dx * dx + dy * dy + dz * dz;
}

#define XRT_PARTICLE_INDEX()                                     \
  const int lblk = XcdBlock(off, n);                             \
  if (lblk < 0) return;                                          \
  const int ent = InstOfBlock(off, n, lblk);                     \
  const ParticleInstDev& c = b.inst[InstOfEntry(off, n, ent)];   \
  const int lb = lblk - off[ent];                                \
  const int i = lb * kBlock + threadIdx.x;                       \
  (void)i;

// ComputeParticleHash_Kernel (SpatialHashGPU.cu:135-140) plus the clear of the instance's
// cell_start table (2 nv entries, two per particle) that FindCellStart relies on
__global__ __launch_bounds__(kBlock) void ParticleHashKernel(ParticleBufs b, const int32_t* off,
                                                             int n, const float4* __restrict__ x) {
  XRT_PARTICLE_INDEX()
  if (i >= c.nv) return;
  const float4 p = x[c.vbeg + i];
  b.keys[c.vbeg + i] = static_cast<uint32_t>(
      HashCoords(IntCoord(p.x, c.spacing), IntCoord(p.y, c.spacing), IntCoord(p.z, c.spacing), c.table));
  b.vals[c.vbeg + i] = i;
  uint32_t* cs = b.cell_start + 2 * static_cast<int64_t>(c.vbeg);
  cs[i] = 0xffffffffu;
  cs[c.nv + i] = 0xffffffffu;
}

// radix pass p: even passes read (keys, vals), odd ones the tmp buffers
__global__ __launch_bounds__(kBlock) void ParticleRadixHistKernel(ParticleBufs b, const int32_t* off,
                                                                  int n, int pass) {
  const int lblk = XcdBlock(off, n);
  if (lblk < 0) return;
  const int ent = InstOfBlock(off, n, lblk);
  const ParticleInstDev& c = b.inst[InstOfEntry(off, n, ent)];
  const uint32_t* keys = (pass & 1 ? b.keys_tmp : b.keys) + c.vbeg;
  RadixHistBody(keys, c.nv, 8 * pass, c.nblocks, b.hist + c.hbeg, lblk - off[ent]);
}

// one workgroup per entry of the tile map
__global__ __launch_bounds__(1024) void ParticleRadixScanKernel(ParticleBufs b, const int32_t* off,
                                                                int n) {
  const ParticleInstDev& c = b.inst[InstOfEntry(off, n, blockIdx.x)];
  BlockExclusiveScan(b.hist + c.hbeg, 256 * c.nblocks, nullptr);
}

__global__ __launch_bounds__(kBlock) void ParticleRadixScatterKernel(ParticleBufs b,
                                                                     const int32_t* off, int n,
                                                                     int pass) {
  const int lblk = XcdBlock(off, n);
  if (lblk < 0) return;
  const int ent = InstOfBlock(off, n, lblk);
  const ParticleInstDev& c = b.inst[InstOfEntry(off, n, ent)];
  const bool odd = pass & 1;
  RadixScatterBody((odd ? b.keys_tmp : b.keys) + c.vbeg, (odd ? b.vals_tmp : b.vals) + c.vbeg,
                   (odd ? b.keys : b.keys_tmp) + c.vbeg, (odd ? b.vals : b.vals_tmp) + c.vbeg,
                   c.nv, 8 * pass, c.nblocks, b.hist + c.hbeg, lblk - off[ent]);
}

// FindCellStart_Kernel (SpatialHashGPU.cu:27-54) over the sorted keys
__global__ __launch_bounds__(kBlock) void ParticleCellKernel(ParticleBufs b, const int32_t* off,
                                                             int n) {
  XRT_PARTICLE_INDEX()
  if (i >= c.nv) return;
  const uint32_t* key = b.keys + c.vbeg;
  uint32_t* cs = b.cell_start + 2 * static_cast<int64_t>(c.vbeg);
  uint32_t* ce = b.cell_end + 2 * static_cast<int64_t>(c.vbeg);
  const uint32_t h = key[i];
  if (i == 0 || h != key[i - 1]) {
    cs[h] = static_cast<uint32_t>(i);
    if (i > 0) ce[key[i - 1]] = static_cast<uint32_t>(i);
  }
  if (i == c.nv - 1) ce[h] = static_cast<uint32_t>(c.nv);
}

// CacheNeighbors_Kernel (SpatialHashGPU.cu:56-97): thread t takes the particle at sorted
// position t; the 27 cells around its cell (x-major), each cell's first max_nb particles,
// those within the cell spacing now and farther apart than the diameter at the start.
// The 27 cell ranges are loaded up front and each cell's candidates kChunkN at a time
// (independent loads in flight); the tests and the writes keep the reference's order.
constexpr int kChunkN = 4;

__global__ __launch_bounds__(kBlock) void ParticleNeighborKernel(ParticleBufs b, const int32_t* off,
                                                                 int n, const float4* __restrict__ x) {
  XRT_PARTICLE_INDEX()
  if (i >= c.nv) return;
  const int32_t* idx = b.vals + c.vbeg;
  const float4* xp = x + c.vbeg;
  const float4* op = b.orig + c.vbeg;
  const uint32_t* cs = b.cell_start + 2 * static_cast<int64_t>(c.vbeg);
  const uint32_t* ce = b.cell_end + 2 * static_cast<int64_t>(c.vbeg);
  uint32_t* list = b.nb + c.nbeg;
  const int id = idx[i];
  const float4 p = xp[id], o = op[id];
  const int32_t ix = IntCoord(p.x, c.spacing), iy = IntCoord(p.y, c.spacing), iz = IntCoord(p.z, c.spacing);
  uint32_t lo[27], hi[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const int32_t h = HashCoords(ix - 1 + q / 9, iy - 1 + (q / 3) % 3, iz - 1 + q % 3, c.table);
    lo[q] = cs[h];
    hi[q] = ce[h];
  }
  const int64_t cap = static_cast<int64_t>(c.nv) * c.max_nb;
  int64_t at = id;
#pragma unroll 1
  for (int q = 0; q < 27; ++q) {
    const uint32_t start = lo[q];
    if (start == 0xffffffffu) continue;
    const uint32_t lim = start + static_cast<uint32_t>(c.max_nb);
    const uint32_t end = hi[q] < lim ? hi[q] : lim;
    for (uint32_t s0 = start; s0 < end; s0 += kChunkN) {
      int jj[kChunkN];
      float4 qq[kChunkN], rr[kChunkN];
#pragma unroll
      for (int u = 0; u < kChunkN; ++u) jj[u] = s0 + u < end ? idx[s0 + u] : id;
#pragma unroll
      for (int u = 0; u < kChunkN; ++u) {
        qq[u] = xp[jj[u]];
        rr[u] = op[jj[u]];
      }
#pragma unroll
      for (int u = 0; u < kChunkN; ++u) {
        if (jj[u] == id) continue;  // the particle itself, or past the cell's end
        if (!(Len2(p.x - qq[u].x, p.y - qq[u].y, p.z - qq[u].z) < c.spacing2)) continue;
        if (!(Len2(o.x - rr[u].x, o.y - rr[u].y, o.z - rr[u].z) > c.diameter2)) continue;
        list[at] = static_cast<uint32_t>(jj[u]);
        at += c.nv;
        if (at >= cap) return;
      }
    }
  }
  list[at] = 0xffffffffu;  // at < cap here
}

// CollideParticles_Kernel (ClothSolverHelper.cu:171-211) with ComputeFriction (:156-169).
// The list ends at the first entry >= nv (the 0xffffffff sentinel; the reference tests
// j > num_particles, the same for every list the neighbor pass writes).  Entries and the
// neighbors' positions are fetched kChunk at a time (independent loads in flight), then
// accumulated in list order, so the sums round exactly as in the one-by-one loop.
constexpr int kChunk = 8;

__global__ __launch_bounds__(kBlock) void ParticleCollideKernel(ParticleBufs b, const int32_t* off,
                                                                int n, const float4* __restrict__ x,
                                                                const float4* __restrict__ x0,
                                                                float mu) {
  XRT_PARTICLE_INDEX()
  if (i >= c.nv) return;
  const float4* xp = x + c.vbeg;
  const float4* x0p = x0 + c.vbeg;
  const uint32_t* list = b.nb + c.nbeg;
  const uint32_t nv = static_cast<uint32_t>(c.nv);
  const float4 p = xp[i], p0 = x0p[i];
  const float vx = p.x - p0.x, vy = p.y - p0.y, vz = p.z - p0.z;
  const float wi = p.w;
  float dx = 0.0f, dy = 0.0f, dz = 0.0f;
  int cnt = 0;
  bool done = false;
  for (int s0 = 0; s0 < c.max_nb && !done; s0 += kChunk) {
    uint32_t jj[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; ++u)
      jj[u] = s0 + u < c.max_nb ? list[static_cast<int64_t>(s0 + u) * c.nv + i] : 0xffffffffu;
    int len = kChunk;  // entries of this chunk before the end of the list
#pragma unroll
    for (int u = kChunk - 1; u >= 0; --u)
      if (jj[u] >= nv) len = u;
    done = len < kChunk;
    float4 q[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; ++u) q[u] = u < len ? xp[jj[u]] : p;
    auto pair = [&](const float4 q, uint32_t j) {
      const float denom = wi + q.w;
      if (denom <= 0.0f) return;
      const float ex = p.x - q.x, ey = p.y - q.y, ez = p.z - q.z;
      const float dist = sqrtf(Len2(ex, ey, ez));
      if (dist >= c.diameter) return;
      const float de = dist + kEps;
      const float gx = ex / de, gy = ey / de, gz = ez / de;
      const float lambda = (dist - c.diameter) / denom;
      const float cx = lambda * gx, cy = lambda * gy, cz = lambda * gz;
      ++cnt;
      dx -= wi * cx; dy -= wi * cy; dz -= wi * cz;
      const float4 q0 = x0p[j];
      const float rx = vx - (q.x - q0.x), ry = vy - (q.y - q0.y), rz = vz - (q.z - q0.z);
      const float cl = sqrtf(Len2(cx, cy, cz));
      if (mu > 0.0f && cl > 0.0f) {
        const float nx = cx / cl, ny = cy / cl, nz = cz / cl;
        const float vn = rx * nx + ry * ny + rz * nz;
        const float tx = rx - nx * vn, ty = ry - ny * vn, tz = rz - nz * vn;
        const float tl = sqrtf(Len2(tx, ty, tz));
        const float r = (cl * mu) / tl;
        const float sc = r < 1.0f ? r : 1.0f;  // MathFunctions::min
        dx += wi * (-tx * sc); dy += wi * (-ty * sc); dz += wi * (-tz * sc);
      }
    };
#pragma unroll
    for (int u = 0; u < kChunk; ++u)
      if (u < len) pair(q[u], jj[u]);
  }
  b.delta[c.vbeg + i] = make_float4(dx, dy, dz, static_cast<float>(cnt));
}

// ApplyDeltas_Kernel (ClothSolverHelper.cu:139-148)
__global__ __launch_bounds__(kBlock) void ParticleApplyKernel(ParticleBufs b, const int32_t* off,
                                                              int n, float4* __restrict__ x,
                                                              float relax) {
  XRT_PARTICLE_INDEX()
  if (i >= c.nv) return;
  const float4 d = b.delta[c.vbeg + i];
  if (d.w > 0.0f) {
    float4 p = x[c.vbeg + i];
    p.x += d.x / d.w * relax;
    p.y += d.y / d.w * relax;
    p.z += d.z / d.w * relax;
    x[c.vbeg + i] = p;
  }
}
// list lengths (entries before the sentinel), one atomic per wave
__global__ __launch_bounds__(kBlock) void ParticleCountKernel(ParticleBufs b, const int32_t* off,
                                                              int n, unsigned long long* total) {
  const int lblk = XcdBlock(off, n);
  if (lblk < 0) return;
  const int ent = InstOfBlock(off, n, lblk);
  const ParticleInstDev& c = b.inst[InstOfEntry(off, n, ent)];
  const int i = (lblk - off[ent]) * kBlock + threadIdx.x;
  unsigned long long len = 0;
  if (i < c.nv) {
    const uint32_t* list = b.nb + c.nbeg;
    while (len < static_cast<unsigned long long>(c.max_nb) &&
           list[static_cast<int64_t>(len) * c.nv + i] < static_cast<uint32_t>(c.nv))
      ++len;
  }
  for (int o = 32; o > 0; o >>= 1) len += __shfl_down(len, o);
  if ((threadIdx.x & 63) == 0 && len) atomicAdd(total, len);
}
}  // namespace

hipError_t LaunchParticleHash(hipStream_t s, const ParticleBufs& b, const DevMap& vmap,
                              const DevMap& tmap, int passes, const float4* x) {
  if (vmap.blocks <= 0) return hipSuccess;
  ParticleHashKernel<<<XcdGrid(vmap.blocks), kBlock, 0, s>>>(b, vmap.off, vmap.n, x);
  for (int pass = 0; pass < passes; ++pass) {
    ParticleRadixHistKernel<<<XcdGrid(tmap.blocks), kSortBlock, 0, s>>>(b, tmap.off, tmap.n, pass);
    ParticleRadixScanKernel<<<tmap.n, 1024, 0, s>>>(b, tmap.off, tmap.n);
    ParticleRadixScatterKernel<<<XcdGrid(tmap.blocks), kSortBlock, 0, s>>>(b, tmap.off, tmap.n, pass);
  }
  ParticleCellKernel<<<XcdGrid(vmap.blocks), kBlock, 0, s>>>(b, vmap.off, vmap.n);
  ParticleNeighborKernel<<<XcdGrid(vmap.blocks), kBlock, 0, s>>>(b, vmap.off, vmap.n, x);
  return hipGetLastError();
}

hipError_t LaunchCollideParticles(hipStream_t s, const ParticleBufs& b, const DevMap& vmap,
                                  float4* x, const float4* x0, float mu, float relax,
                                  hipEvent_t ev_begin, hipEvent_t ev_end) {
  if (vmap.blocks <= 0) return hipSuccess;
  if (ev_begin) {
    const hipError_t e = hipEventRecord(ev_begin, s);
    if (e != hipSuccess) return e;
  }
  ParticleCollideKernel<<<XcdGrid(vmap.blocks), kBlock, 0, s>>>(b, vmap.off, vmap.n, x, x0, mu);
  if (ev_end) {
    const hipError_t e = hipEventRecord(ev_end, s);
    if (e != hipSuccess) return e;
  }
  ParticleApplyKernel<<<XcdGrid(vmap.blocks), kBlock, 0, s>>>(b, vmap.off, vmap.n, x, relax);
  return hipGetLastError();
}

hipError_t LaunchCountNeighbors(hipStream_t s, const ParticleBufs& b, const DevMap& vmap,
                                unsigned long long* total) {
  if (vmap.blocks <= 0) return hipSuccess;
  ParticleCountKernel<<<XcdGrid(vmap.blocks), kBlock, 0, s>>>(b, vmap.off, vmap.n, total);
  return hipGetLastError();
}

}  // namespace xrt
