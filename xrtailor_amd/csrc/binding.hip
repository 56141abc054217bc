// Obstacle skinning, nearest-face records and the binding constraint for gfx950 (built
// with -ffp-contract=off and correctly rounded division / square root: the nearest face is
// an integer output of FP32 comparisons and agrees bit for bit with the FP32 oracle).
#include <hip/hip_runtime.h>

#include "binding.h"
#include "bvh_device.h"
#include "device.h"

namespace xrt {

namespace {
constexpr int kBlock = 256;

__device__ __forceinline__ float Dot3(float ax, float ay, float az, float bx, float by, float bz) {
  return ax * bx + ay * by + az * bz;
}

// MinDist (BVHHelper.cu:114-126): squared distance from p to the box
__device__ __forceinline__ float MinDist(const Box& b, float4 p) {
  const float dx = fminf(b.hi.x, fmaxf(b.lo.x, p.x)) - p.x;
  const float dy = fminf(b.hi.y, fmaxf(b.lo.y, p.y)) - p.y;
  const float dz = fminf(b.hi.z, fmaxf(b.lo.z, p.z)) - p.z;
  return dx * dx + dy * dy + dz * dz;
}

// MinMaxDist (BVHHelper.cu:128-151)
__device__ __forceinline__ float MinMaxDist(const Box& b, float4 p) {
  float rmx = (b.lo.x - p.x) * (b.lo.x - p.x), rMx = (b.hi.x - p.x) * (b.hi.x - p.x);
  float rmy = (b.lo.y - p.y) * (b.lo.y - p.y), rMy = (b.hi.y - p.y) * (b.hi.y - p.y);
  float rmz = (b.lo.z - p.z) * (b.lo.z - p.z), rMz = (b.hi.z - p.z) * (b.hi.z - p.z);
  if ((b.hi.x + b.lo.x) * 0.5f < p.x) { const float t = rmx; rmx = rMx; rMx = t; }
  if ((b.hi.y + b.lo.y) * 0.5f < p.y) { const float t = rmy; rmy = rMy; rMy = t; }
  if ((b.hi.z + b.lo.z) * 0.5f < p.z) { const float t = rmz; rmz = rMz; rMz = t; }
  const float dx = rmx + rMy + rMz;
  const float dy = rMx + rmy + rMz;
  const float dz = rMx + rMy + rmz;
  const float m = dy < dz ? dy : dz;  // MathFunctions::min
  return dx < m ? dx : m;
}

// distance_calculator (BVHHelper.cu:9-103): Ericson's closest point of triangle abc with
// the barycentric weights of its region; returns |p - q|
__device__ float TriNearest(float4 p, float4 a, float4 b, float4 c, float& u, float& v, float& w) {
  const float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
  const float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
  const float e3x = c.x - b.x, e3y = c.y - b.y, e3z = c.z - b.z;
  float qx, qy, qz;
  const float v1x = p.x - a.x, v1y = p.y - a.y, v1z = p.z - a.z;
  const float d1 = Dot3(e1x, e1y, e1z, v1x, v1y, v1z), d2 = Dot3(e2x, e2y, e2z, v1x, v1y, v1z);
  if (d1 <= 0 && d2 <= 0) {
    u = 1; v = 0; w = 0;
    qx = a.x; qy = a.y; qz = a.z;
  } else {
    const float v2x = p.x - b.x, v2y = p.y - b.y, v2z = p.z - b.z;
    const float d3 = Dot3(e1x, e1y, e1z, v2x, v2y, v2z), d4 = Dot3(e2x, e2y, e2z, v2x, v2y, v2z);
    const float vc = d1 * d4 - d3 * d2;
    if (d3 >= 0 && d4 <= d3) {
      u = 0; v = 1; w = 0;
      qx = b.x; qy = b.y; qz = b.z;
    } else if (vc <= 0 && d1 >= 0 && d3 <= 0) {
      const float t = d1 / (d1 - d3);
      qx = a.x + t * e1x; qy = a.y + t * e1y; qz = a.z + t * e1z;
      u = 1 - t; v = t; w = 0;
    } else {
      const float v3x = p.x - c.x, v3y = p.y - c.y, v3z = p.z - c.z;
      const float d5 = Dot3(e1x, e1y, e1z, v3x, v3y, v3z), d6 = Dot3(e2x, e2y, e2z, v3x, v3y, v3z);
      const float vb = d5 * d2 - d1 * d6;
      const float va = d3 * d6 - d5 * d4;
      if (d6 >= 0 && d5 <= d6) {
        u = 0; v = 0; w = 1;
        qx = c.x; qy = c.y; qz = c.z;
      } else if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const float t = d2 / (d2 - d6);
        qx = a.x + t * e2x; qy = a.y + t * e2y; qz = a.z + t * e2z;
        u = 1 - t; v = 0; w = t;
      } else if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        qx = b.x + t * e3x; qy = b.y + t * e3y; qz = b.z + t * e3z;
        u = 0; v = 1 - t; w = t;
      } else {
        const float d = 1 / (va + vb + vc);
        const float sv = vb * d, sw = vc * d;
        qx = a.x + e1x * sv + e2x * sw;
        qy = a.y + e1y * sv + e2y * sw;
        qz = a.z + e1z * sv + e2z * sw;
        u = 1 - sv - sw; v = sv; w = sw;
      }
    }
  }
  const float dx = p.x - qx, dy = p.y - qy, dz = p.z - qz;
  return sqrtf(Dot3(dx, dy, dz, dx, dy, dz));
}

// QueryNearestBarycentricClothToObstacle_Kernel with QueryDevice (BVH.cu:881-948): depth-
// first with a 64-entry stack of (node, squared min distance); children tested left then
// right and pushed in that order; a popped node whose squared min distance exceeds the best
// (unsquared) distance is skipped; a child is kept when its min distance <= its sibling's
// min-max distance; a leaf replaces the best when its distance is <= it.  No other pruning:
// a tighter squared test would skip faces whose FP32 point-triangle distance can still come
// out <= the best (its rounding error is ~ulp(|coordinates|) / distance, far above any fixed
// relative margin at millimetre distances), and the chosen face would then differ.
__global__ __launch_bounds__(kBlock) void NearestBaryKernel(BvhView t, const float4* __restrict__ q,
                                                            int n, int4* __restrict__ out_idx,
                                                            float4* __restrict__ out_uvw,
                                                            int32_t* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = q[i];
  int st_node[64];
  float st_dist[64];
  int sp = 0;
  st_node[sp] = 0;
  st_dist[sp++] = MinDist(t.box[0], p);
  int nearest = -1;
  float best = INFINITY, bu = 0.0f, bv = 0.0f, bw = 0.0f;
  bool overflow = false;
  do {
    const int node = st_node[--sp];
    if (st_dist[sp] > best) continue;
    const int4 nd = t.nodes[node];
    const int kid[2] = {nd.y, nd.z};
    const Box kb0 = t.box[kid[0]], kb1 = t.box[kid[1]];
    const float md[2] = {MinDist(kb0, p), MinDist(kb1, p)};
    const float mm[2] = {MinMaxDist(kb0, p), MinMaxDist(kb1, p)};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (!(md[s] <= mm[1 - s])) continue;
      const int obj = t.nodes[kid[s]].w;
      if (obj >= 0) {
        const int4 f = t.tris[obj];
        float u, v, w;
        const float d = TriNearest(p, t.v[f.x], t.v[f.y], t.v[f.z], u, v, w);
        if (d <= best) {
          best = d;
          nearest = obj;
          bu = u; bv = v; bw = w;
        }
      } else if (sp < 64) {
        st_node[sp] = kid[s];
        st_dist[sp++] = md[s];
      } else {
        overflow = true;
      }
    }
  } while (sp > 0 && !overflow);
  if (overflow || nearest < 0) {
    if (overflow) atomicOr(flag, 1);
    out_idx[i] = make_int4(0, 0, 0, -1);
    out_uvw[i] = make_float4(0.0f, 0.0f, 0.0f, INFINITY);
    return;
  }
  const int4 f = t.tris[nearest];
  out_idx[i] = make_int4(f.x, f.y, f.z, nearest);
  out_uvw[i] = make_float4(bu, bv, bw, best);
}

// GenerateBinding_Kernel (BindingConstraint.cu:20-45)
__global__ __launch_bounds__(kBlock) void BindingGenerateKernel(
    const int32_t* __restrict__ vert, const float* __restrict__ stiffness,
    const float* __restrict__ dist, int n, const float4* __restrict__ cloth_x0,
    const int4* __restrict__ skin_idx, const float4* __restrict__ skin_uvw,
    const float4* __restrict__ obs_x0, BindItem* __restrict__ items) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const int cv = vert[b];
  const float4 pc = cloth_x0[cv];
  const int4 s = skin_idx[cv];
  const float4 w = skin_uvw[cv];
  const float4 pu = obs_x0[s.x], pv = obs_x0[s.y], pw = obs_x0[s.z];
  const float qx = w.x * pu.x + w.y * pv.x + w.z * pw.x;
  const float qy = w.x * pu.y + w.y * pv.y + w.z * pw.y;
  const float qz = w.x * pu.z + w.y * pv.z + w.z * pw.z;
  float d = dist[b];
  if (d < kEps) {
    const float dx = pc.x - qx, dy = pc.y - qy, dz = pc.z - qz;
    d = sqrtf(Dot3(dx, dy, dz, dx, dy, dz));
  }
  BindItem it;
  it.idx = make_int4(cv, s.x, s.y, s.z);
  it.uvwl = make_float4(w.x, w.y, w.z, d);
  it.k = stiffness[b];
  it.pad[0] = it.pad[1] = it.pad[2] = 0.0f;
  items[b] = it;
}

// SolveBinding_Kernel (BindingConstraint.cu:57-90): a cloth vertex farther than L from its
// body point moves back towards it by stiffness x the excess (one thread per binding; the
// bound vertices of an instance are distinct)
__global__ __launch_bounds__(kBlock) void BindingSolveKernel(const BindInstDev* __restrict__ inst,
                                                             const BindItem* __restrict__ items,
                                                             const int32_t* off, int n,
                                                             float4* __restrict__ x) {
  const int lblk = XcdBlock(off, n);
  if (lblk < 0) return;
  const int ent = InstOfBlock(off, n, lblk);
  const BindInstDev& c = inst[InstOfEntry(off, n, ent)];
  const int i = (lblk - off[ent]) * kBlock + threadIdx.x;
  if (i >= c.nb) return;
  const BindItem it = items[c.bbeg + i];
  float4* xp = x + c.vbeg + it.idx.x;
  float4 p = *xp;
  const float im = p.w;
  if (im < kEps) return;
  const float4 pu = c.o_x[it.idx.y], pv = c.o_x[it.idx.z], pw = c.o_x[it.idx.w];
  const float qx = it.uvwl.x * pu.x + it.uvwl.y * pv.x + it.uvwl.z * pw.x;
  const float qy = it.uvwl.x * pu.y + it.uvwl.y * pv.y + it.uvwl.z * pw.y;
  const float qz = it.uvwl.x * pu.z + it.uvwl.y * pv.z + it.uvwl.z * pw.z;
  const float dx = p.x - qx, dy = p.y - qy, dz = p.z - qz;
  const float len = sqrtf(Dot3(dx, dy, dz, dx, dy, dz));
  const float C = len - it.uvwl.w;
  // C < 0: inside the rest distance; len == 0 (then L == 0 and C == 0): no direction, where
  // the reference's normalize would produce NaN
  if (C < 0.0f || !(len > 0.0f)) return;
  const float inv = 1.0f / len;
  const float nx = dx * inv, ny = dy * inv, nz = dz * inv;
  const float s = C / im * it.k;
  const float f = -s * im;
  p.x += f * nx;
  p.y += f * ny;
  p.z += f * nz;
  *xp = p;
}

// lbs (body.cu:39-58): one thread per obstacle vertex, the instance's joint transforms
// staged in LDS
__global__ __launch_bounds__(kBlock) void LbsKernel(const SkinInstDev* __restrict__ inst,
                                                    const float* __restrict__ T,
                                                    const int32_t* off, int n) {
  __shared__ float tr[kMaxJoints * 12];
  const int lblk = XcdBlock(off, n);
  if (lblk < 0) return;
  const int ent = InstOfBlock(off, n, lblk);
  const SkinInstDev& c = inst[InstOfEntry(off, n, ent)];
  for (int k = threadIdx.x; k < 12 * c.nj; k += kBlock) tr[k] = T[c.toff + k];
  __syncthreads();
  const int i = (lblk - off[ent]) * kBlock + threadIdx.x;
  if (i >= c.nv) return;
  const float4 p = c.rest[i];
  float acc[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float a = 0.0f;
    for (int k = c.ptr[i]; k < c.ptr[i + 1]; ++k) {
      const float wk = c.weight[k];
      const float* row = tr + 12 * c.joint[k] + 4 * r;
      a += wk * row[0] * p.x;
      a += wk * row[1] * p.y;
      a += wk * row[2] * p.z;
      a += wk * row[3];
    }
    acc[r] = a;
  }
  c.out[i] = make_float4(acc[0], acc[1], acc[2], p.w);
}
}  // namespace

hipError_t LaunchNearestBary(hipStream_t s, const BvhView& t, const float4* q, int n, int4* idx,
                             float4* uvw, int32_t* flag) {
  if (n <= 0) return hipSuccess;
  NearestBaryKernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(t, q, n, idx, uvw, flag);
  return hipGetLastError();
}

hipError_t LaunchBindingGenerate(hipStream_t s, const int32_t* vert, const float* stiffness,
                                 const float* dist, int n, const float4* cloth_x0,
                                 const int4* skin_idx, const float4* skin_uvw,
                                 const float4* obs_x0, BindItem* items) {
  if (n <= 0) return hipSuccess;
  BindingGenerateKernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(
      vert, stiffness, dist, n, cloth_x0, skin_idx, skin_uvw, obs_x0, items);
  return hipGetLastError();
}

hipError_t LaunchBindingSolve(hipStream_t s, const BindInstDev* inst, const BindItem* items,
                              const DevMap& m, float4* x) {
  if (m.blocks <= 0) return hipSuccess;
  BindingSolveKernel<<<XcdGrid(m.blocks), kBlock, 0, s>>>(inst, items, m.off, m.n, x);
  return hipGetLastError();
}

hipError_t LaunchLbs(hipStream_t s, const SkinInstDev* inst, const float* T, const DevMap& m) {
  if (m.blocks <= 0) return hipSuccess;
  LbsKernel<<<XcdGrid(m.blocks), kBlock, 0, s>>>(inst, T, m.off, m.n);
  return hipGetLastError();
}

}  // namespace xrt
