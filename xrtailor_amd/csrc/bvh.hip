// LBVH kernels (built with -ffp-contract=off: every float result that feeds an integer
// output -- Morton codes, box tests -- rounds exactly as written, matching the FP32 oracle).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bvh.h"
#include "bvh_device.h"

namespace xrt {

namespace {
constexpr int kBlock = 256;

inline int Blocks(int n, int bs = kBlock) { return (n + bs - 1) / bs; }

// order-preserving float <-> uint32 map (for atomic min / max of the root box)
__host__ __device__ __forceinline__ uint32_t EncodeOrdered(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float DecodeOrderedDev(uint32_t u) {
  uint32_t b = (u & 0x80000000u) ? (u ^ 0x80000000u) : ~u;
  return __builtin_bit_cast(float, b);
}

__device__ __forceinline__ uint32_t ExpandBits(uint32_t v) {  // BVH.cu:22-28
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__device__ __forceinline__ uint32_t MortonCode(float x, float y, float z) {  // BVH.cu:32-46
  x = Mn(Mx(x * 1024.0f, 0.0f), 1023.0f);
  y = Mn(Mx(y * 1024.0f, 0.0f), 1023.0f);
  z = Mn(Mx(z * 1024.0f, 0.0f), 1023.0f);
  return ExpandBits(static_cast<uint32_t>(x)) * 4u + ExpandBits(static_cast<uint32_t>(y)) * 2u +
         ExpandBits(static_cast<uint32_t>(z));
}

__device__ __forceinline__ int Clz64(uint64_t v) { return v == 0 ? 64 : __clzll(v); }

// ------------------------------------------------------------------ build
// Each build step is a body over (block, thread) so that the single-tree kernels and the
// multi-tree kernels (every tree of a set per launch) run the same code.
__device__ __forceinline__ void CornerPointsBody(const float4* __restrict__ x0,
                                                 const float4* __restrict__ x,
                                                 const float4* __restrict__ n, float tol, int i,
                                                 float4* __restrict__ v,
                                                 float4* __restrict__ pred) {
  float4 p = x0[i];
  float4 q = x ? x[i] : p;
  if (n) {
    const float4 nn = n[i];
    p.x = p.x + tol * nn.x; p.y = p.y + tol * nn.y; p.z = p.z + tol * nn.z;
    q.x = q.x + tol * nn.x; q.y = q.y + tol * nn.y; q.z = q.z + tol * nn.z;
  }
  p.w = 0.0f;
  q.w = 0.0f;
  v[i] = p;
  pred[i] = q;
}

__global__ __launch_bounds__(kBlock) void CornerPointsKernel(const float4* __restrict__ x0,
                                                             const float4* __restrict__ x,
                                                             const float4* __restrict__ n,
                                                             float tol, int nv,
                                                             float4* __restrict__ v,
                                                             float4* __restrict__ pred) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) CornerPointsBody(x0, x, n, tol, i, v, pred);
}

__device__ __forceinline__ void InitRootBody(uint32_t* root, int k) {
  if (k < 3) root[k] = EncodeOrdered(INFINITY);
  else if (k < 6) root[k] = EncodeOrdered(-INFINITY);
}

__global__ void InitRootKernel(uint32_t* root) { InitRootBody(root, threadIdx.x); }

// InitAABBs_Kernel (BVH.cu:423-462) + the root reduction (BVH.cu:642-647) for faces of
// block `blk`; every thread of the block takes part (block reduction)
__device__ __forceinline__ void LeafBoxBody(const int4* __restrict__ tris,
                                            const float4* __restrict__ v,
                                            const float4* __restrict__ pred, int nf, float scr,
                                            Box* __restrict__ leaf, uint32_t* __restrict__ root,
                                            int blk) {
  __shared__ uint32_t red[6][kBlock / 64];
  const int f = blk * kBlock + threadIdx.x;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (f < nf) {
    const int4 t = tris[f];
    const float4 a = v[t.x], b = v[t.y], c = v[t.z];
    const float4 p = pred[t.x], q = pred[t.y], r = pred[t.z];
    lo[0] = Mn(Mn(a.x, Mn(b.x, c.x)), Mn(p.x, Mn(q.x, r.x)));
    lo[1] = Mn(Mn(a.y, Mn(b.y, c.y)), Mn(p.y, Mn(q.y, r.y)));
    lo[2] = Mn(Mn(a.z, Mn(b.z, c.z)), Mn(p.z, Mn(q.z, r.z)));
    hi[0] = Mx(Mx(a.x, Mx(b.x, c.x)), Mx(p.x, Mx(q.x, r.x)));
    hi[1] = Mx(Mx(a.y, Mx(b.y, c.y)), Mx(p.y, Mx(q.y, r.y)));
    hi[2] = Mx(Mx(a.z, Mx(b.z, c.z)), Mx(p.z, Mx(q.z, r.z)));
    if (scr != 0.0f) {
      const float cx = (lo[0] + hi[0]) * 0.5f, cy = (lo[1] + hi[1]) * 0.5f, cz = (lo[2] + hi[2]) * 0.5f;
      const float d1x = lo[0] - cx, d1y = lo[1] - cy, d1z = lo[2] - cz;
      const float d2x = hi[0] - cx, d2y = hi[1] - cy, d2z = hi[2] - cz;
      const float l1 = sqrtf(d1x * d1x + d1y * d1y + d1z * d1z);
      const float l2 = sqrtf(d2x * d2x + d2y * d2y + d2z * d2z);
      lo[0] = lo[0] + (d1x / l1) * scr; lo[1] = lo[1] + (d1y / l1) * scr; lo[2] = lo[2] + (d1z / l1) * scr;
      hi[0] = hi[0] + (d2x / l2) * scr; hi[1] = hi[1] + (d2y / l2) * scr; hi[2] = hi[2] + (d2z / l2) * scr;
    }
    leaf[f] = Box{make_float4(lo[0], lo[1], lo[2], 0.0f), make_float4(hi[0], hi[1], hi[2], 0.0f)};
  }
  // wave64 reduction of the encoded bounds, then one atomic per wave-slot per block
  uint32_t e[6];
  for (int k = 0; k < 3; ++k) { e[k] = EncodeOrdered(lo[k]); e[3 + k] = EncodeOrdered(hi[k]); }
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 6; ++k) {
      const uint32_t o = __shfl_xor(e[k], off, 64);
      e[k] = k < 3 ? min(e[k], o) : max(e[k], o);
    }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
    for (int k = 0; k < 6; ++k) red[k][w] = e[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    uint32_t m = red[k][0];
    for (int j = 1; j < kBlock / 64; ++j) m = k < 3 ? min(m, red[k][j]) : max(m, red[k][j]);
    if (k < 3) atomicMin(root + k, m); else atomicMax(root + k, m);
  }
}

__global__ __launch_bounds__(kBlock) void LeafBoxKernel(const int4* __restrict__ tris,
                                                        const float4* __restrict__ v,
                                                        const float4* __restrict__ pred, int nf,
                                                        float scr, Box* __restrict__ leaf,
                                                        uint32_t* __restrict__ root) {
  LeafBoxBody(tris, v, pred, nf, scr, leaf, root, blockIdx.x);
}

// ComputeMortons_Kernel (BVH.cu:464-481)
__device__ __forceinline__ void MortonBody(const Box* __restrict__ leaf,
                                           const uint32_t* __restrict__ root, int f,
                                           uint32_t* __restrict__ morton,
                                           int32_t* __restrict__ order) {
  const float rlx = DecodeOrderedDev(root[0]), rly = DecodeOrderedDev(root[1]), rlz = DecodeOrderedDev(root[2]);
  const float rhx = DecodeOrderedDev(root[3]), rhy = DecodeOrderedDev(root[4]), rhz = DecodeOrderedDev(root[5]);
  const Box b = leaf[f];
  float cx = 0.5f * (b.lo.x + b.hi.x), cy = 0.5f * (b.lo.y + b.hi.y), cz = 0.5f * (b.lo.z + b.hi.z);
  cx = cx - rlx; cy = cy - rly; cz = cz - rlz;
  cx = cx / (rhx - rlx); cy = cy / (rhy - rly); cz = cz / (rhz - rlz);
  morton[f] = MortonCode(cx, cy, cz);
  order[f] = f;
}

__global__ __launch_bounds__(kBlock) void MortonKernel(const Box* __restrict__ leaf,
                                                       const uint32_t* __restrict__ root, int nf,
                                                       uint32_t* __restrict__ morton,
                                                       int32_t* __restrict__ order) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < nf) MortonBody(leaf, root, f, morton, order);
}

// -------------------------------------------------------------- radix sort
__global__ __launch_bounds__(kBlock) void RadixHistKernel(const uint32_t* __restrict__ keys, int n,
                                                          int shift, int nblocks,
                                                          uint32_t* __restrict__ hist) {
  RadixHistBody(keys, n, shift, nblocks, hist, blockIdx.x);
}

// single-workgroup exclusive scan of n uint32 (in place); *total receives the sum
__global__ __launch_bounds__(1024) void ExclusiveScanKernel(uint32_t* __restrict__ a, int n,
                                                            int32_t* __restrict__ total) {
  BlockExclusiveScan(a, n, total);
}

__global__ __launch_bounds__(kBlock) void RadixScatterKernel(
    const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    int32_t* __restrict__ vout, int n, int shift, int nblocks, const uint32_t* __restrict__ offs) {
  RadixScatterBody(kin, vin, kout, vout, n, shift, nblocks, offs, blockIdx.x);
}

// ConstructLeafNodes_Kernel (BVH.cu:489-502) + internal-node reset + 64-bit keys
// (m << 32 | object): unique keys reproduce the 32-bit tree when codes are unique and the
// reference's 64-bit fallback when they are not (BVH.cu:671-683)
__device__ __forceinline__ void LeafNodesBody(const uint32_t* __restrict__ morton,
                                              const int32_t* __restrict__ order,
                                              const Box* __restrict__ leaf, int nf, int i,
                                              int4* __restrict__ nodes, Box* __restrict__ box,
                                              uint64_t* __restrict__ code,
                                              int32_t* __restrict__ flags) {
  const int ni = nf - 1;
  const int obj = order[i];
  nodes[ni + i] = make_int4(-1, -1, -1, obj);
  box[ni + i] = leaf[obj];
  code[i] = (static_cast<uint64_t>(morton[i]) << 32) | static_cast<uint32_t>(obj);
  if (i < ni) {
    nodes[i] = make_int4(-1, -1, -1, -1);
    flags[i] = 0;
  } else {
    flags[ni] = 0;  // the topology changes: LevelBottomUpWG lists the depths again
  }
}

__global__ __launch_bounds__(kBlock) void LeafNodesKernel(const uint32_t* __restrict__ morton,
                                                          const int32_t* __restrict__ order,
                                                          const Box* __restrict__ leaf, int nf,
                                                          int4* __restrict__ nodes,
                                                          Box* __restrict__ box,
                                                          uint64_t* __restrict__ code,
                                                          int32_t* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nf) LeafNodesBody(morton, order, leaf, nf, i, nodes, box, code, flags);
}

// determine_range / find_split / ConstructInternalNodes64Bits_Kernel (BVH.cu:102-210,297-317)
__device__ __forceinline__ void KarrasBody(const uint64_t* __restrict__ code, int nf, int idx,
                                           int4* __restrict__ nodes) {
  const int ni = nf - 1;
  int first, last;
  if (idx == 0) {
    first = 0;
    last = nf - 1;
  } else {
    const uint64_t self = code[idx];
    const int L = Clz64(self ^ code[idx - 1]);
    const int R = Clz64(self ^ code[idx + 1]);
    const int d = (R > L) ? 1 : -1;
    const int dmin = min(L, R);
    int lmax = 2, delta = -1;
    int t = idx + d * lmax;
    if (0 <= t && t < nf) delta = Clz64(self ^ code[t]);
    while (delta > dmin) {
      lmax <<= 1;
      t = idx + d * lmax;
      delta = -1;
      if (0 <= t && t < nf) delta = Clz64(self ^ code[t]);
    }
    int l = 0;
    for (int s = lmax >> 1; s > 0; s >>= 1) {
      t = idx + (l + s) * d;
      delta = -1;
      if (0 <= t && t < nf) delta = Clz64(self ^ code[t]);
      if (delta > dmin) l += s;
    }
    const int j = idx + l * d;
    first = d < 0 ? j : idx;
    last = d < 0 ? idx : j;
  }
  int split;
  const uint64_t fc = code[first], lc = code[last];
  if (fc == lc) {
    split = (first + last) >> 1;
  } else {
    const int dn = Clz64(fc ^ lc);
    split = first;
    int stride = last - first;
    do {
      stride = (stride + 1) >> 1;
      const int mid = split + stride;
      if (mid < last && Clz64(fc ^ code[mid]) > dn) split = mid;
    } while (stride > 1);
  }
  int left = split, right = split + 1;
  if (first == split) left += ni;
  if (last == split + 1) right += ni;
  nodes[idx].y = left;
  nodes[idx].z = right;
  nodes[left].x = idx;
  nodes[right].x = idx;
}

__global__ __launch_bounds__(kBlock) void KarrasKernel(const uint64_t* __restrict__ code, int nf,
                                                       int4* __restrict__ nodes) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < nf - 1) KarrasBody(code, nf, idx, nodes);
}


// bottom-up boxes: the second thread to reach a node merges its two children (the
// result is the union of the descendant leaves, as AssignAABBAtomic_Kernel's atomic
// min / max, BVH.cu:540-557, and AssignAABBCAS_Kernel, BVH.cu:515-538).  grow = 1 keeps
// the old box too (Update's grow-only semantics).
__global__ __launch_bounds__(kBlock) void BottomUpKernel(const int4* __restrict__ nodes,
                                                         Box* __restrict__ box,
                                                         int32_t* __restrict__ flags, int nf,
                                                         int grow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  BottomUpBody(nodes, box, flags, nf, grow, i);
}


__global__ __launch_bounds__(kBlock) void ZeroFlagsKernel(int32_t* flags, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = 0;
}

// Build{Internal,External}EscapeIndexTable_Kernel (BVH.cu:212-273), node-indexed
__device__ __forceinline__ void EscapeBody(const int4* __restrict__ nodes, int nn, int k,
                                           int32_t* __restrict__ escape) {
  int cur = k, e = -1;
  for (int level = 0; level < nn; ++level) {
    const int p = nodes[cur].x;
    if (p < 0 || p >= nn) break;
    const int4 pn = nodes[p];
    if (pn.y == cur) {
      e = pn.z;
      break;
    }
    cur = p;
  }
  escape[k] = e;
}

__global__ __launch_bounds__(kBlock) void EscapeKernel(const int4* __restrict__ nodes, int nn,
                                                       int32_t* __restrict__ escape) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nn) EscapeBody(nodes, nn, k, escape);
}

// ---------------------------------------------------------------- refit
// RefitExternalNodeMidstep_Kernel (BVH.cu:770-818): pred = x1 + tol n1, leaf grows
__global__ __launch_bounds__(kBlock) void RefitPredKernel(const float4* __restrict__ x1,
                                                          const float4* __restrict__ n1, float tol,
                                                          int nv, float4* __restrict__ pred) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  RefitPredBody(x1, n1, tol, i, pred);
}

// leaf grows by its 3 pred corners (midstep) or by Face::ComputeBounds(ccd) of x [and x0]
__global__ __launch_bounds__(kBlock) void GrowLeavesKernel(const int4* __restrict__ nodes,
                                                           const int4* __restrict__ tris,
                                                           const float4* __restrict__ pa,
                                                           const float4* __restrict__ pb, int nf,
                                                           Box* __restrict__ box,
                                                           int32_t* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  GrowLeafBody(nodes, tris, pa, pb, nf, i, box, flags);
}

}  // namespace

namespace {
// QueryDeviceStackless point / segment (BVH.cu:957-1021)
__global__ __launch_bounds__(kBlock) void QueryKernel(BvhView t, const float4* __restrict__ p0,
                                                      const float4* __restrict__ p1, int n,
                                                      int32_t* __restrict__ hits,
                                                      int32_t* __restrict__ counts) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const float4 a = p0[q];
  int32_t* out = hits + static_cast<size_t>(q) * kMaxHits;
  int k = 0;
  const int found = p1 ? StacklessQuery<true>(t, a, p1[q], [&](int obj) { out[k++] = obj; })
                       : StacklessQuery<false>(t, a, a, [&](int obj) { out[k++] = obj; });
  counts[q] = found;
}


// CountPairs_Kernel (BVH.cu:1149-1179)
__global__ __launch_bounds__(kBlock) void CountPairsKernel(BvhView l, BvhView r, float th,
                                                           uint32_t* __restrict__ cnt) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= l.nf) return;
  cnt[id] = CountPairsBody(l, r, th, id);
}

// FindPairs_Kernel (BVH.cu:1189-1227): each leaf's pairs fill its range backwards
__global__ __launch_bounds__(kBlock) void FindPairsKernel(BvhView l, BvhView r, float th,
                                                          const uint32_t* __restrict__ cnt,
                                                          const uint32_t* __restrict__ off,
                                                          int2* __restrict__ pairs, int cap) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= l.nf) return;
  FindPairsBody(l, r, th, id, static_cast<int64_t>(off[id]), cnt[id], pairs, cap);
}

// UpdateNodeNormals (xrtailor/src/physics/PhysicsMeshHelper.cu:196-210,282-295) as a
// per-vertex gather over a corner CSR in ascending face order (deterministic)
__global__ __launch_bounds__(kBlock) void VertexNormalsKernel(const float4* __restrict__ x,
                                                              const int4* __restrict__ tris,
                                                              const int32_t* __restrict__ vptr,
                                                              const int32_t* __restrict__ corner,
                                                              int nv, float4* __restrict__ n) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  n[v] = VertexNormalBody(x, tris, vptr, corner, v);
}
}  // namespace

namespace {
// ------------------------------------------------------- multi-tree build
#define XRT_TREE_INDEX()                                          \
  const int lblk = XcdBlock(off, n);                          \
  if (lblk < 0) return;                                           \
  const int k = off[n + 1 + lblk];                                \
  const TreeBuildDev& t = set[k];                                 \
  const int lb = lblk - off[k];                                   \
  const int i = lb * kBlock + threadIdx.x;                        \
  (void)i;

__global__ __launch_bounds__(kBlock) void MultiCornerKernel(const TreeBuildDev* set,
                                                            const int32_t* off, int n) {
  XRT_TREE_INDEX()
  if (i < t.nv) CornerPointsBody(t.x0, nullptr, nullptr, 0.0f, i, t.v, t.pred);
}

__global__ void MultiInitRootKernel(const TreeBuildDev* set) {
  InitRootBody(set[blockIdx.x].root, threadIdx.x);
}

__global__ __launch_bounds__(kBlock) void MultiLeafBoxKernel(const TreeBuildDev* set,
                                                             const int32_t* off, int n) {
  XRT_TREE_INDEX()
  LeafBoxBody(t.tris, t.v, t.pred, t.nf, 0.0f, t.leaf, t.root, lb);
}

__global__ __launch_bounds__(kBlock) void MultiMortonKernel(const TreeBuildDev* set,
                                                            const int32_t* off, int n) {
  XRT_TREE_INDEX()
  if (i < t.nf) MortonBody(t.leaf, t.root, i, t.morton, t.order);
}

// pass p sorts digit p; even passes read (morton, order), odd ones the tmp buffers
__global__ __launch_bounds__(kBlock) void MultiRadixHistKernel(const TreeBuildDev* set,
                                                               const int32_t* off, int n,
                                                               int pass) {
  XRT_TREE_INDEX()
  RadixHistBody(pass & 1 ? t.morton_tmp : t.morton, t.nf, 8 * pass, t.nblocks, t.hist, lb);
}

__global__ __launch_bounds__(1024) void MultiRadixScanKernel(const TreeBuildDev* set) {
  const TreeBuildDev& t = set[blockIdx.x];
  BlockExclusiveScan(t.hist, 256 * t.nblocks, nullptr);
}

__global__ __launch_bounds__(kBlock) void MultiRadixScatterKernel(const TreeBuildDev* set,
                                                                  const int32_t* off, int n,
                                                                  int pass) {
  XRT_TREE_INDEX()
  const bool odd = pass & 1;
  RadixScatterBody(odd ? t.morton_tmp : t.morton, odd ? t.order_tmp : t.order,
                   odd ? t.morton : t.morton_tmp, odd ? t.order : t.order_tmp, t.nf, 8 * pass,
                   t.nblocks, t.hist, lb);
}

__global__ __launch_bounds__(kBlock) void MultiLeafNodesKernel(const TreeBuildDev* set,
                                                               const int32_t* off, int n) {
  XRT_TREE_INDEX()
  if (i < t.nf) LeafNodesBody(t.morton, t.order, t.leaf, t.nf, i, t.nodes, t.box, t.code, t.flags);
}

__global__ __launch_bounds__(kBlock) void MultiKarrasKernel(const TreeBuildDev* set,
                                                            const int32_t* off, int n) {
  XRT_TREE_INDEX()
  if (i < t.nf - 1) KarrasBody(t.code, t.nf, i, t.nodes);
}

__global__ __launch_bounds__(kBlock) void MultiEscapeKernel(const TreeBuildDev* set,
                                                            const int32_t* off, int n) {
  XRT_TREE_INDEX()
  if (i < 2 * t.nf - 1) EscapeBody(t.nodes, 2 * t.nf - 1, i, t.escape);
}
#undef XRT_TREE_INDEX
}  // namespace

hipError_t TreeSet::Setup(const std::vector<Lbvh*>& in, const std::vector<const float4*>& x0,
                          hipStream_t s) {
  trees.clear();
  std::vector<TreeBuildDev> d;
  for (size_t q = 0; q < in.size(); ++q) {
    Lbvh& b = *in[q];
    if (b.nf <= 0) continue;
    TreeBuildDev t;
    t.x0 = x0[q];
    t.nv = b.nv;
    t.nf = b.nf;
    t.nblocks = b.nblocks;
    t.tris = b.tris.ptr;
    t.v = b.v.ptr;
    t.pred = b.pred.ptr;
    t.leaf = b.leaf.ptr;
    t.morton = b.morton.ptr;
    t.morton_tmp = b.morton_tmp.ptr;
    t.order = b.order.ptr;
    t.order_tmp = b.order_tmp.ptr;
    t.code = b.code.ptr;
    t.nodes = b.nodes.ptr;
    t.box = b.box.ptr;
    t.escape = b.escape.ptr;
    t.root = b.root.ptr;
    t.hist = b.hist.ptr;
    t.flags = b.flags.ptr;
    d.push_back(t);
    trees.push_back(&b);
  }
  n = static_cast<int>(d.size());
  std::vector<int32_t> m;
  size_t seg[kKinds];
  for (int kind = 0; kind < kKinds; ++kind) {
    std::vector<int32_t> o(n + 1, 0);
    for (int k = 0; k < n; ++k) {
      const TreeBuildDev& t = d[k];
      const int c = kind == kVert ? Blocks(t.nv)
                  : kind == kFace ? Blocks(t.nf)
                  : kind == kTile ? t.nblocks
                                  : Blocks(2 * t.nf - 1);
      o[k + 1] = o[k] + c;
    }
    blocks[kind] = o[n];
    seg[kind] = m.size();
    m.insert(m.end(), o.begin(), o.end());
    for (int k = 0; k < n; ++k) m.insert(m.end(), static_cast<size_t>(o[k + 1] - o[k]), k);
  }
  if (n == 0) return hipSuccess;
  hipError_t e;
  if ((e = dev.Upload(d, s)) != hipSuccess) return e;
  if ((e = maps.Upload(m, s)) != hipSuccess) return e;
  for (int kind = 0; kind < kKinds; ++kind) off[kind] = maps.ptr + seg[kind];
  return hipStreamSynchronize(s);
}

hipError_t MultiLbvhBuild(TreeSet& set, hipStream_t s) {
  if (set.n == 0) return hipSuccess;
  using K = TreeSet;
  const TreeBuildDev* d = set.dev.ptr;
  const int n = set.n;
  MultiCornerKernel<<<(set.blocks[K::kVert] + 7) / 8 * 8, kBlock, 0, s>>>(d, set.off[K::kVert], n);
  MultiInitRootKernel<<<n, 64, 0, s>>>(d);
  MultiLeafBoxKernel<<<(set.blocks[K::kFace] + 7) / 8 * 8, kBlock, 0, s>>>(d, set.off[K::kFace], n);
  MultiMortonKernel<<<(set.blocks[K::kFace] + 7) / 8 * 8, kBlock, 0, s>>>(d, set.off[K::kFace], n);
  for (int pass = 0; pass < 4; ++pass) {  // the result lands back in (morton, order)
    MultiRadixHistKernel<<<XcdGrid(set.blocks[K::kTile]), kBlock, 0, s>>>(d, set.off[K::kTile], n, pass);
    MultiRadixScanKernel<<<n, 1024, 0, s>>>(d);
    MultiRadixScatterKernel<<<XcdGrid(set.blocks[K::kTile]), kBlock, 0, s>>>(d, set.off[K::kTile], n, pass);
  }
  MultiLeafNodesKernel<<<(set.blocks[K::kFace] + 7) / 8 * 8, kBlock, 0, s>>>(d, set.off[K::kFace], n);
  MultiKarrasKernel<<<(set.blocks[K::kFace] + 7) / 8 * 8, kBlock, 0, s>>>(d, set.off[K::kFace], n);
  MultiEscapeKernel<<<(set.blocks[K::kNode] + 7) / 8 * 8, kBlock, 0, s>>>(d, set.off[K::kNode], n);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  for (Lbvh* b : set.trees) b->active = true;
  return hipSuccess;
}

float DecodeOrdered(uint32_t u) { return DecodeOrderedDev(u); }

BvhView Lbvh::view() const {
  BvhView t;
  t.nf = nf;
  t.nv = nv;
  t.tris = tris.ptr;
  t.v = v.ptr;
  t.pred = pred.ptr;
  t.morton = morton.ptr;
  t.order = order.ptr;
  t.nodes = nodes.ptr;
  t.box = box.ptr;
  t.escape = escape.ptr;
  t.root = root.ptr;
  t.flags = flags.ptr;
  t.counts = reinterpret_cast<uint32_t*>(counts.ptr);
  t.offsets = reinterpret_cast<uint32_t*>(offsets.ptr);
  t.total = total.ptr;
  t.packed = packed.ptr;
  t.lvl = lvl.ptr;
  t.leaf_face = leaf_face.ptr;
  return t;
}

hipError_t Lbvh::Setup(const int32_t* tris_host, int nfaces, int nverts, hipStream_t s) {
  nf = nfaces;
  nv = nverts;
  active = false;
  nblocks = std::max(1, (nf + kSortTile - 1) / kSortTile);
  const int nn = std::max(1, 2 * nf - 1);
  std::vector<int4> t(std::max(nf, 1));
  for (int f = 0; f < nf; ++f) t[f] = make_int4(tris_host[3 * f], tris_host[3 * f + 1], tris_host[3 * f + 2], 0);
  hipError_t e;
  if ((e = tris.Alloc(t.size())) != hipSuccess) return e;
  if ((e = tris.Upload(t.data(), t.size(), s)) != hipSuccess) return e;
  if ((e = v.Alloc(std::max(nv, 1))) != hipSuccess) return e;
  if ((e = pred.Alloc(std::max(nv, 1))) != hipSuccess) return e;
  if ((e = leaf.Alloc(std::max(nf, 1))) != hipSuccess) return e;
  if ((e = morton.Alloc(std::max(nf, 1))) != hipSuccess) return e;
  if ((e = morton_tmp.Alloc(std::max(nf, 1))) != hipSuccess) return e;
  if ((e = order.Alloc(std::max(nf, 1))) != hipSuccess) return e;
  if ((e = order_tmp.Alloc(std::max(nf, 1))) != hipSuccess) return e;
  if ((e = code.Alloc(std::max(nf, 1))) != hipSuccess) return e;
  if ((e = nodes.Alloc(nn)) != hipSuccess) return e;
  if ((e = box.Alloc(nn)) != hipSuccess) return e;
  if ((e = escape.Alloc(nn)) != hipSuccess) return e;
  if ((e = packed.Alloc(nn)) != hipSuccess) return e;
  if ((e = flags.Alloc(std::max(nf, 1))) != hipSuccess) return e;
  // flags[nf - 1]: the level lists of LevelBottomUpWG are not yet valid
  if ((e = hipMemsetAsync(flags.ptr, 0, sizeof(int32_t) * flags.n, s)) != hipSuccess) return e;
  if ((e = lvl.Alloc(4 * static_cast<size_t>(std::max(nf, 1)))) != hipSuccess) return e;
  if ((e = leaf_face.Alloc(std::max(nf, 1))) != hipSuccess) return e;
  if ((e = root.Alloc(6)) != hipSuccess) return e;
  if ((e = hist.Alloc(static_cast<size_t>(256) * nblocks)) != hipSuccess) return e;
  if ((e = counts.Alloc(std::max(nf, 1))) != hipSuccess) return e;
  if ((e = offsets.Alloc(std::max(nf, 1))) != hipSuccess) return e;
  if ((e = total.Alloc(1)) != hipSuccess) return e;
  return hipSuccess;
}

hipError_t LbvhBuild(Lbvh& b, const float4* x0, const float4* x, const float4* n, float tol,
                     int ccd, float scr, hipStream_t s, bool merge) {
  if (b.nf <= 0) {
    b.active = false;
    return hipSuccess;
  }
  const int nf = b.nf;
  CornerPointsKernel<<<Blocks(b.nv), kBlock, 0, s>>>(x0, ccd ? x : nullptr, n, tol, b.nv, b.v.ptr,
                                                     b.pred.ptr);
  InitRootKernel<<<1, 64, 0, s>>>(b.root.ptr);
  LeafBoxKernel<<<Blocks(nf), kBlock, 0, s>>>(b.tris.ptr, b.v.ptr, b.pred.ptr, nf, scr, b.leaf.ptr,
                                              b.root.ptr);
  MortonKernel<<<Blocks(nf), kBlock, 0, s>>>(b.leaf.ptr, b.root.ptr, nf, b.morton.ptr, b.order.ptr);
  // 4 passes of 8 bits; the result lands back in (morton, order)
  uint32_t* kin = b.morton.ptr;
  int32_t* vin = b.order.ptr;
  uint32_t* kout = b.morton_tmp.ptr;
  int32_t* vout = b.order_tmp.ptr;
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 8 * pass;
    RadixHistKernel<<<b.nblocks, kBlock, 0, s>>>(kin, nf, shift, b.nblocks, b.hist.ptr);
    ExclusiveScanKernel<<<1, 1024, 0, s>>>(b.hist.ptr, 256 * b.nblocks, nullptr);
    RadixScatterKernel<<<b.nblocks, kBlock, 0, s>>>(kin, vin, kout, vout, nf, shift, b.nblocks,
                                                    b.hist.ptr);
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  LeafNodesKernel<<<Blocks(nf), kBlock, 0, s>>>(b.morton.ptr, b.order.ptr, b.leaf.ptr, nf,
                                                b.nodes.ptr, b.box.ptr, b.code.ptr, b.flags.ptr);
  if (nf > 1) {
    KarrasKernel<<<Blocks(nf - 1), kBlock, 0, s>>>(b.code.ptr, nf, b.nodes.ptr);
    if (merge) BottomUpKernel<<<Blocks(nf), kBlock, 0, s>>>(b.nodes.ptr, b.box.ptr, b.flags.ptr, nf, 0);
  }
  EscapeKernel<<<Blocks(2 * nf - 1), kBlock, 0, s>>>(b.nodes.ptr, 2 * nf - 1, b.escape.ptr);
  b.active = true;
  return hipGetLastError();
}

hipError_t LbvhRefitMidstep(Lbvh& b, const float4* x1, const float4* n1, float tol, hipStream_t s) {
  if (!b.active) return hipSuccess;
  RefitPredKernel<<<Blocks(b.nv), kBlock, 0, s>>>(x1, n1, tol, b.nv, b.pred.ptr);
  GrowLeavesKernel<<<Blocks(b.nf), kBlock, 0, s>>>(b.nodes.ptr, b.tris.ptr, b.pred.ptr, nullptr,
                                                   b.nf, b.box.ptr, b.flags.ptr);
  if (b.nf > 1)
    BottomUpKernel<<<Blocks(b.nf), kBlock, 0, s>>>(b.nodes.ptr, b.box.ptr, b.flags.ptr, b.nf, 0);
  return hipGetLastError();
}

hipError_t LbvhUpdate(Lbvh& b, const float4* x, const float4* x0, int ccd, hipStream_t s) {
  if (!b.active) return hipSuccess;
  RefitPredKernel<<<Blocks(b.nv), kBlock, 0, s>>>(x, nullptr, 0.0f, b.nv, b.pred.ptr);
  GrowLeavesKernel<<<Blocks(b.nf), kBlock, 0, s>>>(b.nodes.ptr, b.tris.ptr, x, ccd ? x0 : nullptr,
                                                   b.nf, b.box.ptr, b.flags.ptr);
  if (b.nf > 1)
    BottomUpKernel<<<Blocks(b.nf), kBlock, 0, s>>>(b.nodes.ptr, b.box.ptr, b.flags.ptr, b.nf, 1);
  return hipGetLastError();
}

hipError_t LbvhQuery(const Lbvh& b, const float4* p0, const float4* p1, int n, int32_t* hits,
                     int32_t* counts, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (!b.active) return hipMemsetAsync(counts, 0, sizeof(int32_t) * n, s);
  QueryKernel<<<Blocks(n), kBlock, 0, s>>>(b.view(), p0, p1, n, hits, counts);
  return hipGetLastError();
}

hipError_t LbvhCountPairs(Lbvh& lhs, const Lbvh& rhs, float thickness, hipStream_t s) {
  if (!lhs.active || !rhs.active) {
    ZeroFlagsKernel<<<1, 64, 0, s>>>(lhs.total.ptr, 1);
    return hipGetLastError();
  }
  uint32_t* cnt = reinterpret_cast<uint32_t*>(lhs.counts.ptr);
  uint32_t* off = reinterpret_cast<uint32_t*>(lhs.offsets.ptr);
  CountPairsKernel<<<Blocks(lhs.nf), kBlock, 0, s>>>(lhs.view(), rhs.view(), thickness, cnt);
  hipError_t e = hipMemcpyAsync(off, cnt, sizeof(uint32_t) * lhs.nf, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return e;
  ExclusiveScanKernel<<<1, 1024, 0, s>>>(off, lhs.nf, lhs.total.ptr);
  return hipGetLastError();
}

hipError_t LbvhFindPairs(Lbvh& lhs, const Lbvh& rhs, float thickness, int2* pairs, int cap,
                         hipStream_t s) {
  if (!lhs.active || !rhs.active) return hipSuccess;
  FindPairsKernel<<<Blocks(lhs.nf), kBlock, 0, s>>>(
      lhs.view(), rhs.view(), thickness, reinterpret_cast<const uint32_t*>(lhs.counts.ptr),
      reinterpret_cast<const uint32_t*>(lhs.offsets.ptr), pairs, cap);
  return hipGetLastError();
}

hipError_t VertexNormals(const float4* x, const int4* tris, const int32_t* vptr,
                         const int32_t* corner, int nv, float4* n, hipStream_t s) {
  if (nv <= 0) return hipSuccess;
  VertexNormalsKernel<<<Blocks(nv), kBlock, 0, s>>>(x, tris, vptr, corner, nv, n);
  return hipGetLastError();
}

}  // namespace xrt
