// Device-side LBVH helpers shared by the build / query kernels and the contact kernels
// (header-only: every translation unit inlines its own copy).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bvh.h"
#include "device.h"

namespace xrt {

constexpr int kMaxHits = 256;  // LBVH_MAX_BUFFER_SIZE (xrtailor/include/xrtailor/core/Common.cuh:24)

// MathFunctions::min / max (xrtailor/include/xrtailor/math/MathFunctions.cuh:24-51)
__device__ __forceinline__ float Mn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float Mx(float a, float b) { return a > b ? a : b; }

__device__ __forceinline__ Box MergeBox(const Box& a, const Box& b) {
  return Box{make_float4(Mn(a.lo.x, b.lo.x), Mn(a.lo.y, b.lo.y), Mn(a.lo.z, b.lo.z), 0.0f),
             make_float4(Mx(a.hi.x, b.hi.x), Mx(a.hi.y, b.hi.y), Mx(a.hi.z, b.hi.z), 0.0f)};
}

// Bounds::Overlap(p): strict containment (xrtailor/src/physics/broad_phase/Bounds.cu:130-136)
__device__ __forceinline__ bool OverlapPoint(const Box& b, float4 p) {
  return p.x > b.lo.x && p.x < b.hi.x && p.y > b.lo.y && p.y < b.hi.y && p.z > b.lo.z &&
         p.z < b.hi.z;
}

__device__ __forceinline__ bool GetIntersection(float d1, float d2, float4 p1, float4 p2,
                                                float4& hit) {  // Bounds.cu:139-147
  if ((d1 * d2) >= 0.0f) return false;
  if (d1 == d2) return false;
  const float s = -d1 / (d2 - d1);
  hit = make_float4(p1.x + (p2.x - p1.x) * s, p1.y + (p2.y - p1.y) * s, p1.z + (p2.z - p1.z) * s,
                    0.0f);
  return true;
}

__device__ __forceinline__ bool InBox(float4 h, const Box& b, int axis) {  // Bounds.cu:149-158
  if (axis == 1 && h.z > b.lo.z && h.z < b.hi.z && h.y > b.lo.y && h.y < b.hi.y) return true;
  if (axis == 2 && h.z > b.lo.z && h.z < b.hi.z && h.x > b.lo.x && h.x < b.hi.x) return true;
  if (axis == 3 && h.x > b.lo.x && h.x < b.hi.x && h.y > b.lo.y && h.y < b.hi.y) return true;
  return false;
}

// Bounds::Overlap(p1, p2, hit): segment vs box (Bounds.cu:160-191)
__device__ __forceinline__ bool OverlapSegment(const Box& b, float4 p1, float4 p2) {
  if (p2.x < b.lo.x && p1.x < b.lo.x) return false;
  if (p2.x > b.hi.x && p1.x > b.hi.x) return false;
  if (p2.y < b.lo.y && p1.y < b.lo.y) return false;
  if (p2.y > b.hi.y && p1.y > b.hi.y) return false;
  if (p2.z < b.lo.z && p1.z < b.lo.z) return false;
  if (p2.z > b.hi.z && p1.z > b.hi.z) return false;
  if (OverlapPoint(b, p1)) return true;
  float4 h;
  return (GetIntersection(p1.x - b.lo.x, p2.x - b.lo.x, p1, p2, h) && InBox(h, b, 1)) ||
         (GetIntersection(p1.y - b.lo.y, p2.y - b.lo.y, p1, p2, h) && InBox(h, b, 2)) ||
         (GetIntersection(p1.z - b.lo.z, p2.z - b.lo.z, p1, p2, h) && InBox(h, b, 3)) ||
         (GetIntersection(p1.x - b.hi.x, p2.x - b.hi.x, p1, p2, h) && InBox(h, b, 1)) ||
         (GetIntersection(p1.y - b.hi.y, p2.y - b.hi.y, p1, p2, h) && InBox(h, b, 2)) ||
         (GetIntersection(p1.z - b.hi.z, p2.z - b.hi.z, p1, p2, h) && InBox(h, b, 3));
}

// Bounds::Overlap(b, thickness) = Overlap(b.Dilate(thickness)) (Bounds.cu:96-128)
__device__ __forceinline__ bool OverlapDilated(const Box& rhs, const Box& lhs, float th) {
  const float lx = lhs.lo.x - th, ly = lhs.lo.y - th, lz = lhs.lo.z - th;
  const float hx = lhs.hi.x + th, hy = lhs.hi.y + th, hz = lhs.hi.z + th;
  if (rhs.lo.x > hx || rhs.hi.x < lx) return false;
  if (rhs.lo.y > hy || rhs.hi.y < ly) return false;
  if (rhs.lo.z > hz || rhs.hi.z < lz) return false;
  return true;
}

// ------------------------------------------------------------------------------------
// Per-element bodies of the refit / traversal kernels, shared by the single-tree kernels
// (bvh.hip) and the multi-instance ones (contact.hip) so both run the same arithmetic.

// RefitExternalNodeMidstep (BVH.cu:770-818) / Update: pred = x1 + tol n1 (or x)
__device__ __forceinline__ void RefitPredBody(const float4* __restrict__ x1,
                                              const float4* __restrict__ n1, float tol, int i,
                                              float4* __restrict__ pred) {
  const float4 p = x1[i];
  if (n1) {
    const float4 n = n1[i];
    pred[i] = make_float4(p.x + tol * n.x, p.y + tol * n.y, p.z + tol * n.z, 0.0f);
  } else {
    pred[i] = make_float4(p.x, p.y, p.z, 0.0f);
  }
}

// leaf i grows by its 3 corners of pa (midstep) or by Face::ComputeBounds(ccd) of pa [and
// pb]; resets the arrival flag of internal node i for the bottom-up pass that follows
__device__ __forceinline__ void GrowLeafBody(const int4* __restrict__ nodes,
                                             const int4* __restrict__ tris,
                                             const float4* __restrict__ pa,
                                             const float4* __restrict__ pb, int nf, int i,
                                             Box* __restrict__ box, int32_t* __restrict__ flags) {
  const int ni = nf - 1;
  const int4 t = tris[nodes[ni + i].w];
  Box b = box[ni + i];
  float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
  float4 hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
  const int vid[3] = {t.x, t.y, t.z};
  for (int c = 0; c < 3; ++c) {
    const float4 p = pa[vid[c]];
    lo.x = Mn(lo.x, p.x); lo.y = Mn(lo.y, p.y); lo.z = Mn(lo.z, p.z);
    hi.x = Mx(hi.x, p.x); hi.y = Mx(hi.y, p.y); hi.z = Mx(hi.z, p.z);
    if (pb) {
      const float4 q = pb[vid[c]];
      lo.x = Mn(lo.x, q.x); lo.y = Mn(lo.y, q.y); lo.z = Mn(lo.z, q.z);
      hi.x = Mx(hi.x, q.x); hi.y = Mx(hi.y, q.y); hi.z = Mx(hi.z, q.z);
    }
  }
  box[ni + i] = MergeBox(b, Box{lo, hi});
  if (i < ni) flags[i] = 0;
}

// bottom-up boxes from leaf i: the second thread to reach a node merges its two children
// (the union of the descendant leaves, as AssignAABBAtomic_Kernel's atomic min / max,
// BVH.cu:540-557, and AssignAABBCAS_Kernel, BVH.cu:515-538).  grow keeps the old box too
// (Update's grow-only semantics).
__device__ __forceinline__ void BottomUpBody(const int4* __restrict__ nodes, Box* __restrict__ box,
                                             int32_t* __restrict__ flags, int nf, int grow, int i) {
  int node = nf - 1 + i;
  int p = nodes[node].x;
  // a root-ward walk is at most nf - 1 levels; the bound also guards against a malformed tree
  for (int level = 0; p >= 0 && p < nf - 1 && level < nf; ++level) {
    __threadfence();
    const int old = atomicAdd(&flags[p], 1);
    if (old == 0) return;
    __threadfence();
    const int4 pn = nodes[p];
    Box m = MergeBox(box[pn.y], box[pn.z]);
    if (grow) m = MergeBox(m, box[p]);
    box[p] = m;
    node = p;
    p = pn.x;
  }
}

// The same merge for a whole tree by ONE workgroup, level by level: a top-down
// breadth-first pass lists the internal nodes by depth in lvl (entries [0, nf - 1): the
// nodes; entries nf - 1 + d: where depth d starts), then the depths are merged deepest first
// with a workgroup barrier between them.  A node's children are one level deeper, so every
// merge reads finished boxes.  Each thread waits on a handful of loads per level instead of
// a chain of atomics per climbed node (an arrival-counter climb by one workgroup took 1.5 ms
// for the batched workload's largest trees), and the boxes are unions (min / max),
// independent of the order, hence identical to BottomUpBody's.  lvl holds 4 nf ints (the
// order, the depth starts, the depth count at 2 nf - 1, then each listed node's two
// children as int2 from 2 nf on); node 0 is the root; a malformed tree cannot list more than
// nf - 1 nodes.  The lists depend on the topology only, so they are kept while flags[nf - 1]
// (a slot the arrival counters do not use) is 1: every build clears it (LeafNodesBody).
__device__ __forceinline__ void LevelBottomUpWG(const int4* __restrict__ nodes,
                                                Box* __restrict__ box,
                                                int32_t* __restrict__ lvl,
                                                int32_t* __restrict__ flags, int nf, int grow) {
  const int ni = nf - 1;
  if (ni <= 0) return;
  int32_t* order = lvl;
  int32_t* start = lvl + ni;  // start[d], d <= depth count <= ni
  int2* kids = reinterpret_cast<int2*>(lvl + 2 * nf);  // (first, second) child of order[t]
  if (flags[ni] == 1) {       // the same value in every thread: written after the lists
    const int depths = lvl[2 * ni + 1];
    for (int d = depths - 1; d >= 0; --d) {
      const int lo = start[d], hi = start[d + 1];
      for (int t = lo + static_cast<int>(threadIdx.x); t < hi; t += blockDim.x) {
        const int p = order[t];
        const int2 ch = kids[t];  // loaded beside p: one dependent load less per node
        Box m = MergeBox(box[ch.x], box[ch.y]);
        if (grow) m = MergeBox(m, box[p]);
        box[p] = m;
      }
      __syncthreads();
    }
    return;
  }
  __shared__ int s_end;
  if (threadIdx.x == 0) {
    order[0] = 0;
    start[0] = 0;
    start[1] = 1;
    s_end = 1;
  }
  __syncthreads();
  int b = 0, e = 1, depths = 1;  // depths listed; the deepest is [b, e)
  while (b < e) {
    for (int t = b + static_cast<int>(threadIdx.x); t < e; t += blockDim.x) {
      const int4 nd = nodes[order[t]];
      if (nd.y >= 0 && nd.y < ni) {
        const int q = atomicAdd(&s_end, 1);
        if (q < ni) order[q] = nd.y;
      }
      if (nd.z >= 0 && nd.z < ni) {
        const int q = atomicAdd(&s_end, 1);
        if (q < ni) order[q] = nd.z;
      }
    }
    __syncthreads();
    const int next = min(s_end, ni);
    __syncthreads();  // every thread has read s_end before the next depth adds to it
    b = e;
    e = next;
    if (e > b) {
      if (threadIdx.x == 0) start[depths + 1] = e;
      ++depths;
    }
  }
  for (int t = static_cast<int>(threadIdx.x); t < min(e, ni); t += blockDim.x) {
    const int4 nd = nodes[order[t]];
    kids[t] = make_int2(nd.y, nd.z);
  }
  if (threadIdx.x == 0) {
    lvl[2 * ni + 1] = depths;
    flags[ni] = 1;  // every thread read the old value before the first barrier above
  }
  __syncthreads();
  for (int d = depths - 1; d >= 0; --d) {
    const int lo = start[d], hi = start[d + 1];
    for (int t = lo + static_cast<int>(threadIdx.x); t < hi; t += blockDim.x) {
      const int p = order[t];
      const int4 pn = nodes[p];
      Box m = MergeBox(box[pn.y], box[pn.z]);
      if (grow) m = MergeBox(m, box[p]);
      box[p] = m;
    }
    __syncthreads();
  }
}

// CountPairs_Kernel (BVH.cu:1149-1179): rhs leaves overlapping lhs leaf id dilated by th
__device__ __forceinline__ uint32_t CountPairsBody(const BvhView& l, const BvhView& r, float th,
                                                   int id) {
  const Box lb = l.box[l.nf - 1 + id];
  int node = 0;
  uint32_t c = 0;
  const int max_steps = 2 * r.nf;  // each node is visited at most once
  for (int step = 0; node != -1 && step < max_steps; ++step) {
    const int4 nd = r.nodes[node];
    if (OverlapDilated(r.box[node], lb, th)) {
      if (nd.w >= 0) {
        ++c;
        node = r.escape[node];
      } else {
        node = nd.y;
      }
    } else {
      node = r.escape[node];
    }
  }
  return c;
}

// Traversal record of node i for the pair traversal: its box with lo.w = left child
// (internal node) or -(object + 1) (leaf) and hi.w = escape, so that one 32-byte load serves
// a visit.  Written from nodes / box / escape after the boxes change.
__device__ __forceinline__ void PackNodeBody(const BvhView& t, int i) {
  const int4 nd = t.nodes[i];
  Box b = t.box[i];
  b.lo.w = __int_as_float(nd.w >= 0 ? -(nd.w + 1) : nd.y);
  b.hi.w = __int_as_float(t.escape[i]);
  t.packed[i] = b;
}

// the stackless traversal of CountPairsBody / FindPairsBody over packed records: on_leaf
// (rhs object) for every rhs leaf overlapping lb dilated by th, in the same order
template <class OnLeaf>
__device__ __forceinline__ void TraversePacked(const Box* __restrict__ pk, int nf, const Box& lb,
                                               float th, OnLeaf&& on_leaf) {
  int node = 0;
  const int max_steps = 2 * nf;  // each node is visited at most once
  for (int step = 0; node != -1 && step < max_steps; ++step) {
    const Box b = pk[node];
    const int code = __float_as_int(b.lo.w), esc = __float_as_int(b.hi.w);
    if (OverlapDilated(b, lb, th)) {
      if (code < 0) {
        on_leaf(-code - 1);
        node = esc;
      } else {
        node = code;
      }
    } else {
      node = esc;
    }
  }
}

// TraversePacked with the face record of every leaf slot (lf[j], leaf node nf - 1 + j)
// loaded beside the node's record, so a hit's face data costs no dependent load:
// on_leaf(rhs object, its face record) in the same order
template <class OnLeaf>
__device__ __forceinline__ void TraversePackedFace(const Box* __restrict__ pk,
                                                   const int4* __restrict__ lf, int nf,
                                                   const Box& lb, float th, OnLeaf&& on_leaf) {
  const int ni = nf - 1;
  int node = 0;
  const int max_steps = 2 * nf;
  for (int step = 0; node != -1 && step < max_steps; ++step) {
    const Box b = pk[node];
    const int4 f = node >= ni ? lf[node - ni] : make_int4(0, 0, 0, 0);
    const int code = __float_as_int(b.lo.w), esc = __float_as_int(b.hi.w);
    if (OverlapDilated(b, lb, th)) {
      if (code < 0) {
        on_leaf(-code - 1, f);
        node = esc;
      } else {
        node = code;
      }
    } else {
      node = esc;
    }
  }
}

// FindPairs_Kernel (BVH.cu:1189-1227): leaf id's pairs fill [lo, lo + n) backwards; only
// slots below cap are written
__device__ __forceinline__ void FindPairsBody(const BvhView& l, const BvhView& r, float th, int id,
                                              int64_t lo, uint32_t n, int2* __restrict__ pairs,
                                              int64_t cap) {
  const Box lb = l.box[l.nf - 1 + id];
  const int lobj = l.nodes[l.nf - 1 + id].w;
  int64_t index = lo + n - 1;
  int node = 0;
  const int max_steps = 2 * r.nf;
  for (int step = 0; node != -1 && step < max_steps; ++step) {
    const int4 nd = r.nodes[node];
    if (OverlapDilated(r.box[node], lb, th)) {
      if (nd.w >= 0) {
        if (index >= lo && index < cap) pairs[index] = make_int2(lobj, nd.w);
        --index;
        node = r.escape[node];
      } else {
        node = nd.y;
      }
    } else {
      node = r.escape[node];
    }
  }
}

// UpdateNodeNormals (xrtailor/src/physics/PhysicsMeshHelper.cu:196-210,282-295) for vertex
// v as a gather over a corner CSR in ascending face order (deterministic)
__device__ __forceinline__ float4 VertexNormalBody(const float4* __restrict__ x,
                                                   const int4* __restrict__ tris,
                                                   const int32_t* __restrict__ vptr,
                                                   const int32_t* __restrict__ corner, int v) {
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for (int k = vptr[v]; k < vptr[v + 1]; ++k) {
    const int c = corner[k];
    const int4 t = tris[c / 3];
    const int j = c % 3;
    const int a = j == 0 ? t.x : (j == 1 ? t.y : t.z);
    const int b = j == 0 ? t.y : (j == 1 ? t.z : t.x);
    const int d = j == 0 ? t.z : (j == 1 ? t.x : t.y);
    const float4 p = x[a], pb = x[b], pd = x[d];
    const float e0x = pb.x - p.x, e0y = pb.y - p.y, e0z = pb.z - p.z;
    const float e1x = pd.x - p.x, e1y = pd.y - p.y, e1z = pd.z - p.z;
    const float den = (e0x * e0x + e0y * e0y + e0z * e0z) * (e1x * e1x + e1y * e1y + e1z * e1z);
    const float cx = e0y * e1z - e0z * e1y, cy = e0z * e1x - e0x * e1z, cz = e0x * e1y - e0y * e1x;
    sx = sx + cx / den;
    sy = sy + cy / den;
    sz = sz + cz / den;
  }
  const float l = sqrtf(sx * sx + sy * sy + sz * sz);
  return make_float4(sx / l, sy / l, sz / l, 0.0f);
}

// ------------------------------------------------------------ radix sort bodies
// (256-thread workgroups; a tile of kSortTile keys per workgroup)
constexpr int kSortBlock = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortBlock * kSortItems;

// Stable LSD radix sort of (key, value), 8-bit digits: per-tile histograms, one
// exclusive scan over the digit-major table, then a scatter whose in-tile ranks come
// from wave64 ballots (lanes with equal digits), so equal keys keep their input order.
__device__ __forceinline__ void RadixHistBody(const uint32_t* __restrict__ keys, int n, int shift,
                                              int nblocks, uint32_t* __restrict__ hist, int blk) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int base = blk * kSortTile;
  for (int r = 0; r < kSortItems; ++r) {
    const int i = base + r * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * nblocks + blk] = cnt[threadIdx.x];
}

__device__ __forceinline__ void RadixScatterBody(
    const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    int32_t* __restrict__ vout, int n, int shift, int nblocks, const uint32_t* __restrict__ offs,
    int blk) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[kSortBlock / 64][256];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  run[t] = offs[t * nblocks + blk];
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int base = blk * kSortTile;
  for (int r = 0; r < kSortItems; ++r) {
    for (int k = 0; k < kSortBlock / 64; ++k) wc[k][t] = 0u;
    __syncthreads();
    const int i = base + r * kSortBlock + t;
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
    for (int bit = 0; bit < 8; ++bit) {
      const uint64_t bb = __ballot(valid && ((d >> bit) & 1u));
      peers &= ((d >> bit) & 1u) ? bb : ~bb;
    }
    const uint32_t rank = static_cast<uint32_t>(__popcll(peers & lt));
    if (valid && (peers & lt) == 0) wc[w][d] = static_cast<uint32_t>(__popcll(peers));
    __syncthreads();
    {  // per digit: ordered prefix over the waves of this round
      uint32_t acc = run[t];
      for (int k = 0; k < kSortBlock / 64; ++k) {
        const uint32_t c = wc[k][t];
        wc[k][t] = acc;
        acc += c;
      }
      run[t] = acc;
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = wc[w][d] + rank;
      if (pos < static_cast<uint32_t>(n)) {  // always true when histogram and scatter agree
        kout[pos] = key;
        vout[pos] = vin[i];
      }
    }
    __syncthreads();
  }
}

// exclusive scan of n uint32 in place by one 1024-thread workgroup; *total (if given)
// receives the sum
__device__ __forceinline__ void BlockExclusiveScan(uint32_t* __restrict__ a, int n,
                                                   int32_t* __restrict__ total) {
  __shared__ uint32_t part[1024];
  const int t = threadIdx.x;
  const int chunk = (n + 1023) / 1024;
  const int b = t * chunk, e = min(n, b + chunk);
  uint32_t s = 0;
  for (int i = b; i < e; ++i) s += a[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = t > 0 ? part[t - 1] : 0u;
  for (int i = b; i < e; ++i) {
    const uint32_t x = a[i];
    a[i] = run;
    run += x;
  }
  if (t == 1023 && total) *total = static_cast<int32_t>(part[1023]);
}

// QueryDeviceStackless (xrtailor/src/physics/broad_phase/lbvh/BVH.cu:957-1021): visits the
// leaves whose boxes contain p (segment == false) or meet segment p->q, in traversal order,
// calling f(object) for at most kMaxHits of them.
template <bool kSegment, typename F>
__device__ __forceinline__ int StacklessQuery(const BvhView& t, float4 p, float4 q, F&& f) {
  int node = 0, found = 0;
  const int max_steps = 2 * t.nf;  // a stackless pass visits each node at most once
  for (int step = 0; node != -1 && step < max_steps; ++step) {
    if (found >= kMaxHits) break;
    const Box bx = t.box[node];
    const bool hit = kSegment ? OverlapSegment(bx, p, q) : OverlapPoint(bx, p);
    const int4 nd = t.nodes[node];
    if (hit) {
      if (nd.w >= 0) {
        f(nd.w);
        ++found;
        node = t.escape[node];
      } else {
        node = nd.y;
      }
    } else {
      node = t.escape[node];
    }
  }
  return found;
}

// StacklessQuery over the packed traversal records (PackNodeBody): one 32-byte load per
// visit instead of the box, node and escape arrays; the same visits in the same order
template <bool kSegment, class F>
__device__ __forceinline__ int StacklessQueryPacked(const Box* __restrict__ pk, int nf, float4 p,
                                                    float4 q, F&& f) {
  int node = 0, found = 0;
  const int max_steps = 2 * nf;
  for (int step = 0; node != -1 && step < max_steps; ++step) {
    if (found >= kMaxHits) break;
    const Box bx = pk[node];
    const int code = __float_as_int(bx.lo.w), esc = __float_as_int(bx.hi.w);
    const bool hit = kSegment ? OverlapSegment(bx, p, q) : OverlapPoint(bx, p);
    if (hit) {
      if (code < 0) {
        f(-code - 1);
        ++found;
        node = esc;
      } else {
        node = code;
      }
    } else {
      node = esc;
    }
  }
  return found;
}

}  // namespace xrt
