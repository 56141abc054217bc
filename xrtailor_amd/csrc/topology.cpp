#include "topology.h"

#include <algorithm>
#include <cmath>
#include <queue>
#include <unordered_map>

namespace xrt {

namespace {
inline uint64_t EdgeKey(int a, int b) {
  uint32_t lo = static_cast<uint32_t>(std::min(a, b));
  uint32_t hi = static_cast<uint32_t>(std::max(a, b));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
struct V3 {
  double x, y, z;
};
inline V3 Load(const float* p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline V3 Sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 Cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 Normalize(V3 a) {
  double l = std::sqrt(Dot(a, a));
  return {a.x / l, a.y / l, a.z / l};
}
}  // namespace

std::vector<int32_t> BuildEdges(const int32_t* tris, int nt) {
  std::vector<int32_t> edges;
  std::unordered_map<uint64_t, int> seen;
  seen.reserve(static_cast<size_t>(nt) * 2);
  for (int f = 0; f < nt; ++f) {
    for (int l = 0; l < 3; ++l) {
      int a = tris[3 * f + l], b = tris[3 * f + (l + 1) % 3];
      auto it = seen.emplace(EdgeKey(a, b), static_cast<int>(edges.size() / 2));
      if (it.second) {
        edges.push_back(a);
        edges.push_back(b);
      }
    }
  }
  return edges;
}

std::vector<int32_t> BuildBends(const int32_t* tris, int nt, const std::vector<int32_t>& edges) {
  const int ne = static_cast<int>(edges.size() / 2);
  std::unordered_map<uint64_t, int> index;
  index.reserve(static_cast<size_t>(ne) * 2);
  for (int e = 0; e < ne; ++e) index.emplace(EdgeKey(edges[2 * e], edges[2 * e + 1]), e);
  // edge -> faces (ascending face id)
  std::vector<int32_t> ptr(ne + 1, 0);
  std::vector<int32_t> face_edge(static_cast<size_t>(3) * nt);
  for (int f = 0; f < nt; ++f)
    for (int l = 0; l < 3; ++l) {
      int e = index[EdgeKey(tris[3 * f + l], tris[3 * f + (l + 1) % 3])];
      face_edge[3 * f + l] = e;
      ptr[e + 1]++;
    }
  for (int e = 0; e < ne; ++e) ptr[e + 1] += ptr[e];
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1), ef(static_cast<size_t>(3) * nt);
  for (int f = 0; f < nt; ++f)
    for (int l = 0; l < 3; ++l) ef[fill[face_edge[3 * f + l]]++] = f;

  std::vector<int32_t> bends;
  for (int e = 0; e < ne; ++e) {
    const int a = edges[2 * e], b = edges[2 * e + 1];
    for (int i = ptr[e]; i < ptr[e + 1]; ++i)
      for (int j = i + 1; j < ptr[e + 1]; ++j) {
        const int f1 = ef[i], f2 = ef[j];
        int v0 = a, v1 = b;
        bool right = false;
        for (int l = 0; l < 3; ++l)
          if (tris[3 * f1 + l] == v0 && tris[3 * f1 + (l + 1) % 3] == v1) right = true;
        if (!right) std::swap(v0, v1);
        int v2 = -1, v3 = -1;
        for (int l = 0; l < 3; ++l) {
          if (tris[3 * f1 + l] != a && tris[3 * f1 + l] != b) v2 = tris[3 * f1 + l];
          if (tris[3 * f2 + l] != a && tris[3 * f2 + l] != b) v3 = tris[3 * f2 + l];
        }
        bends.insert(bends.end(), {v0, v1, v2, v3});
      }
  }
  return bends;
}

std::vector<float> RestLengths(const float* pos, const std::vector<int32_t>& edges) {
  const size_t ne = edges.size() / 2;
  std::vector<float> out(ne);
  for (size_t e = 0; e < ne; ++e) {
    V3 d = Sub(Load(pos, edges[2 * e]), Load(pos, edges[2 * e + 1]));
    out[e] = static_cast<float>(std::sqrt(Dot(d, d)));
  }
  return out;
}

std::vector<float> RestAngles(const float* pos, const std::vector<int32_t>& bends) {
  const size_t nb = bends.size() / 4;
  std::vector<float> out(nb);
  for (size_t b = 0; b < nb; ++b) {
    V3 v1 = Load(pos, bends[4 * b]), v2 = Load(pos, bends[4 * b + 1]);
    V3 v3 = Load(pos, bends[4 * b + 2]), v4 = Load(pos, bends[4 * b + 3]);
    V3 n1 = Normalize(Cross(Sub(v1, v3), Sub(v2, v3)));
    V3 n2 = Normalize(Cross(Sub(v2, v4), Sub(v1, v4)));
    double d = std::min(1.0, std::max(-1.0, Dot(n1, n2)));
    out[b] = static_cast<float>(std::acos(d));
  }
  return out;
}

uint64_t SplitMix64(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

std::vector<int32_t> ColorConstraints(const int32_t* cons, int arity, int n, int nv,
                                      uint64_t seed, int mode, int* n_colors) {
  *n_colors = 0;
  if (n <= 0) return {};
  // vertex -> constraints
  std::vector<int32_t> vptr(nv + 1, 0);
  for (int i = 0; i < n * arity; ++i) vptr[cons[i] + 1]++;
  for (int v = 0; v < nv; ++v) vptr[v + 1] += vptr[v];
  std::vector<int32_t> vfill(vptr.begin(), vptr.end() - 1), vc(static_cast<size_t>(n) * arity);
  for (int c = 0; c < n; ++c)
    for (int a = 0; a < arity; ++a) vc[vfill[cons[c * arity + a]]++] = c;
  // adjacency (BuildStretchAdjacencyTable / BuildBendAdjacencyTable: shared unknowns)
  std::vector<int32_t> aptr(n + 1, 0), adj, tmp;
  adj.reserve(static_cast<size_t>(n) * 12);
  for (int c = 0; c < n; ++c) {
    tmp.clear();
    for (int a = 0; a < arity; ++a) {
      int v = cons[c * arity + a];
      for (int k = vptr[v]; k < vptr[v + 1]; ++k)
        if (vc[k] != c) tmp.push_back(vc[k]);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    adj.insert(adj.end(), tmp.begin(), tmp.end());
    aptr[c + 1] = static_cast<int32_t>(adj.size());
  }
  std::vector<int32_t> color(n, 0);
  if (mode == kColorGreedy) {
    // first fit in constraint index order: each constraint takes the smallest color no
    // lower-indexed neighbour holds (the index order follows the mesh's face order)
    std::vector<int32_t> stamp(static_cast<size_t>(n) + 2, -1);
    for (int c = 0; c < n; ++c) {
      for (int k = aptr[c]; k < aptr[c + 1]; ++k)
        if (adj[k] < c) stamp[color[adj[k]]] = c;
      int col = 0;
      while (stamp[col] == c) ++col;
      color[c] = col;
    }
  } else {
    int min_deg = 1 << 30;
    for (int c = 0; c < n; ++c) min_deg = std::min(min_deg, aptr[c + 1] - aptr[c]);
    min_deg = std::max(min_deg, 1);
    std::vector<int32_t> pal_ptr(n + 1, 0);
    for (int c = 0; c < n; ++c) pal_ptr[c + 1] = pal_ptr[c] + (aptr[c + 1] - aptr[c]) + 2;
    std::vector<uint8_t> forbid(pal_ptr[n], 0), colored(n, 0), accept(n, 0);
    std::vector<int32_t> vsize(n), act(n);
    std::vector<uint64_t> rs(n), prio(n);
    uint64_t s0 = seed;
    const uint64_t base = SplitMix64(&s0);
    for (int c = 0; c < n; ++c) {
      vsize[c] = (aptr[c + 1] - aptr[c]) / min_deg + 1;  // Initialization_Kernel, Base.cu:19-20
      rs[c] = base + static_cast<uint64_t>(c) * 0xD1B54A32D192ED03ULL;
      uint64_t ps = rs[c] ^ 0x5851F42D4C957F2DULL;
      prio[c] = SplitMix64(&ps);  // conflict ranking: hash priority, ties by index
      act[c] = c;
    }
    auto outranks = [&](int a, int b) { return prio[a] > prio[b] || (prio[a] == prio[b] && a > b); };
    int nact = n;
    while (nact > 0) {
      // tentative coloring (TentativeColoring_Kernel, Base.cu:35-69)
      for (int q = 0; q < nact; ++q) {
        const int c = act[q];
        const uint64_t r = SplitMix64(&rs[c]);
        const double u = (static_cast<double>(r >> 11) + 1.0) * (1.0 / 9007199254740992.0);
        const int vs = vsize[c];
        int lc = static_cast<int>(u * (static_cast<double>(vs) - 1.0 + 0.999999));
        for (int t = 0; t < vs && forbid[pal_ptr[c] + lc]; ++t) lc = (lc + 1) % vs;
        color[c] = lc;
      }
      // conflict resolution on the start-of-round state (ConflictResolution_Kernel, Base.cu:71-118)
      for (int q = 0; q < nact; ++q) {
        const int c = act[q];
        accept[c] = 0;
        bool conflict_uncolored = false, conflict_colored = false, highest = true;
        for (int k = aptr[c]; k < aptr[c + 1]; ++k) {
          const int j = adj[k];
          if (color[j] == color[c]) (colored[j] ? conflict_colored : conflict_uncolored) = true;
          if (!colored[j] && outranks(j, c)) highest = false;
        }
        if (conflict_colored || (conflict_uncolored && !highest)) continue;
        accept[c] = 1;
      }
      for (int q = 0; q < nact; ++q) {
        const int c = act[q];
        if (!accept[c]) continue;
        colored[c] = 1;
        for (int k = aptr[c]; k < aptr[c + 1]; ++k) {
          const int j = adj[k];
          if (color[c] < pal_ptr[j + 1] - pal_ptr[j]) forbid[pal_ptr[j] + color[c]] = 1;
        }
      }
      // feed the hungry (FeedTheHungry_Kernel, Base.cu:120-139); keep the uncolored active
      int na = 0;
      for (int q = 0; q < nact; ++q) {
        const int c = act[q];
        if (colored[c]) continue;
        bool avail = false;
        for (int t = 0; t < vsize[c] && !avail; ++t) avail = !forbid[pal_ptr[c] + t];
        if (!avail && vsize[c] < pal_ptr[c + 1] - pal_ptr[c]) vsize[c]++;
        act[na++] = c;
      }
      nact = na;
    }
  }
  // compact to 0..k-1 so every used color is visited by the color loop
  const int maxc = *std::max_element(color.begin(), color.end());
  std::vector<int32_t> remap(maxc + 1, -1);
  for (int c = 0; c < n; ++c) remap[color[c]] = 0;
  int used = 0;
  for (int k = 0; k <= maxc; ++k)
    if (remap[k] == 0) remap[k] = used++;
  for (int c = 0; c < n; ++c) color[c] = remap[color[c]];
  *n_colors = used;
  return color;
}

ContactTopo BuildContactTopo(const float* pos, int nv, const int32_t* tris, int nt,
                             const std::vector<int32_t>& edges) {
  ContactTopo t;
  const int ne = static_cast<int>(edges.size() / 2);
  std::unordered_map<uint64_t, int> index;
  index.reserve(static_cast<size_t>(ne) * 2);
  for (int e = 0; e < ne; ++e) index.emplace(EdgeKey(edges[2 * e], edges[2 * e + 1]), e);
  t.face_edges.resize(static_cast<size_t>(3) * nt);
  for (int f = 0; f < nt; ++f)
    for (int l = 0; l < 3; ++l)
      t.face_edges[3 * f + l] = index[EdgeKey(tris[3 * f + l], tris[3 * f + (l + 1) % 3])];
  t.ef_ptr.assign(ne + 1, 0);
  for (int i = 0; i < 3 * nt; ++i) t.ef_ptr[t.face_edges[i] + 1]++;
  for (int e = 0; e < ne; ++e) t.ef_ptr[e + 1] += t.ef_ptr[e];
  t.ef.resize(static_cast<size_t>(3) * nt);
  {
    std::vector<int32_t> fill(t.ef_ptr.begin(), t.ef_ptr.end() - 1);
    for (int f = 0; f < nt; ++f)
      for (int l = 0; l < 3; ++l) t.ef[fill[t.face_edges[3 * f + l]]++] = f;
  }
  // vertex -> faces, ascending
  std::vector<int32_t> vptr(nv + 1, 0), vf(static_cast<size_t>(3) * nt);
  for (int i = 0; i < 3 * nt; ++i) vptr[tris[i] + 1]++;
  for (int v = 0; v < nv; ++v) vptr[v + 1] += vptr[v];
  {
    std::vector<int32_t> fill(vptr.begin(), vptr.end() - 1);
    for (int f = 0; f < nt; ++f)
      for (int l = 0; l < 3; ++l) vf[fill[tris[3 * f + l]]++] = f;
  }
  t.rtri.assign(nt, 0u);
  for (int fid = 0; fid < nt; ++fid) {
    uint32_t info = 0;
    for (int i = 0; i < 3; ++i) {
      const int vi = tris[3 * fid + i], vj = tris[3 * fid + (i + 1) % 3];
      if (vf[vptr[vi]] == fid) info |= 1u << i;  // first face of the vertex owns it
      int64_t other = 0xffffffffLL;  // smallest other face on edge (vi, vj)
      for (int a = vptr[vi]; a < vptr[vi + 1] && other == 0xffffffffLL; ++a) {
        const int fa = vf[a];
        if (fa == fid) continue;
        for (int b = vptr[vj]; b < vptr[vj + 1]; ++b)
          if (vf[b] == fa) {
            other = fa;
            break;
          }
      }
      if (static_cast<int64_t>(fid) < other) info |= 1u << (3 + i);  // lower face owns the edge
    }
    t.rtri[fid] = info;
  }
  t.face_n.resize(static_cast<size_t>(3) * nt);
  for (int f = 0; f < nt; ++f) {
    const V3 a = Load(pos, tris[3 * f]), b = Load(pos, tris[3 * f + 1]), c = Load(pos, tris[3 * f + 2]);
    const V3 n = Normalize(Cross(Sub(b, a), Sub(c, a)));
    t.face_n[3 * f] = static_cast<float>(n.x);
    t.face_n[3 * f + 1] = static_cast<float>(n.y);
    t.face_n[3 * f + 2] = static_cast<float>(n.z);
  }
  return t;
}

std::vector<double> Geodesic(const float* pos, int nv, const std::vector<int32_t>& edges, int src) {
  const int ne = static_cast<int>(edges.size() / 2);
  std::vector<int32_t> ptr(nv + 1, 0);
  for (int e = 0; e < ne; ++e) {
    ptr[edges[2 * e] + 1]++;
    ptr[edges[2 * e + 1] + 1]++;
  }
  for (int v = 0; v < nv; ++v) ptr[v + 1] += ptr[v];
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1), nb(static_cast<size_t>(2) * ne);
  for (int e = 0; e < ne; ++e) {
    nb[fill[edges[2 * e]]++] = edges[2 * e + 1];
    nb[fill[edges[2 * e + 1]]++] = edges[2 * e];
  }
  std::vector<double> dist(nv, INFINITY);
  using Item = std::pair<double, int>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  dist[src] = 0.0;
  pq.push({0.0, src});
  while (!pq.empty()) {
    auto [d, v] = pq.top();
    pq.pop();
    if (d > dist[v]) continue;
    for (int k = ptr[v]; k < ptr[v + 1]; ++k) {
      const int u = nb[k];
      V3 diff = Sub(Load(pos, v), Load(pos, u));
      const double w = std::sqrt(Dot(diff, diff));
      if (dist[u] > dist[v] + w) {
        dist[u] = dist[v] + w;
        pq.push({dist[u], u});
      }
    }
  }
  return dist;
}

}  // namespace xrt
