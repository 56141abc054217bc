/*
 * xrt_oracle.c -- serial FP64 restatement of XRTailor's cloth substep (dynamics part).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
 * baseline.  Nothing in the product path links or calls this file.
 *
 * Pinned by (tests/test_oracle_*.py): closed-form free fall with damping, the exact
 * two-particle XPBD update, equal-mass momentum conservation, analytic single-hinge
 * bend angles incl. the flat => pi/2 quirk, finite-difference checks of the bend
 * gradients, brute-force Dijkstra on small grids, coloring validity/determinism.
 */
#include "xrt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846
#define ORC_EPS 1e-6 /* EPSILON (xrtailor/include/xrtailor/core/Scalar.hpp:37) */

/* ------------------------------------------------------------------ vec3 */
typedef struct { double x, y, z; } d3;
static inline d3 mk(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 ld(const double* p, int i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline void st(double* p, int i, d3 a) { p[3 * i] = a.x; p[3 * i + 1] = a.y; p[3 * i + 2] = a.z; }
static inline d3 add(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 sub(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 scl(d3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 crs(d3 a, d3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double len(d3 a) { return sqrt(dot(a, a)); }
static inline d3 nrm(d3 a) { return scl(a, 1.0 / len(a)); }

/* ------------------------------------------------------------- topology */
/* open-addressing map (u64 key -> int) used for edge lookup */
typedef struct { uint64_t* k; int32_t* v; size_t cap; } emap;
static void emap_init(emap* m, size_t n) {
  size_t cap = 16;
  while (cap < 2 * n + 16) cap <<= 1;
  m->cap = cap;
  m->k = (uint64_t*)malloc(cap * sizeof(uint64_t));
  m->v = (int32_t*)malloc(cap * sizeof(int32_t));
  for (size_t i = 0; i < cap; i++) m->k[i] = UINT64_MAX;
}
static void emap_free(emap* m) { free(m->k); free(m->v); }
static inline uint64_t ekey(int a, int b) {
  uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
  return ((uint64_t)hi << 32) | lo;
}
static inline size_t ehash(uint64_t k, size_t cap) {
  k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33;
  return (size_t)k & (cap - 1);
}
/* returns existing value or inserts v and returns -1 */
static int emap_get_or_put(emap* m, uint64_t key, int32_t v) {
  size_t h = ehash(key, m->cap);
  while (m->k[h] != UINT64_MAX) {
    if (m->k[h] == key) return m->v[h];
    h = (h + 1) & (m->cap - 1);
  }
  m->k[h] = key; m->v[h] = v;
  return -1;
}
static int emap_get(const emap* m, uint64_t key) {
  size_t h = ehash(key, m->cap);
  while (m->k[h] != UINT64_MAX) {
    if (m->k[h] == key) return m->v[h];
    h = (h + 1) & (m->cap - 1);
  }
  return -1;
}

int orc_build_edges(const int32_t* tris, int nt, int32_t* edges_out, int max_edges) {
  emap m; emap_init(&m, (size_t)3 * nt);
  int ne = 0;
  for (int f = 0; f < nt; f++) {
    for (int l = 0; l < 3; l++) {
      int a = tris[3 * f + l], b = tris[3 * f + (l + 1) % 3];
      if (emap_get_or_put(&m, ekey(a, b), ne) < 0) {
        if (ne >= max_edges) { emap_free(&m); return -1; }
        edges_out[2 * ne] = a; edges_out[2 * ne + 1] = b; ne++;
      }
    }
  }
  emap_free(&m);
  return ne;
}

int orc_build_bends(const int32_t* tris, int nt, const int32_t* edges, int ne,
                    int32_t* bends_out, int max_bends) {
  /* edge -> list of faces (ascending face id), CSR */
  emap m; emap_init(&m, (size_t)ne);
  for (int e = 0; e < ne; e++) emap_get_or_put(&m, ekey(edges[2 * e], edges[2 * e + 1]), e);
  int32_t* cnt = (int32_t*)calloc((size_t)ne + 1, sizeof(int32_t));
  for (int f = 0; f < nt; f++)
    for (int l = 0; l < 3; l++)
      cnt[emap_get(&m, ekey(tris[3 * f + l], tris[3 * f + (l + 1) % 3])) + 1]++;
  for (int e = 0; e < ne; e++) cnt[e + 1] += cnt[e];
  int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)ne);
  int32_t* ef = (int32_t*)malloc(sizeof(int32_t) * (size_t)(3 * nt + 1));
  memcpy(fill, cnt, sizeof(int32_t) * (size_t)ne);
  for (int f = 0; f < nt; f++)
    for (int l = 0; l < 3; l++) {
      int e = emap_get(&m, ekey(tris[3 * f + l], tris[3 * f + (l + 1) % 3]));
      ef[fill[e]++] = f;
    }
  int nb = 0;
  for (int e = 0; e < ne; e++) {
    int a = edges[2 * e], b = edges[2 * e + 1];
    for (int i = cnt[e]; i < cnt[e + 1]; i++)
      for (int j = i + 1; j < cnt[e + 1]; j++) {
        int f1 = ef[i], f2 = ef[j];
        int v0 = a, v1 = b;
        /* (v0,v1) must appear in f1's cyclic order, else swap (ClothSolverHelper.cpp:39-51) */
        int right = 0;
        for (int l = 0; l < 3; l++)
          if (tris[3 * f1 + l] == v0 && tris[3 * f1 + (l + 1) % 3] == v1) right = 1;
        if (!right) { int t = v0; v0 = v1; v1 = t; }
        int v2 = -1, v3 = -1;
        for (int l = 0; l < 3; l++) {
          int q = tris[3 * f1 + l];
          if (q != a && q != b) v2 = q;
          q = tris[3 * f2 + l];
          if (q != a && q != b) v3 = q;
        }
        if (nb >= max_bends) { nb = -1; goto done; }
        bends_out[4 * nb] = v0; bends_out[4 * nb + 1] = v1;
        bends_out[4 * nb + 2] = v2; bends_out[4 * nb + 3] = v3;
        nb++;
      }
  }
done:
  free(cnt); free(fill); free(ef); emap_free(&m);
  return nb;
}

void orc_rest_lengths(const double* x, const int32_t* edges, int ne, double* out) {
  for (int e = 0; e < ne; e++) out[e] = len(sub(ld(x, edges[2 * e]), ld(x, edges[2 * e + 1])));
}

void orc_rest_angles(const double* x, const int32_t* bends, int nb, double* out) {
  for (int b = 0; b < nb; b++) {
    d3 v1 = ld(x, bends[4 * b]), v2 = ld(x, bends[4 * b + 1]);
    d3 v3 = ld(x, bends[4 * b + 2]), v4 = ld(x, bends[4 * b + 3]);
    d3 n1 = nrm(crs(sub(v1, v3), sub(v2, v3)));
    d3 n2 = nrm(crs(sub(v2, v4), sub(v1, v4)));
    double d = dot(n1, n2);
    if (d < -1) d = -1; else if (d > 1) d = 1;
    out[b] = acos(d);
  }
}

/* ------------------------------------------------------------- coloring */
uint64_t orc_splitmix64(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

int orc_color_constraints(const int32_t* cons, int arity, int n, int nv, uint64_t seed,
                          int mode, int32_t* colors_out) {
  if (n <= 0) return 0;
  /* vertex -> constraints CSR */
  int32_t* vptr = (int32_t*)calloc((size_t)nv + 1, sizeof(int32_t));
  for (int i = 0; i < n * arity; i++) vptr[cons[i] + 1]++;
  for (int v = 0; v < nv; v++) vptr[v + 1] += vptr[v];
  int32_t* vfill = (int32_t*)malloc(sizeof(int32_t) * (size_t)nv);
  memcpy(vfill, vptr, sizeof(int32_t) * (size_t)nv);
  int32_t* vc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n * arity));
  for (int c = 0; c < n; c++)
    for (int a = 0; a < arity; a++) vc[vfill[cons[c * arity + a]]++] = c;
  /* adjacency: distinct constraints sharing a vertex (ascending) */
  int32_t* aptr = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  int32_t cap = 64;
  int32_t* adj = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
  int32_t* tmp = NULL; int tcap = 0;
  aptr[0] = 0;
  for (int c = 0; c < n; c++) {
    int tn = 0;
    for (int a = 0; a < arity; a++) {
      int v = cons[c * arity + a];
      int need = tn + (vptr[v + 1] - vptr[v]);
      if (need > tcap) { tcap = need * 2; tmp = (int32_t*)realloc(tmp, sizeof(int32_t) * (size_t)tcap); }
      for (int k = vptr[v]; k < vptr[v + 1]; k++) if (vc[k] != c) tmp[tn++] = vc[k];
    }
    qsort(tmp, (size_t)tn, sizeof(int32_t), cmp_i32);
    int u = 0;
    for (int k = 0; k < tn; k++) if (u == 0 || tmp[u - 1] != tmp[k]) tmp[u++] = tmp[k];
    if (aptr[c] + u > cap) { while (aptr[c] + u > cap) cap *= 2; adj = (int32_t*)realloc(adj, sizeof(int32_t) * (size_t)cap); }
    memcpy(adj + aptr[c], tmp, sizeof(int32_t) * (size_t)u);
    aptr[c + 1] = aptr[c] + u;
  }
  free(tmp);
  int min_deg = 1 << 30;
  for (int c = 0; c < n; c++) { int d = aptr[c + 1] - aptr[c]; if (d < min_deg) min_deg = d; }
  if (min_deg < 1) min_deg = 1;
  /* palettes: capacity deg+2; forbidden colors tracked separately so that a fed
   * palette never re-offers a color held by a colored neighbour */
  int32_t* pal_ptr = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  pal_ptr[0] = 0;
  for (int c = 0; c < n; c++) pal_ptr[c + 1] = pal_ptr[c] + (aptr[c + 1] - aptr[c]) + 2;
  uint8_t* forbid = (uint8_t*)calloc((size_t)pal_ptr[n], 1);
  int32_t* vsize = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* color = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  uint8_t* colored = (uint8_t*)calloc((size_t)n, 1);
  uint8_t* accept = (uint8_t*)calloc((size_t)n, 1);
  uint64_t* rs = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
  uint64_t s0 = seed;
  uint64_t base = orc_splitmix64(&s0);
  /* conflict ranking: seeded hash priority (ties by index) instead of the bare index, so
   * conflicts resolve in O(log n) rounds rather than a wave along the index order */
  uint64_t* prio = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
  int32_t* act = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  for (int c = 0; c < n; c++) {
    int d = aptr[c + 1] - aptr[c];
    vsize[c] = d / min_deg + 1; /* palette size floor(deg/shrinking)+1 (Base.cu:19-20) */
    color[c] = 0;
    rs[c] = base + (uint64_t)c * 0xD1B54A32D192ED03ULL;
    uint64_t ps = rs[c] ^ 0x5851F42D4C957F2DULL;
    prio[c] = orc_splitmix64(&ps);
    act[c] = c;
  }
  int nact = n;
  if (mode == 0) {
    /* first fit in index order: smallest color no lower-indexed neighbour holds */
    int32_t* stamp = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 2));
    for (int k = 0; k < n + 2; k++) stamp[k] = -1;
    for (int c = 0; c < n; c++) {
      for (int k = aptr[c]; k < aptr[c + 1]; k++)
        if (adj[k] < c) stamp[color[adj[k]]] = c;
      int col = 0;
      while (stamp[col] == c) col++;
      color[c] = col;
    }
    free(stamp);
    nact = 0;
  }
  while (nact > 0) {
    /* tentative coloring (Base.cu:35-69) */
    for (int a = 0; a < nact; a++) {
      int c = act[a];
      uint64_t r = orc_splitmix64(&rs[c]);
      double u = ((double)(r >> 11) + 1.0) * (1.0 / 9007199254740992.0); /* (0,1] */
      int vs = vsize[c];
      int rc = (int)(u * ((double)vs - 1.0 + 0.999999));
      int lc = rc;
      for (int t = 0; t < vs; t++) {
        if (!forbid[pal_ptr[c] + lc]) break;
        lc = (lc + 1) % vs;
      }
      color[c] = lc;
    }
    /* conflict resolution on the start-of-round state (Base.cu:71-118) */
    for (int a = 0; a < nact; a++) {
      int c = act[a];
      accept[c] = 0;
      int conflict_uncolored = 0, conflict_colored = 0, highest = 1;
      for (int k = aptr[c]; k < aptr[c + 1]; k++) {
        int j = adj[k];
        if (color[j] == color[c]) { if (colored[j]) conflict_colored = 1; else conflict_uncolored = 1; }
        if (!colored[j] && (prio[c] < prio[j] || (prio[c] == prio[j] && c < j))) highest = 0;
      }
      if (conflict_colored) continue;
      if (conflict_uncolored && !highest) continue;
      accept[c] = 1;
    }
    for (int a = 0; a < nact; a++) {
      int c = act[a];
      if (!accept[c]) continue;
      colored[c] = 1;
      for (int k = aptr[c]; k < aptr[c + 1]; k++) {
        int j = adj[k];
        int cap_j = pal_ptr[j + 1] - pal_ptr[j];
        if (color[c] < cap_j) forbid[pal_ptr[j] + color[c]] = 1;
      }
    }
    /* feed the hungry (Base.cu:120-139), then drop the newly colored from the active list */
    int na = 0;
    for (int a = 0; a < nact; a++) {
      int c = act[a];
      if (colored[c]) continue;
      int avail = 0;
      for (int t = 0; t < vsize[c]; t++) if (!forbid[pal_ptr[c] + t]) { avail = 1; break; }
      if (!avail && vsize[c] < pal_ptr[c + 1] - pal_ptr[c]) vsize[c]++;
      act[na++] = c;
    }
    nact = na;
  }
  free(prio); free(act);
  /* compact used colors to 0..k-1 (ascending original color) */
  int maxc = 0;
  for (int c = 0; c < n; c++) if (color[c] > maxc) maxc = color[c];
  int32_t* remap = (int32_t*)malloc(sizeof(int32_t) * ((size_t)maxc + 1));
  for (int k = 0; k <= maxc; k++) remap[k] = -1;
  for (int c = 0; c < n; c++) remap[color[c]] = 0;
  int used = 0;
  for (int k = 0; k <= maxc; k++) if (remap[k] == 0) remap[k] = used++;
  for (int c = 0; c < n; c++) colors_out[c] = remap[color[c]];
  free(remap); free(vptr); free(vfill); free(vc); free(aptr); free(adj); free(pal_ptr);
  free(forbid); free(vsize); free(color); free(colored); free(accept); free(rs);
  return used;
}

/* ------------------------------------------------------------- dynamics */
void orc_predict(orc_cloth* c, const orc_params* p, double dt) {
  d3 g = mk(p->gravity[0], p->gravity[1], p->gravity[2]);
  for (int i = 0; i < c->nv; i++) {
    d3 v = ld(c->v, i);
    if (c->inv_mass[i] > ORC_EPS) v = add(v, scl(g, dt));
    st(c->v, i, v);
    st(c->x, i, add(ld(c->x0, i), scl(v, dt)));
  }
}

void orc_stretch_color(orc_cloth* c, int k, double dt, int iter) {
  for (int id = c->s_color_off[k]; id < c->s_color_off[k + 1]; id++) {
    if (iter == 0) c->s_lambda[id] = 0.0;
    int i1 = c->s_idx[2 * id], i2 = c->s_idx[2 * id + 1];
    double rest = c->s_rest[id];
    d3 diff = sub(ld(c->x, i1), ld(c->x, i2));
    double dist = len(diff);
    double w1 = c->inv_mass[i1], w2 = c->inv_mass[i2];
    if (dist != rest && w1 + w2 > 0.0) {
      d3 g = scl(diff, 1.0 / (dist + ORC_EPS));
      double at = c->s_compliance[id] / (dt * dt);
      double C = dist - rest;
      double dl = (-C - at * c->s_lambda[id]) / (w1 + w2 + at);
      d3 corr = scl(g, dl);
      c->s_lambda[id] += dl;
      st(c->x, i1, add(ld(c->x, i1), scl(corr, w1)));
      st(c->x, i2, sub(ld(c->x, i2), scl(corr, w2)));
    }
  }
}

void orc_bend_color(orc_cloth* c, const orc_params* p, int k, double dt) {
  /* bend_model 1: rest angle in the solver's own normal convention (pi - theta0), d in
   * [-1,1] and the PBD step along -q (Mueller 2007); 0: the reference as written */
  const int consistent = p->bend_model != 0;
  const double dlo = consistent ? -1.0 : 0.0;
  const double sgn = consistent ? -1.0 : 1.0;
  for (int id = c->b_color_off[k]; id < c->b_color_off[k + 1]; id++) {
    int i1 = c->b_idx[4 * id], i2 = c->b_idx[4 * id + 1];
    int i3 = c->b_idx[4 * id + 2], i4 = c->b_idx[4 * id + 3];
    double w1 = c->inv_mass[i1], w2 = c->inv_mass[i2], w3 = c->inv_mass[i3], w4 = c->inv_mass[i4];
    d3 p1 = ld(c->x, i1);
    d3 p2 = sub(ld(c->x, i2), p1), p3 = sub(ld(c->x, i3), p1), p4 = sub(ld(c->x, i4), p1);
    d3 c23 = crs(p2, p3), c24 = crs(p2, p4);
    d3 n1 = nrm(c23), n2 = nrm(c24);
    double d = dot(n1, n2);
    int dnan = isnan(d);
    d = d < dlo ? dlo : (d > 1.0 ? 1.0 : d); /* clamp to [0,1] (BasicSolverHelper.cu:36-37) */
    double ang = acos(d);
    double th = consistent ? (ORC_PI - c->b_theta[id]) : c->b_theta[id];
    if (ang < ORC_EPS || dnan || isnan(ang)) continue;
    double l23 = len(c23) + ORC_EPS, l24 = len(c24) + ORC_EPS;
    d3 q3 = scl(add(crs(p2, n2), scl(crs(n1, p2), d)), 1.0 / l23);
    d3 q4 = scl(add(crs(p2, n1), scl(crs(n2, p2), d)), 1.0 / l24);
    d3 q2 = sub(scl(scl(add(crs(p3, n2), scl(crs(n1, p3), d)), 1.0 / l23), -1.0),
                scl(add(crs(p4, n1), scl(crs(n2, p4), d)), 1.0 / l24));
    d3 q1 = scl(add(add(q2, q3), q4), -1.0);
    double denom = c->b_compliance[id] / dt / dt +
                   (w1 * dot(q1, q1) + w2 * dot(q2, q2) + w3 * dot(q3, q3) + w4 * dot(q4, q4));
    if (denom < ORC_EPS) continue;
    double lam = sgn * sqrt(1.0 - d * d) * (ang - th) / denom;
    st(c->x, i1, add(ld(c->x, i1), scl(q1, w1 * lam)));
    st(c->x, i2, add(ld(c->x, i2), scl(q2, w2 * lam)));
    st(c->x, i3, add(ld(c->x, i3), scl(q3, w3 * lam)));
    st(c->x, i4, add(ld(c->x, i4), scl(q4, w4 * lam)));
  }
}

int orc_bend_probe(const double* x12, double* d_raw, double* q12) {
  d3 p1 = ld(x12, 0);
  d3 p2 = sub(ld(x12, 1), p1), p3 = sub(ld(x12, 2), p1), p4 = sub(ld(x12, 3), p1);
  d3 c23 = crs(p2, p3), c24 = crs(p2, p4);
  d3 n1 = nrm(c23), n2 = nrm(c24);
  double d = dot(n1, n2);
  *d_raw = d;
  if (isnan(d)) return 0;
  d = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
  if (acos(d) < ORC_EPS) return 0;
  double l23 = len(c23) + ORC_EPS, l24 = len(c24) + ORC_EPS;
  d3 q3 = scl(add(crs(p2, n2), scl(crs(n1, p2), d)), 1.0 / l23);
  d3 q4 = scl(add(crs(p2, n1), scl(crs(n2, p2), d)), 1.0 / l24);
  d3 q2 = sub(scl(scl(add(crs(p3, n2), scl(crs(n1, p3), d)), 1.0 / l23), -1.0),
              scl(add(crs(p4, n1), scl(crs(n2, p4), d)), 1.0 / l24));
  d3 q1 = scl(add(add(q2, q3), q4), -1.0);
  st(q12, 0, q1); st(q12, 1, q2); st(q12, 2, q3); st(q12, 3, q4);
  return 1;
}

void orc_lra(orc_cloth* c, const orc_params* p) {
  if (c->nslots <= 0) return;
  double* sp = (double*)malloc(sizeof(double) * 3 * (size_t)c->nslots);
  for (int s = 0; s < c->nslots; s++) st(sp, s, ld(c->x, c->lra_slot_vert[s]));
  for (int t = 0; t < c->nv; t++) {
    for (int k = c->lra_ptr[t]; k < c->lra_ptr[t + 1]; k++) {
      double target = c->lra_rest[k] * p->long_range_stretchiness;
      if (c->inv_mass[t] < ORC_EPS && target > 0) continue;
      d3 xt = ld(c->x, t);
      d3 diff = sub(xt, ld(sp, c->lra_slot[k]));
      double dist = len(diff);
      if (dist > target) st(c->x, t, add(xt, add(scl(diff, -1.0), scl(diff, target / dist))));
    }
  }
  free(sp);
}

void orc_ground(orc_cloth* c, const orc_params* p) {
  if (!p->enable_ground) return;
  for (int i = 0; i < c->nv; i++) {
    if (c->inv_mass[i] <= ORC_EPS) continue;
    d3 x = ld(c->x, i);
    if (x.y >= p->ground_height) continue;
    d3 corr = mk(0.0, p->ground_height - x.y, 0.0);
    d3 fr = mk(0, 0, 0);
    double cl = corr.y;
    if (p->ground_friction > 0.0 && cl > 0.0) {
      d3 rel = sub(x, ld(c->x0, i));
      d3 nn = mk(0, 1, 0);
      d3 tan = sub(rel, scl(nn, dot(rel, nn)));
      double tl = len(tan);
      double mx = cl * p->ground_friction;
      double s = (tl > 0.0) ? (mx / tl < 1.0 ? mx / tl : 1.0) : 1.0;
      fr = scl(tan, -s);
    }
    st(c->x, i, add(add(x, corr), fr));
  }
}

void orc_finalize(orc_cloth* c, const orc_params* p, double dt) {
  for (int i = 0; i < c->nv; i++) {
    d3 x = ld(c->x, i);
    d3 raw = scl(sub(x, ld(c->x0, i)), 1.0 / dt);
    st(c->v, i, scl(raw, 1.0 - p->damping * dt));
    st(c->x0, i, x);
  }
}

void orc_substep(orc_cloth* c, const orc_params* p, double dt) {
  orc_predict(c, p, dt);
  for (int it = 0; it < p->num_iterations; it++) {
    for (int k = 0; k < c->s_ncolors; k++) orc_stretch_color(c, k, dt, it);
    if (p->solve_bending)
      for (int k = 0; k < c->b_ncolors; k++) orc_bend_color(c, p, k, dt);
    orc_lra(c, p);
    orc_ground(c, p);
  }
  orc_finalize(c, p, dt);
}

/* ------------------------------------------------------------- geodesic */
typedef struct { double d; int v; } hnode;
static void hpush(hnode* h, int* n, hnode x) {
  int i = (*n)++;
  h[i] = x;
  while (i > 0) {
    int p = (i - 1) / 2;
    if (h[p].d <= h[i].d) break;
    hnode t = h[p]; h[p] = h[i]; h[i] = t; i = p;
  }
}
static hnode hpop(hnode* h, int* n) {
  hnode top = h[0];
  h[0] = h[--(*n)];
  int i = 0;
  for (;;) {
    int l = 2 * i + 1, r = l + 1, m = i;
    if (l < *n && h[l].d < h[m].d) m = l;
    if (r < *n && h[r].d < h[m].d) m = r;
    if (m == i) break;
    hnode t = h[m]; h[m] = h[i]; h[i] = t; i = m;
  }
  return top;
}

void orc_geodesic(const double* x, int nv, const int32_t* edges, int ne, int src, double* dist) {
  int32_t* ptr = (int32_t*)calloc((size_t)nv + 1, sizeof(int32_t));
  for (int e = 0; e < ne; e++) { ptr[edges[2 * e] + 1]++; ptr[edges[2 * e + 1] + 1]++; }
  for (int v = 0; v < nv; v++) ptr[v + 1] += ptr[v];
  int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)nv);
  memcpy(fill, ptr, sizeof(int32_t) * (size_t)nv);
  int32_t* nb = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * ne + 1));
  for (int e = 0; e < ne; e++) {
    nb[fill[edges[2 * e]]++] = edges[2 * e + 1];
    nb[fill[edges[2 * e + 1]]++] = edges[2 * e];
  }
  for (int v = 0; v < nv; v++) dist[v] = INFINITY;
  hnode* h = (hnode*)malloc(sizeof(hnode) * (size_t)(2 * ne + nv + 1));
  int hn = 0;
  dist[src] = 0.0;
  hnode s = {0.0, src};
  hpush(h, &hn, s);
  while (hn > 0) {
    hnode cur = hpop(h, &hn);
    if (cur.d > dist[cur.v]) continue;
    for (int k = ptr[cur.v]; k < ptr[cur.v + 1]; k++) {
      int u = nb[k];
      double w = len(sub(ld(x, cur.v), ld(x, u)));
      if (dist[u] > dist[cur.v] + w) {
        dist[u] = dist[cur.v] + w;
        hnode q = {dist[u], u};
        hpush(h, &hn, q);
      }
    }
  }
  free(ptr); free(fill); free(nb); free(h);
}
