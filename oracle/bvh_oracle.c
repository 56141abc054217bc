/*
 * bvh_oracle.c -- serial FP32-strict restatement of XRTailor's LBVH (build, refit,
 * stackless queries, pair traversal) and of the mesh normals feeding it.
 *
 * TEST INFRASTRUCTURE ONLY (see xrt_oracle.h).  Built with -ffp-contract=off so every
 * float operation rounds exactly as written; the GPU kernels that produce integer results
 * (Morton codes, order, topology, escape tables, hit and pair sets) are built the same way,
 * and the two must agree bit for bit.
 *
 * Pinned by (tests/test_oracle_bvh.py): Morton known answers, std-sort order, a brute-force
 * radix tree from the sorted codes, the brute-force pre-order skip successor for the
 * escape tables, brute-force point / segment / pair overlap sets, unions of leaf boxes.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "xrt_oracle.h"

typedef struct { float x, y, z; } f3;
static inline f3 f3ld(const float* p, int i) { f3 r = {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; return r; }
static inline f3 f3add(f3 a, f3 b) { f3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static inline f3 f3sub(f3 a, f3 b) { f3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline f3 f3scl(f3 a, float s) { f3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static inline float f3dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 f3crs(f3 a, f3 b) {
  f3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  return r;
}
/* MathFunctions::min / max (xrtailor/include/xrtailor/math/MathFunctions.cuh:24-51) */
static inline float mn(float a, float b) { return a < b ? a : b; }
static inline float mx(float a, float b) { return a > b ? a : b; }
static inline float mn3(float a, float b, float c) { return mn(a, mn(b, c)); }
static inline float mx3(float a, float b, float c) { return mx(a, mx(b, c)); }

/* ------------------------------------------------------------- normals */
void orc_vertex_normals(const float* x, int nv, const int32_t* tris, int nt, float* n_out) {
  /* UpdateNodeNormals (xrtailor/src/physics/PhysicsMeshHelper.cu:196-210,282-295):
   * n_v = normalize(sum_f cross(e0,e1)/(|e0|^2 |e1|^2)), e0,e1 the face edges leaving v;
   * summed per vertex in ascending face order */
  int32_t* ptr = (int32_t*)calloc((size_t)nv + 1, sizeof(int32_t));
  for (int i = 0; i < 3 * nt; i++) ptr[tris[i] + 1]++;
  for (int v = 0; v < nv; v++) ptr[v + 1] += ptr[v];
  int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)nv);
  memcpy(fill, ptr, sizeof(int32_t) * (size_t)nv);
  int32_t* corner = (int32_t*)malloc(sizeof(int32_t) * (size_t)(3 * nt + 1));
  for (int f = 0; f < nt; f++)
    for (int j = 0; j < 3; j++) corner[fill[tris[3 * f + j]]++] = 3 * f + j;
  for (int v = 0; v < nv; v++) {
    f3 s = {0.0f, 0.0f, 0.0f};
    for (int k = ptr[v]; k < ptr[v + 1]; k++) {
      int f = corner[k] / 3, j = corner[k] % 3;
      f3 p = f3ld(x, tris[3 * f + j]);
      f3 e0 = f3sub(f3ld(x, tris[3 * f + (j + 1) % 3]), p);
      f3 e1 = f3sub(f3ld(x, tris[3 * f + (j + 2) % 3]), p);
      float den = f3dot(e0, e0) * f3dot(e1, e1);
      f3 c = f3crs(e0, e1);
      f3 t = {c.x / den, c.y / den, c.z / den};
      s = f3add(s, t);
    }
    float l = sqrtf(f3dot(s, s));
    n_out[3 * v] = s.x / l; n_out[3 * v + 1] = s.y / l; n_out[3 * v + 2] = s.z / l;
  }
  free(ptr); free(fill); free(corner);
}

/* ------------------------------------------------------------------ LBVH */
struct orc_bvh {
  int nf, nv;
  int32_t* tris;   /* 3*nf */
  float* v;        /* 3*nv corner points x0 + tol*n */
  float* pred;     /* 3*nv predicted corner points */
  uint32_t* morton; /* nf, sorted order */
  int32_t* order;  /* nf, object id of sorted leaf i */
  int32_t* nodes;  /* 4*(2nf-1): parent, left, right, object (-1 = none) */
  float* box;      /* 6*(2nf-1): lower xyz, upper xyz */
  int32_t* escape; /* 2nf-1 */
  float root[6];
};

static inline uint32_t expand_bits(uint32_t v) { /* BVH.cu:22-28 */
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

uint32_t orc_morton_code(float x, float y, float z) { /* BVH.cu:32-46 */
  x = mn(mx(x * 1024.0f, 0.0f), 1023.0f);
  y = mn(mx(y * 1024.0f, 0.0f), 1023.0f);
  z = mn(mx(z * 1024.0f, 0.0f), 1023.0f);
  return expand_bits((uint32_t)x) * 4u + expand_bits((uint32_t)y) * 2u + expand_bits((uint32_t)z);
}

uint32_t orc_expand_bits(uint32_t v) { return expand_bits(v); }

static inline int clz64(uint64_t v) { return v == 0 ? 64 : __builtin_clzll(v); }

/* determine_range / find_split over 64-bit keys (m << 32 | object) (BVH.cu:57-210) */
static void determine_range(const uint64_t* code, int n, int idx, int* first, int* last) {
  if (idx == 0) { *first = 0; *last = n - 1; return; }
  uint64_t self = code[idx];
  int L = clz64(self ^ code[idx - 1]);
  int R = clz64(self ^ code[idx + 1]);
  int d = (R > L) ? 1 : -1;
  int dmin = L < R ? L : R;
  int lmax = 2, delta = -1;
  int t = idx + d * lmax;
  if (0 <= t && t < n) delta = clz64(self ^ code[t]);
  while (delta > dmin) {
    lmax <<= 1;
    t = idx + d * lmax;
    delta = -1;
    if (0 <= t && t < n) delta = clz64(self ^ code[t]);
  }
  int l = 0;
  for (int s = lmax >> 1; s > 0; s >>= 1) {
    t = idx + (l + s) * d;
    delta = -1;
    if (0 <= t && t < n) delta = clz64(self ^ code[t]);
    if (delta > dmin) l += s;
  }
  int j = idx + l * d;
  if (d < 0) { *first = j; *last = idx; } else { *first = idx; *last = j; }
}

static int find_split(const uint64_t* code, int first, int last) {
  uint64_t fc = code[first], lc = code[last];
  if (fc == lc) return (first + last) >> 1;
  int dn = clz64(fc ^ lc);
  int split = first, stride = last - first;
  do {
    stride = (stride + 1) >> 1;
    int mid = split + stride;
    if (mid < last && clz64(fc ^ code[mid]) > dn) split = mid;
  } while (stride > 1);
  return split;
}

static void box_of_face(const orc_bvh* b, int f, float* out) { /* InitAABBs_Kernel, BVH.cu:423-462 */
  const int32_t* t = b->tris + 3 * f;
  f3 v1 = f3ld(b->v, t[0]), v2 = f3ld(b->v, t[1]), v3 = f3ld(b->v, t[2]);
  f3 p1 = f3ld(b->pred, t[0]), p2 = f3ld(b->pred, t[1]), p3 = f3ld(b->pred, t[2]);
  out[0] = mn(mn3(v1.x, v2.x, v3.x), mn3(p1.x, p2.x, p3.x));
  out[1] = mn(mn3(v1.y, v2.y, v3.y), mn3(p1.y, p2.y, p3.y));
  out[2] = mn(mn3(v1.z, v2.z, v3.z), mn3(p1.z, p2.z, p3.z));
  out[3] = mx(mx3(v1.x, v2.x, v3.x), mx3(p1.x, p2.x, p3.x));
  out[4] = mx(mx3(v1.y, v2.y, v3.y), mx3(p1.y, p2.y, p3.y));
  out[5] = mx(mx3(v1.z, v2.z, v3.z), mx3(p1.z, p2.z, p3.z));
}

static void merge_into(float* dst, const float* a, const float* b) {
  for (int k = 0; k < 3; k++) { dst[k] = mn(a[k], b[k]); dst[3 + k] = mx(a[3 + k], b[3 + k]); }
}

static void fill_boxes(orc_bvh* b, int node) {
  int32_t* nd = b->nodes + 4 * node;
  if (nd[3] >= 0) return;
  fill_boxes(b, nd[1]);
  fill_boxes(b, nd[2]);
  merge_into(b->box + 6 * node, b->box + 6 * nd[1], b->box + 6 * nd[2]);
}

typedef struct { uint32_t m; int32_t i; } mkey;
static int cmp_mkey(const void* a, const void* b) {
  const mkey *x = (const mkey*)a, *y = (const mkey*)b;
  if (x->m != y->m) return x->m < y->m ? -1 : 1;
  return (x->i > y->i) - (x->i < y->i); /* stable: original order on ties */
}

orc_bvh* orc_bvh_create(void) { return (orc_bvh*)calloc(1, sizeof(orc_bvh)); }

void orc_bvh_destroy(orc_bvh* b) {
  if (!b) return;
  free(b->tris); free(b->v); free(b->pred); free(b->morton); free(b->order);
  free(b->nodes); free(b->box); free(b->escape); free(b);
}

/* BVH::UpdateData + BVH::Construct (BVH.cu:382-421,565-573,618-722) */
int orc_bvh_build(orc_bvh* b, const float* x0, const float* x, const float* n, int nv,
                  const int32_t* tris, int nt, float tol, int ccd, float scr) {
  free(b->tris); free(b->v); free(b->pred); free(b->morton); free(b->order);
  free(b->nodes); free(b->box); free(b->escape);
  b->nf = nt; b->nv = nv;
  b->tris = (int32_t*)malloc(sizeof(int32_t) * (size_t)(3 * nt + 1));
  memcpy(b->tris, tris, sizeof(int32_t) * (size_t)(3 * nt));
  b->v = (float*)malloc(sizeof(float) * (size_t)(3 * nv + 1));
  b->pred = (float*)malloc(sizeof(float) * (size_t)(3 * nv + 1));
  for (int i = 0; i < nv; i++) {
    f3 p = f3ld(x0, i);
    f3 q = ccd ? f3ld(x, i) : p;
    if (n) { f3 nn = f3ld(n, i); p = f3add(p, f3scl(nn, tol)); q = f3add(q, f3scl(nn, tol)); }
    b->v[3 * i] = p.x; b->v[3 * i + 1] = p.y; b->v[3 * i + 2] = p.z;
    b->pred[3 * i] = q.x; b->pred[3 * i + 1] = q.y; b->pred[3 * i + 2] = q.z;
  }
  int nn = nt > 0 ? 2 * nt - 1 : 0, ni = nt > 0 ? nt - 1 : 0;
  b->morton = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nt + 1));
  b->order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nt + 1));
  b->nodes = (int32_t*)malloc(sizeof(int32_t) * (size_t)(4 * nn + 4));
  b->box = (float*)malloc(sizeof(float) * (size_t)(6 * nn + 6));
  b->escape = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nn + 1));
  if (nt == 0) return 0;
  float* leaf = (float*)malloc(sizeof(float) * (size_t)(6 * nt));
  float root[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int f = 0; f < nt; f++) {
    float* bx = leaf + 6 * f;
    box_of_face(b, f, bx);
    if (scr != 0.0f) { /* dilation along normalize(corner - center) */
      f3 lo = {bx[0], bx[1], bx[2]}, hi = {bx[3], bx[4], bx[5]};
      f3 c = f3scl(f3add(lo, hi), 0.5f);
      f3 d1 = f3sub(lo, c), d2 = f3sub(hi, c);
      float l1 = sqrtf(f3dot(d1, d1)), l2 = sqrtf(f3dot(d2, d2));
      f3 n1 = {d1.x / l1, d1.y / l1, d1.z / l1}, n2 = {d2.x / l2, d2.y / l2, d2.z / l2};
      lo = f3add(lo, f3scl(n1, scr)); hi = f3add(hi, f3scl(n2, scr));
      bx[0] = lo.x; bx[1] = lo.y; bx[2] = lo.z; bx[3] = hi.x; bx[4] = hi.y; bx[5] = hi.z;
    }
    for (int k = 0; k < 3; k++) { root[k] = mn(root[k], bx[k]); root[3 + k] = mx(root[3 + k], bx[3 + k]); }
  }
  memcpy(b->root, root, sizeof(root));
  mkey* keys = (mkey*)malloc(sizeof(mkey) * (size_t)nt);
  for (int f = 0; f < nt; f++) { /* ComputeMortons_Kernel (BVH.cu:464-481) */
    const float* bx = leaf + 6 * f;
    float cx = 0.5f * (bx[0] + bx[3]), cy = 0.5f * (bx[1] + bx[4]), cz = 0.5f * (bx[2] + bx[5]);
    cx = cx - root[0]; cy = cy - root[1]; cz = cz - root[2];
    cx = cx / (root[3] - root[0]); cy = cy / (root[4] - root[1]); cz = cz / (root[5] - root[2]);
    keys[f].m = orc_morton_code(cx, cy, cz);
    keys[f].i = f;
  }
  qsort(keys, (size_t)nt, sizeof(mkey), cmp_mkey);
  uint64_t* code = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nt);
  for (int i = 0; i < nt; i++) {
    b->morton[i] = keys[i].m; b->order[i] = keys[i].i;
    code[i] = ((uint64_t)keys[i].m << 32) | (uint32_t)keys[i].i;
  }
  for (int k = 0; k < nn; k++) {
    b->nodes[4 * k] = -1; b->nodes[4 * k + 1] = -1; b->nodes[4 * k + 2] = -1; b->nodes[4 * k + 3] = -1;
  }
  for (int i = 0; i < nt; i++) { /* ConstructLeafNodes_Kernel (BVH.cu:489-502) */
    b->nodes[4 * (ni + i) + 3] = keys[i].i;
    memcpy(b->box + 6 * (ni + i), leaf + 6 * keys[i].i, sizeof(float) * 6);
  }
  for (int idx = 0; idx < ni; idx++) { /* ConstructInternalNodes*Bits_Kernel (BVH.cu:275-317) */
    int first, last;
    determine_range(code, nt, idx, &first, &last);
    int g = find_split(code, first, last);
    int l = g, r = g + 1;
    if (first == g) l += ni;
    if (last == g + 1) r += ni;
    b->nodes[4 * idx + 1] = l; b->nodes[4 * idx + 2] = r;
    b->nodes[4 * l] = idx; b->nodes[4 * r] = idx;
  }
  fill_boxes(b, 0);
  for (int k = 0; k < nn; k++) { /* escape tables (BVH.cu:212-273), one node-indexed array */
    int cur = k, e = -1;
    for (;;) {
      int p = b->nodes[4 * cur];
      if (p < 0) { e = -1; break; }
      if (b->nodes[4 * p + 1] == cur) { e = b->nodes[4 * p + 2]; break; }
      cur = p;
    }
    b->escape[k] = e;
  }
  free(leaf); free(keys); free(code);
  return 0;
}

int orc_bvh_num_objects(const orc_bvh* b) { return b->nf; }

void orc_bvh_get(const orc_bvh* b, uint32_t* morton, int32_t* order, int32_t* nodes, float* box,
                 int32_t* escape, float* root) {
  int nn = b->nf > 0 ? 2 * b->nf - 1 : 0;
  if (morton) memcpy(morton, b->morton, sizeof(uint32_t) * (size_t)b->nf);
  if (order) memcpy(order, b->order, sizeof(int32_t) * (size_t)b->nf);
  if (nodes) memcpy(nodes, b->nodes, sizeof(int32_t) * (size_t)(4 * nn));
  if (box) memcpy(box, b->box, sizeof(float) * (size_t)(6 * nn));
  if (escape) memcpy(escape, b->escape, sizeof(int32_t) * (size_t)nn);
  if (root) memcpy(root, b->root, sizeof(float) * 6);
}

const float* orc_bvh_pred(const orc_bvh* b) { return b->pred; }

/* RefitExternalNodeMidstep_Kernel + AssignAABBCAS_Kernel (BVH.cu:515-538,770-818):
 * pred = x1 + tol*n1; leaf box grows by the 3 pred corners; parents = merge(children) */
void orc_bvh_refit_midstep(orc_bvh* b, const float* x1, const float* n1, float tol) {
  if (b->nf == 0) return;
  for (int i = 0; i < b->nv; i++) {
    f3 q = f3add(f3ld(x1, i), f3scl(f3ld(n1, i), tol));
    b->pred[3 * i] = q.x; b->pred[3 * i + 1] = q.y; b->pred[3 * i + 2] = q.z;
  }
  int ni = b->nf - 1;
  for (int i = 0; i < b->nf; i++) {
    float* bx = b->box + 6 * (ni + i);
    const int32_t* t = b->tris + 3 * b->nodes[4 * (ni + i) + 3];
    for (int c = 0; c < 3; c++) {
      f3 p = f3ld(b->pred, t[c]);
      bx[0] = mn(bx[0], p.x); bx[1] = mn(bx[1], p.y); bx[2] = mn(bx[2], p.z);
      bx[3] = mx(bx[3], p.x); bx[4] = mx(bx[4], p.y); bx[5] = mx(bx[5], p.z);
    }
  }
  fill_boxes(b, 0);
}

/* Update_Kernel + AssignAABBAtomic_Kernel (BVH.cu:540-557,820-852): pred = x; leaf box
 * grows by Face::ComputeBounds(ccd) (x, and x0 with ccd); ancestors grow to cover it */
static void grow_up(orc_bvh* b, int node) {
  int32_t* nd = b->nodes + 4 * node;
  if (nd[3] >= 0) return;
  grow_up(b, nd[1]);
  grow_up(b, nd[2]);
  float m[6];
  merge_into(m, b->box + 6 * nd[1], b->box + 6 * nd[2]);
  merge_into(b->box + 6 * node, b->box + 6 * node, m);
}

void orc_bvh_update(orc_bvh* b, const float* x, const float* x0, int ccd) {
  if (b->nf == 0) return;
  memcpy(b->pred, x, sizeof(float) * (size_t)(3 * b->nv));
  int ni = b->nf - 1;
  for (int i = 0; i < b->nf; i++) {
    float* bx = b->box + 6 * (ni + i);
    const int32_t* t = b->tris + 3 * b->nodes[4 * (ni + i) + 3];
    float fb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < 3; c++) {
      f3 p = f3ld(x, t[c]);
      fb[0] = mn(fb[0], p.x); fb[1] = mn(fb[1], p.y); fb[2] = mn(fb[2], p.z);
      fb[3] = mx(fb[3], p.x); fb[4] = mx(fb[4], p.y); fb[5] = mx(fb[5], p.z);
      if (ccd) {
        f3 q = f3ld(x0, t[c]);
        fb[0] = mn(fb[0], q.x); fb[1] = mn(fb[1], q.y); fb[2] = mn(fb[2], q.z);
        fb[3] = mx(fb[3], q.x); fb[4] = mx(fb[4], q.y); fb[5] = mx(fb[5], q.z);
      }
    }
    merge_into(bx, bx, fb);
  }
  grow_up(b, 0);
}

/* ---------------------------------------------------------------- queries */
static int overlap_point(const float* bx, f3 p) { /* Bounds::Overlap(p), strict (Bounds.cu:130-136) */
  return p.x > bx[0] && p.x < bx[3] && p.y > bx[1] && p.y < bx[4] && p.z > bx[2] && p.z < bx[5];
}

static int get_intersection(float d1, float d2, f3 p1, f3 p2, f3* hit) { /* Bounds.cu:139-147 */
  if ((d1 * d2) >= 0.0f) return 0;
  if (d1 == d2) return 0;
  *hit = f3add(p1, f3scl(f3sub(p2, p1), (-d1 / (d2 - d1))));
  return 1;
}
static int in_box(f3 h, const float* bx, int axis) { /* Bounds.cu:149-158 */
  if (axis == 1 && h.z > bx[2] && h.z < bx[5] && h.y > bx[1] && h.y < bx[4]) return 1;
  if (axis == 2 && h.z > bx[2] && h.z < bx[5] && h.x > bx[0] && h.x < bx[3]) return 1;
  if (axis == 3 && h.x > bx[0] && h.x < bx[3] && h.y > bx[1] && h.y < bx[4]) return 1;
  return 0;
}
int orc_overlap_segment(const float* bx, const float* a, const float* c) { /* Bounds.cu:160-191 */
  f3 p1 = {a[0], a[1], a[2]}, p2 = {c[0], c[1], c[2]}, h;
  if (p2.x < bx[0] && p1.x < bx[0]) return 0;
  if (p2.x > bx[3] && p1.x > bx[3]) return 0;
  if (p2.y < bx[1] && p1.y < bx[1]) return 0;
  if (p2.y > bx[4] && p1.y > bx[4]) return 0;
  if (p2.z < bx[2] && p1.z < bx[2]) return 0;
  if (p2.z > bx[5] && p1.z > bx[5]) return 0;
  if (overlap_point(bx, p1)) return 1;
  if ((get_intersection(p1.x - bx[0], p2.x - bx[0], p1, p2, &h) && in_box(h, bx, 1)) ||
      (get_intersection(p1.y - bx[1], p2.y - bx[1], p1, p2, &h) && in_box(h, bx, 2)) ||
      (get_intersection(p1.z - bx[2], p2.z - bx[2], p1, p2, &h) && in_box(h, bx, 3)) ||
      (get_intersection(p1.x - bx[3], p2.x - bx[3], p1, p2, &h) && in_box(h, bx, 1)) ||
      (get_intersection(p1.y - bx[4], p2.y - bx[4], p1, p2, &h) && in_box(h, bx, 2)) ||
      (get_intersection(p1.z - bx[5], p2.z - bx[5], p1, p2, &h) && in_box(h, bx, 3)))
    return 1;
  return 0;
}

/* QueryDeviceStackless point / segment (BVH.cu:957-1021); returns the hit count (<= cap) */
int orc_bvh_query(const orc_bvh* b, const float* p0, const float* p1, int32_t* hits, int cap) {
  if (b->nf == 0) return 0;
  int node = 0, found = 0;
  f3 p = {p0[0], p0[1], p0[2]};
  while (node != -1) {
    if (found >= cap) break;
    const float* bx = b->box + 6 * node;
    int hit = p1 ? orc_overlap_segment(bx, p0, p1) : overlap_point(bx, p);
    int obj = b->nodes[4 * node + 3];
    if (hit) {
      if (obj >= 0) { hits[found++] = obj; node = b->escape[node]; }
      else node = b->nodes[4 * node + 1];
    } else {
      node = b->escape[node];
    }
  }
  return found;
}

static int overlap_dilated(const float* rhs, const float* lhs, float t) { /* Bounds::Overlap(b, t) */
  float lo[3] = {lhs[0] - t, lhs[1] - t, lhs[2] - t};
  float hi[3] = {lhs[3] + t, lhs[4] + t, lhs[5] + t};
  for (int k = 0; k < 3; k++) {
    if (rhs[k] > hi[k]) return 0;
    if (rhs[3 + k] < lo[k]) return 0;
  }
  return 1;
}

/* Traverse(lhs, rhs, thickness): CountPairs_Kernel / inclusive scan / FindPairs_Kernel
 * (BVH.cu:1149-1227,1315-1365).  Pairs of (lhs object, rhs object) in the reference's
 * layout: lhs leaves in sorted order, each leaf's pairs in reverse traversal order.
 * Returns the pair count; writes at most cap pairs. */
long orc_bvh_traverse(const orc_bvh* lhs, const orc_bvh* rhs, float thickness, int32_t* pairs,
                      long cap) {
  if (lhs->nf == 0 || rhs->nf == 0) return 0;
  long total = 0;
  int nil = lhs->nf - 1;
  int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)rhs->nf);
  for (int id = 0; id < lhs->nf; id++) {
    const float* lb = lhs->box + 6 * (nil + id);
    int lobj = lhs->nodes[4 * (nil + id) + 3];
    int node = 0, cnt = 0;
    while (node != -1) {
      int obj = rhs->nodes[4 * node + 3];
      if (overlap_dilated(rhs->box + 6 * node, lb, thickness)) {
        if (obj >= 0) { tmp[cnt++] = obj; node = rhs->escape[node]; }
        else node = rhs->nodes[4 * node + 1];
      } else {
        node = rhs->escape[node];
      }
    }
    for (int k = 0; k < cnt; k++) {
      long slot = total + (cnt - 1 - k);
      if (slot < cap) { pairs[2 * slot] = lobj; pairs[2 * slot + 1] = tmp[k]; }
    }
    total += cnt;
  }
  free(tmp);
  return total;
}

/* ------------------------------------------------- nearest face (setup query) */
/* MinDist / MinMaxDist (xrtailor/src/physics/broad_phase/lbvh/BVHHelper.cu:114-151):
 * squared distances from p to a box and the min-max metric of Roussopoulos et al. */
static float min_dist(const float* bx, f3 p) {
  const float dx = fminf(bx[3], fmaxf(bx[0], p.x)) - p.x;
  const float dy = fminf(bx[4], fmaxf(bx[1], p.y)) - p.y;
  const float dz = fminf(bx[5], fmaxf(bx[2], p.z)) - p.z;
  return dx * dx + dy * dy + dz * dz;
}

static float min_max_dist(const float* bx, f3 p) {
  float rm[3], rM[3];
  const float pp[3] = {p.x, p.y, p.z};
  for (int k = 0; k < 3; k++) {
    rm[k] = (bx[k] - pp[k]) * (bx[k] - pp[k]);
    rM[k] = (bx[3 + k] - pp[k]) * (bx[3 + k] - pp[k]);
    if ((bx[3 + k] + bx[k]) * 0.5f < pp[k]) { float t = rm[k]; rm[k] = rM[k]; rM[k] = t; }
  }
  const float dx = rm[0] + rM[1] + rM[2];
  const float dy = rM[0] + rm[1] + rM[2];
  const float dz = rM[0] + rM[1] + rm[2];
  return mn(dx, mn(dy, dz));
}

/* distance_calculator (BVHHelper.cu:9-103): Ericson's closest point on triangle abc with
 * the barycentric weights of the region it falls in; returns |p - q| */
float orc_tri_nearest(const float* p3, const float* a3, const float* b3, const float* c3,
                      float* uvw) {
  const f3 p = {p3[0], p3[1], p3[2]}, a = {a3[0], a3[1], a3[2]};
  const f3 b = {b3[0], b3[1], b3[2]}, c = {c3[0], c3[1], c3[2]};
  const f3 e1 = f3sub(b, a), e2 = f3sub(c, a), e3 = f3sub(c, b);
  f3 q;
  const f3 v1 = f3sub(p, a);
  const float d1 = f3dot(e1, v1), d2 = f3dot(e2, v1);
  if (d1 <= 0 && d2 <= 0) {
    uvw[0] = 1; uvw[1] = 0; uvw[2] = 0;
    q = a;
    return sqrtf(f3dot(f3sub(p, q), f3sub(p, q)));
  }
  const f3 v2 = f3sub(p, b);
  const float d3 = f3dot(e1, v2), d4 = f3dot(e2, v2);
  if (d3 >= 0 && d4 <= d3) {
    uvw[0] = 0; uvw[1] = 1; uvw[2] = 0;
    q = b;
    return sqrtf(f3dot(f3sub(p, q), f3sub(p, q)));
  }
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    const float v = d1 / (d1 - d3);
    q = f3add(a, f3scl(e1, v));
    uvw[0] = 1 - v; uvw[1] = v; uvw[2] = 0;
    return sqrtf(f3dot(f3sub(p, q), f3sub(p, q)));
  }
  const f3 v3 = f3sub(p, c);
  const float d5 = f3dot(e1, v3), d6 = f3dot(e2, v3);
  if (d6 >= 0 && d5 <= d6) {
    uvw[0] = 0; uvw[1] = 0; uvw[2] = 1;
    q = c;
    return sqrtf(f3dot(f3sub(p, q), f3sub(p, q)));
  }
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) {
    const float w = d2 / (d2 - d6);
    q = f3add(a, f3scl(e2, w));
    uvw[0] = 1 - w; uvw[1] = 0; uvw[2] = w;
    return sqrtf(f3dot(f3sub(p, q), f3sub(p, q)));
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    q = f3add(b, f3scl(e3, w));
    uvw[0] = 0; uvw[1] = 1 - w; uvw[2] = w;
    return sqrtf(f3dot(f3sub(p, q), f3sub(p, q)));
  }
  const float d = 1 / (va + vb + vc);
  const float v = vb * d, w = vc * d;
  q = f3add(f3add(a, f3scl(e1, v)), f3scl(e2, w));
  uvw[0] = 1 - v - w; uvw[1] = v; uvw[2] = w;
  return sqrtf(f3dot(f3sub(p, q), f3sub(p, q)));
}

/* QueryDevice (BVH.cu:881-948): depth-first search with a 64-entry stack of (node, squared
 * min distance), children tested left then right, pushed left then right; a node popped
 * with a squared min distance above the best (unsquared) distance is skipped, a child is
 * kept when its min distance <= the sibling's min-max distance, and a leaf replaces the best
 * when its distance is <= it.  Triangles are the tree's corner points (x0 + tol n).
 * Returns the face, -1 for an empty tree, -2 when the stack would overflow. */
int orc_bvh_nearest(const orc_bvh* b, const float* p3, float* uvw, float* dist) {
  if (b->nf < 2) return -1;
  const f3 p = {p3[0], p3[1], p3[2]};
  int st_node[64];
  float st_dist[64];
  int sp = 0;
  st_node[sp] = 0;
  st_dist[sp++] = min_dist(b->box, p);
  int nearest = -1;
  float best = INFINITY;
  do {
    const int node = st_node[--sp];
    const float md = st_dist[sp];
    if (md > best) continue;
    const int kid[2] = {b->nodes[4 * node + 1], b->nodes[4 * node + 2]};
    const float kmd[2] = {min_dist(b->box + 6 * kid[0], p), min_dist(b->box + 6 * kid[1], p)};
    const float kmm[2] = {min_max_dist(b->box + 6 * kid[0], p), min_max_dist(b->box + 6 * kid[1], p)};
    for (int s = 0; s < 2; s++) {
      if (!(kmd[s] <= kmm[1 - s])) continue;
      const int obj = b->nodes[4 * kid[s] + 3];
      if (obj >= 0) {
        float w[3];
        const int32_t* t = b->tris + 3 * obj;
        const float d = orc_tri_nearest(p3, b->v + 3 * t[0], b->v + 3 * t[1], b->v + 3 * t[2], w);
        if (d <= best) {
          best = d;
          nearest = obj;
          uvw[0] = w[0]; uvw[1] = w[1]; uvw[2] = w[2];
        }
      } else {
        if (sp >= 64) return -2;
        st_node[sp] = kid[s];
        st_dist[sp++] = kmd[s];
      }
    }
  } while (sp > 0);
  *dist = best;
  return nearest;
}
