/*
 * contact_oracle.c -- serial restatement of XRTailor's body-cloth PBD collision and the
 * PBD repulsion self-contact path: FP64 narrow phase and responses, broad phase through
 * the FP32-strict oracle LBVH (bvh_oracle.c).
 *
 * TEST INFRASTRUCTURE ONLY (see xrt_oracle.h).
 *
 * Pinned by (tests/test_oracle_contact.py): Ericson point-triangle / segment-segment
 * routines against dense sampling, Moller-Trumbore against a plane-intersection closed
 * form, R-triangle ownership counts (every vertex and edge owned exactly once), and
 * single-contact closed forms (a vertex under a plane is projected to 1e-3 above it).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "xrt_oracle.h"

#define EPS 1e-6

typedef struct { double x, y, z; } v3;
static inline v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 ldd(const double* p, int i) { return V(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline v3 ldf(const float* p, int i) { return V(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline void std3(double* p, int i, v3 a) { p[3 * i] = a.x; p[3 * i + 1] = a.y; p[3 * i + 2] = a.z; }
static inline v3 ad(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sb(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 sc(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
static inline double dt3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cr(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline double ln(v3 a) { return sqrt(dt3(a, a)); }
static inline v3 nz(v3 a) { return sc(a, 1.0 / ln(a)); }
static inline double mixed(v3 a, v3 b, v3 c) { return dt3(a, cr(b, c)); } /* MathFunctions::Mixed */
static inline double cl(double x, double a, double b) { return x < a ? a : (x > b ? b : x); }

/* ------------------------------------------------------------ topology */
void orc_face_edges(const int32_t* tris, int nt, const int32_t* edges, int ne, int32_t* fe) {
  /* Face::edges[i] = edge (nodes[i], nodes[i+1]) (Face.cu:13-25) */
  for (int f = 0; f < nt; f++)
    for (int l = 0; l < 3; l++) {
      int a = tris[3 * f + l], b = tris[3 * f + (l + 1) % 3];
      fe[3 * f + l] = -1;
      for (int e = 0; e < ne; e++)
        if ((edges[2 * e] == a && edges[2 * e + 1] == b) || (edges[2 * e] == b && edges[2 * e + 1] == a)) {
          fe[3 * f + l] = e;
          break;
        }
    }
}

int orc_edge_faces(const int32_t* fe, int nt, int ne, int32_t* ptr, int32_t* ef) {
  /* BuildEFAdjacency (ClothSolverHelper.cpp:111-125): faces per edge, ascending (std::set) */
  memset(ptr, 0, sizeof(int32_t) * (size_t)(ne + 1));
  for (int i = 0; i < 3 * nt; i++) ptr[fe[i] + 1]++;
  for (int e = 0; e < ne; e++) ptr[e + 1] += ptr[e];
  int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ne + 1));
  memcpy(fill, ptr, sizeof(int32_t) * (size_t)ne);
  for (int f = 0; f < nt; f++)
    for (int l = 0; l < 3; l++) ef[fill[fe[3 * f + l]]++] = f;
  free(fill);
  return ptr[ne];
}

void orc_rtri(const int32_t* tris, int nt, int nv, uint32_t* out) {
  /* RTriBuild_Kernel (xrtailor/src/physics/representative_triangle/RTriangle.cu:19-85) over
   * ascending per-vertex face lists (BuildVFAdjacency, ClothSolverHelper.cu:395-421) */
  int32_t* ptr = (int32_t*)calloc((size_t)nv + 1, sizeof(int32_t));
  for (int i = 0; i < 3 * nt; i++) ptr[tris[i] + 1]++;
  for (int v = 0; v < nv; v++) ptr[v + 1] += ptr[v];
  int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nv + 1));
  memcpy(fill, ptr, sizeof(int32_t) * (size_t)nv);
  int32_t* vf = (int32_t*)malloc(sizeof(int32_t) * (size_t)(3 * nt + 1));
  for (int f = 0; f < nt; f++) /* faces visited ascending -> lists ascending */
    for (int l = 0; l < 3; l++) vf[fill[tris[3 * f + l]]++] = f;
  for (int fid = 0; fid < nt; fid++) {
    uint32_t info = 0;
    for (int i = 0; i < 3; i++) {
      int vi = tris[3 * fid + i], vj = tris[3 * fid + (i + 1) % 3];
      if (vf[ptr[vi]] == fid) info |= 1u << i;
      long of = 0xffffffffL;
      for (int a = ptr[vi]; a < ptr[vi + 1] && of == 0xffffffffL; a++) {
        int fa = vf[a];
        if (fa == fid) continue;
        for (int b = ptr[vj]; b < ptr[vj + 1]; b++)
          if (vf[b] == fa) { of = fa; break; }
      }
      if ((long)fid < of) info |= 1u << (3 + i);
    }
    out[fid] = info;
  }
  free(ptr); free(fill); free(vf);
}

/* ---------------------------------------------------------- narrow phase */
/* NearestPoint: Ericson 5.1.5 ClosestPtPointTriangle (BasicPrimitiveTests.cu:6-100) */
double orc_nearest_point(const double* p3, const double* tri9, double* q3) {
  v3 p = ldd(p3, 0), a = ldd(tri9, 0), b = ldd(tri9, 1), c = ldd(tri9, 2), q;
  v3 e1 = sb(b, a), e2 = sb(c, a), e3 = sb(c, b);
  v3 v1 = sb(p, a);
  double d1 = dt3(e1, v1), d2 = dt3(e2, v1);
  if (d1 <= 0 && d2 <= 0) { q = a; goto done; }
  {
    v3 v2 = sb(p, b);
    double d3 = dt3(e1, v2), d4 = dt3(e2, v2);
    if (d3 >= 0 && d4 <= d3) { q = b; goto done; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) { q = ad(a, sc(e1, d1 / (d1 - d3))); goto done; }
    v3 vv3 = sb(p, c);
    double d5 = dt3(e1, vv3), d6 = dt3(e2, vv3);
    if (d6 >= 0 && d5 <= d6) { q = c; goto done; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) { q = ad(a, sc(e2, d2 / (d2 - d6))); goto done; }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
      q = ad(b, sc(e3, (d4 - d3) / ((d4 - d3) + (d5 - d6))));
      goto done;
    }
    double d = 1 / (va + vb + vc);
    q = ad(a, ad(sc(e1, vb * d), sc(e2, vc * d)));
  }
done:
  std3(q3, 0, q);
  return ln(sb(p, q));
}

/* RayIntersect: Moller-Trumbore, hit if 0 < s < length (BasicPrimitiveTests.cu:242-284) */
double orc_ray_intersect(const double* o3, const double* d3_, double length, const double* tri9) {
  v3 o = ldd(o3, 0), d = ldd(d3_, 0), a = ldd(tri9, 0), b = ldd(tri9, 1), c = ldd(tri9, 2);
  v3 e1 = sb(b, a), e2 = sb(c, a);
  v3 n = cr(d, e2);
  double det = dt3(e1, n);
  if (fabs(det) < EPS) return INFINITY;
  double inv = 1 / det;
  v3 t = sb(o, a);
  double u = dt3(t, n) * inv;
  if (u < 0 || u > 1) return INFINITY;
  v3 q = cr(t, e1);
  double v = dt3(d, q) * inv;
  if (v < 0 || v + u > 1) return INFINITY;
  double s = dt3(e2, q) * inv;
  if (s > 0 && s < length) return s;
  return INFINITY;
}

/* ClosestPtSegmentSegment: Ericson 5.1.9 with EPSILON and t < EPS -> 0
 * (BasicPrimitiveTests.cu:303-361); returns the distance */
double orc_closest_seg_seg(const double* p1_, const double* q1_, const double* p2_,
                           const double* q2_, double* s_out, double* t_out) {
  v3 p1 = ldd(p1_, 0), q1 = ldd(q1_, 0), p2 = ldd(p2_, 0), q2 = ldd(q2_, 0);
  v3 d1 = sb(q1, p1), d2 = sb(q2, p2), r = sb(p1, p2);
  double a = dt3(d1, d1), e = dt3(d2, d2), f = dt3(d2, r), s, t;
  if (a <= EPS && e <= EPS) {
    s = t = 0;
  } else if (a <= EPS) {
    s = 0;
    t = cl(f / e, 0, 1);
  } else {
    double c = dt3(d1, r);
    if (e <= EPS) {
      t = 0;
      s = cl(-c / a, 0, 1);
    } else {
      double b = dt3(d1, d2), den = a * e - b * b;
      s = den > EPS ? cl((b * f - c * e) / den, 0, 1) : 0;
      t = (b * s + f) / e;
      if (t < EPS) { t = 0; s = cl(-c / a, 0, 1); }
      else if (t > 1) { t = 1; s = cl((b - c) / a, 0, 1); }
    }
  }
  *s_out = s; *t_out = t;
  return ln(sb(ad(p1, sc(d1, s)), ad(p2, sc(d2, t))));
}

static v3 face_normal(v3 a, v3 b, v3 c) { return nz(cr(sb(b, a), sb(c, a))); } /* MathFunctions.cu:14-20 */

/* SignedVertexFaceDistance (BasicPrimitiveTests.cu:369-386); w untouched (here: zeroed)
 * when the face is degenerate */
static double signed_vf(v3 x, v3 y0, v3 y1, v3 y2, v3* n, double* w) {
  *n = cr(nz(sb(y1, y0)), nz(sb(y2, y0)));
  if (dt3(*n, *n) < 1e-6) { w[0] = w[1] = w[2] = w[3] = 0; return INFINITY; }
  *n = nz(*n);
  double h = dt3(sb(x, y0), *n);
  double b0 = mixed(sb(y1, x), sb(y2, x), *n);
  double b1 = mixed(sb(y2, x), sb(y0, x), *n);
  double b2 = mixed(sb(y0, x), sb(y1, x), *n);
  w[0] = 1.0;
  w[1] = -b0 / (b0 + b1 + b2);
  w[2] = -b1 / (b0 + b1 + b2);
  w[3] = -b2 / (b0 + b1 + b2);
  return h;
}

/* SignedEdgeEdgeDistance (BasicPrimitiveTests.cu:420-436): normalize(n) is discarded in
 * the reference, so n stays the raw cross product of the unit edge directions */
static double signed_ee(v3 x0, v3 x1, v3 y0, v3 y1, v3* n, double* w) {
  *n = cr(nz(sb(x1, x0)), nz(sb(y1, y0)));
  if (dt3(*n, *n) < 1e-6) { w[0] = w[1] = w[2] = w[3] = 0; return INFINITY; }
  double h = dt3(sb(x0, y0), *n);
  double a0 = mixed(sb(y1, x1), sb(y0, x1), *n), a1 = mixed(sb(y0, x0), sb(y1, x0), *n);
  double b0 = mixed(sb(x0, y1), sb(x1, y1), *n), b1 = mixed(sb(x1, y0), sb(x0, y0), *n);
  w[0] = a0 / (a0 + a1);
  w[1] = a1 / (a0 + a1);
  w[2] = -b0 / (b0 + b1);
  w[3] = -b1 / (b0 + b1);
  return h;
}

double orc_signed_vf(const double* x12, double* n3, double* w4) {
  v3 n;
  double h = signed_vf(ldd(x12, 0), ldd(x12, 1), ldd(x12, 2), ldd(x12, 3), &n, w4);
  std3(n3, 0, n);
  return h;
}
double orc_signed_ee(const double* x12, double* n3, double* w4) {
  v3 n;
  double h = signed_ee(ldd(x12, 0), ldd(x12, 1), ldd(x12, 2), ldd(x12, 3), &n, w4);
  std3(n3, 0, n);
  return h;
}

/* ---------------------------------------------------- body-cloth PBD collision */
/* GenerateVF_DCD_Kernel (PBDCollision.cu:17-43): candidates = leaves strictly containing
 * x; nearest point on the pred triangle; keep the closest */
void orc_vf_dcd_generate(const orc_bvh* obs, const int32_t* obs_tris, const double* x, int nv,
                         int32_t* face_out, double* q_out) {
  int32_t hits[256];
  const float* pred = orc_bvh_pred(obs);
  for (int i = 0; i < nv; i++) {
    float p[3] = {(float)x[3 * i], (float)x[3 * i + 1], (float)x[3 * i + 2]};
    int k = orc_bvh_query(obs, p, NULL, hits, 256);
    double best = 3.4028234663852886e38, bq[3] = {0, 0, 0};
    int bi = -1;
    for (int h = 0; h < k; h++) {
      const int32_t* t = obs_tris + 3 * hits[h];
      double tri[9], q[3];
      for (int c = 0; c < 3; c++) { v3 pc = ldf(pred, t[c]); tri[3 * c] = pc.x; tri[3 * c + 1] = pc.y; tri[3 * c + 2] = pc.z; }
      double d = orc_nearest_point(x + 3 * i, tri, q);
      if (d < best) { best = d; bi = hits[h]; memcpy(bq, q, sizeof(bq)); }
    }
    face_out[i] = bi;
    memcpy(q_out + 3 * i, bq, sizeof(bq));
  }
}

/* GenerateVFCCD_Kernel (PBDCollision.cu:236-270): segment x0 -> x, first Moller-Trumbore hit */
void orc_vf_ccd_generate(const orc_bvh* obs, const int32_t* obs_tris, const double* x0,
                         const double* x, int nv, int32_t* face_out, double* q_out) {
  int32_t hits[256];
  const float* pred = orc_bvh_pred(obs);
  for (int i = 0; i < nv; i++) {
    face_out[i] = -1;
    q_out[3 * i] = q_out[3 * i + 1] = q_out[3 * i + 2] = 0;
    v3 p1 = ldd(x0, i), p2 = ldd(x, i);
    v3 d = sb(p2, p1);
    double len = ln(d);
    if (len < EPS) continue;
    d = sc(d, 1.0 / len);
    float a[3] = {(float)p1.x, (float)p1.y, (float)p1.z}, b[3] = {(float)p2.x, (float)p2.y, (float)p2.z};
    int k = orc_bvh_query(obs, a, b, hits, 256);
    double best = 3.4028234663852886e38;
    for (int h = 0; h < k; h++) {
      const int32_t* t = obs_tris + 3 * hits[h];
      double tri[9], o[3] = {p1.x, p1.y, p1.z}, dd[3] = {d.x, d.y, d.z};
      for (int c = 0; c < 3; c++) { v3 pc = ldf(pred, t[c]); tri[3 * c] = pc.x; tri[3 * c + 1] = pc.y; tri[3 * c + 2] = pc.z; }
      double s = orc_ray_intersect(o, dd, len, tri);
      if (s < best) {
        best = s;
        face_out[i] = hits[h];
        std3(q_out, i, ad(p1, sc(d, s)));
      }
    }
  }
}

/* GenerateEE_DCD_Kernel (PBDCollision.cu:100-150): segment query with the cloth edge,
 * closest segment pair over the R-triangle-owned obstacle edges.  Record per edge:
 * [dist, p2(3), q2(3), s, t, fn(3)] (12 doubles); dist = FLT_MAX when none */
void orc_ee_dcd_generate(const orc_bvh* obs, const int32_t* obs_tris, const uint32_t* obs_rtri,
                         const double* x, const int32_t* edges, int ne, double* rec) {
  int32_t hits[256];
  const float* pred = orc_bvh_pred(obs);
  for (int e = 0; e < ne; e++) {
    double* r = rec + 12 * e;
    for (int k = 0; k < 12; k++) r[k] = 0;
    r[0] = 3.4028234663852886e38;
    const double* pc = x + 3 * edges[2 * e];
    const double* qc = x + 3 * edges[2 * e + 1];
    float a[3] = {(float)pc[0], (float)pc[1], (float)pc[2]}, b[3] = {(float)qc[0], (float)qc[1], (float)qc[2]};
    int k = orc_bvh_query(obs, a, b, hits, 256);
    for (int h = 0; h < k; h++) {
      int f = hits[h];
      const int32_t* t = obs_tris + 3 * f;
      double ov[9];
      for (int c = 0; c < 3; c++) { v3 pp = ldf(pred, t[c]); ov[3 * c] = pp.x; ov[3 * c + 1] = pp.y; ov[3 * c + 2] = pp.z; }
      for (int i = 0; i < 3; i++) {
        if (!((obs_rtri[f] >> (3 + i)) & 1u)) continue;
        int j = (i + 1) % 3;
        double s, tt;
        double dist = orc_closest_seg_seg(pc, qc, ov + 3 * i, ov + 3 * j, &s, &tt);
        if (dist == 0.0) continue; /* the reference uses the distance as the success flag */
        if (dist < r[0]) {
          r[0] = dist;
          memcpy(r + 1, ov + 3 * i, 3 * sizeof(double));
          memcpy(r + 4, ov + 3 * j, 3 * sizeof(double));
          r[7] = s; r[8] = tt;
          v3 fn = face_normal(ldd(ov, 0), ldd(ov, 1), ldd(ov, 2));
          r[9] = fn.x; r[10] = fn.y; r[11] = fn.z;
        }
      }
    }
  }
}

/* SolveVFCCD_Kernel / SolveVFDCD_Kernel (PBDCollision.cu:280-300, 54-76): per vertex push
 * along the pred face normal to the contact point (+1e-3 offset for DCD).  With mu > 0 the
 * build's contact friction: ComputeFriction (ClothSolverHelper.cu:156-169) of that
 * correction on the vertex's substep displacement x - x0 relative to the obstacle face's
 * (the mean of its vertices' (o_x - o_x0) times inv_ns), tangential part only. */
void orc_vf_solve(const float* obs_pred, const int32_t* obs_tris, const int32_t* face,
                  const double* q, double* x, int nv, double offset, const double* x0,
                  const float* o_x0, const float* o_x, double inv_ns, double mu) {
  for (int i = 0; i < nv; i++) {
    if (face[i] < 0) continue;
    const int32_t* t = obs_tris + 3 * face[i];
    v3 n = face_normal(ldf(obs_pred, t[0]), ldf(obs_pred, t[1]), ldf(obs_pred, t[2]));
    v3 xi = ldd(x, i);
    double lam = dt3(n, sb(xi, ldd(q, i))) - offset;
    if (lam >= 0) continue;
    v3 f = {0.0, 0.0, 0.0};
    if (mu > 0) {
      v3 od = {0.0, 0.0, 0.0};
      for (int k = 0; k < 3; k++) od = ad(od, sb(ldf(o_x, t[k]), ldf(o_x0, t[k])));
      od = sc(od, inv_ns / 3.0);
      v3 rel = sb(sb(xi, ldd(x0, i)), od);
      v3 tan = sb(rel, sc(n, dt3(rel, n)));
      double tl = sqrt(dt3(tan, tan));
      if (tl > 0) {
        double s = -lam * mu / tl;
        f = sc(tan, -(s < 1.0 ? s : 1.0));
      }
    }
    std3(x, i, ad(ad(xi, sc(n, -lam)), f));
  }
}

/* SolveEEDCD_Kernel (PBDCollision.cu:155-214): Jacobi deltas on both edge ends */
void orc_ee_solve(const double* rec, const double* x, const int32_t* edges, int ne, double* delta,
                  int32_t* count) {
  for (int e = 0; e < ne; e++) {
    const double* r = rec + 12 * e;
    if (r[0] > 1e6) continue;
    int i1 = edges[2 * e], i2 = edges[2 * e + 1];
    double s = r[7], t = r[8];
    v3 fn = V(r[9], r[10], r[11]);
    v3 c1 = ad(sc(ldd(x, i1), 1 - s), sc(ldd(x, i2), s));
    v3 c2 = ad(sc(V(r[1], r[2], r[3]), 1 - t), sc(V(r[4], r[5], r[6]), t));
    v3 d = sb(c1, c2);
    if (ln(d) < EPS) continue;
    double lam = dt3(d, fn) - 1e-3;
    if (lam > 0) continue;
    std3(delta, i1, ad(ldd(delta, i1), sc(fn, -(1 - s) * lam)));
    std3(delta, i2, ad(ldd(delta, i2), sc(fn, -s * lam)));
    count[i1]++; count[i2]++;
  }
}

/* ApplyDeltas_Kernel (ClothSolverHelper.cu:139-148) / ApplyPBDImpulse_Kernel
 * (PBDRepulsion.cu:299-313, skips count <= EPS) */
void orc_apply_deltas(double* x, double* delta, int32_t* count, int nv, double relax) {
  for (int i = 0; i < nv; i++) {
    if (count[i] > 0) {
      std3(x, i, ad(ldd(x, i), sc(ldd(delta, i), relax / (double)count[i])));
      delta[3 * i] = delta[3 * i + 1] = delta[3 * i + 2] = 0;
      count[i] = 0;
    }
  }
}

/* --------------------------------------------------------- PBD repulsion */
typedef struct {
  const double* x; const double* x0; const double* w; const uint8_t* free_;
  const int32_t* tris; const int32_t* edges; const int32_t* fe;
  const int32_t* ef_ptr; const int32_t* ef; const double* face_n; const uint32_t* rtri;
} rmesh;

static v3 edge_normal(const rmesh* m, int e) { /* Edge::ComputeNormal (Edge.cu:36-50) */
  v3 s = V(0, 0, 0);
  int cnt = 0;
  for (int k = m->ef_ptr[e]; k < m->ef_ptr[e + 1] && cnt < 4; k++, cnt++) {
    const int32_t* t = m->tris + 3 * m->ef[k];
    s = ad(s, face_normal(ldd(m->x, t[0]), ldd(m->x, t[1]), ldd(m->x, t[2])));
  }
  return cnt ? sc(s, 1.0 / cnt) : s;
}

/* InEdge (BasicPrimitiveTests.cu:286-301): rest face normals of edge1's faces */
static int in_edge(double w, const rmesh* m0, int e0, const rmesh* m1, int e1) {
  v3 a = ldd(m0->x, m0->edges[2 * e0]), b = ldd(m0->x, m0->edges[2 * e0 + 1]);
  v3 x = ad(sc(a, 1 - w), sc(b, w));
  int in = 1, cnt = 0;
  v3 n0 = ldd(m1->x, m1->edges[2 * e1]), n1 = ldd(m1->x, m1->edges[2 * e1 + 1]);
  for (int k = m1->ef_ptr[e1]; k < m1->ef_ptr[e1 + 1] && cnt < 4; k++, cnt++) {
    v3 e = sb(n1, n0), fn = ldd(m1->face_n, m1->ef[k]), r = sb(x, n0);
    in &= mixed(e, fn, r) >= 0;
  }
  return in;
}

/* accumulate (AccumulateRepulsions, PBDRepulsion.cu:264-277): skip tiny corrections and
 * non-free nodes (pinned cloth nodes and every obstacle node), so only cloth nodes move */
static int accumulate(double* delta, int32_t* count, const rmesh* const* ms, const int* ids, const v3* corr) {
  int m = 0; /* corners accumulated */
  for (int i = 0; i < 4; i++) {
    if (ln(corr[i]) < EPS) continue;
    if (!ms[i]->free_[ids[i]]) continue;
    std3(delta, ids[i], ad(ldd(delta, ids[i]), corr[i]));
    count[ids[i]]++;
    m |= 1 << i;
  }
  return m;
}

/* EvaluateVFProximity (PBDRepulsion.cu:329-409) */
static void eval_vf(const rmesh* c, int node, const rmesh* o, int face, double h, double* delta,
                    int32_t* count, orc_rep_item* rec) {
  const int32_t* t = o->tris + 3 * face;
  if (c == o && (node == t[0] || node == t[1] || node == t[2])) return;
  double w[4];
  v3 pn;
  /* Proximity(node, face): rest-side normal from x0, flipped if below a free face */
  double d = signed_vf(ldd(c->x0, node), ldd(o->x0, t[0]), ldd(o->x0, t[1]), ldd(o->x0, t[2]), &pn, w);
  int face_free = o->free_[t[0]] && o->free_[t[1]] && o->free_[t[2]];
  if (d < 0.0 && face_free) pn = sc(pn, -1.0);
  v3 fn;
  signed_vf(ldd(c->x, node), ldd(o->x, t[0]), ldd(o->x, t[1]), ldd(o->x, t[2]), &fn, w);
  double mn = -w[1];
  if (-w[2] < mn) mn = -w[2];
  if (-w[3] < mn) mn = -w[3];
  if (!(mn >= 1e-6)) return;
  v3 p1 = ldd(o->x, t[0]);
  v3 n = face_normal(p1, ldd(o->x, t[1]), ldd(o->x, t[2]));
  double side = dt3(n, pn) > 0.0 ? 1.0 : -1.0;
  v3 q = sb(ldd(c->x, node), p1), p2 = sb(ldd(o->x, t[1]), p1), p3 = sb(ldd(o->x, t[2]), p1);
  n = sc(cr(p2, p3), side);
  double c23 = ln(n);
  n = nz(n);
  double C = dt3(q, n) - h;
  if (C > 0) return;
  v3 dcq = n;
  v3 dcp2 = sc(ad(cr(p3, q), sc(cr(n, p3), dt3(n, q))), 1.0 / c23);
  v3 dcp3 = sc(ad(cr(p2, q), sc(cr(n, p2), dt3(n, q))), -1.0 / c23);
  v3 dcp1 = sb(sb(sc(dcq, -1.0), dcp2), dcp3);
  double wq = c->w[node], w1 = o->w[t[0]], w2 = o->w[t[1]], w3 = o->w[t[2]];
  double den = wq * dt3(dcq, dcq) + w1 * dt3(dcp1, dcp1) + w2 * dt3(dcp2, dcp2) + w3 * dt3(dcp3, dcp3);
  if (den < EPS) return;
  double s = C / den;
  v3 corr[4] = {sc(dcq, -wq * s), sc(dcp1, -w1 * s), sc(dcp2, -w2 * s), sc(dcp3, -w3 * s)};
  const rmesh* ms[4] = {c, o, o, o};
  int ids[4] = {node, t[0], t[1], t[2]};
  int m = accumulate(delta, count, ms, ids, corr);
  if (rec && m) {
    memset(rec, 0, sizeof(*rec));
    rec->kind = 0;
    rec->other_cloth = c == o;
    rec->contrib = m;
    for (int i = 0; i < 4; i++) rec->ids[i] = ids[i];
    rec->side = side;
  }
}

/* EvaluateEEProximity (PBDRepulsion.cu:411-468) */
static void eval_ee(const rmesh* c, int e0, const rmesh* o, int e1, double h, double* delta,
                    int32_t* count, orc_rep_item* rec) {
  int a0 = c->edges[2 * e0], a1 = c->edges[2 * e0 + 1], b0 = o->edges[2 * e1], b1 = o->edges[2 * e1 + 1];
  if (c == o && (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1)) return;
  double w[4];
  v3 pn;
  double d = signed_ee(ldd(c->x0, a0), ldd(c->x0, a1), ldd(o->x0, b0), ldd(o->x0, b1), &pn, w);
  int e1_free = o->free_[b0] && o->free_[b1];
  if (!e1_free) pn = edge_normal(o, e1);
  else if (d < 0.0) pn = sc(pn, -1.0);
  double b[4];
  v3 ncur;
  signed_ee(ldd(c->x, a0), ldd(c->x, a1), ldd(o->x, b0), ldd(o->x, b1), &ncur, b);
  double mn = b[0];
  if (b[1] < mn) mn = b[1];
  if (-b[2] < mn) mn = -b[2];
  if (-b[3] < mn) mn = -b[3];
  if (!(mn >= 1e-6 && in_edge(b[1], c, e0, o, e1) && in_edge(-b[3], o, e1, c, e0))) return;
  v3 dd = ad(ad(sc(ldd(c->x, a0), b[0]), sc(ldd(c->x, a1), b[1])),
             ad(sc(ldd(o->x, b0), b[2]), sc(ldd(o->x, b1), b[3])));
  double C = dt3(dd, pn) - h;
  if (C > 0) return;
  b[2] *= -1; b[3] *= -1;
  v3 g[4] = {sc(pn, b[0]), sc(pn, b[1]), sc(pn, -b[2]), sc(pn, -b[3])};
  double wq[4] = {c->w[a0], c->w[a1], o->w[b0], o->w[b1]};
  double den = 0;
  for (int i = 0; i < 4; i++) den += wq[i] * b[i] * b[i];
  if (den < EPS) return;
  double s = C / den;
  v3 corr[4];
  for (int i = 0; i < 4; i++) corr[i] = sc(g[i], -s * wq[i]);
  const rmesh* ms[4] = {c, c, o, o};
  int ids[4] = {a0, a1, b0, b1};
  int m = accumulate(delta, count, ms, ids, corr);
  if (rec && m) {
    memset(rec, 0, sizeof(*rec));
    rec->kind = 1;
    rec->other_cloth = c == o;
    rec->contrib = m;
    for (int i = 0; i < 4; i++) rec->ids[i] = ids[i];
    rec->pn_x0 = e1_free;
    rec->pn_sign = d < 0.0 ? -1.0 : 1.0;
    rec->pn[0] = pn.x; rec->pn[1] = pn.y; rec->pn[2] = pn.z;
  }
}

/* EvaluateRep_Kernel (PBDRepulsion.cu:470-501) over pairs (cloth face, other face);
 * kind[p] = 1: other face is a cloth face, 0: an obstacle face */
static long repulsion(const orc_rmesh_in* cl_in, const orc_rmesh_in* ob_in, const int32_t* pairs,
                      const int32_t* kind, long np, double h, double* delta, int32_t* count,
                      orc_rep_item* items, long cap) {
  rmesh c = {cl_in->x, cl_in->x0, cl_in->w, cl_in->free_, cl_in->tris, cl_in->edges, cl_in->fe,
             cl_in->ef_ptr, cl_in->ef, cl_in->face_n, cl_in->rtri};
  rmesh o = {0};
  if (ob_in)
    o = (rmesh){ob_in->x, ob_in->x0, ob_in->w, ob_in->free_, ob_in->tris, ob_in->edges, ob_in->fe,
                ob_in->ef_ptr, ob_in->ef, ob_in->face_n, ob_in->rtri};
  long n = 0;
  orc_rep_item scratch;
  for (long p = 0; p < np; p++) {
    int f0 = pairs[2 * p], f1 = pairs[2 * p + 1];
    const rmesh* other = kind[p] ? &c : &o;
    uint32_t r0 = c.rtri[f0], r1 = other->rtri[f1];
    for (int i = 0; i < 3; i++) {
      if (!((r0 >> i) & 1u)) continue;
      orc_rep_item* rec = items ? (n < cap ? items + n : &scratch) : NULL;
      if (rec) rec->contrib = 0;
      eval_vf(&c, c.tris[3 * f0 + i], other, f1, h, delta, count, rec);
      if (rec && rec->contrib) n++;
    }
    for (int i = 0; i < 3; i++) {
      if (!((r0 >> (3 + i)) & 1u)) continue;
      for (int j = 0; j < 3; j++) {
        if (!((r1 >> (3 + j)) & 1u)) continue;
        orc_rep_item* rec = items ? (n < cap ? items + n : &scratch) : NULL;
        if (rec) rec->contrib = 0;
        eval_ee(&c, c.fe[3 * f0 + i], other, other->fe[3 * f1 + j], h, delta, count, rec);
        if (rec && rec->contrib) n++;
      }
    }
  }
  return n;
}

void orc_repulsion(const orc_rmesh_in* cl_in, const orc_rmesh_in* ob_in, const int32_t* pairs,
                   const int32_t* kind, long np, double h, double* delta, int32_t* count) {
  repulsion(cl_in, ob_in, pairs, kind, np, h, delta, count, NULL, 0);
}

long orc_repulsion_record(const orc_rmesh_in* cl_in, const orc_rmesh_in* ob_in,
                          const int32_t* pairs, const int32_t* kind, long np, double h,
                          double* delta, int32_t* count, orc_rep_item* items, long cap) {
  return repulsion(cl_in, ob_in, pairs, kind, np, h, delta, count, items, cap);
}
