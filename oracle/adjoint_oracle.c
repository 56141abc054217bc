/*
 * adjoint_oracle.c -- FP64 reverse-mode products (vector-Jacobian) of the substep's
 * operators, the host reference gradient for the library's adjoint kernels.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP adjoint.  Nothing in the product
 * path links or calls this file.
 *
 * Method: every operator is restated element by element in dual numbers (value plus one
 * tangent) following the forward oracle's own formulas (xrt_oracle.c, contact_oracle.c);
 * an element's Jacobian is assembled one input direction at a time and its transpose is
 * applied to the output adjoint.  This is independent of the hand-derived reverse
 * formulas of the library.  The assembly (operator order, multiplier carry, predict /
 * finalize) is pinned against central finite differences of whole oracle rollouts
 * (tests/test_oracle_adjoint.py).
 *
 * Each function reads the operator's INPUT state (c->x = positions before the operator,
 * c->x0 = start of the substep) and replaces the adjoint of the operator's output with
 * the adjoint of its input.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "xrt_oracle.h"

#define ORC_PI 3.14159265358979323846
#define ORC_EPS 1e-6 /* EPSILON (xrtailor/include/xrtailor/core/Scalar.hpp:37) */
#define MAXIO 24

/* ----------------------------------------------------------- dual numbers */
typedef struct { double v, d; } dn;
typedef struct { dn x, y, z; } dv3;

static inline dn K(double v) { dn r = {v, 0.0}; return r; }
static inline dn dadd(dn a, dn b) { dn r = {a.v + b.v, a.d + b.d}; return r; }
static inline dn dsub(dn a, dn b) { dn r = {a.v - b.v, a.d - b.d}; return r; }
static inline dn dneg(dn a) { dn r = {-a.v, -a.d}; return r; }
static inline dn dmul(dn a, dn b) { dn r = {a.v * b.v, a.d * b.v + a.v * b.d}; return r; }
static inline dn ddiv(dn a, dn b) {
  dn r = {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)};
  return r;
}
/* sqrt and acos take the zero subgradient where their slope is infinite (sqrt(0),
 * acos(+-1)): there the bend's d = n1.n2 sits at a fold (|d| = 1) whose own derivative
 * dd/dx vanishes, so the product is zero in the limit */
static inline dn dsqrt(dn a) {
  double s = sqrt(a.v);
  dn r = {s, (a.d == 0.0 || a.v <= 0.0) ? 0.0 : a.d / (2.0 * s)};
  return r;
}
static inline dn dacos(dn a) {
  double q = 1.0 - a.v * a.v;
  dn r = {acos(a.v), (a.d == 0.0 || q <= 0.0) ? 0.0 : -a.d / sqrt(q)};
  return r;
}
static inline dv3 V3(dn x, dn y, dn z) { dv3 r = {x, y, z}; return r; }
static inline dv3 vadd(dv3 a, dv3 b) { return V3(dadd(a.x, b.x), dadd(a.y, b.y), dadd(a.z, b.z)); }
static inline dv3 vsub(dv3 a, dv3 b) { return V3(dsub(a.x, b.x), dsub(a.y, b.y), dsub(a.z, b.z)); }
static inline dv3 vscl(dv3 a, dn s) { return V3(dmul(a.x, s), dmul(a.y, s), dmul(a.z, s)); }
static inline dn vdot(dv3 a, dv3 b) {
  return dadd(dadd(dmul(a.x, b.x), dmul(a.y, b.y)), dmul(a.z, b.z));
}
static inline dv3 vcrs(dv3 a, dv3 b) {
  return V3(dsub(dmul(a.y, b.z), dmul(a.z, b.y)), dsub(dmul(a.z, b.x), dmul(a.x, b.z)),
            dsub(dmul(a.x, b.y), dmul(a.y, b.x)));
}
static inline dn vlen(dv3 a) { return dsqrt(vdot(a, a)); }
static inline dv3 vnrm(dv3 a) { return vscl(a, ddiv(K(1.0), vlen(a))); }
static inline dv3 vin(const dn* in, int k) { return V3(in[k], in[k + 1], in[k + 2]); }
static inline void vout(dn* out, int k, dv3 a) { out[k] = a.x; out[k + 1] = a.y; out[k + 2] = a.z; }

/* out_bar -> in_bar for one element: J^T ybar, J assembled column by column */
typedef void (*elem_fn)(const dn* in, dn* out, const void* ctx);
static void elem_vjp(elem_fn f, const void* ctx, const double* x, int nin, const double* ybar,
                     int nout, double* xbar) {
  dn in[MAXIO], out[MAXIO];
  for (int k = 0; k < nin; k++) {
    for (int j = 0; j < nin; j++) { in[j].v = x[j]; in[j].d = (j == k) ? 1.0 : 0.0; }
    f(in, out, ctx);
    double s = 0.0;
    for (int j = 0; j < nout; j++) s += ybar[j] * out[j].d;
    xbar[k] = s;
  }
}

static inline void ld3(const double* p, int i, double* o) {
  o[0] = p[3 * i]; o[1] = p[3 * i + 1]; o[2] = p[3 * i + 2];
}
static inline void st3(double* p, int i, const double* o) {
  p[3 * i] = o[0]; p[3 * i + 1] = o[1]; p[3 * i + 2] = o[2];
}

/* ---------------------------------------------------------------- stretch */
typedef struct { double w1, w2, rest, dt; int iter; } stretch_ctx;

/* in: a[3] b[3] lambda compliance; out: a'[3] b'[3] lambda'  (orc_stretch_color) */
static void stretch_elem(const dn* in, dn* out, const void* vctx) {
  const stretch_ctx* c = (const stretch_ctx*)vctx;
  dv3 a = vin(in, 0), b = vin(in, 3);
  dn lam = c->iter == 0 ? K(0.0) : in[6];
  dv3 diff = vsub(a, b);
  dn dist = vlen(diff);
  if (dist.v != c->rest && c->w1 + c->w2 > 0.0) {
    dv3 g = vscl(diff, ddiv(K(1.0), dadd(dist, K(ORC_EPS))));
    dn at = ddiv(in[7], K(c->dt * c->dt));
    dn C = dsub(dist, K(c->rest));
    dn dl = ddiv(dsub(dneg(C), dmul(at, lam)), dadd(K(c->w1 + c->w2), at));
    dv3 corr = vscl(g, dl);
    lam = dadd(lam, dl);
    a = vadd(a, vscl(corr, K(c->w1)));
    b = vsub(b, vscl(corr, K(c->w2)));
  }
  vout(out, 0, a);
  vout(out, 3, b);
  out[6] = lam;
}

void orc_adj_stretch_color(const orc_cloth* c, int k, double dt, int iter, const double* lam_pre,
                           double* gx, double* glam, double* gs) {
  for (int id = c->s_color_off[k]; id < c->s_color_off[k + 1]; id++) {
    int i1 = c->s_idx[2 * id], i2 = c->s_idx[2 * id + 1];
    stretch_ctx ctx = {c->inv_mass[i1], c->inv_mass[i2], c->s_rest[id], dt, iter};
    double x[8], yb[7], xb[8];
    ld3(c->x, i1, x); ld3(c->x, i2, x + 3);
    x[6] = lam_pre[id]; x[7] = c->s_compliance[id];
    ld3(gx, i1, yb); ld3(gx, i2, yb + 3);
    yb[6] = glam[id];
    elem_vjp(stretch_elem, &ctx, x, 8, yb, 7, xb);
    st3(gx, i1, xb); st3(gx, i2, xb + 3);
    glam[id] = iter == 0 ? 0.0 : xb[6];
    gs[id] += xb[7];
  }
}

/* ----------------------------------------------------------------- LRA */
typedef struct { double target; } lra_ctx;

/* in: x_t[3] s[3]; out: x_t'[3] (one slot's projection in orc_lra, active branch) */
static void lra_elem(const dn* in, dn* out, const void* vctx) {
  const lra_ctx* c = (const lra_ctx*)vctx;
  dv3 xt = vin(in, 0), sp = vin(in, 3);
  dv3 diff = vsub(xt, sp);
  dn dist = vlen(diff);
  vout(out, 0, vadd(vadd(xt, vscl(diff, K(-1.0))), vscl(diff, ddiv(K(c->target), dist))));
}

/* SolveGeodesicLRA (orc_lra) reversed: each target replays its slots in registration order
 * from the pre-LRA state (c->x), then walks them backwards; the slot snapshot's adjoint is
 * summed per slot and added to the slot vertex's (the snapshot is x at the pass start). */
void orc_adj_lra(const orc_cloth* c, const orc_params* p, double* gx) {
  if (c->nslots <= 0) return;
  double* sp = (double*)malloc(sizeof(double) * 3 * (size_t)c->nslots);
  double* gsl = (double*)calloc(3 * (size_t)c->nslots, sizeof(double));
  int maxk = 0;
  for (int t = 0; t < c->nv; t++)
    if (c->lra_ptr[t + 1] - c->lra_ptr[t] > maxk) maxk = c->lra_ptr[t + 1] - c->lra_ptr[t];
  double* xs = (double*)malloc(sizeof(double) * 3 * (size_t)(maxk + 1));
  for (int s = 0; s < c->nslots; s++) ld3(c->x, c->lra_slot_vert[s], sp + 3 * s);
  for (int t = 0; t < c->nv; t++) {
    const int kb = c->lra_ptr[t], ke = c->lra_ptr[t + 1];
    if (kb == ke) continue;
    ld3(c->x, t, xs);
    for (int k = kb; k < ke; k++) {  /* forward replay: xs[j] = x_t before slot j */
      double* a = xs + 3 * (k - kb);
      double* b = a + 3;
      const double* sk = sp + 3 * c->lra_slot[k];
      const double target = c->lra_rest[k] * p->long_range_stretchiness;
      const double d[3] = {a[0] - sk[0], a[1] - sk[1], a[2] - sk[2]};
      const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      const int skip = (c->inv_mass[t] < ORC_EPS && target > 0) || !(dist > target);
      for (int q = 0; q < 3; q++) b[q] = skip ? a[q] : a[q] - d[q] + d[q] * (target / dist);
    }
    double g[3];
    ld3(gx, t, g);
    for (int k = ke - 1; k >= kb; k--) {
      const double* a = xs + 3 * (k - kb);
      const double* sk = sp + 3 * c->lra_slot[k];
      const double target = c->lra_rest[k] * p->long_range_stretchiness;
      const double d[3] = {a[0] - sk[0], a[1] - sk[1], a[2] - sk[2]};
      const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if ((c->inv_mass[t] < ORC_EPS && target > 0) || !(dist > target)) continue;
      lra_ctx ctx = {target};
      double in[6] = {a[0], a[1], a[2], sk[0], sk[1], sk[2]}, ib[6];
      elem_vjp(lra_elem, &ctx, in, 6, g, 3, ib);
      g[0] = ib[0]; g[1] = ib[1]; g[2] = ib[2];
      double* gs = gsl + 3 * c->lra_slot[k];
      gs[0] += ib[3]; gs[1] += ib[4]; gs[2] += ib[5];
    }
    st3(gx, t, g);
  }
  for (int s = 0; s < c->nslots; s++) {
    double* gv = gx + 3 * (size_t)c->lra_slot_vert[s];
    gv[0] += gsl[3 * s]; gv[1] += gsl[3 * s + 1]; gv[2] += gsl[3 * s + 2];
  }
  free(xs);
  free(gsl);
  free(sp);
}

/* ------------------------------------------------------------------- bend */
typedef struct { double w[4], theta, dt; int model; } bend_ctx;

/* in: x1..x4 [12] compliance; out: x1'..x4' [12]  (orc_bend_color) */
static void bend_elem(const dn* in, dn* out, const void* vctx) {
  const bend_ctx* c = (const bend_ctx*)vctx;
  const int consistent = c->model != 0;
  const double dlo = consistent ? -1.0 : 0.0, sgn = consistent ? -1.0 : 1.0;
  for (int j = 0; j < 12; j++) out[j] = in[j];
  dv3 x1 = vin(in, 0);
  dv3 p2 = vsub(vin(in, 3), x1), p3 = vsub(vin(in, 6), x1), p4 = vsub(vin(in, 9), x1);
  dv3 c23 = vcrs(p2, p3), c24 = vcrs(p2, p4);
  dv3 n1 = vnrm(c23), n2 = vnrm(c24);
  dn d = vdot(n1, n2);
  if (isnan(d.v)) return;
  if (d.v < dlo) d = K(dlo);
  if (d.v > 1.0) d = K(1.0);
  dn ang = dacos(d);
  double th = consistent ? (ORC_PI - c->theta) : c->theta;
  if (ang.v < ORC_EPS || isnan(ang.v)) return;
  dn i23 = ddiv(K(1.0), dadd(vlen(c23), K(ORC_EPS)));
  dn i24 = ddiv(K(1.0), dadd(vlen(c24), K(ORC_EPS)));
  dv3 q3 = vscl(vadd(vcrs(p2, n2), vscl(vcrs(n1, p2), d)), i23);
  dv3 q4 = vscl(vadd(vcrs(p2, n1), vscl(vcrs(n2, p2), d)), i24);
  dv3 q2 = vsub(vscl(vscl(vadd(vcrs(p3, n2), vscl(vcrs(n1, p3), d)), i23), K(-1.0)),
                vscl(vadd(vcrs(p4, n1), vscl(vcrs(n2, p4), d)), i24));
  dv3 q1 = vscl(vadd(vadd(q2, q3), q4), K(-1.0));
  dn denom = ddiv(in[12], K(c->dt * c->dt));
  denom = dadd(denom, dmul(K(c->w[0]), vdot(q1, q1)));
  denom = dadd(denom, dmul(K(c->w[1]), vdot(q2, q2)));
  denom = dadd(denom, dmul(K(c->w[2]), vdot(q3, q3)));
  denom = dadd(denom, dmul(K(c->w[3]), vdot(q4, q4)));
  if (denom.v < ORC_EPS) return;
  dn s = dsqrt(dsub(K(1.0), dmul(d, d)));
  dn lam = ddiv(dmul(K(sgn), dmul(s, dsub(ang, K(th)))), denom);
  dv3 q[4] = {q1, q2, q3, q4};
  for (int k = 0; k < 4; k++)
    vout(out, 3 * k, vadd(vin(in, 3 * k), vscl(q[k], dmul(K(c->w[k]), lam))));
}

void orc_adj_bend_color(const orc_cloth* c, const orc_params* p, int k, double dt, double* gx,
                        double* gb) {
  for (int id = c->b_color_off[k]; id < c->b_color_off[k + 1]; id++) {
    const int32_t* ix = c->b_idx + 4 * id;
    bend_ctx ctx;
    ctx.theta = c->b_theta[id];
    ctx.dt = dt;
    ctx.model = p->bend_model;
    double x[13], yb[12], xb[13];
    for (int j = 0; j < 4; j++) {
      ctx.w[j] = c->inv_mass[ix[j]];
      ld3(c->x, ix[j], x + 3 * j);
      ld3(gx, ix[j], yb + 3 * j);
    }
    x[12] = c->b_compliance[id];
    elem_vjp(bend_elem, &ctx, x, 13, yb, 12, xb);
    for (int j = 0; j < 4; j++) st3(gx, ix[j], xb + 3 * j);
    gb[id] += xb[12];
  }
}

/* ----------------------------------------------------------------- ground */
typedef struct { double h, mu; } ground_ctx;

/* in: x[3] x0[3]; out: x'[3]  (orc_ground for a free vertex below the plane) */
static void ground_elem(const dn* in, dn* out, const void* vctx) {
  const ground_ctx* c = (const ground_ctx*)vctx;
  dv3 x = vin(in, 0), x0 = vin(in, 3);
  dn cl = dsub(K(c->h), x.y);
  dv3 corr = V3(K(0.0), cl, K(0.0));
  dv3 fr = V3(K(0.0), K(0.0), K(0.0));
  if (c->mu > 0.0 && cl.v > 0.0) {
    dv3 rel = vsub(x, x0);
    dv3 nn = V3(K(0.0), K(1.0), K(0.0));
    dv3 tan = vsub(rel, vscl(nn, vdot(rel, nn)));
    dn tl2 = vdot(tan, tan);
    dn s = K(1.0);
    if (tl2.v > 0.0) {
      dn tl = dsqrt(tl2);
      dn mx = dmul(cl, K(c->mu));
      if (mx.v / tl.v < 1.0) s = ddiv(mx, tl);
    }
    fr = vscl(tan, dneg(s));
  }
  vout(out, 0, vadd(vadd(x, corr), fr));
}

void orc_adj_ground(const orc_cloth* c, const orc_params* p, double* gx, double* gx0) {
  if (!p->enable_ground) return;
  ground_ctx ctx = {p->ground_height, p->ground_friction};
  for (int i = 0; i < c->nv; i++) {
    if (c->inv_mass[i] <= ORC_EPS) continue;
    if (c->x[3 * i + 1] >= p->ground_height) continue;
    double x[6], yb[3], xb[6];
    ld3(c->x, i, x); ld3(c->x0, i, x + 3);
    ld3(gx, i, yb);
    elem_vjp(ground_elem, &ctx, x, 6, yb, 3, xb);
    st3(gx, i, xb);
    gx0[3 * i] += xb[3]; gx0[3 * i + 1] += xb[4]; gx0[3 * i + 2] += xb[5];
  }
}

/* ------------------------------------------------------ body-cloth contacts */
typedef struct { double n[3], q[3], offset; } vf_ctx;

/* in: x[3]; out: x'[3]  (orc_vf_solve; the record's face and q fixed) */
static void vf_elem(const dn* in, dn* out, const void* vctx) {
  const vf_ctx* c = (const vf_ctx*)vctx;
  dv3 x = vin(in, 0);
  dv3 n = V3(K(c->n[0]), K(c->n[1]), K(c->n[2]));
  dv3 q = V3(K(c->q[0]), K(c->q[1]), K(c->q[2]));
  dn lam = dsub(vdot(n, vsub(x, q)), K(c->offset));
  if (lam.v < 0) x = vadd(x, vscl(n, dneg(lam)));
  vout(out, 0, x);
}

static void face_normal_f(const float* pred, const int32_t* t, double* n) {
  double a[3], b[3], cc[3];
  for (int k = 0; k < 3; k++) {
    a[k] = pred[3 * t[0] + k]; b[k] = pred[3 * t[1] + k]; cc[k] = pred[3 * t[2] + k];
  }
  double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double e2[3] = {cc[0] - a[0], cc[1] - a[1], cc[2] - a[2]};
  double x = e1[1] * e2[2] - e1[2] * e2[1], y = e1[2] * e2[0] - e1[0] * e2[2],
         z = e1[0] * e2[1] - e1[1] * e2[0];
  double l = sqrt(x * x + y * y + z * z);
  n[0] = x / l; n[1] = y / l; n[2] = z / l;
}

void orc_adj_vf(const float* obs_pred, const int32_t* obs_tris, const int32_t* face, const double* q,
                const double* x_pre, int nv, double offset, double* gx) {
  for (int i = 0; i < nv; i++) {
    if (face[i] < 0) continue;
    vf_ctx ctx;
    face_normal_f(obs_pred, obs_tris + 3 * face[i], ctx.n);
    ld3(q, i, ctx.q);
    ctx.offset = offset;
    double x[3], yb[3], xb[3];
    ld3(x_pre, i, x);
    ld3(gx, i, yb);
    elem_vjp(vf_elem, &ctx, x, 3, yb, 3, xb);
    st3(gx, i, xb);
  }
}

typedef struct { const double* r; } ee_ctx;

/* in: x_i[3] x_j[3]; out: delta_i[3] delta_j[3]  (orc_ee_solve for an active record) */
static void ee_elem(const dn* in, dn* out, const void* vctx) {
  const double* r = ((const ee_ctx*)vctx)->r;
  const double s = r[7], t = r[8];
  dv3 fn = V3(K(r[9]), K(r[10]), K(r[11]));
  dv3 c1 = vadd(vscl(vin(in, 0), K(1 - s)), vscl(vin(in, 3), K(s)));
  dv3 c2 = V3(K((1 - t) * r[1] + t * r[4]), K((1 - t) * r[2] + t * r[5]),
              K((1 - t) * r[3] + t * r[6]));
  dv3 d = vsub(c1, c2);
  dn lam = dsub(vdot(d, fn), K(1e-3));
  vout(out, 0, vscl(fn, dmul(K(-(1 - s)), lam)));
  vout(out, 3, vscl(fn, dmul(K(-s), lam)));
}

static int ee_active(const double* r, const double* x, int i1, int i2) {
  if (r[0] > 1e6) return 0;
  const double s = r[7], t = r[8];
  double d[3], dd = 0.0, lam = 0.0;
  for (int k = 0; k < 3; k++) {
    double c1 = (1 - s) * x[3 * i1 + k] + s * x[3 * i2 + k];
    double c2 = (1 - t) * r[1 + k] + t * r[4 + k];
    d[k] = c1 - c2;
    dd += d[k] * d[k];
    lam += d[k] * r[9 + k];
  }
  if (sqrt(dd) < ORC_EPS) return 0;
  return (lam - 1e-3) <= 0;
}

/* reverse of orc_ee_solve + orc_apply_deltas(relax): x' = x + relax * delta / count */
void orc_adj_ee(const double* rec12, const double* x_pre, const int32_t* edges, int ne, int nv,
                double relax, double* gx) {
  int32_t* cnt = (int32_t*)calloc((size_t)nv, sizeof(int32_t));
  for (int e = 0; e < ne; e++)
    if (ee_active(rec12 + 12 * e, x_pre, edges[2 * e], edges[2 * e + 1])) {
      cnt[edges[2 * e]]++;
      cnt[edges[2 * e + 1]]++;
    }
  double* add = (double*)calloc(3 * (size_t)nv, sizeof(double));
  for (int e = 0; e < ne; e++) {
    const int i1 = edges[2 * e], i2 = edges[2 * e + 1];
    if (!ee_active(rec12 + 12 * e, x_pre, i1, i2)) continue;
    ee_ctx ctx = {rec12 + 12 * e};
    double x[6], yb[6], xb[6];
    ld3(x_pre, i1, x); ld3(x_pre, i2, x + 3);
    for (int k = 0; k < 3; k++) {
      yb[k] = gx[3 * i1 + k] * relax / cnt[i1];
      yb[3 + k] = gx[3 * i2 + k] * relax / cnt[i2];
    }
    elem_vjp(ee_elem, &ctx, x, 6, yb, 6, xb);
    for (int k = 0; k < 3; k++) { add[3 * i1 + k] += xb[k]; add[3 * i2 + k] += xb[3 + k]; }
  }
  for (int k = 0; k < 3 * nv; k++) gx[k] += add[k];
  free(add);
  free(cnt);
}

/* ----------------------------------------------------------- PBD repulsion */
typedef struct { double side, h, w[4]; } vfr_ctx;

/* in: node[3] face corners y0 y1 y2 [3 each]; out: corrections of node, y0, y1, y2
 * (EvaluateVFProximity's response with the recorded orientation, contact_oracle.c eval_vf) */
static void vfr_elem(const dn* in, dn* out, const void* vctx) {
  const vfr_ctx* c = (const vfr_ctx*)vctx;
  dv3 p1 = vin(in, 3);
  dv3 q = vsub(vin(in, 0), p1), p2 = vsub(vin(in, 6), p1), p3 = vsub(vin(in, 9), p1);
  dv3 n = vscl(vcrs(p2, p3), K(c->side));
  dn c23 = vlen(n);
  n = vnrm(n);
  dn C = dsub(vdot(q, n), K(c->h));
  dn nq = vdot(n, q);
  dv3 dc[4];
  dc[0] = n;
  dc[2] = vscl(vadd(vcrs(p3, q), vscl(vcrs(n, p3), nq)), ddiv(K(1.0), c23));
  dc[3] = vscl(vadd(vcrs(p2, q), vscl(vcrs(n, p2), nq)), ddiv(K(-1.0), c23));
  dc[1] = vsub(vsub(vscl(dc[0], K(-1.0)), dc[2]), dc[3]);
  dn den = K(0.0);
  for (int i = 0; i < 4; i++) den = dadd(den, dmul(K(c->w[i]), vdot(dc[i], dc[i])));
  dn s = ddiv(C, den);
  for (int i = 0; i < 4; i++) vout(out, 3 * i, vscl(dc[i], dmul(K(-c->w[i]), s)));
}

typedef struct { int pn_x0; double pn_sign, pn[3], h, w[4]; } eer_ctx;

static inline dn dmixed(dv3 a, dv3 b, dv3 c) { return vdot(a, vcrs(b, c)); }

/* in: x_a0 x_a1 x_b0 x_b1 (current) then the same four at the start of the substep [3
 * each]; out: corrections of a0, a1, b0, b1 (EvaluateEEProximity's response, eval_ee) */
static void eer_elem(const dn* in, dn* out, const void* vctx) {
  const eer_ctx* c = (const eer_ctx*)vctx;
  dv3 x0 = vin(in, 0), x1 = vin(in, 3), y0 = vin(in, 6), y1 = vin(in, 9);
  dv3 pn;
  if (c->pn_x0)
    pn = vscl(vcrs(vnrm(vsub(vin(in, 15), vin(in, 12))), vnrm(vsub(vin(in, 21), vin(in, 18)))),
              K(c->pn_sign));
  else
    pn = V3(K(c->pn[0]), K(c->pn[1]), K(c->pn[2]));
  dv3 n = vcrs(vnrm(vsub(x1, x0)), vnrm(vsub(y1, y0)));
  dn a0 = dmixed(vsub(y1, x1), vsub(y0, x1), n), a1 = dmixed(vsub(y0, x0), vsub(y1, x0), n);
  dn b0 = dmixed(vsub(x0, y1), vsub(x1, y1), n), b1 = dmixed(vsub(x1, y0), vsub(x0, y0), n);
  dn w[4] = {ddiv(a0, dadd(a0, a1)), ddiv(a1, dadd(a0, a1)), dneg(ddiv(b0, dadd(b0, b1))),
             dneg(ddiv(b1, dadd(b0, b1)))};
  dv3 dd = vadd(vadd(vscl(x0, w[0]), vscl(x1, w[1])), vadd(vscl(y0, w[2]), vscl(y1, w[3])));
  dn C = dsub(vdot(dd, pn), K(c->h));
  dn den = K(0.0);
  for (int i = 0; i < 4; i++) den = dadd(den, dmul(K(c->w[i]), dmul(w[i], w[i])));
  dn s = ddiv(C, den);
  for (int i = 0; i < 4; i++) vout(out, 3 * i, vscl(pn, dmul(w[i], dmul(K(-c->w[i]), s))));
}

void orc_adj_repulsion(const orc_rep_item* items, long n, const orc_rmesh_in* cloth,
                       const orc_rmesh_in* obstacle, const int32_t* count, int nv, double h,
                       double relax, double* gx, double* gx0) {
  double* add = (double*)calloc(3 * (size_t)nv, sizeof(double));
  double* add0 = (double*)calloc(3 * (size_t)nv, sizeof(double));
  for (long k = 0; k < n; k++) {
    const orc_rep_item* it = items + k;
    const orc_rmesh_in* other = it->other_cloth ? cloth : obstacle;
    /* corner i's mesh: VF node (cloth) + face (other); EE edge a (cloth) + edge b (other) */
    const orc_rmesh_in* ms[4] = {cloth, other, other, other};
    if (it->kind == 1) ms[1] = cloth;
    double x[24], yb[12], xb[24];
    int nin = it->kind == 0 ? 12 : 24;
    double w[4];
    for (int i = 0; i < 4; i++) {
      ld3(ms[i]->x, it->ids[i], x + 3 * i);
      if (it->kind == 1) ld3(ms[i]->x0, it->ids[i], x + 12 + 3 * i);
      w[i] = ms[i]->w[it->ids[i]];
      for (int j = 0; j < 3; j++)
        yb[3 * i + j] = ((it->contrib >> i) & 1) ? gx[3 * it->ids[i] + j] * relax / count[it->ids[i]]
                                                 : 0.0;
    }
    if (it->kind == 0) {
      vfr_ctx ctx = {it->side, h, {w[0], w[1], w[2], w[3]}};
      elem_vjp(vfr_elem, &ctx, x, nin, yb, 12, xb);
    } else {
      eer_ctx ctx = {it->pn_x0, it->pn_sign, {it->pn[0], it->pn[1], it->pn[2]}, h,
                     {w[0], w[1], w[2], w[3]}};
      elem_vjp(eer_elem, &ctx, x, nin, yb, 12, xb);
    }
    for (int i = 0; i < 4; i++) {
      if (ms[i] != cloth) continue;  /* obstacle nodes are constants */
      for (int j = 0; j < 3; j++) {
        add[3 * it->ids[i] + j] += xb[3 * i + j];
        if (it->kind == 1) add0[3 * it->ids[i] + j] += xb[12 + 3 * i + j];
      }
    }
  }
  for (int k = 0; k < 3 * nv; k++) {
    gx[k] += add[k];
    gx0[k] += add0[k];
  }
  free(add);
  free(add0);
}
