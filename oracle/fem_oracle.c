/*
 * fem_oracle.c -- serial FP64 restatement of the quality mode's FEM constraints: the
 * orthotropic Green-strain triangle constraint and the isometric (quadratic) bending
 * constraint (xrtailor/src/physics/dynamics/FEMHelper.cu:10-392), with glm's column-major
 * storage spelt out index by index.
 *
 * TEST INFRASTRUCTURE ONLY (see xrt_oracle.h).
 *
 * Pinned by (tests/test_oracle_fem.py): the strain energy is zero at rest and under rigid
 * motions, its gradient (the constraint gradient H) matches central differences of the
 * energy, uniaxial stretch gives the closed-form Green strain energy; the bending weights K
 * satisfy sum_i K_i x_i = 0 for any planar hinge (the cotangent identity), the energy is
 * 0.5 coeff |sum_i K_i x_i|^2 and its gradient matches central differences; one XPBD bending
 * step lowers the energy.
 */
#include <math.h>

#include "xrt_oracle.h"

#define EPS 1e-6

typedef struct { double x, y, z; } d3;
static d3 L3(const double* p, int i) { d3 r = {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; return r; }
static d3 sub3(d3 a, d3 b) { d3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 cross3(d3 a, d3 b) {
  d3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  return r;
}
static d3 norm3(d3 a) {
  const double l = sqrt(dot3(a, a));
  d3 r = {a.x / l, a.y / l, a.z / l};
  return r;
}
static double comp(d3 a, int c) { return c == 0 ? a.x : (c == 1 ? a.y : a.z); }

/* ComputeStrainMeta (FEMHelper.cu:10-40): area and glm::inverse of the 2x2 rest matrix,
 * inv = [inv[0][0], inv[0][1], inv[1][0], inv[1][1]] (column, row); identity when the rest
 * triangle is degenerate.  Returns 0 for a degenerate triangle. */
int orc_strain_init(const double* x, const int32_t* t, double* area, double* inv) {
  const d3 p0 = L3(x, t[0]), p1 = L3(x, t[1]), p2 = L3(x, t[2]);
  const d3 n0 = cross3(sub3(p1, p0), sub3(p2, p0));
  *area = 0.5 * sqrt(dot3(n0, n0));
  const d3 a01 = norm3(sub3(p1, p0));
  const d3 a02 = norm3(cross3(n0, a01));
  const double q[3][2] = {{dot3(p0, a02), dot3(p0, a01)}, {dot3(p1, a02), dot3(p1, a01)},
                          {dot3(p2, a02), dot3(p2, a01)}};
  const double P00 = q[0][0] - q[2][0], P10 = q[0][1] - q[2][1];  /* P[col][row] */
  const double P01 = q[1][0] - q[2][0], P11 = q[1][1] - q[2][1];
  const double det = P00 * P11 - P10 * P01;
  if (fabs(det) > EPS) {
    inv[0] = P11 / det;   /* [0][0] */
    inv[1] = -P01 / det;  /* [0][1] */
    inv[2] = -P10 / det;  /* [1][0] */
    inv[3] = P00 / det;   /* [1][1] */
    return 1;
  }
  inv[0] = 1; inv[1] = 0; inv[2] = 0; inv[3] = 1;
  return 0;
}

/* the orthotropic elasticity entries ComputeStrainCorrection builds (FEMHelper.cu:103-111):
 * C[0][0], C[0][1], C[1][0], C[1][1], C[2][2] from (E_x, E_y, E_shear, nu_xy, nu_yx) */
static void elasticity(const double* E, double* C) {
  const double den = 1.0 - E[3] * E[4];
  C[0] = E[0] / den;
  C[1] = E[0] * E[4] / den;
  C[2] = E[1] * E[3] / den;
  C[3] = E[1] / den;
  C[4] = E[2];
}

/* energy and constraint gradients g (3 x 3) of one strain element (FEMHelper.cu:115-158) */
double orc_strain_energy(const double* x, const int32_t* t, double area, const double* inv,
                         const double* E, double* g) {
  double C[5];
  elasticity(E, C);
  const d3 p0 = L3(x, t[0]), p1 = L3(x, t[1]), p2 = L3(x, t[2]);
  const d3 p13 = sub3(p0, p2), p23 = sub3(p1, p2);
  double F[3][2];
  for (int c = 0; c < 3; c++) {
    F[c][0] = comp(p13, c) * inv[0] + comp(p23, c) * inv[2];
    F[c][1] = comp(p13, c) * inv[1] + comp(p23, c) * inv[3];
  }
  const double e00 = 0.5 * (F[0][0] * F[0][0] + F[1][0] * F[1][0] + F[2][0] * F[2][0] - 1.0);
  const double e11 = 0.5 * (F[0][1] * F[0][1] + F[1][1] * F[1][1] + F[2][1] * F[2][1] - 1.0);
  const double e01 = 0.5 * (F[0][0] * F[0][1] + F[1][0] * F[1][1] + F[2][0] * F[2][1]);
  const double s00 = C[0] * e00 + C[1] * e11;
  const double s11 = C[2] * e00 + C[3] * e11;
  const double s01 = C[4] * e01;
  const double psi = 0.5 * (e00 * s00 + e01 * s01 + e01 * s01 + e11 * s11);
  for (int c = 0; c < 3; c++) {
    const double pk0 = s00 * F[c][0] + s01 * F[c][1];  /* (stress * F)[c][0] */
    const double pk1 = s01 * F[c][0] + s11 * F[c][1];  /* (stress * F)[c][1] */
    const double h0 = (inv[0] * pk0 + inv[1] * pk1) * area;  /* (inv^T * PK * area)[c][0] */
    const double h1 = (inv[2] * pk0 + inv[3] * pk1) * area;
    g[c] = h0;
    g[3 + c] = h1;
    g[6 + c] = -h0 - h1;
  }
  return area * psi;
}

/* SolveStrain over the elements [beg, end) (one color: no shared vertices) */
void orc_strain_solve(double* x, const double* w, const int32_t* tris, const double* area,
                      const double* inv, int beg, int end, const double* E) {
  for (int e = beg; e < end; e++) {
    const int32_t* t = tris + 3 * e;
    double g[9];
    const double energy = orc_strain_energy(x, t, area[e], inv + 4 * e, E, g);
    double sum = 0.0;
    for (int v = 0; v < 3; v++)
      sum += w[t[v]] * (g[3 * v] * g[3 * v] + g[3 * v + 1] * g[3 * v + 1] + g[3 * v + 2] * g[3 * v + 2]);
    if (!(fabs(sum) > EPS)) continue;
    const double s = energy / sum;
    for (int v = 0; v < 3; v++) {
      if (!(w[t[v]] > EPS)) continue;
      for (int c = 0; c < 3; c++) x[3 * t[v] + c] += -(s * w[t[v]]) * g[3 * v + c];
    }
  }
}

static double cot(d3 v, d3 u) {
  const d3 c = cross3(v, u);
  return dot3(v, u) / sqrt(dot3(c, c));
}

/* ComputeHessianEnergy (FEMHelper.cu:221-245): Q = coeff K K^T over the hinge (v0, v1 the
 * shared edge, v2 and v3 opposite) */
void orc_iso_init(const double* x, const int32_t* b, double* K, double* coeff) {
  const d3 p0 = L3(x, b[0]), p1 = L3(x, b[1]), p2 = L3(x, b[2]), p3 = L3(x, b[3]);
  const d3 e0 = sub3(p1, p0), e1 = sub3(p2, p0), e2 = sub3(p3, p0), e3 = sub3(p2, p1),
           e4 = sub3(p3, p1);
  const d3 me0 = {-e0.x, -e0.y, -e0.z};
  const double c01 = cot(e0, e1), c02 = cot(e0, e2), c03 = cot(me0, e3), c04 = cot(me0, e4);
  const d3 n0 = cross3(sub3(p1, p0), sub3(p2, p0)), n1 = cross3(sub3(p0, p1), sub3(p3, p1));
  const double A0 = 0.5 * sqrt(dot3(n0, n0)), A1 = 0.5 * sqrt(dot3(n1, n1));
  *coeff = 3.0 / (A0 + A1);
  K[0] = c03 + c04;
  K[1] = c01 + c02;
  K[2] = -c01 - c03;
  K[3] = -c02 - c04;
}

/* 0.5 x^T Q x and its gradient Q x (4 x 3) in the reference's summation order */
double orc_iso_energy(const double* x, const int32_t* b, const double* K, double coeff,
                      double* grad) {
  double energy = 0.0;
  for (int k = 0; k < 4; k++)
    for (int j = 0; j < 4; j++)
      energy += coeff * (K[k] * K[j]) * dot3(L3(x, b[k]), L3(x, b[j]));
  energy *= 0.5;
  for (int j = 0; j < 12; j++) grad[j] = 0.0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int c = 0; c < 3; c++) grad[3 * j + c] += coeff * (K[i] * K[j]) * x[3 * b[i] + c];
  return energy;
}

/* SolveIsometricBending over the elements [beg, end) (FEMHelper.cu:317-388): XPBD with
 * alpha~ = 1 / (stiffness dt^2); lambda reset at iteration 0 */
void orc_iso_solve(double* x, const double* w, const int32_t* bidx, const double* K,
                   const double* coeff, double* lambda, int beg, int end, int iter, double dt,
                   double stiffness) {
  for (int e = beg; e < end; e++) {
    if (iter == 0) lambda[e] = 0.0;
    const int32_t* b = bidx + 4 * e;
    double d[12];
    const double energy = orc_iso_energy(x, b, K + 4 * e, coeff[e], d);
    if (fabs(energy) < EPS) continue;
    double div = 0.0;
    for (int j = 0; j < 4; j++)
      div += w[b[j]] * (d[3 * j] * d[3 * j] + d[3 * j + 1] * d[3 * j + 1] + d[3 * j + 2] * d[3 * j + 2]);
    const double at = 1.0 / (stiffness * dt * dt);
    const double dl = (-energy - at * lambda[e]) / (div + at);
    for (int j = 0; j < 4; j++)
      for (int c = 0; c < 3; c++) x[3 * b[j] + c] += dl * w[b[j]] * d[3 * j + c];
    lambda[e] += dl;
  }
}
