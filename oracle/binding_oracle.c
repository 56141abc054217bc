/*
 * binding_oracle.c -- serial FP64 restatement of the obstacle skinning and the cloth-to-body
 * binding around the substep: linear blend skinning (xrtailor/src/runtime/rag_doll/smpl/
 * cuda/body.cu:39-58) and BindingConstraint::Generate / Solve (xrtailor/src/physics/
 * dynamics/BindingConstraint.cu:20-99).  The nearest-face query that feeds the binding is
 * FP32-strict and lives with the tree in bvh_oracle.c.
 *
 * TEST INFRASTRUCTURE ONLY (see xrt_oracle.h).
 *
 * Pinned by (tests/test_oracle_binding.py): LBS against rigid closed forms (identity, a
 * translation, a rotation about an axis, two joints with complementary weights blending
 * two rigid motions); the binding against closed forms (a vertex farther than L from its
 * body point lands at L along the same ray, one inside stays, stiffness scales the step,
 * pinned vertices stay); the query against a brute-force nearest face over every face.
 */
#include <math.h>

#include "xrt_oracle.h"

#define EPS 1e-6

void orc_lbs(const double* rest, int nv, const int32_t* ptr, const int32_t* joint,
             const double* weight, const double* T, double* out) {
  for (int v = 0; v < nv; v++)
    for (int i = 0; i < 3; i++) {
      double acc = 0.0;
      for (int k = ptr[v]; k < ptr[v + 1]; k++) {
        const double* row = T + 12 * joint[k] + 4 * i;
        for (int j = 0; j < 3; j++) acc += weight[k] * row[j] * rest[3 * v + j];
        acc += weight[k] * row[3];
      }
      out[3 * v + i] = acc;
    }
}

static void body_point(const double* obs, const int32_t* t3, const double* uvw, double* q) {
  for (int a = 0; a < 3; a++)
    q[a] = uvw[0] * obs[3 * t3[0] + a] + uvw[1] * obs[3 * t3[1] + a] + uvw[2] * obs[3 * t3[2] + a];
}

void orc_binding_generate(const double* cloth_x0, const double* obs_x0, const int32_t* vert,
                          const int32_t* tri3, const double* uvw, const double* dist_in, int n,
                          double* L) {
  for (int b = 0; b < n; b++) {
    double q[3];
    body_point(obs_x0, tri3 + 3 * b, uvw + 3 * b, q);
    double d = dist_in[b];
    if (d < EPS) {
      const double* p = cloth_x0 + 3 * vert[b];
      d = sqrt((p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) +
               (p[2] - q[2]) * (p[2] - q[2]));
    }
    L[b] = d;
  }
}

void orc_binding_solve(double* x, const double* w, const double* obs_x, const int32_t* vert,
                       const int32_t* tri3, const double* uvw, const double* L,
                       const double* stiffness, int n) {
  for (int b = 0; b < n; b++) {
    const int v = vert[b];
    if (w[v] < EPS) continue;
    double q[3];
    body_point(obs_x, tri3 + 3 * b, uvw + 3 * b, q);
    double d[3] = {x[3 * v] - q[0], x[3 * v + 1] - q[1], x[3 * v + 2] - q[2]};
    const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double C = len - L[b];
    if (C < 0 || len <= 0) continue;
    const double s = C / w[v] * stiffness[b];
    for (int a = 0; a < 3; a++) x[3 * v + a] += -s * w[v] * (d[a] / len);
  }
}
